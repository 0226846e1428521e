// xg_pattern.hip -- sparsity pattern of the global matrix on the device.
//
// NGSolve's BilinearForm::Assemble (the caller of CalcElementMatrix,
// xfem/symboliccutbfi.cpp:37-46) builds the matrix graph from the element
// dof tables before the element matrices are added.  Here the graph is
// built on the GPU: every coupling (row, col) of every element becomes one
// 64-bit key row << 32 | col, the keys are radix-sorted and made unique
// (hipCUB, rocPRIM underneath), the unique columns are the CSR column
// indices in ascending order per row, and the row pointers are the starts
// of the rows (a count per row, then an exclusive scan).  Work: 8 B x 2 per
// coupling through the sort, ndof^2 couplings per element.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "xg_internal.hpp"

using namespace xg;

namespace {

__global__ __launch_bounds__(256) void k_pair_keys(long long nent, int ndof, const int32_t *__restrict__ dofs, int n,
                                                   uint64_t *__restrict__ keys, int32_t *err) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nent) return;
  const int nn = ndof * ndof;
  const long long e = q / nn;
  const int ab = (int)(q - e * nn), a = ab / ndof, b = ab - a * ndof;
  const int r = dofs[e * ndof + a], c = dofs[e * ndof + b];
  if (r >= n || c >= n) atomicOr(err, 8);
  // skipped dofs (< 0) and out-of-range ones sort to the end as all-ones keys
  keys[q] = (r < 0 || c < 0 || r >= n || c >= n) ? ~0ull : ((uint64_t)(uint32_t)r << 32) | (uint32_t)c;
}

// unique keys -> column indices and a count per row; the trailing all-ones
// key (skipped couplings) is not an entry
__global__ __launch_bounds__(256) void k_unpack_keys(const uint64_t *__restrict__ ukeys, const int32_t *__restrict__ nu,
                                                     int32_t *__restrict__ col_idx, long long cap,
                                                     int32_t *__restrict__ row_cnt, int32_t *__restrict__ nnz,
                                                     int32_t *err) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int cnt = *nu;
  const uint64_t last = cnt > 0 ? ukeys[cnt - 1] : 0ull;
  const int n_ent = (cnt > 0 && last == ~0ull) ? cnt - 1 : cnt;
  if (i == 0) {
    *nnz = n_ent;
    if ((long long)n_ent > cap) atomicOr(err, 4);
  }
  if (i >= n_ent || i >= cap) return;
  const uint64_t k = ukeys[i];
  col_idx[i] = (int32_t)(uint32_t)(k & 0xffffffffu);
  atomicAdd(&row_cnt[(int)(k >> 32)], 1);
}

__global__ void k_fill_i32(int32_t *p, long long n, int32_t v) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" int xg_csr_pattern(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, int n, int32_t *row_ptr,
                              int32_t *col_idx, long long cap, int32_t *nnz) {
  if (!ctx || nel < 0 || ndof < 1 || n < 0 || cap < 0) return set_error(XG_ERR_ARG, "xg_csr_pattern: bad argument");
  if (!row_ptr || !nnz || (nel > 0 && (!dofs || (cap > 0 && !col_idx))))
    return set_error(XG_ERR_ARG, "xg_csr_pattern: null pointer");
  const long long nent = (long long)nel * ndof * ndof;
  if (nent > 0x7fffffffLL) return set_error(XG_ERR_ARG, "xg_csr_pattern: more than 2^31 couplings");
  hipStream_t st = ctx->stream;
  // row counts in row_ptr[1..n] (then scanned in place), row_ptr[0] = 0
  k_fill_i32<<<(unsigned)((n + 1 + 255) / 256), 256, 0, st>>>(row_ptr, (long long)n + 1, 0);
  XG_LAUNCH_CHECK();
  if (nent == 0) {
    k_fill_i32<<<1, 1, 0, st>>>(nnz, 1, 0);
    XG_LAUNCH_CHECK();
    return XG_OK;
  }
  const int N = (int)nent;
  // workspace: keys, sorted keys, unique count, temp storage of the three
  // hipCUB passes (sort, unique, scan), sized by their dry runs
  size_t t_sort = 0, t_uniq = 0, t_scan = 0;
  XG_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, t_sort, (uint64_t *)nullptr, (uint64_t *)nullptr, N, 0, 64, st));
  XG_CHECK(hipcub::DeviceSelect::Unique(nullptr, t_uniq, (uint64_t *)nullptr, (uint64_t *)nullptr, (int32_t *)nullptr,
                                        N, st));
  XG_CHECK(hipcub::DeviceScan::InclusiveSum(nullptr, t_scan, (int32_t *)nullptr, (int32_t *)nullptr, n, st));
  const size_t tmp = std::max(t_sort, std::max(t_uniq, t_scan));
  const size_t o_a = 0, o_b = align256(sizeof(uint64_t) * (size_t)N), o_n = o_b + align256(sizeof(uint64_t) * (size_t)N);
  const size_t o_t = o_n + 256;
  int rc = ensure_ws(ctx, o_t + tmp);
  if (rc != XG_OK) return rc;
  char *ws = (char *)ctx->ws;
  uint64_t *ka = (uint64_t *)(ws + o_a), *kb = (uint64_t *)(ws + o_b);
  int32_t *nu = (int32_t *)(ws + o_n);
  void *tp = ws + o_t;
  size_t tsz = tmp;
  k_pair_keys<<<(unsigned)((nent + 255) / 256), 256, 0, st>>>(nent, ndof, dofs, n, ka, ctx->d_err);
  XG_LAUNCH_CHECK();
  XG_CHECK(hipcub::DeviceRadixSort::SortKeys(tp, tsz, ka, kb, N, 0, 64, st));
  tsz = tmp;
  XG_CHECK(hipcub::DeviceSelect::Unique(tp, tsz, kb, ka, nu, N, st));
  k_unpack_keys<<<(unsigned)((nent + 255) / 256), 256, 0, st>>>(ka, nu, col_idx, cap, row_ptr + 1, nnz, ctx->d_err);
  XG_LAUNCH_CHECK();
  // row_ptr[r + 1] = number of entries in rows 0..r (inclusive scan of the
  // counts in place), row_ptr[0] = 0
  tsz = tmp;
  if (n > 0) XG_CHECK(hipcub::DeviceScan::InclusiveSum(tp, tsz, row_ptr + 1, row_ptr + 1, n, st));
  return XG_OK;
}
