// xg_moment.hip -- fused cut rule + element matrices in moment form.
//
// Same quadrature as k_pipeline (xg_elmat.hip): for every (element, domain
// type) the straight cut rule is built in FP64 (cut points
// straightcutrule.cpp:50, decomposition :131-196, sub-simplex volumes
// :66-72, interface factor :558-578) and mapped point by point
// (x = ow V0 + sum xi_m V_{m+1}, :80-100).  Instead of the B^T W B
// contraction of symboliccutbfi.cpp:229-287, each lane accumulates the
// homogeneous barycentric moments m_alpha = sum_i w_i lam(x_i)^alpha,
// |alpha| = 2P, in FP64, and the mass and stiffness entries are formed from
// them by the generated identities of moments_table.h:
//
//     M_ab = sum_alpha C^M_ab,alpha m_alpha
//     K_ab = sum_{d<=e} H_de sum_beta C^K_ab,de,beta mk_beta,  H = F^{-1} F^{-T}
//
// This is the same quadrature sum, reassociated through the polynomial
// identities phi_a phi_b = sum C lam^alpha, which hold at every point; it
// holds for any rule, not only exact ones.  All arithmetic is FP64 (CDNA4
// issues v_fma_f64 at the FP32 rate), the output is cast to T at the store.
//
// Mapping: one lane per (element, domain type); a block holds 256 elements
// of one domain type, so the number of base-rule points per sub-simplex is
// block-uniform.  The decomposition table and the base rules are staged in
// LDS once per block.  Each lane gathers its element into LDS, then the
// block sorts its elements by the number of sub-simplices their rule sums
// (0..3, xg_moment_core.hpp make_plan) with a counting sort in LDS, so that a
// wave's lanes run the same number of point loops instead of all running the
// largest; the lanes then take the elements in that order from LDS (no second
// gather).  Occupancy is set by the VGPR count of the moment accumulators.
// (One lane per element with all its domain types in turn shares the gather
// and the cut points, but halves the number of waves: on the 64^3 workload
// that left most SIMDs with a single wave and was 25 % slower.)
#include "xg_device.hpp"
#include "xg_elmat_core.hpp"
#include "xg_internal.hpp"
#include "xg_moment_core.hpp"

using namespace xg;

namespace {

using namespace xg::mcore;

template <int D, int P, typename T, int KIND>
__global__ __launch_bounds__(256) void k_moment_pipeline(int nel, const int32_t *__restrict__ elems,
                                                         const int32_t *__restrict__ el2v,
                                                         const double *__restrict__ coords,
                                                         const double *__restrict__ lset, int order, int ndt,
                                                         int dts, int mask, T *__restrict__ out,
                                                         int32_t *err) {
  constexpr int N = ndof(D, P);
  constexpr int NG = (D + 1) * (D + 1);  // geometry doubles per element: D+1 vertices (D coords) + D+1 values
  extern __shared__ char smem[];
  __shared__ int cnt[4];
  __shared__ int perm[256];
  __shared__ double geo[NG][256];
  const int tid = threadIdx.x;
  if (tid < 4) cnt[tid] = 0;
  const int ntiles = (nel + 255) >> 8;
  const int t = blockIdx.x / ntiles;  // domain-type slot (block-uniform)
  const int base = (blockIdx.x - t * ntiles) * 256;
  const int dt = (dts >> (2 * t)) & 3;
  const int nitems = min(256, nel - base);
  // phase 0: each lane gathers its own element (issued before the table
  // staging, whose loads and barrier then overlap the gather's dependent
  // round trips), stores it into LDS and computes its sort key, the number
  // of sub-simplices its rule sums
  // (unconditional: lanes past the end load the last element again, so no
  // branch end waits for the loads before the staging)
  double V[D + 1][3], l[D + 1];
  {
    const int ic = min(base + tid, nel - 1);
    const int e = elems ? elems[ic] : ic;
    ecore::load_geom<D>(el2v, coords, lset, e, V, l);
  }
  const MomLds S = stage_lds<D, P>(smem, order);  // block barrier: before any exit
  int key = 0, pos = 0;
  if (tid < nitems) {
#pragma unroll
    for (int k = 0; k <= D; ++k) {
#pragma unroll
      for (int d = 0; d < D; ++d) geo[k * D + d][tid] = V[k][d];
      geo[(D + 1) * D + k][tid] = l[k];
    }
    key = make_plan<D>(S, l, dt).ns;
    pos = atomicAdd(&cnt[key], 1);
  }
  __syncthreads();
  if (tid < nitems) {
    int off = 0;
#pragma unroll
    for (int k = 0; k < 3; ++k) off += (k < key) ? cnt[k] : 0;
    perm[off + pos] = tid;
  }
  __syncthreads();
  if (tid >= nitems) return;  // no cross-lane operations below
  // phase 1: the lanes take the elements in sorted order, from LDS
  const int li = perm[tid];
  const int i = base + li;
#pragma unroll
  for (int k = 0; k <= D; ++k) {
#pragma unroll
    for (int d = 0; d < 3; ++d) V[k][d] = d < D ? geo[k * D + d][li] : 0.0;
    l[k] = geo[(D + 1) * D + k][li];
  }
  double F[D * D], Finv[D * D], absdet;
  jacobian<D>(V, F, Finv, absdet);
  bool bad;
  if constexpr (KIND == 0)
    bad = mcore::moment_elmat<D, P, T>(S, l, Finv, absdet, dt, mask, out + ((size_t)i * ndt + t) * (size_t)(N * N));
  else
    bad = mcore::moment_elvec<D, P, T>(S, l, Finv, absdet, dt, out + ((size_t)i * ndt + t) * (size_t)N);
  if (bad) atomicOr(err, 1);
}

// element matrices from given moments (one lane per element): the rule is
// whatever produced the moments (subdivision rule, learned rule)
template <int D, int P, typename T>
__global__ __launch_bounds__(256) void k_elmat_from_moments(int nel, int ndt, const int32_t *__restrict__ elems,
                                                            const int32_t *__restrict__ el2v,
                                                            const double *__restrict__ coords,
                                                            const double *__restrict__ mom, int mask,
                                                            T *__restrict__ out) {
  // thread g = (element g / ndt, domain type g % ndt): moments and matrices
  // are stored element-major, ndt consecutive sets per element
  using MM = Moments<D, P>;
  constexpr int N = ndof(D, P);
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)nel * ndt) return;
  const int i = (int)(g / ndt);
  double Finv[D * D], absdet = 1.0;
#pragma unroll
  for (int k = 0; k < D * D; ++k) Finv[k] = (k % (D + 1) == 0) ? 1.0 : 0.0;
  if (coords) {
    const int e = elems ? elems[i] : i;
    double V[D + 1][3], F[D * D];
    ecore::load_geom<D>(el2v, coords, nullptr, e, V, nullptr);
    jacobian<D>(V, F, Finv, absdet);
  }
  double m[MM::NM];
#pragma unroll
  for (int k = 0; k < MM::NM; ++k) m[k] = mom[(size_t)g * MM::NM + k];
  double mk[MM::NK];
  MM::lower(m, mk);
  // Gram matrix of the physical barycentric gradients: grad lam_i = row i
  // of F^{-1} (i < D), grad lam_D = -sum of the others
  double Gr[D + 1][D];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    double sacc = 0.0;
#pragma unroll
    for (int i = 0; i < D; ++i) {
      Gr[i][r] = Finv[i * D + r];
      sacc += Finv[i * D + r];
    }
    Gr[D][r] = -sacc;
  }
  double Hs[(D + 1) * (D + 2) / 2];
  int h = 0;
#pragma unroll
  for (int i = 0; i <= D; ++i)
#pragma unroll
    for (int j = i; j <= D; ++j) {
      double acc = 0.0;
#pragma unroll
      for (int r = 0; r < D; ++r) acc = fma(Gr[i][r], Gr[j][r], acc);
      Hs[h++] = acc;
    }
  const double sm = (mask & 1) ? absdet : 0.0, sk = (mask & 2) ? absdet : 0.0;
  MM::template forms<T>(m, mk, Hs, sm, sk, out + (size_t)g * (N * N));
}

// learned moments m = base + corr (FP64 + FP32 correction) and their check:
// the degree-2 moments of m, obtained through sum lam = 1 (lower, once per
// degree step), against the plain rule's exact degree-2 moments ref2
template <int P>
__global__ __launch_bounds__(256) void k_learned_check(int nel, const double *__restrict__ base,
                                                       const float *__restrict__ corr, const double *__restrict__ ref2,
                                                       double thr, double *__restrict__ out, double *__restrict__ rel,
                                                       uint8_t *__restrict__ accept) {
  using MM = Moments<3, P>;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  double m[MM::NM];
#pragma unroll
  for (int k = 0; k < MM::NM; ++k) {
    m[k] = base[(size_t)i * MM::NM + k] + (double)corr[(size_t)i * MM::NM + k];
    out[(size_t)i * MM::NM + k] = m[k];
  }
  double m2[10];
  if constexpr (P == 1) {
#pragma unroll
    for (int k = 0; k < 10; ++k) m2[k] = m[k];
  } else if constexpr (P == 2) {
    MM::lower(m, m2);
  } else {
    double m4[Moments<3, 2>::NM];
    MM::lower(m, m4);
    Moments<3, 2>::lower(m4, m2);
  }
  double num = 0.0, den = 0.0;
#pragma unroll
  for (int k = 0; k < 10; ++k) {
    const double r = ref2[(size_t)i * 10 + k];
    num += (m2[k] - r) * (m2[k] - r);
    den += r * r;
  }
  const double q = den > 0.0 ? sqrt(num / den) : (num > 0.0 ? INFINITY : 0.0);
  if (rel) rel[i] = q;
  accept[i] = q <= thr ? 1 : 0;
}

template <int D, int P, typename T>
int launch_from_moments(xg_ctx *ctx, int nel, int ndt, const int32_t *elems, const int32_t *el2v,
                        const double *coords, const double *mom, int mask, void *out) {
  const long long blocks = ((long long)nel * ndt + 255) / 256;
  if (blocks > 0x7fffffffLL) return set_error(XG_ERR_ARG, "xg_elmat_moments: too many elements");
  k_elmat_from_moments<D, P, T><<<(unsigned)blocks, 256, 0, ctx->stream>>>(nel, ndt, elems, el2v, coords, mom,
                                                                           mask, (T *)out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int D, int P, typename T, int KIND = 0>
int launch(xg_ctx *ctx, int order, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
           const double *lset, int ndt, int dts, int mask, void *out) {
  const long long blocks = (long long)ndt * ((nel + 255) / 256);
  if (blocks == 0) return XG_OK;
  if (blocks > 0x7fffffffLL) return set_error(XG_ERR_ARG, "xg_cut_pipeline: too many elements");
  k_moment_pipeline<D, P, T, KIND><<<(unsigned)blocks, 256, mom_lds_bytes(D, P, order), ctx->stream>>>(nel, elems, el2v, coords, lset, order,
                                                                      ndt, dts, mask, (T *)out, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

}  // namespace

namespace xg {

int launch_moment_pipeline(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems,
                           const int32_t *el2v, const double *coords, const double *lset, int ndt, int dts,
                           int mask, bool fp64_out, void *out) {
#define XG_MCALL(DD, PP)                                                                                   \
  if (D == DD && p == PP)                                                                                  \
    return fp64_out ? launch<DD, PP, double>(ctx, order, nel, elems, el2v, coords, lset, ndt, dts, mask, out) \
                    : launch<DD, PP, float>(ctx, order, nel, elems, el2v, coords, lset, ndt, dts, mask, out);
  XG_MCALL(1, 1) XG_MCALL(1, 2) XG_MCALL(1, 3)
  XG_MCALL(2, 1) XG_MCALL(2, 2) XG_MCALL(2, 3)
  XG_MCALL(3, 1) XG_MCALL(3, 2) XG_MCALL(3, 3)
#undef XG_MCALL
  return set_error(XG_ERR_ARG, "xg_cut_pipeline: bad D or p");
}

int elmat_from_moments(xg_ctx *ctx, int p, int nel, int ndt, const int32_t *elems, const int32_t *el2v,
                       const double *coords, const double *mom, int mask, bool fp64_out, void *out) {
  if (p == 1)
    return fp64_out ? launch_from_moments<3, 1, double>(ctx, nel, ndt, elems, el2v, coords, mom, mask, out)
                    : launch_from_moments<3, 1, float>(ctx, nel, ndt, elems, el2v, coords, mom, mask, out);
  if (p == 2)
    return fp64_out ? launch_from_moments<3, 2, double>(ctx, nel, ndt, elems, el2v, coords, mom, mask, out)
                    : launch_from_moments<3, 2, float>(ctx, nel, ndt, elems, el2v, coords, mom, mask, out);
  if (p == 3)
    return fp64_out ? launch_from_moments<3, 3, double>(ctx, nel, ndt, elems, el2v, coords, mom, mask, out)
                    : launch_from_moments<3, 3, float>(ctx, nel, ndt, elems, el2v, coords, mom, mask, out);
  return set_error(XG_ERR_ARG, "element matrices from moments: bad p");
}

int launch_moment_elvec(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems, const int32_t *el2v,
                        const double *coords, const double *lset, int ndt, int dts, bool fp64_out, void *out) {
#define XG_VCALL(DD, PP)                                                                                          \
  if (D == DD && p == PP)                                                                                         \
    return fp64_out ? launch<DD, PP, double, 1>(ctx, order, nel, elems, el2v, coords, lset, ndt, dts, 1, out)    \
                    : launch<DD, PP, float, 1>(ctx, order, nel, elems, el2v, coords, lset, ndt, dts, 1, out);
  XG_VCALL(1, 1) XG_VCALL(1, 2) XG_VCALL(1, 3)
  XG_VCALL(2, 1) XG_VCALL(2, 2) XG_VCALL(2, 3)
  XG_VCALL(3, 1) XG_VCALL(3, 2) XG_VCALL(3, 3)
#undef XG_VCALL
  return set_error(XG_ERR_ARG, "xg_cut_elvec: bad D or p");
}

}  // namespace xg

extern "C" int xg_elmat_moments(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
                                const double *coords, const double *moments, int mask, int precision, void *out) {
  if (!ctx || D < 1 || D > 3 || p < 1 || p > 3 || nel < 0 || mask < 1 || mask > 3)
    return set_error(XG_ERR_ARG, "xg_elmat_moments: bad argument");
  if (precision != XG_PREC_FP32 && precision != XG_PREC_FP64)
    return set_error(XG_ERR_UNSUPPORTED, "xg_elmat_moments: FP32 or FP64 output");
  if (nel == 0) return XG_OK;
  if (!moments || !out || (coords && !el2v)) return set_error(XG_ERR_ARG, "xg_elmat_moments: null pointer");
  const bool f64 = precision == XG_PREC_FP64;
#define XG_ECALL(DD, PP)                                                                          \
  if (D == DD && p == PP)                                                                         \
    return f64 ? launch_from_moments<DD, PP, double>(ctx, nel, 1, elems, el2v, coords, moments, mask, out) \
               : launch_from_moments<DD, PP, float>(ctx, nel, 1, elems, el2v, coords, moments, mask, out);
  XG_ECALL(1, 1) XG_ECALL(1, 2) XG_ECALL(1, 3)
  XG_ECALL(2, 1) XG_ECALL(2, 2) XG_ECALL(2, 3)
  XG_ECALL(3, 1) XG_ECALL(3, 2) XG_ECALL(3, 3)
#undef XG_ECALL
  return set_error(XG_ERR_ARG, "xg_elmat_moments: bad D or p");
}

extern "C" int xg_cut_elvec(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems, const int32_t *el2v,
                            const double *coords, const double *lset_v, int ndt, const int *dts, int precision,
                            void *out) {
  if (!ctx || D < 1 || D > 3 || p < 1 || p > 3 || nel < 0 || ndt < 1 || ndt > 3 || !dts || order < 0 ||
      order / 2 + 1 > XG_RULE_NMAX)
    return set_error(XG_ERR_ARG, "xg_cut_elvec: bad argument");
  if (precision != XG_PREC_FP32 && precision != XG_PREC_FP64)
    return set_error(XG_ERR_UNSUPPORTED, "xg_cut_elvec: FP32 or FP64 output");
  int packed = 0;
  for (int t = 0; t < ndt; ++t) {
    if (dts[t] < 0 || dts[t] > 2) return set_error(XG_ERR_ARG, "xg_cut_elvec: bad domain type");
    packed |= dts[t] << (2 * t);
  }
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !lset_v || !out) return set_error(XG_ERR_ARG, "xg_cut_elvec: null pointer");
  return xg::launch_moment_elvec(ctx, D, p, order, nel, elems, el2v, coords, lset_v, ndt, packed,
                                 precision == XG_PREC_FP64, out);
}

// ---- global assembly (NGSolve's BilinearForm::Assemble / LinearForm::Assemble
// scatter): FP64 atomic adds into a CSR matrix / a vector
namespace {

template <typename T>
__global__ __launch_bounds__(256) void k_assemble_csr(long long nent, int ndof, const int32_t *__restrict__ dofs,
                                                      const T *__restrict__ elmat,
                                                      const int32_t *__restrict__ row_ptr,
                                                      const int32_t *__restrict__ col_idx, double scale,
                                                      double *values, int32_t *err) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nent) return;
  const int nn = ndof * ndof;
  const long long e = q / nn;
  const int ab = (int)(q - e * nn), a = ab / ndof, b = ab - a * ndof;
  const int r = dofs[e * ndof + a], c = dofs[e * ndof + b];
  if (r < 0 || c < 0) return;  // dof not in the space (e.g. Dirichlet-free dofs)
  int lo = row_ptr[r], hi = row_ptr[r + 1];
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (col_idx[mid] <= c) lo = mid;
    else hi = mid;
  }
  if (lo >= row_ptr[r + 1] || col_idx[lo] != c) {
    atomicOr(err, 4);  // entry missing from the sparsity pattern
    return;
  }
  atomicAdd(&values[lo], scale * (double)elmat[q]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_assemble_vec(long long nent, const int32_t *__restrict__ dofs,
                                                      const T *__restrict__ elvec, double *out) {
  const long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nent) return;
  const int r = dofs[q];
  if (r >= 0) atomicAdd(&out[r], (double)elvec[q]);
}

}  // namespace

extern "C" int xg_assemble_csr_scaled(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, const void *elmat,
                                      int precision, double scale, const int32_t *row_ptr, const int32_t *col_idx,
                                      double *values) {
  if (!ctx || nel < 0 || ndof < 1 || ndof > 64 || (precision != XG_PREC_FP32 && precision != XG_PREC_FP64))
    return set_error(XG_ERR_ARG, "xg_assemble_csr: bad argument");
  if (nel == 0) return XG_OK;
  if (!dofs || !elmat || !row_ptr || !col_idx || !values) return set_error(XG_ERR_ARG, "xg_assemble_csr: null pointer");
  const long long nent = (long long)nel * ndof * ndof;
  const long long blocks = (nent + 255) / 256;
  if (blocks > 0x7fffffffLL) return set_error(XG_ERR_ARG, "xg_assemble_csr: too many entries");
  if (precision == XG_PREC_FP64)
    k_assemble_csr<double><<<(unsigned)blocks, 256, 0, ctx->stream>>>(nent, ndof, dofs, (const double *)elmat, row_ptr,
                                                                      col_idx, scale, values, ctx->d_err);
  else
    k_assemble_csr<float><<<(unsigned)blocks, 256, 0, ctx->stream>>>(nent, ndof, dofs, (const float *)elmat, row_ptr,
                                                                     col_idx, scale, values, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_assemble_csr(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, const void *elmat, int precision,
                               const int32_t *row_ptr, const int32_t *col_idx, double *values) {
  return xg_assemble_csr_scaled(ctx, nel, ndof, dofs, elmat, precision, 1.0, row_ptr, col_idx, values);
}

extern "C" int xg_assemble_vec(xg_ctx *ctx, int nel, int ndof, const int32_t *dofs, const void *elvec, int precision,
                               double *out) {
  if (!ctx || nel < 0 || ndof < 1 || (precision != XG_PREC_FP32 && precision != XG_PREC_FP64))
    return set_error(XG_ERR_ARG, "xg_assemble_vec: bad argument");
  if (nel == 0) return XG_OK;
  if (!dofs || !elvec || !out) return set_error(XG_ERR_ARG, "xg_assemble_vec: null pointer");
  const long long nent = (long long)nel * ndof;
  const long long blocks = (nent + 255) / 256;
  if (precision == XG_PREC_FP64)
    k_assemble_vec<double><<<(unsigned)blocks, 256, 0, ctx->stream>>>(nent, dofs, (const double *)elvec, out);
  else
    k_assemble_vec<float><<<(unsigned)blocks, 256, 0, ctx->stream>>>(nent, dofs, (const float *)elvec, out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_learned_moments(xg_ctx *ctx, int nel, int p, const double *base, const float *corr,
                                  const double *ref2, double thr, double *out, double *rel, uint8_t *accept) {
  if (!ctx || nel < 0 || p < 1 || p > 3) return set_error(XG_ERR_ARG, "xg_learned_moments: bad argument");
  if (nel == 0) return XG_OK;
  if (!base || !corr || !ref2 || !out || !accept) return set_error(XG_ERR_ARG, "xg_learned_moments: null pointer");
  const int blocks = (nel + 255) / 256;
  if (p == 1) k_learned_check<1><<<blocks, 256, 0, ctx->stream>>>(nel, base, corr, ref2, thr, out, rel, accept);
  if (p == 2) k_learned_check<2><<<blocks, 256, 0, ctx->stream>>>(nel, base, corr, ref2, thr, out, rel, accept);
  if (p == 3) k_learned_check<3><<<blocks, 256, 0, ctx->stream>>>(nel, base, corr, ref2, thr, out, rel, accept);
  XG_LAUNCH_CHECK();
  return XG_OK;
}
