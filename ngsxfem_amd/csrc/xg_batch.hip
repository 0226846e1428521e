// xg_batch.hip -- many meshes of uneven size and mixed polynomial order in one
// captured launch sequence: marking + compaction over the concatenated
// meshes, a per-mesh plan, and one persistent fused rule + element-matrix
// kernel per polynomial degree with dynamic work distribution.
//
// No direct counterpart in the reference (it assembles one mesh per
// BilinearForm::Assemble, threaded by TaskManager, utils/ngsxstd.cpp:3-31);
// per element the work is CutInformation::Update (xfem/cutinfo.cpp:61-133)
// and CalcElementMatrix on the straight cut rule
// (xfem/symboliccutbfi.cpp:76-304).
//
// Device-side plan (all sizes fixed at capture, so the whole sequence is a
// hipGraph replayed per run; no memset/memcpy nodes):
//   k_mark / scan / k_compact over all elements -> cut list (ascending ids);
//   k_mesh_counts: cut elements per mesh (popcount of the IF bits of the
//     mesh's element range);
//   k_plan: per-mesh start in the cut list and output offset, meshes grouped
//     by degree p with a virtual index space per group, work counters reset;
//   k_batch<P>: persistent waves grab `grab` cut elements at a time from the
//     group's counter (atomic, lane 0), find the mesh by binary search over
//     the group's prefix array in LDS and run the fused pipeline.
#include <string>
#include <vector>

#include "xg_device.hpp"
#include "xg_elmat_core.hpp"
#include "xg_internal.hpp"
#include "xg_moment_core.hpp"

using namespace xg;
using namespace xg::ecore;

struct xg_batch {
  int D = 3, nmesh = 0, ne = 0, ndt = 0, dts = 0, mask = 3;
  std::vector<int32_t> h_off, h_p;
  const int32_t *el2v = nullptr;
  const double *coords = nullptr, *lset = nullptr;
  // device plan
  int32_t *d_off = nullptr, *d_p = nullptr;
  uint32_t *d_bits = nullptr;
  int32_t *d_scr = nullptr, *d_cut = nullptr, *d_ncut = nullptr;
  int32_t *d_cnt = nullptr, *d_start = nullptr, *d_vstart = nullptr;
  int32_t *d_cls_mesh = nullptr, *d_cls_nm = nullptr, *d_cls_total = nullptr, *d_counter = nullptr;
  long long *d_outoff = nullptr;
  // tuning per degree p = 1..3
  int waves[4] = {0, 2, 2, 2}, bpc[4] = {0, 4, 4, 4}, grab[4] = {0, 1, 1, 1};
  int method[4] = {0, 0, 0, 0};  // per p: 0 MFMA contraction, XG_METHOD_MOMENTS moment form
  int ncu = 256;
  // graph
  hipStream_t stream = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;
  float *g_out = nullptr;
  long long g_cap = 0;
};

namespace {

constexpr int MAXMESH = 1024;

__global__ void k_mesh_counts(int nmesh, const int32_t *__restrict__ off, const uint32_t *__restrict__ if_bits,
                              int32_t *__restrict__ cnt) {
  __shared__ int s[4];
  const int m = blockIdx.x;
  if (m >= nmesh) return;
  const int a = off[m], b = off[m + 1];
  int c = 0;
  for (int w = (a >> 5) + threadIdx.x; w <= ((b - 1) >> 5) && b > a; w += blockDim.x) {
    uint32_t bits = if_bits[w];
    const int lo = w << 5;
    if (lo < a) bits &= ~0u << (a - lo);
    if (lo + 32 > b) bits &= (b - lo) >= 32 ? ~0u : ((1u << (b - lo)) - 1u);
    c += __popc(bits);
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += s[k];
    cnt[m] = t;
  }
}

// single thread: the plan is O(nmesh)
__global__ void k_plan(int nmesh, int ndt, const int32_t *__restrict__ p, const int32_t *__restrict__ cnt,
                       int32_t *__restrict__ start, long long *__restrict__ outoff, int32_t *__restrict__ vstart,
                       int32_t *__restrict__ cls_mesh, int32_t *__restrict__ cls_nm,
                       int32_t *__restrict__ cls_total, int32_t *__restrict__ counter) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int s = 0;
  long long o = 0;
  int vs[4] = {0, 0, 0, 0}, nm[4] = {0, 0, 0, 0};
  for (int m = 0; m < nmesh; ++m) {
    const int pm = p[m];
    const int nd = pm == 1 ? 4 : (pm == 2 ? 10 : 20);
    start[m] = s;
    outoff[m] = o;
    vstart[m] = vs[pm];
    cls_mesh[(pm - 1) * nmesh + nm[pm]] = m;
    nm[pm]++;
    s += cnt[m];
    o += (long long)cnt[m] * ndt * nd * nd;
    vs[pm] += cnt[m];
  }
  start[nmesh] = s;
  outoff[nmesh] = o;
  for (int q = 1; q <= 3; ++q) {
    cls_nm[q - 1] = nm[q];
    cls_total[q - 1] = vs[q];
    counter[q - 1] = 0;
  }
}

template <int P, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_batch(int order, int ndt, int dts, int mask, int grab,
                                                      const int32_t *__restrict__ el2v,
                                                      const double *__restrict__ coords,
                                                      const double *__restrict__ lset,
                                                      const int32_t *__restrict__ cut,
                                                      const int32_t *__restrict__ start,
                                                      const long long *__restrict__ outoff,
                                                      const int32_t *__restrict__ vstart,
                                                      const int32_t *__restrict__ cls_mesh,
                                                      const int32_t *__restrict__ cls_nm,
                                                      const int32_t *__restrict__ cls_total, int32_t *counter,
                                                      int nmesh, float *__restrict__ out, long long cap,
                                                      int32_t *err) {
  constexpr int D = 3;
  using T = float;
  using C = Cfg<D, P, T>;
  extern __shared__ char smem[];
  __shared__ StageLds<D, T> s_stage[WAVES];
  __shared__ double s_rule[2][PIPE_NQMAX][4];
  __shared__ int s_vs[MAXMESH + 1];
  __shared__ int s_mesh[MAXMESH];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  T *sB = reinterpret_cast<T *>(smem) + (size_t)wv * (C::LDS_PER_WAVE / sizeof(T));
  T *sW = sB + C::SB_SIZE;
  for (int k = lane; k < C::SB_SIZE; k += 64) sB[k] = T(0);
  const int nq_vol = rule_cnt(D, order), nq_if = rule_cnt(D - 1, order);
  const int off_vol = rule_off(D, order), off_if = rule_off(D - 1, order);
  for (int k = threadIdx.x; k < nq_vol + nq_if; k += blockDim.x) {
    const int which = k < nq_vol ? 0 : 1;
    const int kk = which ? k - nq_vol : k;
    const double *r = xg_rule_data_d[(which ? off_if : off_vol) + kk];
#pragma unroll
    for (int c = 0; c < 4; ++c) s_rule[which][kk][c] = r[c];
  }
  const int nm = cls_nm[P - 1];
  for (int k = threadIdx.x; k < nm; k += blockDim.x) {
    const int m = cls_mesh[(P - 1) * nmesh + k];
    s_mesh[k] = m;
    s_vs[k] = vstart[m];
  }
  if (threadIdx.x == 0) s_vs[nm] = cls_total[P - 1];
  __syncthreads();
  const int total = cls_total[P - 1];
  constexpr int NN = C::N * C::N;
  while (true) {
    int k0 = 0;
    if (lane == 0) k0 = atomicAdd(&counter[P - 1], grab);
    k0 = __shfl(k0, 0, 64);
    k0 = __builtin_amdgcn_readfirstlane(k0);
    if (k0 >= total) break;
    const int k1 = min(k0 + grab, total);
    for (int j = k0; j < k1; ++j) {
      // mesh of virtual index j: last k with s_vs[k] <= j (meshes with no cut
      // elements share their start with the next one; upper_bound skips them)
      int lo = 0, hi = nm;  // s_vs[lo] <= j < s_vs[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_vs[mid] <= j) lo = mid;
        else hi = mid;
      }
      const int m = s_mesh[lo];
      const int r = j - s_vs[lo];
      const int e = cut[start[m] + r];
      const long long obase = outoff[m] + (long long)r * ndt * NN;
      if (obase + (long long)ndt * NN > cap) {
        if (lane == 0) atomicOr(err, 4);
        continue;
      }
      double V[D + 1][3], l[D + 1];
      load_geom<D>(el2v, coords, lset, e, V, l);
      double F[D * D], Finv[D * D], absdet;
      jacobian<D>(V, F, Finv, absdet);
      for (int t = 0; t < ndt; ++t) {
        const int dt = (dts >> (2 * t)) & 3;
        int nq = 0, nv1 = 0;
        bool bad = false;
        const int npts = build_stage<D, T>(&s_stage[wv], l, Finv, dt, nq_vol, nq_if, nq, nv1, bad);
        if (bad && lane == 0) atomicOr(err, 1);
        StagedSrc<D, T> src{&s_stage[wv], s_rule[dt == IF ? 1 : 0], nq, nv1};
        elmat_one<D, P, T>(src, npts, Finv, absdet, mask, sB, sW, out + obase + (long long)t * NN);
      }
    }
  }
}

// moment form (xg_moment_core.hpp): one lane per cut element, a wave grabs
// 64 * grab consecutive virtual indices of its degree group
template <int P, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_batch_mom(int order, int ndt, int dts, int mask, int grab,
                                                          const int32_t *__restrict__ el2v,
                                                          const double *__restrict__ coords,
                                                          const double *__restrict__ lset,
                                                          const int32_t *__restrict__ cut,
                                                          const int32_t *__restrict__ start,
                                                          const long long *__restrict__ outoff,
                                                          const int32_t *__restrict__ vstart,
                                                          const int32_t *__restrict__ cls_mesh,
                                                          const int32_t *__restrict__ cls_nm,
                                                          const int32_t *__restrict__ cls_total, int32_t *counter,
                                                          int nmesh, float *__restrict__ out, long long cap,
                                                          int32_t *err) {
  constexpr int D = 3;
  constexpr int NN = ndof(D, P) * ndof(D, P);
  __shared__ int s_vs[MAXMESH + 1];
  __shared__ int s_mesh[MAXMESH];
  extern __shared__ char smem[];
  const mcore::MomLds S = mcore::stage_lds<D, P>(smem, order);
  const int lane = threadIdx.x & 63;
  const int nm = cls_nm[P - 1];
  for (int k = threadIdx.x; k < nm; k += blockDim.x) {
    const int m = cls_mesh[(P - 1) * nmesh + k];
    s_mesh[k] = m;
    s_vs[k] = vstart[m];
  }
  if (threadIdx.x == 0) s_vs[nm] = cls_total[P - 1];
  __syncthreads();
  const int total = cls_total[P - 1];
  const int chunk = 64 * grab;
  while (true) {
    int k0 = 0;
    if (lane == 0) k0 = atomicAdd(&counter[P - 1], chunk);
    k0 = __shfl(k0, 0, 64);
    k0 = __builtin_amdgcn_readfirstlane(k0);
    if (k0 >= total) break;
    for (int g = 0; g < grab; ++g) {
      const int j = k0 + g * 64 + lane;
      if (j >= total) break;
      int lo = 0, hi = nm;  // s_vs[lo] <= j < s_vs[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (s_vs[mid] <= j) lo = mid;
        else hi = mid;
      }
      const int m = s_mesh[lo];
      const int r = j - s_vs[lo];
      const int e = cut[start[m] + r];
      const long long obase = outoff[m] + (long long)r * ndt * NN;
      if (obase + (long long)ndt * NN > cap) {
        atomicOr(err, 4);
        continue;
      }
      double V[D + 1][3], l[D + 1];
      load_geom<D>(el2v, coords, lset, e, V, l);
      double F[D * D], Finv[D * D], absdet;
      jacobian<D>(V, F, Finv, absdet);
      bool bad = false;
      for (int t = 0; t < ndt; ++t) {
        const int dt = (dts >> (2 * t)) & 3;
        bad |= mcore::moment_elmat<D, P, float>(S, l, Finv, absdet, dt, mask, out + obase + (long long)t * NN);
      }
      if (bad) atomicOr(err, 1);
    }
  }
}

template <int P, int W>
int launch_batch_mom(xg_batch *b, hipStream_t st, int32_t *err, int order, float *out, long long cap) {
  const int grid = b->bpc[P] * b->ncu;
  k_batch_mom<P, W><<<grid, 64 * W, mcore::mom_lds_bytes(3, P, order), st>>>(order, b->ndt, b->dts, b->mask, b->grab[P], b->el2v, b->coords,
                                             b->lset, b->d_cut, b->d_start, b->d_outoff, b->d_vstart,
                                             b->d_cls_mesh, b->d_cls_nm, b->d_cls_total, b->d_counter, b->nmesh,
                                             out, cap, err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int P, int W>
int launch_batch(xg_batch *b, hipStream_t st, int32_t *err, int order, float *out, long long cap) {
  using C = Cfg<3, P, float>;
  const size_t lds = C::LDS_PER_WAVE * W;
  const int grid = b->bpc[P] * b->ncu;
  k_batch<P, W><<<grid, 64 * W, lds, st>>>(order, b->ndt, b->dts, b->mask, b->grab[P], b->el2v, b->coords, b->lset,
                                           b->d_cut, b->d_start, b->d_outoff, b->d_vstart, b->d_cls_mesh,
                                           b->d_cls_nm, b->d_cls_total, b->d_counter, b->nmesh, out, cap, err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

int launch_degree(xg_batch *b, int P, hipStream_t st, int32_t *err, float *out, long long cap) {
  const int order = 2 * P;
  const int W = b->waves[P];
  if (b->method[P] == XG_METHOD_MOMENTS) {
    if (P == 1) return W == 1 ? launch_batch_mom<1, 1>(b, st, err, order, out, cap) : launch_batch_mom<1, 2>(b, st, err, order, out, cap);
    if (P == 2) return W == 1 ? launch_batch_mom<2, 1>(b, st, err, order, out, cap) : launch_batch_mom<2, 2>(b, st, err, order, out, cap);
    return W == 1 ? launch_batch_mom<3, 1>(b, st, err, order, out, cap) : launch_batch_mom<3, 2>(b, st, err, order, out, cap);
  }
  if (P == 1) return W == 1 ? launch_batch<1, 1>(b, st, err, order, out, cap) : launch_batch<1, 2>(b, st, err, order, out, cap);
  if (P == 2) return W == 1 ? launch_batch<2, 1>(b, st, err, order, out, cap) : launch_batch<2, 2>(b, st, err, order, out, cap);
  return W == 1 ? launch_batch<3, 1>(b, st, err, order, out, cap) : launch_batch<3, 2>(b, st, err, order, out, cap);
}

// the whole batch on stream st: marking, plan, three persistent kernels
int enqueue(xg_batch *b, hipStream_t st, int32_t *err, float *out, long long cap, int only_p = 0) {
  int rc = mark_compact(st, err, b->D, b->ne, b->el2v, b->lset, nullptr, nullptr, b->d_bits, b->d_scr, b->d_cut,
                        b->d_ncut);
  if (rc) return rc;
  const int nw = (b->ne + 31) / 32;
  k_mesh_counts<<<b->nmesh, 256, 0, st>>>(b->nmesh, b->d_off, b->d_bits + (size_t)4 * nw, b->d_cnt);
  k_plan<<<1, 64, 0, st>>>(b->nmesh, b->ndt, b->d_p, b->d_cnt, b->d_start, b->d_outoff, b->d_vstart, b->d_cls_mesh,
                           b->d_cls_nm, b->d_cls_total, b->d_counter);
  XG_LAUNCH_CHECK();
  if (!out) return XG_OK;
  for (int P = 3; P >= 1; --P) {  // heaviest first
    if (only_p && P != only_p) continue;
    rc = launch_degree(b, P, st, err, out, cap);
    if (rc) return rc;
  }
  return XG_OK;
}

void free_graph(xg_batch *b) {
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->g_out = nullptr;
}

template <typename T>
int dmalloc(T **p, size_t n) {
  XG_CHECK(hipMalloc((void **)p, sizeof(T) * (n ? n : 1)));
  return XG_OK;
}

}  // namespace

extern "C" int xg_batch_create(xg_ctx *ctx, int D, int nmesh, const int32_t *elem_off, const int32_t *mesh_p,
                               int ndt, const int *dts, int mask, const int32_t *el2v, const double *coords,
                               const double *lset, xg_batch **out) {
  if (!ctx || D != 3 || nmesh < 1 || nmesh > MAXMESH || !elem_off || !mesh_p || ndt < 1 || ndt > 3 || !dts ||
      mask < 1 || mask > 3 || !el2v || !coords || !lset || !out)
    return set_error(XG_ERR_ARG, "xg_batch_create: bad argument (D == 3, 1 <= nmesh <= 1024)");
  xg_batch *b = new xg_batch();
  b->D = D;
  b->nmesh = nmesh;
  b->ndt = ndt;
  b->mask = mask;
  for (int t = 0; t < ndt; ++t) {
    if (dts[t] < 0 || dts[t] > 2) {
      delete b;
      return set_error(XG_ERR_ARG, "xg_batch_create: bad domain type");
    }
    b->dts |= dts[t] << (2 * t);
  }
  b->h_off.assign(elem_off, elem_off + nmesh + 1);
  b->h_p.assign(mesh_p, mesh_p + nmesh);
  for (int m = 0; m < nmesh; ++m) {
    if (b->h_off[m + 1] < b->h_off[m] || b->h_p[m] < 1 || b->h_p[m] > 3) {
      delete b;
      return set_error(XG_ERR_ARG, "xg_batch_create: bad mesh offsets or degree");
    }
  }
  b->ne = b->h_off[nmesh];
  b->el2v = el2v;
  b->coords = coords;
  b->lset = lset;
  int dev = 0;
  hipDeviceProp_t prop;
  XG_CHECK(hipGetDevice(&dev));
  XG_CHECK(hipGetDeviceProperties(&prop, dev));
  b->ncu = prop.multiProcessorCount;
  const int nw = (b->ne + 31) / 32;
  int rc = 0;
  rc |= dmalloc(&b->d_off, nmesh + 1);
  rc |= dmalloc(&b->d_p, nmesh);
  rc |= dmalloc(&b->d_bits, (size_t)8 * nw);
  rc |= dmalloc(&b->d_scr, mark_scratch_ints(b->ne));
  rc |= dmalloc(&b->d_cut, b->ne);
  rc |= dmalloc(&b->d_ncut, 1);
  rc |= dmalloc(&b->d_cnt, nmesh);
  rc |= dmalloc(&b->d_start, nmesh + 1);
  rc |= dmalloc(&b->d_vstart, nmesh);
  rc |= dmalloc(&b->d_cls_mesh, (size_t)3 * nmesh);
  rc |= dmalloc(&b->d_cls_nm, 3);
  rc |= dmalloc(&b->d_cls_total, 3);
  rc |= dmalloc(&b->d_counter, 3);
  rc |= dmalloc(&b->d_outoff, nmesh + 1);
  if (rc) {
    delete b;
    return XG_ERR_HIP;
  }
  XG_CHECK(hipMemcpy(b->d_off, b->h_off.data(), sizeof(int32_t) * (nmesh + 1), hipMemcpyHostToDevice));
  XG_CHECK(hipMemcpy(b->d_p, b->h_p.data(), sizeof(int32_t) * nmesh, hipMemcpyHostToDevice));
  XG_CHECK(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  *out = b;
  return XG_OK;
}

extern "C" int xg_batch_destroy(xg_batch *b) {
  if (!b) return XG_OK;
  if (b->stream) hipStreamSynchronize(b->stream);
  free_graph(b);
  void *ptrs[] = {b->d_off, b->d_p, b->d_bits, b->d_scr, b->d_cut, b->d_ncut, b->d_cnt, b->d_start,
                  b->d_vstart, b->d_cls_mesh, b->d_cls_nm, b->d_cls_total, b->d_counter, b->d_outoff};
  for (void *p : ptrs)
    if (p) hipFree(p);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return XG_OK;
}

extern "C" int xg_batch_tune(xg_batch *b, int p, int waves_per_block, int blocks_per_cu, int grab) {
  if (!b || p < 1 || p > 3 || (waves_per_block != 1 && waves_per_block != 2) || blocks_per_cu < 1 ||
      blocks_per_cu > 64 || grab < 1 || grab > 64)
    return set_error(XG_ERR_ARG, "xg_batch_tune: bad argument");
  b->waves[p] = waves_per_block;
  b->bpc[p] = blocks_per_cu;
  b->grab[p] = grab;
  free_graph(b);
  return XG_OK;
}

extern "C" int xg_batch_method(xg_batch *b, int p, int method) {
  if (!b || p < 0 || p > 3 || (method != 0 && method != XG_METHOD_MOMENTS))
    return set_error(XG_ERR_ARG, "xg_batch_method: p in 0..3, method 0 (MFMA) or XG_METHOD_MOMENTS");
  for (int q = 1; q <= 3; ++q)
    if (p == 0 || p == q) b->method[q] = method;
  free_graph(b);
  return XG_OK;
}

// marking + plan only (on the context's stream); afterwards
// xg_batch_info gives the output size
extern "C" int xg_batch_plan(xg_ctx *ctx, xg_batch *b) {
  if (!ctx || !b) return set_error(XG_ERR_ARG, "xg_batch_plan: null");
  return enqueue(b, ctx->stream, ctx->d_err, nullptr, 0);
}

// host copies (after a plan / run and a sync): cut elements per mesh
// (nullable) and output offsets per mesh (nmesh + 1, floats; nullable)
extern "C" int xg_batch_info(xg_ctx *ctx, xg_batch *b, int32_t *cut_per_mesh, long long *out_off) {
  if (!ctx || !b) return set_error(XG_ERR_ARG, "xg_batch_info: null");
  XG_CHECK(hipStreamSynchronize(ctx->stream));
  XG_CHECK(hipStreamSynchronize(b->stream));
  if (cut_per_mesh)
    XG_CHECK(hipMemcpy(cut_per_mesh, b->d_cnt, sizeof(int32_t) * b->nmesh, hipMemcpyDeviceToHost));
  if (out_off)
    XG_CHECK(hipMemcpy(out_off, b->d_outoff, sizeof(long long) * (b->nmesh + 1), hipMemcpyDeviceToHost));
  return XG_OK;
}

// device pointer to the cut element ids (ascending, grouped by mesh)
extern "C" const int32_t *xg_batch_cut_list(xg_batch *b) { return b ? b->d_cut : nullptr; }

// one full batch: marking, plan and the element matrices into out (cap
// floats).  use_graph: capture the sequence once (per out buffer / tuning)
// and replay it.  only_p (0 = all): restrict to one degree (tuning).
extern "C" int xg_batch_run(xg_ctx *ctx, xg_batch *b, float *out, long long cap, int use_graph, int only_p) {
  if (!ctx || !b || !out || cap < 0 || only_p < 0 || only_p > 3) return set_error(XG_ERR_ARG, "xg_batch_run: bad argument");
  if (!use_graph || only_p) return enqueue(b, ctx->stream, ctx->d_err, out, cap, only_p);
  if (!b->exec || b->g_out != out || b->g_cap != cap) {
    free_graph(b);
    XG_CHECK(hipStreamSynchronize(ctx->stream));
    XG_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(b, b->stream, ctx->d_err, out, cap);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return check_hip(e, "hipStreamEndCapture");
    b->graph = g;
    XG_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_out = out;
    b->g_cap = cap;
  }
  // order the graph after prior work on the context's stream and make the
  // context's stream wait for it
  hipEvent_t ev;
  XG_CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  XG_CHECK(hipEventRecord(ev, ctx->stream));
  XG_CHECK(hipStreamWaitEvent(b->stream, ev, 0));
  XG_CHECK(hipGraphLaunch(b->exec, b->stream));
  XG_CHECK(hipEventRecord(ev, b->stream));
  XG_CHECK(hipStreamWaitEvent(ctx->stream, ev, 0));
  XG_CHECK(hipEventDestroy(ev));
  return XG_OK;
}
