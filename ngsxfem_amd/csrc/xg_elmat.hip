// xg_elmat.hip -- element matrices B^T W B on MFMA, on a materialised rule or
// fused with the straight-cut rule construction.
//
// Reference: xfem/symboliccutbfi.cpp:76-304 (T_CalcElementMatrixAdd, SIMD
// path): for the proxy pairs (u, v) and (grad u, grad v),
//   elmat += B_test^T diag(w_i |det J_i|) B_trial          (AddABt :272-287)
// i.e. elmat = sum_i w_i |J| (phi phi^T + grad phi grad phi^T).
//
// Mapping on CDNA4 (one wavefront per element):
//  * quadrature points are produced lane-parallel (lane = point, chunks of CH
//    points); each lane evaluates the degree-P Lagrange shapes and physical
//    gradients of its point and stores them in LDS as B[c][dof][point]
//    (c = 0: value, 1..D: gradient components), padded to 16 dofs per block
//    and CH+1 points per row so both the stores (lane = point) and the MFMA
//    operand loads (lane = (c, dof)) are bank-conflict free;
//  * the contraction over (point, component) runs on the matrix cores:
//    v_mfma_{f32,f64}_16x16x4 with K = the 4 components of one point, so one
//    MFMA per point per 16x16 dof block (blocks bi <= bj only; the lower
//    blocks are mirrored at store time).  The B operand is the A operand
//    scaled by the lane's point weight, so no second LDS read is needed.
//  * FP64 is kept for the geometry (cut points, sub-simplex determinants,
//    weights); the shapes are evaluated in the accumulation type T.
#include "xg_device.hpp"
#include "xg_elmat_core.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

using namespace xg::ecore;

// SPLIT = 1: split-BF16 products (T = float, elmat_one_bf16x3)
template <int D, int P, typename T, int SPLIT = 0>
__global__ __launch_bounds__(128) void k_pipeline(int nel, const int32_t *__restrict__ elems,
                                                  const int32_t *__restrict__ el2v,
                                                  const double *__restrict__ coords,
                                                  const double *__restrict__ lset, int order, int ndt, int dts,
                                                  int mask, T *__restrict__ out, int32_t *err) {
  using C = Cfg<D, P, T>;
  extern __shared__ char smem[];
  __shared__ StageLds<D, T> s_stage[C::WAVES];
  __shared__ double s_rule[2][PIPE_NQMAX][4];  // base rules of dimension D and D-1
  const int wv = threadIdx.x >> 6;
  T *sB = reinterpret_cast<T *>(smem) + (size_t)wv * (C::LDS_PER_WAVE / sizeof(T));
  T *sW = sB + C::SB_SIZE;
  const int lane = threadIdx.x & 63;
  // zero the padding rows once (dofs >= N never get written)
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
  __syncthreads();  // the only block barrier: before the per-wave loops
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  constexpr int NN = C::N * C::N;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
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
      if constexpr (SPLIT) elmat_one_bf16x3<D, P>(src, npts, Finv, absdet, mask, sB, sW, out + ((size_t)i * ndt + t) * NN);
      else elmat_one<D, P, T>(src, npts, Finv, absdet, mask, sB, sW, out + ((size_t)i * ndt + t) * NN);
    }
  }
}

template <typename T>
struct CsrSrc {
  const double4 *pts;
  int base;
  __device__ void point(int g, double *x, double &w) const {
    const double4 p = pts[base + g];
    x[0] = p.x;
    x[1] = p.y;
    x[2] = p.z;
    w = p.w;
  }
};

template <int D, int P, typename T, int SPLIT = 0>
__global__ __launch_bounds__(128) void k_elmat_csr(int nel, const int32_t *__restrict__ elems,
                                                   const int32_t *__restrict__ el2v,
                                                   const double *__restrict__ coords,
                                                   const int32_t *__restrict__ offsets,
                                                   const double4 *__restrict__ pts, int mask,
                                                   T *__restrict__ out) {
  using C = Cfg<D, P, T>;
  extern __shared__ char smem[];
  const int wv = threadIdx.x >> 6;
  T *sB = reinterpret_cast<T *>(smem) + (size_t)wv * (C::LDS_PER_WAVE / sizeof(T));
  T *sW = sB + C::SB_SIZE;
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < C::SB_SIZE; k += 64) sB[k] = T(0);
  wave_sync();
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  constexpr int NN = C::N * C::N;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    double V[D + 1][3];
    load_geom<D>(el2v, coords, nullptr, e, V, nullptr);
    double F[D * D], Finv[D * D], absdet;
    jacobian<D>(V, F, Finv, absdet);
    CsrSrc<T> src{pts, offsets[i]};
    if constexpr (SPLIT) elmat_one_bf16x3<D, P>(src, offsets[i + 1] - offsets[i], Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
    else elmat_one<D, P, T>(src, offsets[i + 1] - offsets[i], Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
  }
}

// Fixed-node rule: the degree-Q Lagrange lattice nodes of the reference tet
// with per-element weights (learned rule or contracted subdivision rule).
template <int Q, typename T>
struct WeightSrc {
  const double *w;
  __device__ void point(int g, double *x, double &wt) const {
    constexpr Lattice<3, Q> L{};
    x[0] = (double)L.a[g][0] / Q;
    x[1] = (double)L.a[g][1] / Q;
    x[2] = (double)L.a[g][2] / Q;
    wt = w[g];
  }
};

template <int P, int Q, typename T, int SPLIT = 0>
__global__ __launch_bounds__(128) void k_elmat_weights(int nel, const int32_t *__restrict__ elems,
                                                       const int32_t *__restrict__ el2v,
                                                       const double *__restrict__ coords,
                                                       const double *__restrict__ weights, int mask,
                                                       T *__restrict__ out) {
  using C = Cfg<3, P, T>;
  constexpr int NQ = ndof(3, Q);
  extern __shared__ char smem[];
  const int wv = threadIdx.x >> 6;
  T *sB = reinterpret_cast<T *>(smem) + (size_t)wv * (C::LDS_PER_WAVE / sizeof(T));
  T *sW = sB + C::SB_SIZE;
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < C::SB_SIZE; k += 64) sB[k] = T(0);
  wave_sync();
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  constexpr int NN = C::N * C::N;
  for (int i = wid; i < nel; i += nwaves) {
    double Finv[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, absdet = 1.0;
    if (coords) {
      const int e = elems ? elems[i] : i;
      double V[4][3], F[9];
      load_geom<3>(el2v, coords, nullptr, e, V, nullptr);
      jacobian<3>(V, F, Finv, absdet);
    }
    WeightSrc<Q, T> src{weights + (size_t)i * NQ};
    if constexpr (SPLIT) elmat_one_bf16x3<3, P>(src, NQ, Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
    else elmat_one<3, P, T>(src, NQ, Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
  }
}

int grid_waves(int nel) {
  long long blocks = ((long long)nel + 1) / 2;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

template <int D, int P, typename T, int SPLIT = 0>
int launch_pipeline(xg_ctx *ctx, int order, int nel, const int32_t *elems, const int32_t *el2v,
                    const double *coords, const double *lset, int ndt, int dts, int mask, void *out) {
  using C = Cfg<D, P, T>;
  const size_t lds = C::LDS_PER_WAVE * C::WAVES;
  k_pipeline<D, P, T, SPLIT><<<grid_waves(nel), 64 * C::WAVES, lds, ctx->stream>>>(
      nel, elems, el2v, coords, lset, order, ndt, dts, mask, (T *)out, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int D, int P, typename T, int SPLIT = 0>
int launch_csr(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
               const int32_t *offsets, const double *pts, int mask, void *out) {
  using C = Cfg<D, P, T>;
  const size_t lds = C::LDS_PER_WAVE * C::WAVES;
  k_elmat_csr<D, P, T, SPLIT><<<grid_waves(nel), 64 * C::WAVES, lds, ctx->stream>>>(
      nel, elems, el2v, coords, offsets, reinterpret_cast<const double4 *>(pts), mask, (T *)out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

#define XG_DISPATCH(D, p, T, CALL)                  \
  do {                                              \
    if (D == 1 && p == 1) return CALL(1, 1, T);     \
    if (D == 1 && p == 2) return CALL(1, 2, T);     \
    if (D == 1 && p == 3) return CALL(1, 3, T);     \
    if (D == 2 && p == 1) return CALL(2, 1, T);     \
    if (D == 2 && p == 2) return CALL(2, 2, T);     \
    if (D == 2 && p == 3) return CALL(2, 3, T);     \
    if (D == 3 && p == 1) return CALL(3, 1, T);     \
    if (D == 3 && p == 2) return CALL(3, 2, T);     \
    if (D == 3 && p == 3) return CALL(3, 3, T);     \
  } while (0)

}  // namespace

extern "C" int xg_cut_pipeline(xg_ctx *ctx, int D, int p, int order, int nel, const int32_t *elems,
                               const int32_t *el2v, const double *coords, const double *lset_v, int ndt,
                               const int *dts, int mask, int precision, void *out) {
  if (!ctx || D < 1 || D > 3 || p < 1 || p > 3 || nel < 0 || ndt < 1 || ndt > 3 || !dts || order < 0 ||
      order / 2 + 1 > XG_RULE_NMAX || mask < 1 || mask > 3)
    return set_error(XG_ERR_ARG, "xg_cut_pipeline: bad argument");
  const bool moments = (precision & XG_METHOD_MOMENTS) != 0;
  precision &= ~XG_METHOD_MOMENTS;
  if (!moments && xg_rule_cnt[D][order / 2 + 1] > PIPE_NQMAX)
    return set_error(XG_ERR_UNSUPPORTED, "xg_cut_pipeline: order too high for the fused kernel (use xg_elmat)");
  int packed = 0;
  for (int t = 0; t < ndt; ++t) {
    if (dts[t] < 0 || dts[t] > 2) return set_error(XG_ERR_ARG, "xg_cut_pipeline: bad domain type");
    packed |= dts[t] << (2 * t);
  }
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !lset_v || !out) return set_error(XG_ERR_ARG, "xg_cut_pipeline: null pointer");
  if (moments) {
    if (precision != XG_PREC_FP32 && precision != XG_PREC_FP64)
      return set_error(XG_ERR_UNSUPPORTED, "xg_cut_pipeline: moment form writes FP32 or FP64");
    return launch_moment_pipeline(ctx, D, p, order, nel, elems, el2v, coords, lset_v, ndt, packed, mask,
                                  precision == XG_PREC_FP64, out);
  }
#define CALL_P(DD, PP, TT) launch_pipeline<DD, PP, TT>(ctx, order, nel, elems, el2v, coords, lset_v, ndt, packed, mask, out)
  if (precision == XG_PREC_FP32) XG_DISPATCH(D, p, float, CALL_P);
  if (precision == XG_PREC_FP64) XG_DISPATCH(D, p, double, CALL_P);
#undef CALL_P
#define CALL_S(DD, PP, TT) launch_pipeline<DD, PP, TT, 1>(ctx, order, nel, elems, el2v, coords, lset_v, ndt, packed, mask, out)
  if (precision == XG_PREC_BF16X3) XG_DISPATCH(D, p, float, CALL_S);
#undef CALL_S
  return set_error(XG_ERR_UNSUPPORTED, "xg_cut_pipeline: unsupported precision");
}

template <int P, int Q, typename T, int SPLIT = 0>
int launch_weights(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
                   const double *weights, int mask, void *out) {
  using C = Cfg<3, P, T>;
  const size_t lds = C::LDS_PER_WAVE * C::WAVES;
  k_elmat_weights<P, Q, T, SPLIT><<<grid_waves(nel), 64 * C::WAVES, lds, ctx->stream>>>(nel, elems, el2v, coords, weights,
                                                                                mask, (T *)out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_elmat_weights(xg_ctx *ctx, int p, int q, int nel, const int32_t *elems, const int32_t *el2v,
                                const double *coords, const double *weights, int mask, int precision, void *out) {
  if (!ctx || p < 1 || p > 2 || q != 4 || nel < 0 || mask < 1 || mask > 3)
    return set_error(XG_ERR_ARG, "xg_elmat_weights: bad argument (p <= 2, q == 4)");
  if (nel == 0) return XG_OK;
  if (!weights || !out || (coords && !el2v)) return set_error(XG_ERR_ARG, "xg_elmat_weights: null pointer");
  if (precision == XG_PREC_FP32)
    return p == 1 ? launch_weights<1, 4, float>(ctx, nel, elems, el2v, coords, weights, mask, out)
                  : launch_weights<2, 4, float>(ctx, nel, elems, el2v, coords, weights, mask, out);
  if (precision == XG_PREC_FP64)
    return p == 1 ? launch_weights<1, 4, double>(ctx, nel, elems, el2v, coords, weights, mask, out)
                  : launch_weights<2, 4, double>(ctx, nel, elems, el2v, coords, weights, mask, out);
  if (precision == XG_PREC_BF16X3)
    return p == 1 ? launch_weights<1, 4, float, 1>(ctx, nel, elems, el2v, coords, weights, mask, out)
                  : launch_weights<2, 4, float, 1>(ctx, nel, elems, el2v, coords, weights, mask, out);
  return set_error(XG_ERR_UNSUPPORTED, "xg_elmat_weights: unsupported precision");
}

extern "C" int xg_elmat(xg_ctx *ctx, int D, int p, int nel, const int32_t *elems, const int32_t *el2v,
                        const double *coords, const int32_t *offsets, const double *pts, int mask,
                        int precision, void *out) {
  if (!ctx || D < 1 || D > 3 || p < 1 || p > 3 || nel < 0 || mask < 1 || mask > 3)
    return set_error(XG_ERR_ARG, "xg_elmat: bad argument");
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts || !out) return set_error(XG_ERR_ARG, "xg_elmat: null pointer");
#define CALL_C(DD, PP, TT) launch_csr<DD, PP, TT>(ctx, nel, elems, el2v, coords, offsets, pts, mask, out)
  if (precision == XG_PREC_FP32) XG_DISPATCH(D, p, float, CALL_C);
  if (precision == XG_PREC_FP64) XG_DISPATCH(D, p, double, CALL_C);
#undef CALL_C
#define CALL_CS(DD, PP, TT) launch_csr<DD, PP, TT, 1>(ctx, nel, elems, el2v, coords, offsets, pts, mask, out)
  if (precision == XG_PREC_BF16X3) XG_DISPATCH(D, p, float, CALL_CS);
#undef CALL_CS
  return set_error(XG_ERR_UNSUPPORTED, "xg_elmat: unsupported precision");
}
