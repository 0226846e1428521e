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
#include "xg_internal.hpp"

using namespace xg;

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace {

template <typename T>
struct Mfma;
template <>
struct Mfma<float> {
  using acc_t = floatx4;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma<double> {
  using acc_t = doublex4;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int D, int P, typename T>
struct Cfg {
  static constexpr int N = ndof(D, P);
  static constexpr int NB = (N + 15) / 16;
  static constexpr int CH = (sizeof(T) == 4) ? (NB == 1 ? 64 : 32) : (NB == 1 ? 32 : 16);
  static constexpr int PAD = CH + 1;
  static constexpr int WAVES = 2;  // 128-thread blocks, <= 35 KB LDS per block
  static constexpr size_t LDS_PER_WAVE = sizeof(T) * (4 * NB * 16 * PAD + CH);
};

// Point source: either a materialised CSR rule or the straight cut rule built
// on the fly.  prepare() sets up the element, count() gives the number of
// points, point() gives reference coordinates and the reference weight.
template <int D>
struct FusedSrc {
  double l[D + 1];
  Decomp dc;
  double P[D + 1 + 4][3];
  int uncut, nq, off, Ds, npts, bad;
  double ifac;
  __device__ void prepare(const double *lv, const double *Finv, int dt, int order) {
#pragma unroll
    for (int k = 0; k <= D; ++k) l[k] = lv[k];
    const int eldom = check_if_straight_cut<D>(l);
    npts = 0;
    bad = 0;
    ifac = 1.0;
    if (eldom != IF) {
      uncut = 1;
      Ds = D;
      nq = rule_cnt(D, order);
      off = rule_off(D, order);
      npts = (eldom == dt) ? nq : 0;
      return;
    }
    uncut = 0;
    if (!decompose<D>(l, dt, dc)) {  // the reference throws here
      bad = 1;
      return;
    }
    cut_point_array<D>(l, dc, P);
    Ds = dc.nv - 1;
    nq = rule_cnt(Ds, order);
    off = rule_off(Ds, order);
    npts = dc.ns * nq;
    if (dt == IF) ifac = interface_factor<D>(l, Finv);
  }
  __device__ void point(int g, double *x, double &w) const {
    if (uncut) {
      const double *r = xg_rule_data_d[off + g];
      x[0] = r[0];
      x[1] = D > 1 ? r[1] : 0.0;
      x[2] = D > 2 ? r[2] : 0.0;
      w = r[3];
      return;
    }
    const int s = g / nq, k = g - s * nq;
    const int8_t *idx = dc.v[s];
    const double tf = simplex_volume(dc.nv, idx, P);
    const double *r = xg_rule_data_d[off + k];
    double ow = 1.0;
    for (int m = 0; m < Ds; ++m) ow -= r[m];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * P[idx[0]][d];
    for (int m = 0; m < Ds; ++m)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] += r[m] * P[idx[m + 1]][d];
    w = r[3] * tf;
    w = w * ifac;
  }
};

// one element (or one (element, dt) pair): accumulate and store
template <int D, int P, typename T, typename SRC>
__device__ __forceinline__ void elmat_one(const SRC &src, int npts, const double *Finv, double absdet,
                                          int mask, T *sB, T *sW, T *out) {
  using C = Cfg<D, P, T>;
  using acc_t = typename Mfma<T>::acc_t;
  constexpr int N = C::N, NB = C::NB, CH = C::CH, PAD = C::PAD;
  const int lane = threadIdx.x & 63;
  const int c = lane >> 4, ii = lane & 15;
  acc_t acc[NB][NB];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) acc[a][b] = acc_t{0, 0, 0, 0};
  T FinvT[D * D];
#pragma unroll
  for (int k = 0; k < D * D; ++k) FinvT[k] = (T)Finv[k];
  const bool m_mass = mask & 1, m_lap = mask & 2;

  for (int c0 = 0; c0 < npts; c0 += CH) {
    const int nch = min(CH, npts - c0);
    if (lane < CH) {
      const int g = c0 + lane;
      T phi[N], dphi[N][D];
      T w = T(0);
      if (g < npts) {
        double x[3], wd;
        src.point(g, x, wd);
        w = (T)(wd * absdet);
        T lam[D + 1];
        T s = T(1);
#pragma unroll
        for (int m = 0; m < D; ++m) {
          lam[m] = (T)x[m];
          s -= lam[m];
        }
        lam[D] = s;
        h1_shape<D, P, T>(lam, phi, dphi);
      } else {
#pragma unroll
        for (int n = 0; n < N; ++n) {
          phi[n] = T(0);
#pragma unroll
          for (int d = 0; d < D; ++d) dphi[n][d] = T(0);
        }
      }
      sW[lane] = w;
#pragma unroll
      for (int n = 0; n < N; ++n) {
        sB[(0 * NB * 16 + n) * PAD + lane] = m_mass ? phi[n] : T(0);
#pragma unroll
        for (int r = 0; r < D; ++r) {
          T gr = T(0);
#pragma unroll
          for (int cc = 0; cc < D; ++cc) gr += FinvT[cc * D + r] * dphi[n][cc];
          sB[((1 + r) * NB * 16 + n) * PAD + lane] = m_lap ? gr : T(0);
        }
      }
    }
    wave_sync();
    for (int pp = 0; pp < nch; ++pp) {
      const T w = sW[pp];
      T av[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) av[b] = sB[(c * NB * 16 + b * 16 + ii) * PAD + pp];
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = bi; bj < NB; ++bj) acc[bi][bj] = Mfma<T>::run(av[bi], av[bj] * w, acc[bi][bj]);
    }
    wave_sync();
  }
  // D[row][col]: lane holds rows 4*(lane/16)+v, col lane%16
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi; bj < NB; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi = bi * 16 + 4 * c + v, gj = bj * 16 + ii;
        if (gi < N && gj < N) {
          out[gi * N + gj] = acc[bi][bj][v];
          if (bi < bj) out[gj * N + gi] = acc[bi][bj][v];
        }
      }
}

template <int D>
__device__ __forceinline__ void load_geom(const int32_t *__restrict__ el2v, const double *__restrict__ coords,
                                          const double *__restrict__ lset, int e, double (*V)[3], double *l) {
#pragma unroll
  for (int k = 0; k <= D; ++k) {
    const int v = el2v[(size_t)e * (D + 1) + k];
    if (l) l[k] = lset[v];
#pragma unroll
    for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
  }
}

template <int D, int P, typename T>
__global__ __launch_bounds__(128) void k_pipeline(int nel, const int32_t *__restrict__ elems,
                                                  const int32_t *__restrict__ el2v,
                                                  const double *__restrict__ coords,
                                                  const double *__restrict__ lset, int order, int ndt, int dts,
                                                  int mask, T *__restrict__ out, int32_t *err) {
  using C = Cfg<D, P, T>;
  extern __shared__ char smem[];
  const int wv = threadIdx.x >> 6;
  T *sB = reinterpret_cast<T *>(smem) + (size_t)wv * (C::LDS_PER_WAVE / sizeof(T));
  T *sW = sB + 4 * C::NB * 16 * C::PAD;
  const int lane = threadIdx.x & 63;
  // zero the padding rows once (dofs >= N never get written)
  for (int k = lane; k < 4 * C::NB * 16 * C::PAD; k += 64) sB[k] = T(0);
  wave_sync();
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
      FusedSrc<D> src;
      src.prepare(l, Finv, dt, order);
      if (src.bad && (threadIdx.x & 63) == 0) atomicOr(err, 1);
      elmat_one<D, P, T>(src, src.npts, Finv, absdet, mask, sB, sW, out + ((size_t)i * ndt + t) * NN);
    }
  }
}

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

template <int D, int P, typename T>
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
  T *sW = sB + 4 * C::NB * 16 * C::PAD;
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < 4 * C::NB * 16 * C::PAD; k += 64) sB[k] = T(0);
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
    CsrSrc src{pts, offsets[i]};
    elmat_one<D, P, T>(src, offsets[i + 1] - offsets[i], Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
  }
}

int grid_waves(int nel) {
  long long blocks = ((long long)nel + 1) / 2;
  if (blocks > 16384) blocks = 16384;
  if (blocks < 1) blocks = 1;
  return (int)blocks;
}

template <int D, int P, typename T>
int launch_pipeline(xg_ctx *ctx, int order, int nel, const int32_t *elems, const int32_t *el2v,
                    const double *coords, const double *lset, int ndt, int dts, int mask, void *out) {
  using C = Cfg<D, P, T>;
  const size_t lds = C::LDS_PER_WAVE * C::WAVES;
  k_pipeline<D, P, T><<<grid_waves(nel), 64 * C::WAVES, lds, ctx->stream>>>(
      nel, elems, el2v, coords, lset, order, ndt, dts, mask, (T *)out, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int D, int P, typename T>
int launch_csr(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
               const int32_t *offsets, const double *pts, int mask, void *out) {
  using C = Cfg<D, P, T>;
  const size_t lds = C::LDS_PER_WAVE * C::WAVES;
  k_elmat_csr<D, P, T><<<grid_waves(nel), 64 * C::WAVES, lds, ctx->stream>>>(
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
  int packed = 0;
  for (int t = 0; t < ndt; ++t) {
    if (dts[t] < 0 || dts[t] > 2) return set_error(XG_ERR_ARG, "xg_cut_pipeline: bad domain type");
    packed |= dts[t] << (2 * t);
  }
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !lset_v || !out) return set_error(XG_ERR_ARG, "xg_cut_pipeline: null pointer");
#define CALL_P(DD, PP, TT) launch_pipeline<DD, PP, TT>(ctx, order, nel, elems, el2v, coords, lset_v, ndt, packed, mask, out)
  if (precision == XG_PREC_FP32) XG_DISPATCH(D, p, float, CALL_P);
  if (precision == XG_PREC_FP64) XG_DISPATCH(D, p, double, CALL_P);
#undef CALL_P
  return set_error(XG_ERR_UNSUPPORTED, "xg_cut_pipeline: unsupported precision");
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
  return set_error(XG_ERR_UNSUPPORTED, "xg_elmat: unsupported precision");
}
