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
  // output row of accumulator element v in lane group c = lane / 16
  static __device__ __forceinline__ int row(int c, int v) { return 4 * c + v; }
};
template <>
struct Mfma<double> {
  using acc_t = doublex4;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // the FP64 variant interleaves the rows: element v of lane group c is row c + 4 v
  static __device__ __forceinline__ int row(int c, int v) { return c + 4 * v; }
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

// Point sources.  point(g, x, w) gives the reference coordinates x (in the
// accumulation type T) and the reference weight w of point g; elmat_one
// multiplies by |det F|.
//
// StagedSrc: the straight cut rule built on the fly.  The topology comes from
// the compile-time decomposition table (wave-uniform sign mask -> scalar
// loads), the FP64 cut points, sub-simplex volumes and weight scales are
// computed once per (element, dt) by a few lanes and staged in LDS, together
// with the base rules; each lane then maps its point with
// x = ow V0 + sum_m xi_m V_{m+1} (straightcutrule.cpp:80-100) in precision T.
constexpr int PIPE_NQMAX = 125;  // base rule points (order <= 8 tets)

template <int D, typename T>
struct StageLds {
  double P[D + 5][3];  // reference vertices, then cut points (FP64)
  double V[3][4][3];   // sub-simplex vertices (mapping order, FP64)
  double wsc[3];       // volume * interface factor per sub-simplex
};

template <int D, typename T>
struct StagedSrc {
  StageLds<D, T> *st;
  const double (*R)[4];  // base rule (xi0, xi1, xi2, w), in LDS
  int nq, nv1;      // points per sub-simplex, vertices per sub-simplex
  // point mapping in FP64 (the barycentric lam_D = 1 - sum xi of the element
  // is formed from these coordinates; FP32 here costs ~1e-6 for P3)
  __device__ void point(int g, double *x, double &w) const {
    const int s = (g >= nq) + (g >= 2 * nq);
    const int k = g - s * nq;
    const double *r = R[k];
    double ow = 1.0;
    for (int m = 0; m + 1 < nv1; ++m) ow -= r[m];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * st->V[s][0][d];
    for (int m = 0; m + 1 < nv1; ++m)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] += r[m] * st->V[s][m + 1][d];
    w = r[3] * st->wsc[s];
  }
};

__device__ __forceinline__ void wave_fence() { wave_sync(); }

// Builds the stage for (element with level set l, dt); returns the number of
// points (0 = no rule).  Wave-uniform.  *bad set when the reference throws.
template <int D, typename T>
__device__ int build_stage(StageLds<D, T> *st, const double *l, const double *Finv, int dt, int nq_vol, int nq_if,
                           int &nq, int &nv1, bool &bad) {
  const int lane = threadIdx.x & 63;
  const int eldom = check_if_straight_cut<D>(l);
  bad = false;
  if (eldom != IF) {
    if (eldom != dt) return 0;
    // uncut: the base rule itself = one "sub-simplex" (origin, e_1, .., e_D)
    if (lane < (D + 1) * 3) {
      const int m = lane / 3, d = lane % 3;
      st->V[0][m][d] = (m > 0 && d == m - 1) ? 1.0 : 0.0;
    }
    if (lane == 0) st->wsc[0] = 1.0;
    nq = nq_vol;
    nv1 = D + 1;
    wave_fence();
    return nq;
  }
  if (!cut_is_valid<D>(l)) {
    bad = true;
    return 0;
  }
  const int mask = sign_mask<D>(l);
  const DecompEntry &E = xg_decomp_d.e[D - 1][dt][mask];
  const int ns = E.ns, nv = E.nv, nc = E.nc;
  if (!E.ok) {
    bad = true;
    return 0;
  }
  // reference vertices and cut points (FP64, straightcutrule.cpp:50)
  if (lane < D + 1) {
#pragma unroll
    for (int d = 0; d < 3; ++d) st->P[lane][d] = (d == lane && lane < D) ? 1.0 : 0.0;
  } else if (lane < D + 1 + nc) {
    const int c = lane - (D + 1);
    const int i = E.ce[c][0], j = E.ce[c][1];
    double li = 0.0, lj = 0.0;
#pragma unroll
    for (int k = 0; k <= D; ++k) {
      li = (k == i) ? l[k] : li;
      lj = (k == j) ? l[k] : lj;
    }
    const double t = li / (li - lj);
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const double pi = (d == i && i < D) ? 1.0 : 0.0;
      const double pj = (d == j && j < D) ? 1.0 : 0.0;
      st->P[lane][d] = pi + t * (pj - pi);
    }
  }
  wave_fence();
  double ifac = 1.0;
  if (dt == IF) ifac = interface_factor<D>(l, Finv);
  if (lane < ns) {
    const double tf = simplex_volume(nv, E.v[lane], st->P);
    st->wsc[lane] = tf * ifac;
  }
  if (lane < ns * nv * 3) {
    const int s = lane / (nv * 3), m = (lane / 3) % nv, d = lane % 3;
    st->V[s][m][d] = st->P[E.v[s][m]][d];
  }
  wave_fence();
  nq = (dt == IF) ? nq_if : nq_vol;
  nv1 = nv;
  return ns * nq;
}

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
        double x[3], wr;
        src.point(g, x, wr);
        w = (T)(wr * absdet);
        T lam[D + 1];
        double s = 1.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
          lam[m] = (T)x[m];
          s -= x[m];
        }
        lam[D] = (T)s;
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
    // contraction over the chunk's points, 8 at a time: all LDS operand loads
    // of a group are issued before its MFMAs (points past npts carry zero
    // weight and zero shape rows, so rounding nch up to 8 adds nothing)
    const int nch8 = min(CH, (nch + 7) & ~7);
    for (int p0 = 0; p0 < nch8; p0 += 8) {
      T av[8][NB], wv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        wv[u] = sW[p0 + u];
#pragma unroll
        for (int b = 0; b < NB; ++b) av[u][b] = sB[(c * NB * 16 + b * 16 + ii) * PAD + p0 + u];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int bi = 0; bi < NB; ++bi)
#pragma unroll
          for (int bj = bi; bj < NB; ++bj) acc[bi][bj] = Mfma<T>::run(av[u][bi], av[u][bj] * wv[u], acc[bi][bj]);
    }
    wave_sync();
  }
  // D[row][col]: lane holds col lane%16 and rows Mfma<T>::row(lane/16, v)
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi; bj < NB; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi = bi * 16 + Mfma<T>::row(c, v), gj = bj * 16 + ii;
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
  __shared__ StageLds<D, T> s_stage[C::WAVES];
  __shared__ double s_rule[2][PIPE_NQMAX][4];  // base rules of dimension D and D-1
  const int wv = threadIdx.x >> 6;
  T *sB = reinterpret_cast<T *>(smem) + (size_t)wv * (C::LDS_PER_WAVE / sizeof(T));
  T *sW = sB + 4 * C::NB * 16 * C::PAD;
  const int lane = threadIdx.x & 63;
  // zero the padding rows once (dofs >= N never get written)
  for (int k = lane; k < 4 * C::NB * 16 * C::PAD; k += 64) sB[k] = T(0);
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
      elmat_one<D, P, T>(src, npts, Finv, absdet, mask, sB, sW, out + ((size_t)i * ndt + t) * NN);
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
    CsrSrc<T> src{pts, offsets[i]};
    elmat_one<D, P, T>(src, offsets[i + 1] - offsets[i], Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
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

template <int P, int Q, typename T>
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
  T *sW = sB + 4 * C::NB * 16 * C::PAD;
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < 4 * C::NB * 16 * C::PAD; k += 64) sB[k] = T(0);
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
    elmat_one<3, P, T>(src, NQ, Finv, absdet, mask, sB, sW, out + (size_t)i * NN);
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
  if (xg_rule_cnt[D][order / 2 + 1] > PIPE_NQMAX)
    return set_error(XG_ERR_UNSUPPORTED, "xg_cut_pipeline: order too high for the fused kernel (use xg_elmat)");
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

template <int P, int Q, typename T>
int launch_weights(xg_ctx *ctx, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
                   const double *weights, int mask, void *out) {
  using C = Cfg<3, P, T>;
  const size_t lds = C::LDS_PER_WAVE * C::WAVES;
  k_elmat_weights<P, Q, T><<<grid_waves(nel), 64 * C::WAVES, lds, ctx->stream>>>(nel, elems, el2v, coords, weights,
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
  return set_error(XG_ERR_UNSUPPORTED, "xg_elmat: unsupported precision");
}
