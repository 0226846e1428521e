// xg_elmat_core.hpp -- device building blocks of the element-matrix kernels
// (shared by the single-mesh kernels in xg_elmat.hip and the batched
// persistent kernels in xg_batch.hip).  See xg_elmat.hip for the mapping.
#pragma once
#include "xg_device.hpp"

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef double doublex4 __attribute__((ext_vector_type(4)));

namespace xg {
namespace ecore {


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
  static constexpr int ROWS = N;                      // LDS rows per component (real dofs only)
  static constexpr int SB_SIZE = 4 * ROWS * PAD;      // shape/gradient table per wave
  static constexpr int WAVES = 2;  // 128-thread blocks, <= 35 KB LDS per block
  static constexpr size_t LDS_PER_WAVE = sizeof(T) * (SB_SIZE + CH);
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
  double dacc[NB][NB][sizeof(T) == 4 ? 4 : 1];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      acc[a][b] = acc_t{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < (sizeof(T) == 4 ? 4 : 1); ++v) dacc[a][b][v] = 0.0;
    }
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
        sB[(0 * N + n) * PAD + lane] = m_mass ? phi[n] : T(0);
#pragma unroll
        for (int r = 0; r < D; ++r) {
          T gr = T(0);
#pragma unroll
          for (int cc = 0; cc < D; ++cc) gr += FinvT[cc * D + r] * dphi[n][cc];
          sB[((1 + r) * N + n) * PAD + lane] = m_lap ? gr : T(0);
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
        for (int b = 0; b < NB; ++b) {
          const int dof = b * 16 + ii;  // dofs >= N are zero rows of the MFMA tile
          av[u][b] = dof < N ? sB[(c * N + dof) * PAD + p0 + u] : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int bi = 0; bi < NB; ++bi)
#pragma unroll
          for (int bj = bi; bj < NB; ++bj) acc[bi][bj] = Mfma<T>::run(av[u][bi], av[u][bj] * wv[u], acc[bi][bj]);
      if constexpr (sizeof(T) == 4) {
        // the FP32 sums of a group of 8 points join FP64 accumulators, so
        // the rounding of long point sums does not grow with the rule size
#pragma unroll
        for (int bi = 0; bi < NB; ++bi)
#pragma unroll
          for (int bj = bi; bj < NB; ++bj)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              dacc[bi][bj][v] += (double)acc[bi][bj][v];
              acc[bi][bj][v] = 0.0f;
            }
      }
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
          const T val = (sizeof(T) == 4) ? (T)dacc[bi][bj][v] : acc[bi][bj][v];
          out[gi * N + gj] = val;
          if (bi < bj) out[gj * N + gi] = val;
        }
      }
}

// ---------------------------------------------------------------- split BF16
// x = hi + mid + lo, each a BF16 (8 significant bits, round to nearest
// even); the two subtractions are exact in FP32, so hi + mid + lo carries 24
// bits of x
typedef short shortx8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint32_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u >> 16;
}
__device__ __forceinline__ float bf16_value(uint32_t h) { return __uint_as_float(h << 16); }

__device__ __forceinline__ void split3(const float *x, shortx8 &hi, shortx8 &mid, shortx8 &lo) {
#pragma unroll
  for (int u = 0; u < 8; ++u) {
    const uint32_t h = bf16_bits(x[u]);
    const float r1 = x[u] - bf16_value(h);
    const uint32_t m = bf16_bits(r1);
    const float r2 = r1 - bf16_value(m);
    hi[u] = (short)h;
    mid[u] = (short)m;
    lo[u] = (short)bf16_bits(r2);
  }
}

__device__ __forceinline__ floatx4 mfma_bf16(const shortx8 &a, const shortx8 &b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c,
                                                 0, 0, 0);
}

// elmat_one with split-BF16 products on v_mfma_f32_16x16x32_bf16: K = 32 =
// 4 components x 8 points, so one MFMA covers a whole group of 8 points.
// Each lane stores sqrt(|w|) * (value or gradient component) in FP32 (the
// weight sign in sW), the contraction lane (dof, component) splits its 8
// values into hi / mid / lo BF16 vectors and runs the six products
// hi.hi + hi.mid + mid.hi + hi.lo + mid.mid + lo.hi (the dropped ones are
// below 2^-26 relative) with FP32 accumulation; the B operand is the A
// operand with the sign of the points of negative weight flipped.  Each
// group's FP32 sums join FP64 accumulators.  Output FP32.
template <int D, int P, typename SRC>
__device__ __forceinline__ void elmat_one_bf16x3(const SRC &src, int npts, const double *Finv, double absdet,
                                                 int mask, float *sB, float *sW, float *out) {
  using C = Cfg<D, P, float>;
  constexpr int N = C::N, NB = C::NB, CH = C::CH, PAD = C::PAD;
  const int lane = threadIdx.x & 63;
  const int c = lane >> 4, ii = lane & 15;
  floatx4 acc[NB][NB];
  double dacc[NB][NB][4];
#pragma unroll
  for (int a = 0; a < NB; ++a)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      acc[a][b] = floatx4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; ++v) dacc[a][b][v] = 0.0;
    }
  float FinvT[D * D];
#pragma unroll
  for (int k = 0; k < D * D; ++k) FinvT[k] = (float)Finv[k];
  const bool m_mass = mask & 1, m_lap = mask & 2;

  for (int c0 = 0; c0 < npts; c0 += CH) {
    const int nch = min(CH, npts - c0);
    if (lane < CH) {
      const int g = c0 + lane;
      float phi[N], dphi[N][D];
      float sw = 0.0f, sgn = 1.0f;
      if (g < npts) {
        double x[3], wr;
        src.point(g, x, wr);
        const double wp = wr * absdet;
        sgn = wp < 0.0 ? -1.0f : 1.0f;
        sw = (float)sqrt(fabs(wp));
        float lam[D + 1];
        double s = 1.0;
#pragma unroll
        for (int m = 0; m < D; ++m) {
          lam[m] = (float)x[m];
          s -= x[m];
        }
        lam[D] = (float)s;
        h1_shape<D, P, float>(lam, phi, dphi);
      } else {
#pragma unroll
        for (int n = 0; n < N; ++n) {
          phi[n] = 0.0f;
#pragma unroll
          for (int d = 0; d < D; ++d) dphi[n][d] = 0.0f;
        }
      }
      sW[lane] = sgn;
#pragma unroll
      for (int n = 0; n < N; ++n) {
        sB[(0 * N + n) * PAD + lane] = m_mass ? sw * phi[n] : 0.0f;
#pragma unroll
        for (int r = 0; r < D; ++r) {
          float gr = 0.0f;
#pragma unroll
          for (int cc = 0; cc < D; ++cc) gr += FinvT[cc * D + r] * dphi[n][cc];
          sB[((1 + r) * N + n) * PAD + lane] = m_lap ? sw * gr : 0.0f;
        }
      }
    }
    wave_sync();
    const int nch8 = min(CH, (nch + 7) & ~7);
    for (int p0 = 0; p0 < nch8; p0 += 8) {
      // sign flips of the group's negative-weight points (BF16 sign bit)
      short flip[8];
      bool anyneg = false;
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const bool neg = sW[p0 + u] < 0.0f;
        flip[u] = neg ? (short)0x8000 : (short)0;
        anyneg |= neg;
      }
      shortx8 ah[NB], am[NB], al[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        const int dof = b * 16 + ii;  // dofs >= N are zero rows of the MFMA tile
        float xv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) xv[u] = dof < N ? sB[(c * N + dof) * PAD + p0 + u] : 0.0f;
        split3(xv, ah[b], am[b], al[b]);
      }
#pragma unroll
      for (int bi = 0; bi < NB; ++bi)
#pragma unroll
        for (int bj = bi; bj < NB; ++bj) {
          shortx8 bh = ah[bj], bm = am[bj], bl = al[bj];
          if (anyneg) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              bh[u] ^= flip[u];
              bm[u] ^= flip[u];
              bl[u] ^= flip[u];
            }
          }
          floatx4 t = acc[bi][bj];
          t = mfma_bf16(ah[bi], bh, t);
          t = mfma_bf16(ah[bi], bm, t);
          t = mfma_bf16(am[bi], bh, t);
          t = mfma_bf16(ah[bi], bl, t);
          t = mfma_bf16(am[bi], bm, t);
          t = mfma_bf16(al[bi], bh, t);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            dacc[bi][bj][v] += (double)t[v];
            t[v] = 0.0f;
          }
          acc[bi][bj] = t;
        }
    }
    wave_sync();
  }
#pragma unroll
  for (int bi = 0; bi < NB; ++bi)
#pragma unroll
    for (int bj = bi; bj < NB; ++bj)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int gi = bi * 16 + 4 * c + v, gj = bj * 16 + ii;
        if (gi < N && gj < N) {
          const float val = (float)dacc[bi][bj][v];
          out[gi * N + gj] = val;
          if (bi < bj) out[gj * N + gi] = val;
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


}  // namespace ecore
}  // namespace xg
