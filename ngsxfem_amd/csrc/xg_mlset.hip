// xg_mlset.hip -- integration rules and element marking for several level
// sets on simplices.
//
// Reference: cutint/mlsetintegration.cpp:8-158
//            (CreateMultiLevelsetCutIntegrationRule),
//            cutint/straightcutrule.cpp:679-800 (StraightCutsIntegrationRule,
//            simplex branch), :503-555 (LevelsetWrapper), :131-196
//            (LevelsetCutSimplex::Decompose),
//            xfem/cutinfo.cpp:286-415 (MultiLevelsetCutInformation).
//
// One wave per element.  For each domain tuple (one domain type per level
// set) that the element is compatible with, the element is cut by the level
// sets that cut it, one after the other: the current sub-simplices (all of
// one dimension) sit in LDS, one per lane; each lane evaluates the next level
// set at its simplex's vertices (the reference's LevelsetWrapper, clamped to
// 1e-14 away from zero), looks its decomposition up in the same compile-time
// table as the single-level-set kernels (indexed by the simplex dimension,
// the target domain type and the sign pattern) and writes its children to
// the other LDS buffer at the offset of a wave prefix sum.  The final
// simplices get the base rule of their dimension, lane-parallel over
// (simplex, point); codimension 1, 2, 3 weights follow straightcutrule.cpp:
// 749-796.  Points are written in the reference's order (tuple, simplex,
// point), so they agree point by point with the CPU restatement.
#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

constexpr int MAXLS = 8;     // level sets
constexpr int MAXDOM = 64;   // domain tuples
constexpr int SCAP = 256;    // sub-simplices per LDS buffer
constexpr int ERR_CAP = 4;   // sticky error bits (xg_sync)
constexpr int ERR_DEGEN = 1;

struct MlsetArgs {
  const double *lset[MAXLS];
  int M, K;
  int8_t dtt[MAXDOM][MAXLS];
};

// simplex buffers of one wave, in dynamic LDS: two buffers of scap
// sub-simplices each (scap = min(3^M, SCAP): every cut by one level set
// leaves at most three children per simplex), then two counters.  Sizing by
// M keeps the block small enough for many waves per CU; a fixed 2 x 256
// simplices (49 KB) allowed three.
__host__ __device__ inline size_t ml_lds_bytes(int scap) {
  return sizeof(double) * 2 * (size_t)scap * 12 + sizeof(int) * 2;
}

inline int ml_scap(int M) {
  int c = 1;
  for (int j = 0; j < M && c < SCAP; ++j) c *= 3;
  return c < SCAP ? c : SCAP;
}

enum { MODE_COUNT = 0, MODE_FILL = 1, MODE_WSUM = 2 };

__device__ __forceinline__ double lsw_eval(const double *c, const double *p) {
  // LevelsetWrapper::operator() for a linear level set (the bilinear
  // coefficients vanish): terms in its loop order
  double v = 0.0;
  v += c[0];
  v += c[3] * p[2];
  v += c[2] * p[1];
  v += c[1] * p[0];
  return fabs(v) > 1e-14 ? v : 1e-14;  // update_initial_coefs
}

// exclusive prefix sum of v over the wave; total in *tot
__device__ __forceinline__ int wave_scan(int v, int *tot) {
  const int lane = threadIdx.x & 63;
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  *tot = __shfl(x, 63, 64);
  return x - v;
}

template <int D, int MODE>
__global__ __launch_bounds__(64) void k_mlset(int nel, const int32_t *__restrict__ elems,
                                              const int32_t *__restrict__ el2v, const double *__restrict__ coords,
                                              MlsetArgs A, int order, int32_t *__restrict__ offsets,
                                              double4 *__restrict__ pts, double *__restrict__ wsum,
                                              int32_t *err, int scap) {
  extern __shared__ double ml_smem[];
  double(*const buf_a)[4][3] = reinterpret_cast<double(*)[4][3]>(ml_smem);
  double(*const buf_b)[4][3] = buf_a + scap;
  int *const lcnt = reinterpret_cast<int *>(buf_b + scap);
  const int lane = threadIdx.x;
  for (int i = blockIdx.x; i < nel; i += gridDim.x) {
    const int e = elems ? elems[i] : i;
    const int32_t *ev = el2v + (size_t)e * (D + 1);
    double lv[MAXLS][D + 1];
    int ldt[MAXLS];
    for (int j = 0; j < A.M; ++j) {
#pragma unroll
      for (int k = 0; k <= D; ++k) lv[j][k] = A.lset[j][ev[k]];
      ldt[j] = straight_cut(lv[j], D + 1);
    }
    double F[D * D], Finv[D * D], absdet = 1.0;
    if (MODE != MODE_COUNT) {
      double V[D + 1][3];
#pragma unroll
      for (int k = 0; k <= D; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)ev[k] * D + d] : 0.0;
      jacobian<D>(V, F, Finv, absdet);
    }
    int n = 0;  // points so far (wave-uniform)
    double ws = 0.0;
    const int base = (MODE == MODE_FILL) ? offsets[i] : 0;
    bool bad = false;
    for (int t = 0; t < A.K; ++t) {
      bool compatible = true, cut = false;
      int cl[MAXLS], nc = 0, nif = 0;
      for (int j = 0; j < A.M; ++j) {
        const int dt = A.dtt[t][j];
        if (ldt[j] != IF && ldt[j] != dt) compatible = false;
        else if (ldt[j] == IF) cut = true;
      }
      if (!compatible) continue;
      if (!cut) {
        // the base rule of the element
        const int nq = rule_cnt(D, order);
        const double(*r)[4] = &xg_rule_data_d[rule_off(D, order)];
        if (MODE == MODE_FILL)
          for (int q = lane; q < nq; q += 64)
            pts[base + n + q] = make_double4(D > 0 ? r[q][0] : 0.0, D > 1 ? r[q][1] : 0.0, D > 2 ? r[q][2] : 0.0,
                                             r[q][3]);
        if (MODE == MODE_WSUM && lane == 0)
          for (int q = 0; q < nq; ++q) ws += r[q][3];
        n += nq;
        continue;
      }
      for (int j = 0; j < A.M; ++j)
        if (ldt[j] == IF) {
          cl[nc++] = j;
          nif += (A.dtt[t][j] == IF) ? 1 : 0;
        }
      // level 0: the element cut by the first cutting level set (its
      // element values, no clamp)
      int dim = D, ncur = 0;
      double(*cur)[4][3] = buf_a;
      double(*nxt)[4][3] = buf_b;
      if (lane == 0) {
        double P[4][3];
        for (int k = 0; k <= D; ++k)
          for (int d = 0; d < 3; ++d) P[k][d] = (d == k && k < D) ? 1.0 : 0.0;
        bool b0 = false;
        const int dt0 = A.dtt[t][cl[0]];
        // zeros: the reference throws for a zero next to a negative value
        bool hz = false, hn = false;
        for (int k = 0; k <= D; ++k) {
          hz |= lv[cl[0]][k] == 0.0;
          hn |= lv[cl[0]][k] < 0.0;
        }
        if (hz && hn) b0 = true;
        lcnt[0] = b0 ? 0 : children(D, P, lv[cl[0]], dt0, cur, b0);
        lcnt[1] = b0;
      }
      __syncthreads();
      ncur = lcnt[0];
      bad |= lcnt[1] != 0;
      __syncthreads();
      if (A.dtt[t][cl[0]] == IF) --dim;
      bool over = false;
      for (int c = 1; c < nc; ++c) {
        const int j = cl[c];
        const int dt = A.dtt[t][j];
        // LevelsetWrapper coefficients from the element values
        double co[4];
        co[0] = lv[j][D];
        co[1] = D >= 1 ? lv[j][0] - lv[j][D] : 0.0;
        co[2] = D >= 2 ? lv[j][1 < D ? 1 : 0] - lv[j][D] : 0.0;
        co[3] = D >= 3 ? lv[j][2 < D ? 2 : 0] - lv[j][D] : 0.0;
        int nn = 0;
        for (int c0 = 0; c0 < ncur; c0 += 64) {
          const int s = c0 + lane;
          double ch[3][4][3];
          int nch = 0;
          if (s < ncur) {
            double P[4][3], v[4];
            for (int k = 0; k <= dim; ++k) {
              for (int d = 0; d < 3; ++d) P[k][d] = cur[s][k][d];
              v[k] = lsw_eval(co, P[k]);
            }
            const int ds = straight_cut(v, dim + 1);
            if (ds == IF && dim == 0) {
              bad = true;  // a point cannot be cut (only a NaN value gets here)
            } else if (ds == IF) {
              bool b = false;
              nch = children(dim, P, v, dt, ch, b);
              bad |= b;
            } else if (ds == dt) {
              for (int k = 0; k <= dim; ++k)
                for (int d = 0; d < 3; ++d) ch[0][k][d] = P[k][d];
              nch = 1;
            }
          }
          int tot = 0;
          const int off = wave_scan(nch, &tot);
          if (nn + tot > scap) {
            over = true;
            break;
          }
          const int cdim = (dt == IF) ? dim - 1 : dim;
          for (int q = 0; q < nch; ++q)
            for (int k = 0; k <= cdim; ++k)
              for (int d = 0; d < 3; ++d) nxt[nn + off + q][k][d] = ch[q][k][d];
          nn += tot;
        }
        __syncthreads();
        if (over) break;
        double(*sw)[4][3] = cur;
        cur = nxt;
        nxt = sw;
        ncur = nn;
        if (dt == IF) --dim;
      }
      if (over) {
        if (lane == 0) atomicOr(err, ERR_CAP);
        ncur = 0;
      }
      // codimension factor (straightcutrule.cpp:749-796)
      double fac = 1.0;
      if (MODE != MODE_COUNT && nif > 0) {
        int ifs[3], k2 = 0;
        for (int c = 0; c < nc && k2 < 3; ++c)
          if (A.dtt[t][cl[c]] == IF) ifs[k2++] = cl[c];
        if (nif == 1) {
          fac = interface_factor<D>(lv[ifs[0]], Finv);
        } else if (nif == 2 && D == 2) {
          fac = 1.0 / absdet;
        } else if (nif == 2 && D == 3) {
          double n0[3], n1[3];
          for (int m = 0; m < 3; ++m) {
            n0[m] = lv[ifs[0]][m < D ? m : 0] - lv[ifs[0]][D];
            n1[m] = lv[ifs[1]][m < D ? m : 0] - lv[ifs[1]][D];
          }
          const double l0 = norm3(n0), l1 = norm3(n1);
          for (int m = 0; m < 3; ++m) {
            n0[m] /= l0;
            n1[m] /= l1;
          }
          double cp[3] = {n0[1] * n1[2] - n0[2] * n1[1], n0[2] * n1[0] - n0[0] * n1[2], n0[0] * n1[1] - n0[1] * n1[0]};
          const double cpf = 1.0 / norm3(cp);
          double v[3];
          for (int r = 0; r < 3; ++r) {
            double acc = 0.0;
            for (int q = 0; q < 3; ++q) acc += F[(r < D ? r : 0) * D + (q < D ? q : 0)] * cp[q];
            v[r] = cpf * acc;
          }
          fac = norm3(v) / absdet;
        } else if (nif == 3 && D == 3) {
          fac = 1.0 / absdet;
        } else {
          bad = true;  // "Codim possibilities available: 2D: 0,1,2; 3D: 0,1,2,3"
        }
      }
      // the base rule of dimension dim on each final simplex
      const int nq = rule_cnt(dim, order);
      const double(*r)[4] = &xg_rule_data_d[rule_off(dim, order)];
      if (MODE == MODE_FILL || MODE == MODE_WSUM) {
        for (int pq = lane; pq < ncur * nq; pq += 64) {
          const int s = pq / nq, q = pq - s * nq;
          const double(*P)[3] = cur[s];
          const double vol = simplex_measure(dim, P);
          double ow = 1.0;
          for (int m = 0; m < dim; ++m) ow -= r[q][m];
          double x[3];
          for (int d = 0; d < 3; ++d) x[d] = ow * P[0][d];
          for (int m = 0; m < dim; ++m)
            for (int d = 0; d < 3; ++d) x[d] += r[q][m] * P[m + 1][d];
          double w = r[q][3] * vol;
          if (nif > 0) w = w * fac;
          if (MODE == MODE_FILL) pts[base + n + pq] = make_double4(x[0], x[1], x[2], w);
          else ws += w;
        }
      }
      n += ncur * nq;
      __syncthreads();  // the buffers are reused by the next tuple
    }
    if (bad) atomicOr(err, ERR_DEGEN);
    if (MODE == MODE_COUNT && lane == 0) offsets[i] = n;
    if (MODE == MODE_WSUM) {
      // wave sum of the lanes' partial weight sums
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) ws += __shfl_xor(ws, o, 64);
      if (lane == 0) wsum[i] = ws;
    }
  }
}

// MultiLevelsetCutInformation::UpdateElementsOfDomainType: bit e = the
// element's tuple of CheckIfStraightCut values is one of the domain tuples
template <int D>
__global__ __launch_bounds__(256) void k_mlset_of_type(int ne, const int32_t *__restrict__ el2v, MlsetArgs A,
                                                       uint32_t *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  bool hit = false;
  if (e < ne) {
    int ldt[MAXLS];
    for (int j = 0; j < A.M; ++j) {
      double l[D + 1];
#pragma unroll
      for (int k = 0; k <= D; ++k) l[k] = A.lset[j][el2v[(size_t)e * (D + 1) + k]];
      ldt[j] = straight_cut(l, D + 1);
    }
    for (int t = 0; t < A.K; ++t) {
      bool m = true;
      for (int j = 0; j < A.M; ++j) m &= A.dtt[t][j] == ldt[j];
      hit |= m;
    }
  }
  const unsigned long long m = __ballot(hit);
  if ((threadIdx.x & 63) == 0) {
    const int w = (e - (int)(threadIdx.x & 63)) >> 5;
    const int nw = (ne + 31) >> 5;
    if (w < nw) out[w] = (uint32_t)(m & 0xffffffffull);
    if (w + 1 < nw) out[w + 1] = (uint32_t)(m >> 32);
  }
}

__global__ void k_bits_positive(int ne, const double *__restrict__ ws, uint32_t *out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const bool hit = e < ne && ws[e] > 0.0;
  const unsigned long long m = __ballot(hit);
  if ((threadIdx.x & 63) == 0) {
    const int w = (e - (int)(threadIdx.x & 63)) >> 5;
    const int nw = (ne + 31) >> 5;
    if (w < nw) out[w] = (uint32_t)(m & 0xffffffffull);
    if (w + 1 < nw) out[w + 1] = (uint32_t)(m >> 32);
  }
}

int make_args(int M, const double *const *lsets, int K, const int8_t *dtts, MlsetArgs &A, const char *who) {
  if (M < 1 || M > MAXLS || K < 1 || K > MAXDOM || !lsets || !dtts)
    return set_error(XG_ERR_ARG, std::string(who) + ": 1 <= M <= 8 level sets, 1 <= K <= 64 domain tuples");
  A.M = M;
  A.K = K;
  for (int j = 0; j < MAXLS; ++j) A.lset[j] = j < M ? lsets[j] : nullptr;
  for (int j = 0; j < M; ++j)
    if (!lsets[j]) return set_error(XG_ERR_ARG, std::string(who) + ": null level set");
  for (int t = 0; t < MAXDOM; ++t)
    for (int j = 0; j < MAXLS; ++j) {
      const int v = (t < K && j < M) ? dtts[t * M + j] : 0;
      if (v < 0 || v > 2) return set_error(XG_ERR_ARG, std::string(who) + ": domain types are NEG, POS, IF");
      A.dtt[t][j] = (int8_t)v;
    }
  return XG_OK;
}

int grid_for(int nel) { return nel < 65536 ? nel : 65536; }

template <int MODE>
int launch_ml(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v, const double *coords,
              const MlsetArgs &A, int order, int32_t *offsets, double *pts, double *wsum) {
  const int g = grid_for(nel);
  double4 *o = reinterpret_cast<double4 *>(pts);
  const int scap = ml_scap(A.M);
  const size_t lds = ml_lds_bytes(scap);
  switch (D) {
    case 1: k_mlset<1, MODE><<<g, 64, lds, ctx->stream>>>(nel, elems, el2v, coords, A, order, offsets, o, wsum, ctx->d_err, scap); break;
    case 2: k_mlset<2, MODE><<<g, 64, lds, ctx->stream>>>(nel, elems, el2v, coords, A, order, offsets, o, wsum, ctx->d_err, scap); break;
    default: k_mlset<3, MODE><<<g, 64, lds, ctx->stream>>>(nel, elems, el2v, coords, A, order, offsets, o, wsum, ctx->d_err, scap); break;
  }
  XG_LAUNCH_CHECK();
  return XG_OK;
}

}  // namespace

extern "C" int xg_mlset_rule_count(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v, int M,
                                   const double *const *lsets, int K, const int8_t *dtts, int order,
                                   int32_t *offsets) {
  if (!ctx || D < 1 || D > 3 || nel < 0 || !offsets || order < 0 || order / 2 + 1 > XG_RULE_NMAX)
    return set_error(XG_ERR_ARG, "xg_mlset_rule_count: bad argument");
  MlsetArgs A;
  int r = make_args(M, lsets, K, dtts, A, "xg_mlset_rule_count");
  if (r) return r;
  if (nel == 0) {
    XG_CHECK(hipMemsetAsync(offsets, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  if (!el2v) return set_error(XG_ERR_ARG, "xg_mlset_rule_count: null mesh");
  r = launch_ml<MODE_COUNT>(ctx, D, nel, elems, el2v, nullptr, A, order, offsets, nullptr, nullptr);
  if (r) return r;
  return scan_exclusive_i32(ctx, offsets, nel, true);
}

extern "C" int xg_mlset_rule_fill(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                  const double *coords, int M, const double *const *lsets, int K,
                                  const int8_t *dtts, int order, const int32_t *offsets, double *pts) {
  if (!ctx || D < 1 || D > 3 || nel < 0 || !offsets || order < 0 || order / 2 + 1 > XG_RULE_NMAX)
    return set_error(XG_ERR_ARG, "xg_mlset_rule_fill: bad argument");
  MlsetArgs A;
  int r = make_args(M, lsets, K, dtts, A, "xg_mlset_rule_fill");
  if (r) return r;
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !pts) return set_error(XG_ERR_ARG, "xg_mlset_rule_fill: null pointer");
  return launch_ml<MODE_FILL>(ctx, D, nel, elems, el2v, coords, A, order, const_cast<int32_t *>(offsets), pts,
                              nullptr);
}

extern "C" int xg_mlset_mark(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *coords, int M,
                             const double *const *lsets, int K, const int8_t *dtts, uint32_t *of_type,
                             uint32_t *contrib) {
  if (!ctx || D < 1 || D > 3 || ne < 0) return set_error(XG_ERR_ARG, "xg_mlset_mark: bad argument");
  MlsetArgs A;
  int r = make_args(M, lsets, K, dtts, A, "xg_mlset_mark");
  if (r) return r;
  if (ne == 0) return XG_OK;
  if (!el2v || (contrib && !coords)) return set_error(XG_ERR_ARG, "xg_mlset_mark: null pointer");
  const int g = (ne + 255) / 256;
  if (of_type) {
    switch (D) {
      case 1: k_mlset_of_type<1><<<g, 256, 0, ctx->stream>>>(ne, el2v, A, of_type); break;
      case 2: k_mlset_of_type<2><<<g, 256, 0, ctx->stream>>>(ne, el2v, A, of_type); break;
      default: k_mlset_of_type<3><<<g, 256, 0, ctx->stream>>>(ne, el2v, A, of_type); break;
    }
    XG_LAUNCH_CHECK();
  }
  if (contrib) {
    // weight sums of the order-0 rules in the context scratch
    r = ensure_ws(ctx, sizeof(double) * (size_t)ne);
    if (r) return r;
    double *ws = (double *)ctx->ws;
    r = launch_ml<MODE_WSUM>(ctx, D, ne, nullptr, el2v, coords, A, 0, nullptr, nullptr, ws);
    if (r) return r;
    k_bits_positive<<<g, 256, 0, ctx->stream>>>(ne, ws, contrib);
    XG_LAUNCH_CHECK();
  }
  return XG_OK;
}
