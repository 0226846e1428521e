// xg_subdiv.hip -- subdivision cut rule for non-P1 (here P2) level sets on
// tetrahedra, fused with the fixed-node weight accumulation.
//
// Reference: cutint/xintegration.cpp:617-752 (lattice cut check), :755-1063
// (MakeQuadRule: 8-way refinement, uncut fill), :1084-1388 (CutSimplex<3>::
// MakeQuad: tet+prism / prism+prism split of a cut leaf), :223-242 and
// :289-326 (FillSimplexWithRule / FillSimplexCoDim1WithRule), :1698-1716
// (interface weight transform), xdecompose.hpp:34-85 (leaf check with the
// relative 1e-14 threshold), :129-146 (prism -> 3 tets).
//
// The rule of a P2-cut tet at level L has thousands of points, so it is
// never materialised: every point (x, w) is contracted on the fly into the
// weights of the degree-Q Lagrange basis psi_k on the reference tet,
//     W_k = sum_points w * psi_k(x),
// which integrate any polynomial of degree <= Q exactly (the element
// matrices of P2 elements need Q = 4).
//
// Mapping: one wavefront per element, depth-first traversal with the frame
// stack in LDS (wave-uniform control flow).  Lanes work in parallel on the
// lattice points of a cut check (up to C(2^L+3, 3) = 969 P2 evaluations at
// L = 4) and on the quadrature points of a fill; each lane keeps its own
// partial W (NEG and IF) in registers, reduced across lanes at the end.
// FP64 throughout, in the same operation order as the reference so the
// cut decisions are identical.
#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

constexpr int WPB = 2;  // waves per block
constexpr int MAXL = XG_LAT_LMAX;

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct Frame {
  double V[4][3];
  int lvl;
  int child;
};

template <int LO>
__device__ __forceinline__ double lset_eval(const double *c, const double *x) {
  constexpr int NL = ndof(3, LO);
  double lam[4];
  double s = 1.0;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    lam[m] = x[m];
    s -= x[m];
  }
  lam[3] = s;
  double phi[NL];
  h1_shape<3, LO, double>(lam, phi, (double(*)[3]) nullptr);
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < NL; ++k) v += phi[k] * c[k];
  return v;
}

template <int Q>
struct Acc {
  static constexpr int NQ = ndof(3, Q);
  double neg[NQ];
  double itf[NQ];
  __device__ void zero() {
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      neg[k] = 0.0;
      itf[k] = 0.0;
    }
  }
  __device__ __forceinline__ void add(int r, const double *x, double w) {
    double lam[4];
    double s = 1.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      lam[m] = x[m];
      s -= x[m];
    }
    lam[3] = s;
    double psi[NQ];
    h1_shape<3, Q, double>(lam, psi, (double(*)[3]) nullptr);
    if (r == NEG) {
#pragma unroll
      for (int k = 0; k < NQ; ++k) neg[k] += w * psi[k];
    } else {
#pragma unroll
      for (int k = 0; k < NQ; ++k) itf[k] += w * psi[k];
    }
  }
};

struct Geo {
  const double *Finv;  // 3x3 or identity when null
  double absdet;
  // w * | absdet F^{-T} n | / absdet   (xintegration.cpp:1698-1716)
  __device__ double if_weight(double w, const double *n) const {
    double v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += (Finv ? Finv[j * 3 + i] : (i == j ? 1.0 : 0.0)) * n[j];
      v[i] = absdet * acc;
    }
    return w * norm3(v) / absdet;
  }
};

// FillSimplexWithRule<3> (tf = Measure*6) or the uncut fill (tf = |det|):
// lane-parallel over the base rule points
template <int Q>
__device__ void fill_tet(Acc<Q> &acc, const double *p0, const double *p1, const double *p2, const double *p3,
                         int r, int order, bool measure6) {
  if (r == POS) return;  // POS weights are not accumulated
  double a[3], b[3], c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a[d] = p1[d] - p0[d];
    b[d] = p2[d] - p0[d];
    c[d] = p3[d] - p0[d];
  }
  const double tf = measure6 ? (fabs(det3(a, b, c)) / 6.0) * 6.0 : fabs(det3(a, b, c));
  const int nq = rule_cnt(3, order), off = rule_off(3, order);
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < nq; k += 64) {
    const double *rr = xg_rule_data_d[off + k];
    double ow = 1.0;
#pragma unroll
    for (int m = 0; m < 3; ++m) ow -= rr[m];
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * p0[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] += rr[0] * p1[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] += rr[1] * p2[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] += rr[2] * p3[d];
    const double w = measure6 ? rr[3] * tf : 1.0 * 1.0 * rr[3] * tf;
    acc.add(r, x, w);
  }
}

// FillSimplexCoDim1WithRule<3>
template <int Q>
__device__ void fill_tri(Acc<Q> &acc, const Geo &geo, const double *s0, const double *s1, const double *s2,
                         const double *pospoint, int order) {
  double a[3], b[3], c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a[d] = s1[d] - s0[d];
    b[d] = s2[d] - s0[d];
  }
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
  const double tf = norm3(c);
  const double na = norm3(a), nb = norm3(b);
  const double rel = na > nb ? na : nb;
  if (tf < 1e-14 * rel) return;
#pragma unroll
  for (int d = 0; d < 3; ++d) c[d] /= tf;
  double dd[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) dd[d] = pospoint[d] - s0[d];
  if (dd[0] * c[0] + dd[1] * c[1] + dd[2] * c[2] >= 0)
#pragma unroll
    for (int d = 0; d < 3; ++d) c[d] *= -1.0;
  const int nq = rule_cnt(2, order), off = rule_off(2, order);
  const int lane = threadIdx.x & 63;
  for (int k = lane; k < nq; k += 64) {
    const double *rr = xg_rule_data_d[off + k];
    double ow = 1.0;
    ow -= rr[0];
    ow -= rr[1];
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * s0[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] += rr[0] * s1[d];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] += rr[1] * s2[d];
    acc.add(IF, x, geo.if_weight(rr[3] * tf, c));
  }
}

// NumericalIntegrationStrategy::CheckIfCut, lane-parallel over the lattice
template <int LO>
__device__ int lattice_check(const double *cl, const double (*V)[3], int lvl) {
  const int lane = threadIdx.x & 63;
  const int np1 = 1 << lvl;
  const double dx = 1.0 / np1;
  const int off = xg_lat_off_d[lvl], cnt = xg_lat_cnt_d[lvl];
  bool hp = false, hn = false;
  for (int base = 0; base < cnt; base += 64) {
    const int k = base + lane;
    if (k < cnt) {
      const unsigned pk = xg_lat_d[off + k];
      const int I[3] = {(int)(pk & 255u), (int)((pk >> 8) & 255u), (int)(pk >> 16)};
      double pos[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) pos[d] = V[0][d];
#pragma unroll
      for (int j = 0; j < 3; ++j)
#pragma unroll
        for (int d = 0; d < 3; ++d) pos[d] += I[j] * dx * (V[j + 1][d] - V[0][d]);
      const double v = lset_eval<LO>(cl, pos);
      if (v >= 0.0) hp = true;
      else hn = true;
    }
    const bool anyp = __ballot(hp) != 0ull, anyn = __ballot(hn) != 0ull;
    if (anyp && anyn) return IF;
  }
  return (__ballot(hp) != 0ull) ? POS : NEG;
}

// Simplex<3>::CheckIfCut (relative 1e-14 threshold)
template <int LO>
__device__ int leaf_check(const double *cl, const double (*V)[3]) {
  bool haspos = false, hasneg = false;
  double sumneg = 0.0, sumpos = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double v = lset_eval<LO>(cl, V[i]);
    if (v >= 0.0) {
      sumpos += v;
      haspos = true;
    } else {
      sumneg -= v;
      hasneg = true;
    }
  }
  const double s = sumpos + sumneg;
  sumpos /= s;
  sumneg /= s;
  if (sumpos < 1e-14) haspos = false;
  if (sumneg < 1e-14) hasneg = false;
  if (haspos && hasneg) return IF;
  if (haspos) return POS;
  if (hasneg) return NEG;
  return -1;
}

// CutSimplex<3>::MakeQuad.  Vertex pool: 0..3 leaf vertices, 4..7 cut points.
template <int LO, int Q>
__device__ bool cut_leaf(Acc<Q> &acc, const Geo &geo, const double *cl, const double (*V)[3], int order) {
  constexpr int edge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  double pool[8][3];
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int d = 0; d < 3; ++d) pool[j][d] = V[j][d];
  int posp[8], negp[8], npos = 0, nneg = 0, ncut = 0;
  int posv[4], negv[4], nposv = 0, nnegv = 0;
  int v2c1[4] = {-1, -1, -1, -1}, v2c2[4] = {-1, -1, -1, -1};
  double vv[4];
  bool zero[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    zero[j] = false;
    vv[j] = lset_eval<LO>(cl, V[j]);
    if (vv[j] > 0) {
      posp[npos++] = j;
      posv[nposv++] = j;
    } else if (vv[j] < 0) {
      negp[nneg++] = j;
      negv[nnegv++] = j;
    } else {
      posp[npos++] = j;
      posv[nposv++] = j;
      zero[j] = true;
    }
  }
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    const int lv = edge[j][0], rv = edge[j][1];
    const double vl = vv[lv], vr = vv[rv];
    bool hascut = (vl * vr < 0);
    if (zero[lv] && vr < 0) hascut = true;
    if (zero[rv] && vl < 0) hascut = true;
    if (hascut) {
      if (ncut >= 4) return false;
      const double cp = vl / (vl - vr);
#pragma unroll
      for (int d = 0; d < 3; ++d) pool[4 + ncut][d] = (1 - cp) * V[lv][d] + cp * V[rv][d];
      if (v2c1[lv] == -1) v2c1[lv] = ncut; else v2c2[lv] = ncut;
      if (v2c1[rv] == -1) v2c1[rv] = ncut; else v2c2[rv] = ncut;
      ncut++;
    }
  }
  if (ncut == 3) {
    const bool minor_pos = nneg > npos;
    int *minorg = minor_pos ? posp : negp;
    int nmin = minor_pos ? npos : nneg;
    const int dt_minor = minor_pos ? POS : NEG;
    const bool major_pos = nneg <= npos;
    int *majorg = major_pos ? posp : negp;
    int nmaj = major_pos ? npos : nneg;
    const int dt_major = major_pos ? POS : NEG;
    const int *majv = (nnegv > nposv) ? negv : posv;
    if (nmin != 1 || nmaj != 3) return false;
    int g4[4] = {minorg[0], 4, 5, 6};
    fill_tet<Q>(acc, pool[g4[0]], pool[g4[1]], pool[g4[2]], pool[g4[3]], dt_minor, order, true);
    int g6[6] = {majorg[0], majorg[1], majorg[2], 4 + v2c1[majv[0]], 4 + v2c1[majv[1]], 4 + v2c1[majv[2]]};
    for (int i = 0; i < 3; ++i)
      fill_tet<Q>(acc, pool[g6[i]], pool[g6[i + 1]], pool[g6[i + 2]], pool[g6[i + 3]], dt_major, order, true);
    fill_tri<Q>(acc, geo, pool[4], pool[5], pool[6], pool[posp[0]], order);
    return true;
  }
  if (ncut == 4) {
    if (nposv != 2 || nnegv != 2) return false;
    for (int side = 0; side < 2; ++side) {
      const int *own = side == 0 ? posp : negp;
      const int *ownv = side == 0 ? posv : negv;
      const int *othv = side == 0 ? negv : posv;
      int pr[6];
      pr[0] = own[0];
      const int idxn = ownv[0];
      const int c1 = v2c1[idxn], c2 = v2c2[idxn];
      pr[1] = 4 + c1;
      pr[2] = 4 + c2;
      pr[3] = own[1];
      int c3 = -1;
      for (int l = 0; l < 4; ++l)
        if (c1 != l && c2 != l) c3 = l;
      int c4 = 6 - c3 - c2 - c1;
      if ((v2c1[othv[0]] == c1 && v2c2[othv[0]] == c4) || (v2c1[othv[0]] == c2 && v2c2[othv[0]] == c3)) {
        const int tt = c3;
        c3 = c4;
        c4 = tt;
      }
      pr[4] = 4 + c3;
      pr[5] = 4 + c4;
      for (int i = 0; i < 3; ++i)
        fill_tet<Q>(acc, pool[pr[i]], pool[pr[i + 1]], pool[pr[i + 2]], pool[pr[i + 3]], side == 0 ? POS : NEG,
                    order, true);
    }
    int diag1, diag2, ndiag1, ndiag2;
    if (v2c1[negv[0]] == v2c1[posv[0]] || v2c2[negv[0]] == v2c1[posv[0]]) {
      diag1 = v2c1[posv[0]];
      ndiag1 = v2c2[posv[0]];
    } else {
      diag1 = v2c2[posv[0]];
      ndiag1 = v2c1[posv[0]];
    }
    if (v2c1[negv[1]] == v2c1[posv[1]] || v2c2[negv[1]] == v2c1[posv[1]]) {
      diag2 = v2c1[posv[1]];
      ndiag2 = v2c2[posv[1]];
    } else {
      diag2 = v2c2[posv[1]];
      ndiag2 = v2c1[posv[1]];
    }
    fill_tri<Q>(acc, geo, pool[4 + diag1], pool[4 + ndiag1], pool[4 + diag2], pool[posp[0]], order);
    fill_tri<Q>(acc, geo, pool[4 + diag2], pool[4 + ndiag2], pool[4 + diag1], pool[posp[0]], order);
    return true;
  }
  return false;
}

template <int LO, int Q>
__global__ __launch_bounds__(64 * WPB) void k_subdiv(int nel, const double *__restrict__ lset_nodal, int L,
                                                     int order, const double *__restrict__ finv_all,
                                                     double *__restrict__ out_neg, double *__restrict__ out_if,
                                                     uint8_t *__restrict__ dom_out, int32_t *err) {
  constexpr int NL = ndof(3, LO);
  constexpr int NQ = ndof(3, Q);
  __shared__ Frame stk[WPB][MAXL + 1];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(blockIdx.x * WPB + wv);
  const int nwaves = gridDim.x * WPB;
  Frame *S = stk[wv];
  for (int e = wid; e < nel; e += nwaves) {
    double cl[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) cl[k] = lset_nodal[(size_t)e * NL + k];
    double Fi[9];
    Geo geo;
    geo.Finv = nullptr;
    geo.absdet = 1.0;
    if (finv_all) {
#pragma unroll
      for (int k = 0; k < 9; ++k) Fi[k] = finv_all[(size_t)e * 9 + k];
      const double dinv = Fi[0] * (Fi[4] * Fi[8] - Fi[5] * Fi[7]) - Fi[1] * (Fi[3] * Fi[8] - Fi[5] * Fi[6]) +
                          Fi[2] * (Fi[3] * Fi[7] - Fi[4] * Fi[6]);
      geo.absdet = 1.0 / fabs(dinv);
      geo.Finv = Fi;
    }
    Acc<Q> acc;
    acc.zero();
    bool ok = true;
    int dom = -1;
    // root frame
    if (lane == 0) {
      for (int j = 0; j < 4; ++j)
        for (int d = 0; d < 3; ++d) S[0].V[j][d] = (j == d) ? 1.0 : 0.0;
      S[0].lvl = L;
      S[0].child = -1;
    }
    wave_sync();
    int sp = 1;
    while (sp > 0) {
      Frame &f = S[sp - 1];
      const int lvl = f.lvl;
      if (f.child < 0) {
        // first visit
        double V[4][3];
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int d = 0; d < 3; ++d) V[j][d] = f.V[j][d];
        const int dt_self = lattice_check<LO>(cl, V, lvl);
        if (sp == 1) dom = dt_self;
        if (dt_self != IF) {
          fill_tet<Q>(acc, V[0], V[1], V[2], V[3], dt_self, order, false);
          --sp;
          continue;
        }
        if (lvl == 0) {
          const int dts = leaf_check<LO>(cl, V);
          if (dts < 0) ok = false;
          else if (dts == IF) ok = cut_leaf<LO, Q>(acc, geo, cl, V, order) && ok;
          else fill_tet<Q>(acc, V[0], V[1], V[2], V[3], dts, order, true);
          --sp;
          continue;
        }
        wave_sync();
        if (lane == 0) f.child = 0;
        wave_sync();
      }
      const int ch = f.child;
      if (ch >= 8) {
        --sp;
        continue;
      }
      // child vertices (xintegration.cpp:872-914)
      constexpr double baryc[10][4] = {
          {0.0, 0.0, 0.0, 1.0}, {0.5, 0.0, 0.0, 0.5}, {1.0, 0.0, 0.0, 0.0}, {0.0, 0.5, 0.0, 0.5},
          {0.5, 0.5, 0.0, 0.0}, {0.0, 1.0, 0.0, 0.0}, {0.0, 0.0, 0.5, 0.5}, {0.5, 0.0, 0.5, 0.0},
          {0.0, 0.5, 0.5, 0.0}, {0.0, 0.0, 1.0, 0.0}};
      constexpr int tets[8][4] = {{1, 2, 4, 7}, {3, 4, 5, 8}, {6, 7, 8, 9}, {3, 4, 8, 6},
                                  {3, 4, 1, 6}, {0, 1, 3, 6}, {8, 6, 1, 7}, {8, 4, 1, 7}};
      double nv[4][3];
      for (int j = 0; j < 4; ++j) {
        const int t = tets[ch][j];
#pragma unroll
        for (int c = 0; c < 3; ++c) nv[j][c] = 0.0;
#pragma unroll
        for (int d = 0; d < 4; ++d)
#pragma unroll
          for (int c = 0; c < 3; ++c) nv[j][c] += baryc[t][d] * f.V[d][c];
      }
      wave_sync();
      if (lane == 0) {
        f.child = ch + 1;
        Frame &g = S[sp];
        for (int j = 0; j < 4; ++j)
          for (int c = 0; c < 3; ++c) g.V[j][c] = nv[j][c];
        g.lvl = lvl - 1;
        g.child = -1;
      }
      wave_sync();
      ++sp;
    }
    // reduce the per-lane partial weights
#pragma unroll
    for (int k = 0; k < NQ; ++k) {
      double a = acc.neg[k], b = acc.itf[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        b += __shfl_xor(b, o, 64);
      }
      if (lane == 0) {
        if (out_neg) out_neg[(size_t)e * NQ + k] = a;
        if (out_if) out_if[(size_t)e * NQ + k] = b;
      }
    }
    if (lane == 0) {
      if (dom_out) dom_out[e] = (uint8_t)dom;
      if (!ok) atomicOr(err, 2);
    }
    wave_sync();
  }
}

}  // namespace

extern "C" int xg_subdiv_weights(xg_ctx *ctx, int nel, const double *lset_nodal, int lset_order, int L,
                                 int order, int q, const double *finv, double *out_neg, double *out_if,
                                 uint8_t *dom) {
  if (!ctx || nel < 0 || L < 0 || L > XG_LAT_LMAX || order < 0 || order / 2 + 1 > XG_RULE_NMAX ||
      (lset_order != 1 && lset_order != 2) || q != 4)
    return set_error(XG_ERR_ARG, "xg_subdiv_weights: bad argument (L <= 5, lset_order 1|2, q == 4)");
  if (nel == 0) return XG_OK;
  if (!lset_nodal) return set_error(XG_ERR_ARG, "xg_subdiv_weights: null level set");
  long long blocks = ((long long)nel + WPB - 1) / WPB;
  if (blocks > 65536) blocks = 65536;
  if (lset_order == 2)
    k_subdiv<2, 4><<<(int)blocks, 64 * WPB, 0, ctx->stream>>>(nel, lset_nodal, L, order, finv, out_neg, out_if,
                                                               dom, ctx->d_err);
  else
    k_subdiv<1, 4><<<(int)blocks, 64 * WPB, 0, ctx->stream>>>(nel, lset_nodal, L, order, finv, out_neg, out_if,
                                                               dom, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}
