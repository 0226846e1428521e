// xg_subdiv.hip -- subdivision cut rule for non-P1 (here P2) level sets on
// tetrahedra, contracted on the fly into FP64 barycentric moments.
//
// Reference: cutint/xintegration.cpp:617-752 (lattice cut check), :755-1063
// (MakeQuadRule: 8-way refinement, uncut fill), :1084-1388 (CutSimplex<3>::
// MakeQuad: tet+prism / prism+prism split of a cut leaf), :223-242 and
// :289-326 (FillSimplexWithRule / FillSimplexCoDim1WithRule), :1698-1716
// (interface weight transform), xdecompose.hpp:34-85 (leaf check with the
// relative 1e-14 threshold), :129-146 (prism -> 3 tets).
//
// The rule of a P2-cut tet at level L has thousands of points, so it is never
// materialised: every point (x, w) adds w lam(x)^alpha, |alpha| = 2 PM, to
// FP64 moment accumulators (moments_table.h).  The element matrices of
// degree-PM elements are linear in these moments, and the weights of the
// degree-4 Lagrange basis on the reference tet (fixed-node rule) are
// W = nodal34 m for PM = 2.
//
// Mapping: one wavefront per element.
//  1. the level set is evaluated once at every point of the finest lattice
//     (2^L subdivisions per edge, C(2^L+3, 3) points: 969 at L = 4) into LDS.
//     Every point of every lattice check and leaf check of the refinement
//     tree is a point of this lattice, and its FP64 coordinates are exact
//     dyadic fractions, so the cached values are the ones a direct
//     evaluation at the node's own lattice gives;
//  2. the tree is walked breadth first.  Nodes carry integer lattice
//     coordinates and are named by their path of child indices.  The
//     candidates of a level (the 8 children of each cut node above) are
//     checked lane-parallel over (candidate, lattice point) pairs with LDS
//     atomic ORs of the sign bits; cut candidates go to the next level's
//     list, uncut NEG candidates to a fill list, cut leaves to a leaf list
//     (wave compaction with ballot + popcount, wave-uniform counts);
//  3. fills run lane-parallel over (node, base-rule point) pairs; cut leaves
//     run one per lane (tet+prism / prism+prism pieces);
//  4. the per-lane accumulators are reduced across the wave at the end.
// FP64 throughout; the cut decisions use the same values and comparisons as
// the reference, so they are identical to the oracle's.
#include <cstdlib>

#include "moments_table.h"
#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

constexpr int SD_CH = 256;  // candidates per chunk

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------- layout
__host__ __device__ constexpr int sd_nlat(int L) {
  return ((1 << L) + 1) * ((1 << L) + 2) * ((1 << L) + 3) / 6;
}
__host__ __device__ constexpr int sd_maxcut(int L) { return L >= 1 ? 1 << (3 * (L - 1)) : 1; }

struct SdLds {
  double *lat;      // sd_nlat(L)
  double *filltf;   // SD_CH
  double *red;      // 2 * 84
  double (*rule_vol)[4], (*rule_if)[4];  // base rules of order `order`
  double (*pool)[8][3];                  // 64 cut leaves x 8 points
  double *negtf, *ifw;                   // 192 / 128 piece weight factors
  uint32_t *negrec, *ifrec;              // 192 / 128 piece records
  uint32_t *nodeV;  // SD_CH * 4 packed vertices (x | y << 8 | z << 16)
  uint32_t *flags;  // SD_CH
  uint16_t *cur, *nxt;  // sd_maxcut(L)
  uint16_t *fillq;      // SD_CH: candidate | measure6 << 15
  uint16_t *leafq;      // SD_CH
};

__host__ __device__ inline size_t sd_align(size_t x) { return (x + 15) & ~size_t(15); }

__host__ __device__ inline size_t sd_bytes(int L, int nq_vol, int nq_if, SdLds *s, char *base) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = sd_align(o + bytes);
    return at;
  };
  const size_t a_lat = take(sizeof(double) * sd_nlat(L));
  const size_t a_ftf = take(sizeof(double) * SD_CH);
  const size_t a_red = take(sizeof(double) * 2 * 84);
  const size_t a_rv = take(sizeof(double) * 4 * nq_vol);
  const size_t a_ri = take(sizeof(double) * 4 * nq_if);
  const size_t a_pool = take(sizeof(double) * 64 * 8 * 3);
  const size_t a_ntf = take(sizeof(double) * 192);
  const size_t a_ifw = take(sizeof(double) * 128);
  const size_t a_nrec = take(sizeof(uint32_t) * 192);
  const size_t a_irec = take(sizeof(uint32_t) * 128);
  const size_t a_nv = take(sizeof(uint32_t) * SD_CH * 4);
  const size_t a_fl = take(sizeof(uint32_t) * SD_CH);
  const size_t a_cur = take(sizeof(uint16_t) * sd_maxcut(L));
  const size_t a_nxt = take(sizeof(uint16_t) * sd_maxcut(L));
  const size_t a_fq = take(sizeof(uint16_t) * SD_CH);
  const size_t a_lq = take(sizeof(uint16_t) * SD_CH);
  if (s) {
    s->lat = (double *)(base + a_lat);
    s->filltf = (double *)(base + a_ftf);
    s->red = (double *)(base + a_red);
    s->rule_vol = (double(*)[4])(base + a_rv);
    s->rule_if = (double(*)[4])(base + a_ri);
    s->pool = (double(*)[8][3])(base + a_pool);
    s->negtf = (double *)(base + a_ntf);
    s->ifw = (double *)(base + a_ifw);
    s->negrec = (uint32_t *)(base + a_nrec);
    s->ifrec = (uint32_t *)(base + a_irec);
    s->nodeV = (uint32_t *)(base + a_nv);
    s->flags = (uint32_t *)(base + a_fl);
    s->cur = (uint16_t *)(base + a_cur);
    s->nxt = (uint16_t *)(base + a_nxt);
    s->fillq = (uint16_t *)(base + a_fq);
    s->leafq = (uint16_t *)(base + a_lq);
  }
  return o;
}

// global lattice index of (a, b, c), a + b + c <= n, in the order of the
// level-L lattice table (c outermost, then b, then a)
__device__ __forceinline__ int lat_index(int a, int b, int c, int n) {
  const int M = n + 1;
  const int tc = (M * (M + 1) * (M + 2) - (M - c) * (M - c + 1) * (M - c + 2)) / 6;
  const int m = n - c;
  return tc + b * (m + 1) - (b * (b - 1)) / 2 + a;
}

__device__ __forceinline__ uint32_t pack3(int x, int y, int z) {
  return (uint32_t)x | ((uint32_t)y << 8) | ((uint32_t)z << 16);
}
__device__ __forceinline__ void unpack3(uint32_t p, int *v) {
  v[0] = (int)(p & 255u);
  v[1] = (int)((p >> 8) & 255u);
  v[2] = (int)((p >> 16) & 255u);
}

// children of the 8-way refinement (xintegration.cpp:872-914): the 10
// points of the refined tet in halves of the parent's barycentric
// coordinates, and the 8 child tets as point lists
__device__ __forceinline__ void child_vertices(const int (*W)[3], int ch, int (*C)[3]) {
  constexpr int BC2[10][4] = {{0, 0, 0, 2}, {1, 0, 0, 1}, {2, 0, 0, 0}, {0, 1, 0, 1}, {1, 1, 0, 0},
                              {0, 2, 0, 0}, {0, 0, 1, 1}, {1, 0, 1, 0}, {0, 1, 1, 0}, {0, 0, 2, 0}};
  constexpr int TETS[8][4] = {{1, 2, 4, 7}, {3, 4, 5, 8}, {6, 7, 8, 9}, {3, 4, 8, 6},
                              {3, 4, 1, 6}, {0, 1, 3, 6}, {8, 6, 1, 7}, {8, 4, 1, 7}};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int t = TETS[0][j];
#pragma unroll
    for (int q = 1; q < 8; ++q) t = (ch == q) ? TETS[q][j] : t;
    int b[4];
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      int v = BC2[0][d];
#pragma unroll
      for (int q = 1; q < 10; ++q) v = (t == q) ? BC2[q][d] : v;
      b[d] = v;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) C[j][c] = (b[0] * W[0][c] + b[1] * W[1][c] + b[2] * W[2][c] + b[3] * W[3][c]) / 2;
  }
}

// vertices of the node with the given path (depth digits, first digit most
// significant), root (n,0,0), (0,n,0), (0,0,n), (0,0,0) = reference vertices
__device__ __forceinline__ void node_vertices(int path, int depth, int n, int (*W)[3]) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int c = 0; c < 3; ++c) W[j][c] = (j == c) ? n : 0;
  for (int s = depth - 1; s >= 0; --s) {
    int C[4][3];
    child_vertices(W, (path >> (3 * s)) & 7, C);
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int c = 0; c < 3; ++c) W[j][c] = C[j][c];
  }
}

struct Geo {
  const double *Finv;  // 3x3 or identity when null
  double absdet;
  // |absdet F^{-T} n| / absdet (xintegration.cpp:1698-1716)
  __device__ double if_factor(const double *nrm) const {
    double v[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) acc += (Finv ? Finv[j * 3 + i] : (i == j ? 1.0 : 0.0)) * nrm[j];
      v[i] = absdet * acc;
    }
    return norm3(v) / absdet;
  }
};

// base-rule points of one tet mapped by x = ow p0 + sum xi_m p_{m+1}
template <int PM>
__device__ __forceinline__ void accum_tet(double *m, const double *p0, const double *p1, const double *p2,
                                          const double *p3, double tf, const double (*rule)[4], int k) {
  const double *rr = rule[k];
  double ow = 1.0;
#pragma unroll
  for (int q = 0; q < 3; ++q) ow -= rr[q];
  double x[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) x[d] = ow * p0[d] + rr[0] * p1[d] + rr[1] * p2[d] + rr[2] * p3[d];
  double lam[4];
  lam[0] = x[0];
  lam[1] = x[1];
  lam[2] = x[2];
  lam[3] = 1.0 - x[0] - x[1] - x[2];
  Moments<3, PM>::accum(lam, rr[3] * tf, m);
}

constexpr double sd_inv_factorial(int n) { return n <= 1 ? 1.0 : sd_inv_factorial(n - 1) / n; }

// exact moments of one tet from its vertices (p <= 2: Moments::hstep, the
// complete homogeneous recursion over the 4 vertices' barycentrics) added
// with the signed weight factor tf = +-6|T| (negative: a complemented
// leaf's piece is subtracted, by negating the sums around the update); the
// barycentrics are scaled by (|tf| / (2p + 3)!)^(1/2p) so the recursion's
// top degree lands in m weighted, and the alpha! factors are applied once
// per element (Moments::hfinish).  4 x 140 FMAs for a degree-4 tet against
// 14 points x 73 operations with the base rule; the result is that rule's
// point sum to rounding.
template <int PM>
__device__ __forceinline__ void htet(double *m, const double *p0, const double *p1, const double *p2,
                                     const double *p3, double tf) {
  using MM = Moments<3, PM>;
  double h[MM::NH];
#pragma unroll
  for (int k = 0; k < MM::NH; ++k) h[k] = 0.0;
  double c = sqrt(fabs(tf) * sd_inv_factorial(2 * PM + 3));
  if constexpr (PM == 2) c = sqrt(c);
  const double sg = tf < 0.0 ? -1.0 : 1.0;
#pragma unroll
  for (int k = 0; k < MM::NM; ++k) m[k] *= sg;
  const double *P[4] = {p0, p1, p2, p3};
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double l[4] = {c * P[v][0], c * P[v][1], c * P[v][2], c * (1.0 - P[v][0] - P[v][1] - P[v][2])};
    MM::hstep(l, h, m);
  }
#pragma unroll
  for (int k = 0; k < MM::NM; ++k) m[k] *= sg;
}

__device__ __forceinline__ double tet_tf(const double *p0, const double *p1, const double *p2, const double *p3,
                                         bool measure6) {
  double a[3], b[3], c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a[d] = p1[d] - p0[d];
    b[d] = p2[d] - p0[d];
    c[d] = p3[d] - p0[d];
  }
  const double v = fabs(det3(a, b, c));
  return measure6 ? (v / 6.0) * 6.0 : v;
}

// 4-bit fields (vertex -> cut index, 0xF = none)
__device__ __forceinline__ int nib(uint32_t x, int v) { return (int)((x >> (4 * v)) & 15u); }
__device__ __forceinline__ uint32_t set_nib(uint32_t x, int v, int val) {
  return (x & ~(15u << (4 * v))) | ((uint32_t)val << (4 * v));
}
// k-th set bit (ascending) of a 4-bit mask
__device__ __forceinline__ int kth_bit(int mask, int k) {
  int r = 0;
#pragma unroll
  for (int j = 3; j >= 0; --j) {
    const int below = __popc(mask & ((1 << j) - 1));
    if (((mask >> j) & 1) && below == k) r = j;
  }
  return r;
}

// The pieces of a cut leaf (CutSimplex<3>::MakeQuad, xintegration.cpp:
// 1084-1388): the point pool (0..3 leaf vertices, 4..7 cut points in edge
// order) and index lists of the NEG volume tets (3 bits per vertex) and of
// the interface triangles (3 vertices + the first POS vertex, which orients
// the normal).  Returns false where the reference throws.
struct LeafPieces {
  int nneg, nif;
  uint32_t neg[3];  // i0 | i1 << 3 | i2 << 6 | i3 << 9
  uint32_t itf[2];  // i0 | i1 << 3 | i2 << 6 | ipos << 9
};

__device__ __forceinline__ uint32_t pack4i(int a, int b, int c, int d) {
  return (uint32_t)a | ((uint32_t)b << 3) | ((uint32_t)c << 6) | ((uint32_t)d << 9);
}

__device__ bool leaf_topology(const double (*V)[3], const double *vv, double (*pool)[3], LeafPieces &P) {
  constexpr int edge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  P.nneg = 0;
  P.nif = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      pool[j][d] = V[j][d];
      pool[4 + j][d] = 0.0;
    }
  int posm = 0, negm = 0, zerom = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (vv[j] > 0) posm |= 1 << j;
    else if (vv[j] < 0) negm |= 1 << j;
    else {
      posm |= 1 << j;  // a zero vertex value counts as POS (xintegration.cpp:1159-1164)
      zerom |= 1 << j;
    }
  }
  const int npos = __popc(posm), nneg = __popc(negm);
  int ncut = 0;
  uint32_t v2c1 = 0xFFFFu, v2c2 = 0xFFFFu;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const int lv = edge[e][0], rv = edge[e][1];
    const double vl = vv[lv], vr = vv[rv];
    bool hascut = (vl * vr < 0);
    if (((zerom >> lv) & 1) && vr < 0) hascut = true;
    if (((zerom >> rv) & 1) && vl < 0) hascut = true;
    if (hascut) {
      if (ncut >= 4) return false;
      const double cp = vl / (vl - vr);
      double q[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) q[d] = (1 - cp) * V[lv][d] + cp * V[rv][d];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t == ncut)
#pragma unroll
          for (int d = 0; d < 3; ++d) pool[4 + t][d] = q[d];
      if (nib(v2c1, lv) == 15) v2c1 = set_nib(v2c1, lv, ncut); else v2c2 = set_nib(v2c2, lv, ncut);
      if (nib(v2c1, rv) == 15) v2c1 = set_nib(v2c1, rv, ncut); else v2c2 = set_nib(v2c2, rv, ncut);
      ncut++;
    }
  }
  if (ncut == 3) {
    const bool minor_pos = nneg > npos;
    const int nmin = minor_pos ? npos : nneg;
    if (nmin != 1 || 4 - nmin != 3) return false;
    const int minm = minor_pos ? posm : negm, majm = minor_pos ? negm : posm;
    if (!minor_pos) {
      P.neg[0] = pack4i(kth_bit(minm, 0), 4, 5, 6);
      P.nneg = 1;
    } else {
      const int m0 = kth_bit(majm, 0), m1 = kth_bit(majm, 1), m2 = kth_bit(majm, 2);
      const int g3 = 4 + nib(v2c1, m0), g4 = 4 + nib(v2c1, m1), g5 = 4 + nib(v2c1, m2);
      P.neg[0] = pack4i(m0, m1, m2, g3);
      P.neg[1] = pack4i(m1, m2, g3, g4);
      P.neg[2] = pack4i(m2, g3, g4, g5);
      P.nneg = 3;
    }
    P.itf[0] = pack4i(4, 5, 6, kth_bit(posm, 0));
    P.nif = 1;
    return true;
  }
  if (ncut == 4) {
    if (npos != 2 || nneg != 2) return false;
    const int n0 = kth_bit(negm, 0), n1 = kth_bit(negm, 1), p0 = kth_bit(posm, 0), p1 = kth_bit(posm, 1);
    // NEG prism (the reference's second side: own = NEG vertices)
    const int c1 = nib(v2c1, n0), c2 = nib(v2c2, n0);
    int c3 = -1;
#pragma unroll
    for (int l = 0; l < 4; ++l)
      if (c1 != l && c2 != l) c3 = l;
    int c4 = 6 - c3 - c2 - c1;
    if ((nib(v2c1, p0) == c1 && nib(v2c2, p0) == c4) || (nib(v2c1, p0) == c2 && nib(v2c2, p0) == c3)) {
      const int tt = c3;
      c3 = c4;
      c4 = tt;
    }
    const int q1 = 4 + c1, q2 = 4 + c2, q4 = 4 + c3, q5 = 4 + c4;
    P.neg[0] = pack4i(n0, q1, q2, n1);
    P.neg[1] = pack4i(q1, q2, n1, q4);
    P.neg[2] = pack4i(q2, n1, q4, q5);
    P.nneg = 3;
    int diag1, diag2, ndiag1, ndiag2;
    if (nib(v2c1, n0) == nib(v2c1, p0) || nib(v2c2, n0) == nib(v2c1, p0)) {
      diag1 = nib(v2c1, p0);
      ndiag1 = nib(v2c2, p0);
    } else {
      diag1 = nib(v2c2, p0);
      ndiag1 = nib(v2c1, p0);
    }
    if (nib(v2c1, n1) == nib(v2c1, p1) || nib(v2c2, n1) == nib(v2c1, p1)) {
      diag2 = nib(v2c1, p1);
      ndiag2 = nib(v2c2, p1);
    } else {
      diag2 = nib(v2c2, p1);
      ndiag2 = nib(v2c1, p1);
    }
    P.itf[0] = pack4i(4 + diag1, 4 + ndiag1, 4 + diag2, p0);
    P.itf[1] = pack4i(4 + diag2, 4 + ndiag2, 4 + diag1, p0);
    P.nif = 2;
    return true;
  }
  return false;
}

// exclusive prefix sum over the wave (and the total)
__device__ __forceinline__ int wave_excl_scan(int x, int lane, int &total) {
  int incl = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  total = __shfl(incl, 63, 64);
  return incl - x;
}

// Simplex<3>::CheckIfCut with the relative 1e-14 threshold
__device__ __forceinline__ int leaf_check(const double *vv) {
  bool haspos = false, hasneg = false;
  double sumneg = 0.0, sumpos = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double v = vv[i];
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

// wave-level append: lanes with `want` get consecutive slots from base
__device__ __forceinline__ int wave_slot(bool want, int base, int lane, int &count) {
  const unsigned long long bal = __ballot(want);
  const int pos = base + __popcll(bal & ((1ull << lane) - 1ull));
  count = __popcll(bal);
  return pos;
}

// OUT 0: moments (NM per element), 1: degree-4 fixed-node weights (PM == 2)
template <int LO, int PM, int WHICH, int OUT>
__global__ __launch_bounds__(64) void k_subdiv(int nel, const double *__restrict__ lset_nodal, int L, int order,
                                               const double *__restrict__ finv_all, double *__restrict__ out_neg,
                                               double *__restrict__ out_if, uint8_t *__restrict__ dom_out,
                                               int32_t *err) {
  constexpr int NL = ndof(3, LO);
  constexpr int NM = Moments<3, PM>::NM;
  extern __shared__ char smem[];
  SdLds S;
  // the moments are integrals of degree-2 PM polynomials over simplices:
  // once the order reaches 2 PM every exact rule gives them, so the minimal
  // symmetric rule of degree 2 PM (tet 24 instead of 64 points, triangle 12
  // instead of 16 for PM = 3) replaces the conical one of the order
  const bool minimal = order >= 2 * PM;
  const int vol_nq = minimal ? xg_minrule_cnt_d[3][PM] : rule_cnt(3, order);
  const int if_nq = minimal ? xg_minrule_cnt_d[2][PM] : rule_cnt(2, order);
  sd_bytes(L, vol_nq, if_nq, &S, smem);
  const int lane = threadIdx.x & 63;
  const int n = 1 << L;
  const int nlat = sd_nlat(L);
  const int loff = xg_lat_off_d[L];
  {
    const double(*src_v)[4] = minimal ? &xg_minrule_data_d[xg_minrule_off_d[3][PM]] : &xg_rule_data_d[rule_off(3, order)];
    const double(*src_i)[4] = minimal ? &xg_minrule_data_d[xg_minrule_off_d[2][PM]] : &xg_rule_data_d[rule_off(2, order)];
    for (int k = lane; k < 4 * (vol_nq + if_nq); k += 64) {
      const int q = k >> 2, c = k & 3;
      if (q < vol_nq) S.rule_vol[q][c] = src_v[q][c];
      else S.rule_if[q - vol_nq][c] = src_i[q - vol_nq][c];
    }
    wave_sync();
  }
  for (int e = blockIdx.x; e < nel; e += gridDim.x) {
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
    // 1. level set on the finest lattice
    const double inv_n = 1.0 / n;
    for (int q = lane; q < nlat; q += 64) {
      const unsigned pk = xg_lat_d[loff + q];
      const double x[3] = {(double)(pk & 255u) * inv_n, (double)((pk >> 8) & 255u) * inv_n,
                           (double)(pk >> 16) * inv_n};
      S.lat[q] = lagrange_eval3<LO>(cl, x);
    }
    double mneg[NM], mif[NM];
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      mneg[k] = 0.0;
      mif[k] = 0.0;
    }
    bool ok = true;
    int dom = -1;
    int ncur = 1;  // level 0: the root (path 0)
    wave_sync();
    for (int k = 0; k <= L && ncur > 0; ++k) {
      const int r = L - k;
      const int ncand = (k == 0) ? 1 : 8 * ncur;
      const int np = xg_lat_cnt_d[r], poff = xg_lat_off_d[r];
      int nnxt = 0;
      for (int cbase = 0; cbase < ncand; cbase += SD_CH) {
        const int nc = min(SD_CH, ncand - cbase);
        // candidate vertices, cleared flags
        for (int c = lane; c < nc; c += 64) {
          const int gc = cbase + c;
          const int path = (k == 0) ? 0 : (int)S.cur[gc >> 3] * 8 + (gc & 7);
          int W[4][3];
          node_vertices(path, k, n, W);
#pragma unroll
          for (int j = 0; j < 4; ++j) S.nodeV[c * 4 + j] = pack3(W[j][0], W[j][1], W[j][2]);
          S.flags[c] = 0u;
        }
        wave_sync();
        // lattice checks: sign bits of every (candidate, lattice point) pair
        for (int q = lane; q < nc * np; q += 64) {
          const int c = q / np, t = q - c * np;
          const unsigned pk = xg_lat_d[poff + t];
          const int I0 = (int)(pk & 255u), I1 = (int)((pk >> 8) & 255u), I2 = (int)(pk >> 16);
          int W0[3], W1[3], W2[3], W3[3];
          unpack3(S.nodeV[c * 4 + 0], W0);
          unpack3(S.nodeV[c * 4 + 1], W1);
          unpack3(S.nodeV[c * 4 + 2], W2);
          unpack3(S.nodeV[c * 4 + 3], W3);
          int P[3];
#pragma unroll
          for (int d = 0; d < 3; ++d)
            P[d] = W0[d] + (I0 * (W1[d] - W0[d]) + I1 * (W2[d] - W0[d]) + I2 * (W3[d] - W0[d])) / (1 << r);
          const double v = S.lat[lat_index(P[0], P[1], P[2], n)];
          atomicOr(&S.flags[c], v >= 0.0 ? 1u : 2u);
        }
        wave_sync();
        // classification and compaction
        int nfill = 0, nleaf = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
          const int c = c0 + lane;
          const bool valid = c < nc;
          const uint32_t f = valid ? S.flags[c] : 0u;
          const int cls = (f == 3u) ? IF : ((f == 1u) ? POS : NEG);
          if (k == 0) dom = __shfl(cls, 0, 64);
          int W[4][3] = {};
          double vv[4] = {0.0, 0.0, 0.0, 0.0};
          if (valid) {
#pragma unroll
            for (int j = 0; j < 4; ++j) unpack3(S.nodeV[c * 4 + j], W[j]);
          }
          int lc = NEG;
          bool leafcut = false;
          const bool isleafIF = valid && cls == IF && r == 0;
          if (isleafIF) {
#pragma unroll
            for (int j = 0; j < 4; ++j) vv[j] = S.lat[lat_index(W[j][0], W[j][1], W[j][2], n)];
            lc = leaf_check(vv);
            if (lc < 0) ok = false;
            leafcut = lc == IF;
          }
          const bool fill_uncut = valid && cls == NEG;
          const bool fill_leaf = isleafIF && lc == NEG;
          const bool wantfill = (WHICH & 1) && (fill_uncut || fill_leaf);
          const bool wantnext = valid && cls == IF && r > 0;
          int cnt;
          const int sf = wave_slot(wantfill, nfill, lane, cnt);
          if (wantfill) {
            double V[4][3];
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
              for (int d = 0; d < 3; ++d) V[j][d] = W[j][d] * inv_n;
            S.fillq[sf] = (uint16_t)(c | (fill_leaf ? 0x8000 : 0));
            S.filltf[sf] = tet_tf(V[0], V[1], V[2], V[3], fill_leaf);
          }
          nfill += cnt;
          const int sl = wave_slot(leafcut, nleaf, lane, cnt);
          if (leafcut) S.leafq[sl] = (uint16_t)c;
          nleaf += cnt;
          const int sn = wave_slot(wantnext, nnxt, lane, cnt);
          if (wantnext) S.nxt[sn] = (uint16_t)((k == 0) ? 0 : (int)S.cur[(cbase + c) >> 3] * 8 + ((cbase + c) & 7));
          nnxt += cnt;
        }
        wave_sync();
        // fills: (node, base-rule point) pairs
        if (WHICH & 1) {
          for (int q = lane; q < nfill * vol_nq; q += 64) {
            const int fi = q / vol_nq, t = q - fi * vol_nq;
            const int c = S.fillq[fi] & 0x7fff;
            double V[4][3];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              int w[3];
              unpack3(S.nodeV[c * 4 + j], w);
#pragma unroll
              for (int d = 0; d < 3; ++d) V[j][d] = w[d] * inv_n;
            }
            accum_tet<PM>(mneg, V[0], V[1], V[2], V[3], S.filltf[fi], S.rule_vol, t);
          }
        }
        // cut leaves, 64 at a time: topology one lane per leaf (point pool
        // and piece lists into LDS), then the pieces' base-rule points
        // lane-parallel over (piece, point) pairs
        for (int l0 = 0; l0 < nleaf; l0 += 64) {
          const int l = l0 + lane;
          LeafPieces PC;
          PC.nneg = 0;
          PC.nif = 0;
          if (l < nleaf) {
            const int c = S.leafq[l];
            double V[4][3], vv[4], pool[8][3];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              int w[3];
              unpack3(S.nodeV[c * 4 + j], w);
#pragma unroll
              for (int d = 0; d < 3; ++d) V[j][d] = w[d] * inv_n;
              vv[j] = S.lat[lat_index(w[0], w[1], w[2], n)];
            }
            if (!leaf_topology(V, vv, pool, PC)) {
              ok = false;
              PC.nneg = 0;
              PC.nif = 0;
            }
#pragma unroll
            for (int j = 0; j < 8; ++j)
#pragma unroll
              for (int d = 0; d < 3; ++d) S.pool[lane][j][d] = pool[j][d];
          }
          if (!(WHICH & 1)) PC.nneg = 0;
          if (!(WHICH & 2)) PC.nif = 0;
          int tot_neg, tot_if;
          const int bneg = wave_excl_scan(PC.nneg, lane, tot_neg);
          const int bif = wave_excl_scan(PC.nif, lane, tot_if);
#pragma unroll
          for (int k = 0; k < 3; ++k)
            if (k < PC.nneg) S.negrec[bneg + k] = (uint32_t)lane | (PC.neg[k] << 8);
#pragma unroll
          for (int k = 0; k < 2; ++k)
            if (k < PC.nif) S.ifrec[bif + k] = (uint32_t)lane | (PC.itf[k] << 8);
          wave_sync();
          // weight factors per piece (FillSimplexWithRule: (|det|/6)*6;
          // FillSimplexCoDim1WithRule: |cross| and the interface transform)
          for (int k = lane; k < tot_neg; k += 64) {
            const uint32_t r = S.negrec[k];
            const int sl = r & 63;
            const double(*pl)[3] = S.pool[sl];
            S.negtf[k] = tet_tf(pl[(r >> 8) & 7], pl[(r >> 11) & 7], pl[(r >> 14) & 7], pl[(r >> 17) & 7], true);
          }
          for (int k = lane; k < tot_if; k += 64) {
            const uint32_t r = S.ifrec[k];
            const double(*pl)[3] = S.pool[r & 63];
            const double *s0 = pl[(r >> 8) & 7], *s1 = pl[(r >> 11) & 7], *s2 = pl[(r >> 14) & 7];
            const double *pp = pl[(r >> 17) & 7];
            double a[3], b[3], cc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              a[d] = s1[d] - s0[d];
              b[d] = s2[d] - s0[d];
            }
            cc[0] = a[1] * b[2] - a[2] * b[1];
            cc[1] = a[2] * b[0] - a[0] * b[2];
            cc[2] = a[0] * b[1] - a[1] * b[0];
            const double tf = norm3(cc);
            const double na = norm3(a), nb = norm3(b);
            const double rel = na > nb ? na : nb;
            double wf = 0.0;
            if (!(tf < 1e-14 * rel)) {
#pragma unroll
              for (int d = 0; d < 3; ++d) cc[d] /= tf;
              const double dd0 = pp[0] - s0[0], dd1 = pp[1] - s0[1], dd2 = pp[2] - s0[2];
              if (dd0 * cc[0] + dd1 * cc[1] + dd2 * cc[2] >= 0)
#pragma unroll
                for (int d = 0; d < 3; ++d) cc[d] *= -1.0;
              wf = tf * geo.if_factor(cc);
            }
            S.ifw[k] = wf;
          }
          wave_sync();
          for (int q = lane; q < tot_neg * vol_nq; q += 64) {
            const int k = q / vol_nq, t = q - k * vol_nq;
            const uint32_t r = S.negrec[k];
            const double(*pl)[3] = S.pool[r & 63];
            accum_tet<PM>(mneg, pl[(r >> 8) & 7], pl[(r >> 11) & 7], pl[(r >> 14) & 7], pl[(r >> 17) & 7],
                          S.negtf[k], S.rule_vol, t);
          }
          for (int q = lane; q < tot_if * if_nq; q += 64) {
            const int k = q / if_nq, t = q - k * if_nq;
            const uint32_t r = S.ifrec[k];
            const double(*pl)[3] = S.pool[r & 63];
            const double *s0 = pl[(r >> 8) & 7], *s1 = pl[(r >> 11) & 7], *s2 = pl[(r >> 14) & 7];
            const double *rr = S.rule_if[t];
            const double ow = 1.0 - rr[0] - rr[1];
            double lam[4];
#pragma unroll
            for (int d = 0; d < 3; ++d) lam[d] = ow * s0[d] + rr[0] * s1[d] + rr[1] * s2[d];
            lam[3] = 1.0 - lam[0] - lam[1] - lam[2];
            Moments<3, PM>::accum(lam, rr[3] * S.ifw[k], mif);
          }
          wave_sync();
        }
        wave_sync();
      }
      // next level
      uint16_t *tmp = S.cur;
      S.cur = S.nxt;
      S.nxt = tmp;
      ncur = nnxt;
    }
    // wave reduction of the per-lane moments
#pragma unroll
    for (int k = 0; k < NM; ++k) {
      double a = mneg[k], b = mif[k];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        if (WHICH & 1) a += __shfl_xor(a, o, 64);
        if (WHICH & 2) b += __shfl_xor(b, o, 64);
      }
      if (lane == 0) {
        S.red[k] = a;
        S.red[NM + k] = b;
      }
    }
    wave_sync();
    if (OUT == 0) {
      for (int k = lane; k < NM; k += 64) {
        if ((WHICH & 1) && out_neg) out_neg[(size_t)e * NM + k] = S.red[k];
        if ((WHICH & 2) && out_if) out_if[(size_t)e * NM + k] = S.red[NM + k];
      }
    } else {
      if (lane < 35) {
        double a = 0.0, b = 0.0;
        for (int q = 0; q < 35; ++q) {
          a += nodal34[lane][q] * S.red[q];
          b += nodal34[lane][q] * S.red[NM + q];
        }
        if ((WHICH & 1) && out_neg) out_neg[(size_t)e * 35 + lane] = a;
        if ((WHICH & 2) && out_if) out_if[(size_t)e * 35 + lane] = b;
      }
    }
    const bool allok = __ballot(!ok) == 0ull;
    if (lane == 0) {
      if (dom_out) dom_out[e] = (uint8_t)dom;
      if (!allok) atomicOr(err, 2);
    }
    wave_sync();
  }
}

template <int LO, int PM, int WHICH, int OUT>
int launch_subdiv(xg_ctx *ctx, int nel, const double *lset_nodal, int L, int order, const double *finv,
                  double *out_neg, double *out_if, uint8_t *dom) {
  const bool minimal = order >= 2 * PM;  // as in k_subdiv
  const size_t lds = sd_bytes(L, minimal ? xg_minrule_cnt[3][PM] : xg_rule_cnt[3][order / 2 + 1],
                              minimal ? xg_minrule_cnt[2][PM] : xg_rule_cnt[2][order / 2 + 1], nullptr, nullptr);
  if (lds > 65536)
    XG_CHECK(hipFuncSetAttribute((const void *)k_subdiv<LO, PM, WHICH, OUT>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  long long blocks = nel;
  if (blocks > 262144) blocks = 262144;
  k_subdiv<LO, PM, WHICH, OUT><<<(int)blocks, 64, lds, ctx->stream>>>(nel, lset_nodal, L, order, finv, out_neg,
                                                                     out_if, dom, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

// p = 3: NEG and IF accumulators do not fit together (2 x 84 FP64 per
// lane), so the two parts run as separate passes
template <int LO, int PM>
int subdiv_moments_impl(xg_ctx *ctx, int nel, const double *lset_nodal, int L, int order, const double *finv,
                        double *out_neg, double *out_if, uint8_t *dom) {
  if constexpr (PM < 3) {
    return launch_subdiv<LO, PM, 3, 0>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  } else {
    int rc = XG_OK;
    if (out_neg || !out_if) rc = launch_subdiv<LO, PM, 1, 0>(ctx, nel, lset_nodal, L, order, finv, out_neg, nullptr, dom);
    if (rc == XG_OK && out_if)
      rc = launch_subdiv<LO, PM, 2, 0>(ctx, nel, lset_nodal, L, order, finv, nullptr, out_if,
                                       (out_neg || !out_if) ? nullptr : dom);
    return rc;
  }
}

// ======================================================================
// Exact-rule subdivision kernel (order >= 2 PM).
//
// Once the base rules integrate degree-2 PM polynomials exactly, the moments
// of every uncut node of the refinement tree are the exact integrals of
// lam^alpha over that node, independent of the level set: they are read from
// a table (one row per node of depth <= 5, built once per context) instead of
// being summed point by point.  The same table gives the complement of a cut
// leaf whose kept side holds three of its vertices: that part (the
// reference's prism of three tets, xintegration.cpp:1164-1238) has the moments
// of the whole leaf minus those of the one-vertex corner tet on the other
// side, so one tet is integrated instead of three.  Only the finest-lattice
// signs are kept in LDS (one bit per point); the FP64 values at the vertices
// of a cut leaf are evaluated again, with the same arithmetic, when the leaf
// is cut.  Node vertices come from a table indexed by (depth, path).  One
// accumulator set per lane: a launch integrates one volume side (SIDE = NEG or
// POS) and / or the interface (WHICH bit 0 / bit 1).
constexpr int SDX_DMAX = XG_LAT_LMAX;                                   // table depth
constexpr int SDX_NODES = ((1 << (3 * (SDX_DMAX + 1))) - 1) / 7;        // 37449
constexpr int SDX_LATQ = 214;  // entries of the level-0..3 lattice offset tables (4 + 10 + 35 + 165)
constexpr int SDX_CH = 128;                                             // candidates per chunk
__host__ __device__ constexpr int sdx_lvl_off(int k) { return ((1 << (3 * k)) - 1) / 7; }

// vertices of every node of depth <= 5 in level-5 lattice units (root
// (32,0,0), (0,32,0), (0,0,32), (0,0,0)), packed x | y << 8 | z << 16
__global__ void k_sdx_node_verts(uint4 *__restrict__ out) {
  const int id = blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= SDX_NODES) return;
  int k = 0;
  while (id >= sdx_lvl_off(k + 1)) ++k;
  int W[4][3];
  node_vertices(id - sdx_lvl_off(k), k, 1 << SDX_DMAX, W);
  out[id] = make_uint4(pack3(W[0][0], W[0][1], W[0][2]), pack3(W[1][0], W[1][1], W[1][2]),
                       pack3(W[2][0], W[2][1], W[2][2]), pack3(W[3][0], W[3][1], W[3][2]));
}

// exact degree-2PM moments of every node (minimal exact rule, FP64): one wave
// per node, rows of NM values
template <int PM>
__global__ __launch_bounds__(64) void k_sdx_node_moments(const uint4 *__restrict__ verts, double *__restrict__ out) {
  constexpr int NM = Moments<3, PM>::NM;
  const int id = blockIdx.x, lane = threadIdx.x & 63;
  const uint4 pv = verts[id];
  const uint32_t pk[4] = {pv.x, pv.y, pv.z, pv.w};
  double V[4][3];
  constexpr double inv = 1.0 / (1 << SDX_DMAX);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int w[3];
    unpack3(pk[j], w);
#pragma unroll
    for (int d = 0; d < 3; ++d) V[j][d] = w[d] * inv;
  }
  const double tf = tet_tf(V[0], V[1], V[2], V[3], false);
  const int nq = xg_minrule_cnt_d[3][PM];
  const double(*rule)[4] = &xg_minrule_data_d[xg_minrule_off_d[3][PM]];
  double m[NM];
#pragma unroll
  for (int a = 0; a < NM; ++a) m[a] = 0.0;
  for (int q = lane; q < nq; q += 64) accum_tet<PM>(m, V[0], V[1], V[2], V[3], tf, rule, q);
#pragma unroll
  for (int a = 0; a < NM; ++a) {
    double s = m[a];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0) out[(size_t)id * NM + a] = s;
  }
}

struct SdxLds {
  uint32_t *latbits;                     // sign bits (v >= 0) of the finest lattice
  double *red;                           // 2 NM
  double (*rule_vol)[4], (*rule_if)[4];  // minimal exact rules of degree 2 PM
  double (*pool)[8][3];                  // 64 cut leaves x (4 vertices, 4 cut points)
  double *voltf, *ifw;                   // 192 / 128 piece weight factors (signed)
  uint32_t *volrec, *ifrec;              // 192 / 128 piece records
  uint4 *nodeV;                          // SDX_CH candidate vertices (level-L units)
  uint32_t *flags;                       // SDX_CH
  uint16_t *cur, *nxt;                   // sd_maxcut(L) paths
  uint16_t *fillq;                       // 2 SDX_CH table rows
  uint16_t *leafq;                       // SDX_CH
  uint32_t *latq;                        // lattice offsets of levels 0..3 (the pair checks)
  double *cl;                            // 20: the element's level-set nodal values
  double *fi;                            // 10: F^{-1} (row major) and |det F|
  double *tab;                           // 128: table row sums, moments lane and lane + 64
};

__host__ __device__ inline size_t sdx_bytes(int L, int nq_vol, int nq_if, int NM, SdxLds *s, char *base) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = sd_align(o + bytes);
    return at;
  };
  const size_t a_lb = take(sizeof(uint32_t) * 2 * ((sd_nlat(L) + 63) / 64));
  const size_t a_red = take(sizeof(double) * 2 * NM);
  const size_t a_rv = take(sizeof(double) * 4 * nq_vol);
  const size_t a_ri = take(sizeof(double) * 4 * nq_if);
  const size_t a_cp = take(sizeof(double) * 64 * 8 * 3);
  const size_t a_vtf = take(sizeof(double) * 192);
  const size_t a_ifw = take(sizeof(double) * 128);
  const size_t a_vrec = take(sizeof(uint32_t) * 192);
  const size_t a_irec = take(sizeof(uint32_t) * 128);
  const size_t a_nv = take(sizeof(uint4) * SDX_CH);
  const size_t a_fl = take(sizeof(uint32_t) * SDX_CH);
  const size_t a_cur = take(sizeof(uint16_t) * sd_maxcut(L));
  const size_t a_nxt = take(sizeof(uint16_t) * sd_maxcut(L));
  const size_t a_fq = take(sizeof(uint16_t) * 2 * SDX_CH);
  const size_t a_lq = take(sizeof(uint16_t) * SDX_CH);
  const size_t a_lo = take(sizeof(uint32_t) * SDX_LATQ);
  const size_t a_cl = take(sizeof(double) * 20);
  const size_t a_fi = take(sizeof(double) * 10);
  const size_t a_tab = take(sizeof(double) * 128);
  if (s) {
    s->latq = (uint32_t *)(base + a_lo);
    s->cl = (double *)(base + a_cl);
    s->fi = (double *)(base + a_fi);
    s->tab = (double *)(base + a_tab);
    s->latbits = (uint32_t *)(base + a_lb);
    s->red = (double *)(base + a_red);
    s->rule_vol = (double(*)[4])(base + a_rv);
    s->rule_if = (double(*)[4])(base + a_ri);
    s->pool = (double(*)[8][3])(base + a_cp);
    s->voltf = (double *)(base + a_vtf);
    s->ifw = (double *)(base + a_ifw);
    s->volrec = (uint32_t *)(base + a_vrec);
    s->ifrec = (uint32_t *)(base + a_irec);
    s->nodeV = (uint4 *)(base + a_nv);
    s->flags = (uint32_t *)(base + a_fl);
    s->cur = (uint16_t *)(base + a_cur);
    s->nxt = (uint16_t *)(base + a_nxt);
    s->fillq = (uint16_t *)(base + a_fq);
    s->leafq = (uint16_t *)(base + a_lq);
  }
  return o;
}

// LDS of k_sdx_leaves alone: the arrays it uses (the walk's candidate and
// lattice arrays left out), about 19 KB at p = 3, so 8 one-wave blocks fit
// in a CU's 160 KB (2 waves on every SIMD; the whole SdxLds, 26 KB, allowed
// 6)
__host__ __device__ inline size_t sdx_leaf_bytes(int nq_vol, int nq_if, int NM, SdxLds *s, char *base) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = sd_align(o + bytes);
    return at;
  };
  const size_t a_red = take(sizeof(double) * 2 * NM);
  const size_t a_rv = take(sizeof(double) * 4 * nq_vol);
  const size_t a_ri = take(sizeof(double) * 4 * nq_if);
  const size_t a_cp = take(sizeof(double) * 64 * 8 * 3);
  const size_t a_vtf = take(sizeof(double) * 192);
  const size_t a_ifw = take(sizeof(double) * 128);
  const size_t a_vrec = take(sizeof(uint32_t) * 192);
  const size_t a_irec = take(sizeof(uint32_t) * 128);
  const size_t a_cl = take(sizeof(double) * 20);
  const size_t a_fi = take(sizeof(double) * 10);
  if (s) {
    *s = SdxLds{};
    s->red = (double *)(base + a_red);
    s->rule_vol = (double(*)[4])(base + a_rv);
    s->rule_if = (double(*)[4])(base + a_ri);
    s->pool = (double(*)[8][3])(base + a_cp);
    s->voltf = (double *)(base + a_vtf);
    s->ifw = (double *)(base + a_ifw);
    s->volrec = (uint32_t *)(base + a_vrec);
    s->ifrec = (uint32_t *)(base + a_irec);
    s->cl = (double *)(base + a_cl);
    s->fi = (double *)(base + a_fi);
  }
  return o;
}

// The pieces of a cut leaf for the kept volume side SIDE (CutSimplex<3>::
// MakeQuad, xintegration.cpp:1084-1388), as index lists into
// [leaf vertices 0..3, cut points 4..7]: `nvol` tets with sign `vsign`
// (-1: the listed corner tet is subtracted from the whole leaf, whose table
// row the caller adds, `comp`), and the interface triangles (3 points + the
// first POS vertex, which orients the normal).  Returns false where the
// reference throws.
struct LeafPieces2 {
  int nvol, nif;
  bool comp;
  uint32_t vol[3];
  uint32_t itf[2];
};

template <int SIDE>
__device__ bool leaf_pieces(const double (*V)[3], const double *vv, double (*cp)[3], LeafPieces2 &P) {
  constexpr int edge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  P.nvol = 0;
  P.nif = 0;
  P.comp = false;
#pragma unroll
  for (int j = 0; j < 4; ++j)
#pragma unroll
    for (int d = 0; d < 3; ++d) cp[j][d] = 0.0;
  int posm = 0, negm = 0, zerom = 0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    if (vv[j] > 0) posm |= 1 << j;
    else if (vv[j] < 0) negm |= 1 << j;
    else {
      posm |= 1 << j;  // a zero vertex value counts as POS (xintegration.cpp:1159-1164)
      zerom |= 1 << j;
    }
  }
  const int npos = __popc(posm), nneg = __popc(negm);
  int ncut = 0;
  uint32_t v2c1 = 0xFFFFu, v2c2 = 0xFFFFu;
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const int lv = edge[e][0], rv = edge[e][1];
    const double vl = vv[lv], vr = vv[rv];
    bool hascut = (vl * vr < 0);
    if (((zerom >> lv) & 1) && vr < 0) hascut = true;
    if (((zerom >> rv) & 1) && vl < 0) hascut = true;
    if (hascut) {
      if (ncut >= 4) return false;
      const double c = vl / (vl - vr);
      double q[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) q[d] = (1 - c) * V[lv][d] + c * V[rv][d];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        if (t == ncut)
#pragma unroll
          for (int d = 0; d < 3; ++d) cp[t][d] = q[d];
      if (nib(v2c1, lv) == 15) v2c1 = set_nib(v2c1, lv, ncut); else v2c2 = set_nib(v2c2, lv, ncut);
      if (nib(v2c1, rv) == 15) v2c1 = set_nib(v2c1, rv, ncut); else v2c2 = set_nib(v2c2, rv, ncut);
      ncut++;
    }
  }
  if (ncut == 3) {
    const bool minor_pos = nneg > npos;
    const int nmin = minor_pos ? npos : nneg;
    if (nmin != 1) return false;
    const int minm = minor_pos ? posm : negm;
    // the one-vertex corner tet [minor vertex, cut 0, cut 1, cut 2] is the
    // minor side; the major side (the reference's prism) is leaf - corner
    P.vol[0] = pack4i(kth_bit(minm, 0), 4, 5, 6);
    P.nvol = 1;
    P.comp = (SIDE == POS) ? !minor_pos : minor_pos;
    P.itf[0] = pack4i(4, 5, 6, kth_bit(posm, 0));
    P.nif = 1;
    return true;
  }
  if (ncut == 4) {
    if (npos != 2 || nneg != 2) return false;
    // the kept side's prism (the reference's side loop, own = kept vertices,
    // oth = the other side's first vertex for the orientation switch)
    const int ownm = (SIDE == POS) ? posm : negm, othm = (SIDE == POS) ? negm : posm;
    const int o0 = kth_bit(ownm, 0), o1 = kth_bit(ownm, 1), t0 = kth_bit(othm, 0);
    const int c1 = nib(v2c1, o0), c2 = nib(v2c2, o0);
    int c3 = -1;
#pragma unroll
    for (int l = 0; l < 4; ++l)
      if (c1 != l && c2 != l) c3 = l;
    int c4 = 6 - c3 - c2 - c1;
    if ((nib(v2c1, t0) == c1 && nib(v2c2, t0) == c4) || (nib(v2c1, t0) == c2 && nib(v2c2, t0) == c3)) {
      const int tt = c3;
      c3 = c4;
      c4 = tt;
    }
    const int q1 = 4 + c1, q2 = 4 + c2, q4 = 4 + c3, q5 = 4 + c4;
    P.vol[0] = pack4i(o0, q1, q2, o1);
    P.vol[1] = pack4i(q1, q2, o1, q4);
    P.vol[2] = pack4i(q2, o1, q4, q5);
    P.nvol = 3;
    const int n0 = kth_bit(negm, 0), n1 = kth_bit(negm, 1), p0 = kth_bit(posm, 0), p1 = kth_bit(posm, 1);
    int diag1, diag2, ndiag1, ndiag2;
    if (nib(v2c1, n0) == nib(v2c1, p0) || nib(v2c2, n0) == nib(v2c1, p0)) {
      diag1 = nib(v2c1, p0);
      ndiag1 = nib(v2c2, p0);
    } else {
      diag1 = nib(v2c2, p0);
      ndiag1 = nib(v2c1, p0);
    }
    if (nib(v2c1, n1) == nib(v2c1, p1) || nib(v2c2, n1) == nib(v2c1, p1)) {
      diag2 = nib(v2c1, p1);
      ndiag2 = nib(v2c2, p1);
    } else {
      diag2 = nib(v2c2, p1);
      ndiag2 = nib(v2c1, p1);
    }
    P.itf[0] = pack4i(4 + diag1, 4 + ndiag1, 4 + diag2, p0);
    P.itf[1] = pack4i(4 + diag2, 4 + ndiag2, 4 + diag1, p0);
    P.nif = 2;
    return true;
  }
  return false;
}

__device__ __forceinline__ void unpack_uint4(const uint4 &pv, int shift, int (*W)[3]) {
  const uint32_t pk[4] = {pv.x, pv.y, pv.z, pv.w};
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    unpack3(pk[j], W[j]);
#pragma unroll
    for (int d = 0; d < 3; ++d) W[j][d] >>= shift;
  }
}

__device__ __forceinline__ bool lat_bit(const uint32_t *bits, int idx) { return (bits[idx >> 5] >> (idx & 31)) & 1u; }

// Sum of the wave's per-lane arrays m[0..NM) into dst[0..NM), through LDS
// scratch of at least 23 x 65 doubles: 23 entries per round, lane kk of the
// round sums column kk (columns padded to 65 doubles so the summing lanes
// hit distinct banks).  Cheaper than 6 shuffle steps per entry.  Every lane
// of the wave calls it; it ends with dst written and visible to the wave.
template <int NM>
__device__ __forceinline__ void wave_sum_lds(const double *m, double *scr, double *dst, int lane,
                                             double *dst_hi = nullptr, int split = 64) {
  constexpr int RB = 23;
#pragma unroll
  for (int k0 = 0; k0 < NM; k0 += RB) {
#pragma unroll
    for (int kk = 0; kk < RB; ++kk)
      if (k0 + kk < NM) scr[kk * 65 + lane] = m[k0 + kk < NM ? k0 + kk : 0];
    wave_sync();
    if (lane < RB && k0 + lane < NM) {
      if (!dst_hi) {
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
        for (int j = 0; j < 64; j += 4) {
          s0 += scr[lane * 65 + j];
          s1 += scr[lane * 65 + j + 1];
          s2 += scr[lane * 65 + j + 2];
          s3 += scr[lane * 65 + j + 3];
        }
        dst[k0 + lane] = (s0 + s1) + (s2 + s3);
      } else {
        // two sums: lanes below `split` into dst, the others into dst_hi
        double s0 = 0.0, s1 = 0.0, h0 = 0.0, h1 = 0.0;
        for (int j = 0; j < 64; j += 2) {
          const double v0 = scr[lane * 65 + j], v1 = scr[lane * 65 + j + 1];
          if (j < split) s0 += v0; else h0 += v0;
          if (j + 1 < split) s1 += v1; else h1 += v1;
        }
        dst[k0 + lane] = s0 + s1;
        dst_hi[k0 + lane] = h0 + h1;
      }
    }
    wave_sync();
  }
}

// centred affine degree-6 sums (Moments<3,3>::accum_aff of y = lam - 1/4,
// NS sets of 84 in red) -> homogeneous moments in place: lane computes the
// columns lane and lane + 64 of the dense FP32-exact map xg_aff33
template <int NS>
__device__ __forceinline__ void aff33_to_hom(double *red, int lane) {
  constexpr int NM = Moments<3, 3>::NM;
  double lo[NS], hi[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) lo[s] = hi[s] = 0.0;
#pragma unroll 4
  for (int r = 0; r < NM; ++r) {
    const float2 c = *reinterpret_cast<const float2 *>(&xg_aff33[r][lane][0]);
    const double c0 = c.x, c1 = c.y;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      const double n = red[s * NM + r];
      lo[s] = fma(c0, n, lo[s]);
      hi[s] = fma(c1, n, hi[s]);
    }
  }
  wave_sync();
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    red[s * NM + lane] = lo[s];
    if (lane + 64 < NM) red[s * NM + lane + 64] = hi[s];
  }
  wave_sync();
}

// stats (nullable, int32 x 4 per element): table rows added, volume tets and
// interface triangles integrated point by point, cut leaves
template <int LO, int PM, int SIDE, int WHICH>
__global__ __launch_bounds__(64, 2) void k_subdiv_exact(int nel, const int32_t *__restrict__ nel_dev,
                                                     const int32_t *__restrict__ sub,
                                                     const int32_t *__restrict__ elems,
                                                     const double *__restrict__ lset_nodal,
                                                     const int32_t *__restrict__ el2v,
                                                     const double *__restrict__ coords,
                                                     const double *__restrict__ finv_all, int L,
                                                     const uint4 *__restrict__ nverts,
                                                     const double *__restrict__ nmom, double *__restrict__ out_vol,
                                                     double *__restrict__ out_if, int out_stride,
                                                     uint8_t *__restrict__ dom_out, int32_t *__restrict__ stats,
                                                     int stats_stride, int32_t *err) {
  constexpr int NL = ndof(3, LO);
  constexpr int NM = Moments<3, PM>::NM;
  // minimal exact rules of degree 2 PM: tets 4 / 14 / 24 points, triangles
  // 3 / 6 / 12, processed in blocks of SDX_VB / SDX_IB points per lane task
  constexpr int SDX_VB = PM == 1 ? 4 : (PM == 2 ? 7 : 8), SDX_NVB = PM == 1 ? 1 : (PM == 2 ? 2 : 3);
  constexpr int SDX_IB = PM == 1 ? 3 : (PM == 2 ? 6 : 4), SDX_NIB = PM == 1 ? 1 : (PM == 2 ? 1 : 3);
  // (nel_dev: the element count from device memory, at most nel; a block
  // with no element leaves before staging the rules)
  const int n_el = nel_dev ? min(*nel_dev, nel) : nel;
  if ((int)blockIdx.x >= n_el) return;
  extern __shared__ char smem[];
  SdxLds S;
  const int vol_nq = xg_minrule_cnt_d[3][PM];
  const int if_nq = xg_minrule_cnt_d[2][PM];
  sdx_bytes(L, vol_nq, if_nq, NM, &S, smem);
  const int lane = threadIdx.x & 63;
  const int n = 1 << L;
  const int nlat = sd_nlat(L);
  const int loff = xg_lat_off_d[L];
  const int shift = SDX_DMAX - L;
  {
    const double(*src_v)[4] = &xg_minrule_data_d[xg_minrule_off_d[3][PM]];
    const double(*src_i)[4] = &xg_minrule_data_d[xg_minrule_off_d[2][PM]];
    for (int k = lane; k < 4 * (vol_nq + if_nq); k += 64) {
      const int q = k >> 2, c = k & 3;
      if (q < vol_nq) S.rule_vol[q][c] = src_v[q][c];
      else S.rule_if[q - vol_nq][c] = src_i[q - vol_nq][c];
    }
    for (int k = lane; k < SDX_LATQ; k += 64) S.latq[k] = xg_lat_d[k];
    wave_sync();
  }
  const double inv_n = 1.0 / n;
  for (int j = blockIdx.x; j < n_el; j += gridDim.x) {
    const int i = sub ? sub[j] : j;
    const int e = elems ? elems[i] : i;
    // the element's nodal values and geometry live in LDS (the lanes'
    // registers hold the moment sums)
    if (lane < NL) S.cl[lane] = lset_nodal[(size_t)e * NL + lane];
    if (lane < 128) S.tab[lane] = 0.0;
    if (lane + 64 < 128) S.tab[lane + 64] = 0.0;
    bool have_geo = false;
    if (WHICH & 2) {
      if (el2v && coords) {
        have_geo = true;
        if (lane == 0) {
          double V[4][3], F[9], Fi[9], ad;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int v = el2v[(size_t)e * 4 + k];
#pragma unroll
            for (int d = 0; d < 3; ++d) V[k][d] = coords[(size_t)v * 3 + d];
          }
          jacobian<3>(V, F, Fi, ad);
#pragma unroll
          for (int k = 0; k < 9; ++k) S.fi[k] = Fi[k];
          S.fi[9] = ad;
        }
      } else if (finv_all) {
        have_geo = true;
        if (lane == 0) {
          double Fi[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) Fi[k] = finv_all[(size_t)i * 9 + k];
          const double dinv = Fi[0] * (Fi[4] * Fi[8] - Fi[5] * Fi[7]) - Fi[1] * (Fi[3] * Fi[8] - Fi[5] * Fi[6]) +
                              Fi[2] * (Fi[3] * Fi[7] - Fi[4] * Fi[6]);
#pragma unroll
          for (int k = 0; k < 9; ++k) S.fi[k] = Fi[k];
          S.fi[9] = 1.0 / fabs(dinv);
        }
      }
    }
    wave_sync();
    const double *cl = S.cl;
    // 1. signs of the level set on the finest lattice (one ballot per 64 points)
    for (int q0 = 0; q0 < nlat; q0 += 64) {
      const int q = q0 + lane;
      bool nonneg = false;
      if (q < nlat) {
        const unsigned pk = xg_lat_d[loff + q];
        const double x[3] = {(double)(pk & 255u) * inv_n, (double)((pk >> 8) & 255u) * inv_n,
                             (double)(pk >> 16) * inv_n};
        nonneg = lagrange_eval3<LO>(cl, x) >= 0.0;
      }
      const unsigned long long b = __ballot(nonneg);
      if (lane == 0) {
        S.latbits[(q0 >> 5)] = (uint32_t)b;
        S.latbits[(q0 >> 5) + 1] = (uint32_t)(b >> 32);
      }
    }
    double mvol[(WHICH & 1) ? NM : 1], mif[(WHICH & 2) ? NM : 1];
#pragma unroll
    for (int k = 0; k < ((WHICH & 1) ? NM : 1); ++k) mvol[k] = 0.0;
#pragma unroll
    for (int k = 0; k < ((WHICH & 2) ? NM : 1); ++k) mif[k] = 0.0;
    int st_rows = 0, st_vol = 0, st_if = 0, st_leaf = 0;
    bool ok = true;
    int dom = -1;
    int ncur = 1;
    wave_sync();
    for (int k = 0; k <= L && ncur > 0; ++k) {
      const int r = L - k;
      const int ncand = (k == 0) ? 1 : 8 * ncur;
      const int np = xg_lat_cnt_d[r], poff = xg_lat_off_d[r];
      const int koff = sdx_lvl_off(k);
      int nnxt = 0;
      for (int cbase = 0; cbase < ncand; cbase += SDX_CH) {
        const int nc = min(SDX_CH, ncand - cbase);
        for (int c = lane; c < nc; c += 64) {
          const int gc = cbase + c;
          const int path = (k == 0) ? 0 : (int)S.cur[gc >> 3] * 8 + (gc & 7);
          S.nodeV[c] = nverts[koff + path];
          S.flags[c] = 0u;
        }
        wave_sync();
        // lattice checks above the leaves: sign bits of every (candidate,
        // lattice point) pair (a leaf's lattice is its 4 vertices, read below)
        if (r > 0) {
          for (int q = lane; q < nc * np; q += 64) {
            const int c = q / np, t = q - c * np;
            const unsigned pk = (r <= 3) ? S.latq[poff + t] : xg_lat_d[poff + t];
            const int I0 = (int)(pk & 255u), I1 = (int)((pk >> 8) & 255u), I2 = (int)(pk >> 16);
            int W[4][3];
            unpack_uint4(S.nodeV[c], shift, W);
            int P[3];
#pragma unroll
            for (int d = 0; d < 3; ++d)
              P[d] = W[0][d] + ((I0 * (W[1][d] - W[0][d]) + I1 * (W[2][d] - W[0][d]) + I2 * (W[3][d] - W[0][d])) >> r);
            atomicOr(&S.flags[c], lat_bit(S.latbits, lat_index(P[0], P[1], P[2], n)) ? 1u : 2u);
          }
          wave_sync();
        }
        // classification and compaction
        int nfill = 0, nleaf = 0;
        for (int c0 = 0; c0 < nc; c0 += 64) {
          const int c = c0 + lane;
          const bool valid = c < nc;
          int W[4][3] = {};
          if (valid) unpack_uint4(S.nodeV[c], shift, W);
          uint32_t f = 0u;
          if (valid) {
            if (r > 0) f = S.flags[c];
            else
#pragma unroll
              for (int j = 0; j < 4; ++j) f |= lat_bit(S.latbits, lat_index(W[j][0], W[j][1], W[j][2], n)) ? 1u : 2u;
          }
          const int cls = (f == 3u) ? IF : ((f == 1u) ? POS : NEG);
          if (k == 0) dom = __shfl(cls, 0, 64);
          int lc = -2;
          if (valid && cls == IF && r == 0) {
            double vv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const double x[3] = {(double)W[j][0] * inv_n, (double)W[j][1] * inv_n, (double)W[j][2] * inv_n};
              vv[j] = lagrange_eval3<LO>(cl, x);
            }
            lc = leaf_check(vv);
            if (lc < 0) ok = false;
          }
          const int path = (k == 0 || !valid) ? 0 : (int)S.cur[(cbase + c) >> 3] * 8 + ((cbase + c) & 7);
          const bool wantfill = (WHICH & 1) && valid && (cls == SIDE || lc == SIDE);
          const bool leafcut = valid && lc == IF;
          const bool wantnext = valid && cls == IF && r > 0;
          int cnt;
          const int sf = wave_slot(wantfill, nfill, lane, cnt);
          if (wantfill) S.fillq[sf] = (uint16_t)(koff + path);
          nfill += cnt;
          const int sl = wave_slot(leafcut, nleaf, lane, cnt);
          if (leafcut) S.leafq[sl] = (uint16_t)c;
          nleaf += cnt;
          const int sn = wave_slot(wantnext, nnxt, lane, cnt);
          if (wantnext) S.nxt[sn] = (uint16_t)path;
          nnxt += cnt;
        }
        wave_sync();
        st_leaf += nleaf;
        // cut leaves, 64 at a time: pieces one lane per leaf (its 4 vertices
        // and up to 4 cut points into the LDS pool), then the pieces' base-
        // rule points lane-parallel over (piece, block of points) tasks
        for (int l0 = 0; l0 < nleaf; l0 += 64) {
          const int l = l0 + lane;
          LeafPieces2 PC;
          PC.nvol = 0;
          PC.nif = 0;
          PC.comp = false;
          if (l < nleaf) {
            const int c = S.leafq[l];
            int W[4][3];
            unpack_uint4(S.nodeV[c], shift, W);
            double V[4][3], vv[4], cpt[4][3];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
#pragma unroll
              for (int d = 0; d < 3; ++d) V[j][d] = W[j][d] * inv_n;
              vv[j] = lagrange_eval3<LO>(cl, V[j]);
            }
            if (!leaf_pieces<SIDE>(V, vv, cpt, PC)) {
              ok = false;
              PC.nvol = 0;
              PC.nif = 0;
              PC.comp = false;
            }
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                S.pool[lane][j][d] = V[j][d];
                S.pool[lane][4 + j][d] = cpt[j][d];
              }
          }
          if (!(WHICH & 1)) {
            PC.nvol = 0;
            PC.comp = false;
          }
          if (!(WHICH & 2)) PC.nif = 0;
          // complement: the whole leaf's table row joins this chunk's rows
          {
            int cnt;
            const int sf = wave_slot(PC.comp, nfill, lane, cnt);
            if (PC.comp) {
              const int gc = cbase + S.leafq[l];
              S.fillq[sf] = (uint16_t)(koff + ((k == 0) ? 0 : (int)S.cur[gc >> 3] * 8 + (gc & 7)));
            }
            nfill += cnt;
          }
          int tot_vol, tot_if;
          const int bvol = wave_excl_scan(PC.nvol, lane, tot_vol);
          const int bif = wave_excl_scan(PC.nif, lane, tot_if);
#pragma unroll
          for (int q = 0; q < 3; ++q)
            if (q < PC.nvol) S.volrec[bvol + q] = (uint32_t)lane | (PC.vol[q] << 8) | (PC.comp ? 0x80u : 0u);
#pragma unroll
          for (int q = 0; q < 2; ++q)
            if (q < PC.nif) S.ifrec[bif + q] = (uint32_t)lane | (PC.itf[q] << 8);
          st_vol += tot_vol;
          st_if += tot_if;
          wave_sync();
          // weight factors: FillSimplexWithRule (|det|/6)*6, negative for a
          // subtracted corner; FillSimplexCoDim1WithRule |cross| and the
          // interface transform
          for (int q = lane; q < tot_vol; q += 64) {
            const uint32_t rec = S.volrec[q];
            const double(*pl)[3] = S.pool[rec & 63];
            const double tf = tet_tf(pl[(rec >> 8) & 7], pl[(rec >> 11) & 7], pl[(rec >> 14) & 7],
                                     pl[(rec >> 17) & 7], true);
            S.voltf[q] = (rec & 0x80u) ? -tf : tf;
          }
          for (int q = lane; q < tot_if; q += 64) {
            const uint32_t rec = S.ifrec[q];
            const double(*pl)[3] = S.pool[rec & 63];
            const double *s0 = pl[(rec >> 8) & 7], *s1 = pl[(rec >> 11) & 7], *s2 = pl[(rec >> 14) & 7];
            const double *pp = pl[(rec >> 17) & 7];
            double a[3], b[3], cc[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              a[d] = s1[d] - s0[d];
              b[d] = s2[d] - s0[d];
            }
            cc[0] = a[1] * b[2] - a[2] * b[1];
            cc[1] = a[2] * b[0] - a[0] * b[2];
            cc[2] = a[0] * b[1] - a[1] * b[0];
            const double tf = norm3(cc);
            const double na = norm3(a), nb = norm3(b);
            const double rel = na > nb ? na : nb;
            double wf = 0.0;
            if (!(tf < 1e-14 * rel)) {
#pragma unroll
              for (int d = 0; d < 3; ++d) cc[d] /= tf;
              const double dd0 = pp[0] - s0[0], dd1 = pp[1] - s0[1], dd2 = pp[2] - s0[2];
              if (dd0 * cc[0] + dd1 * cc[1] + dd2 * cc[2] >= 0)
#pragma unroll
                for (int d = 0; d < 3; ++d) cc[d] *= -1.0;
              Geo geo;
              geo.Finv = have_geo ? S.fi : nullptr;
              geo.absdet = have_geo ? S.fi[9] : 1.0;
              wf = tf * geo.if_factor(cc);
            }
            S.ifw[q] = wf;
          }
          wave_sync();
          // tasks (piece, block of SDX_VB points): the piece's vertices are
          // read once per block
          if (WHICH & 1) {
            for (int q = lane; q < tot_vol * SDX_NVB; q += 64) {
              const int pc = q / SDX_NVB, t0 = (q - pc * SDX_NVB) * SDX_VB;
              const uint32_t rec = S.volrec[pc];
              const double(*pl)[3] = S.pool[rec & 63];
              double p0[3], p1[3], p2[3], p3[3];
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                p0[d] = pl[(rec >> 8) & 7][d];
                p1[d] = pl[(rec >> 11) & 7][d];
                p2[d] = pl[(rec >> 14) & 7][d];
                p3[d] = pl[(rec >> 17) & 7][d];
              }
              const double tf = S.voltf[pc];
              for (int t = t0; t < t0 + SDX_VB; ++t) accum_tet<PM>(mvol, p0, p1, p2, p3, tf, S.rule_vol, t);
            }
          }
          if (WHICH & 2) {
            for (int q = lane; q < tot_if * SDX_NIB; q += 64) {
              const int pc = q / SDX_NIB, t0 = (q - pc * SDX_NIB) * SDX_IB;
              const uint32_t rec = S.ifrec[pc];
              const double(*pl)[3] = S.pool[rec & 63];
              double s0[3], s1[3], s2[3];
#pragma unroll
              for (int d = 0; d < 3; ++d) {
                s0[d] = pl[(rec >> 8) & 7][d];
                s1[d] = pl[(rec >> 11) & 7][d];
                s2[d] = pl[(rec >> 14) & 7][d];
              }
              const double wf = S.ifw[pc];
              for (int t = t0; t < t0 + SDX_IB; ++t) {
                const double *rr = S.rule_if[t];
                const double ow = 1.0 - rr[0] - rr[1];
                double lam[4];
#pragma unroll
                for (int d = 0; d < 3; ++d) lam[d] = ow * s0[d] + rr[0] * s1[d] + rr[1] * s2[d];
                lam[3] = 1.0 - lam[0] - lam[1] - lam[2];
                Moments<3, PM>::accum(lam, rr[3] * wf, mif);
              }
            }
          }
          wave_sync();
        }
        // table rows of the chunk: uncut nodes of the kept side and the
        // complemented leaves, summed lane-parallel over the moments (lane
        // holds moments lane and lane + 64), four rows per round trip
        if constexpr ((WHICH & 1) != 0) {
          st_rows += nfill;
          int f = 0;
          for (; f + 4 <= nfill; f += 4) {
            const int r0 = S.fillq[f], r1 = S.fillq[f + 1], r2 = S.fillq[f + 2], r3 = S.fillq[f + 3];
            double a0 = 0, a1 = 0, a2 = 0, a3 = 0, b0 = 0, b1 = 0, b2 = 0, b3 = 0;
            if (lane < NM) {
              a0 = nmom[(size_t)r0 * NM + lane];
              a1 = nmom[(size_t)r1 * NM + lane];
              a2 = nmom[(size_t)r2 * NM + lane];
              a3 = nmom[(size_t)r3 * NM + lane];
            }
            if (lane + 64 < NM) {
              b0 = nmom[(size_t)r0 * NM + 64 + lane];
              b1 = nmom[(size_t)r1 * NM + 64 + lane];
              b2 = nmom[(size_t)r2 * NM + 64 + lane];
              b3 = nmom[(size_t)r3 * NM + 64 + lane];
            }
            S.tab[lane] += (a0 + a1) + (a2 + a3);
            S.tab[64 + lane] += (b0 + b1) + (b2 + b3);
          }
          for (; f < nfill; ++f) {
            const size_t row = (size_t)S.fillq[f] * NM;
            if (lane < NM) S.tab[lane] += nmom[row + lane];
            if (lane + 64 < NM) S.tab[64 + lane] += nmom[row + 64 + lane];
          }
        }
        wave_sync();
      }
      uint16_t *tmp = S.cur;
      S.cur = S.nxt;
      S.nxt = tmp;
      ncur = nnxt;
    }
    // wave reduction of the per-lane moments (through LDS: the point pool is
    // free now), plus the table sums
    if constexpr ((WHICH & 1) != 0) wave_sum_lds<NM>(mvol, &S.pool[0][0][0], S.red, lane);
    if constexpr ((WHICH & 2) != 0) wave_sum_lds<NM>(mif, &S.pool[0][0][0], S.red + NM, lane);
    if (WHICH & 1) {
      if (lane < NM) S.red[lane] += S.tab[lane];
      if (lane + 64 < NM) S.red[64 + lane] += S.tab[64 + lane];
    }
    wave_sync();
    for (int k = lane; k < NM; k += 64) {
      if ((WHICH & 1) && out_vol) out_vol[(size_t)i * out_stride + k] = S.red[k];
      if ((WHICH & 2) && out_if) out_if[(size_t)i * out_stride + k] = S.red[NM + k];
    }
    const bool allok = __ballot(!ok) == 0ull;
    if (lane == 0) {
      if (dom_out) dom_out[i] = (uint8_t)dom;
      if (stats) {
        stats[(size_t)i * stats_stride + 0] = st_rows;
        stats[(size_t)i * stats_stride + 1] = st_vol;
        stats[(size_t)i * stats_stride + 2] = st_if;
        stats[(size_t)i * stats_stride + 3] = st_leaf;
      }
      if (!allok) atomicOr(err, 2);
    }
    wave_sync();
  }
}

template <int LO, int PM, int SIDE, int WHICH>
int launch_subdiv_exact(xg_ctx *ctx, int nel, const int32_t *elems, const double *lset_nodal, const int32_t *el2v,
                        const double *coords, const double *finv, int L, double *out_vol, double *out_if,
                        int out_stride, uint8_t *dom, int32_t *stats, int stats_stride,
                        const int32_t *sub = nullptr, const int32_t *nel_dev = nullptr) {
  constexpr int NM = Moments<3, PM>::NM;
  const size_t lds = sdx_bytes(L, xg_minrule_cnt[3][PM], xg_minrule_cnt[2][PM], NM, nullptr, nullptr);
  if (lds > 65536)
    XG_CHECK(hipFuncSetAttribute((const void *)k_subdiv_exact<LO, PM, SIDE, WHICH>,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  // (a device-side count: a grid of 2048 waves strides over it)
  const int cap_blocks = nel_dev ? 2048 : (1 << 22);
  const int blocks = nel < cap_blocks ? nel : cap_blocks;
  KernelTimer tm(ctx, "subdiv_exact");
  k_subdiv_exact<LO, PM, SIDE, WHICH><<<blocks, 64, lds, ctx->stream>>>(
      nel, nel_dev, sub, elems, lset_nodal, el2v, coords, finv, L, (const uint4 *)ctx->sdx_verts, ctx->sdx_mom[PM], out_vol,
      out_if, out_stride, dom, stats, stats_stride, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

// node tables of the exact-rule kernel (once per context and degree)
int ensure_sdx_tables(xg_ctx *ctx, int PM) {
  if (!ctx->sdx_verts) {
    XG_CHECK(hipMalloc(&ctx->sdx_verts, sizeof(uint4) * SDX_NODES));
    k_sdx_node_verts<<<(SDX_NODES + 255) / 256, 256, 0, ctx->stream>>>((uint4 *)ctx->sdx_verts);
    XG_LAUNCH_CHECK();
  }
  if (!ctx->sdx_mom[PM]) {
    const int NM = PM == 1 ? Moments<3, 1>::NM : (PM == 2 ? Moments<3, 2>::NM : Moments<3, 3>::NM);
    XG_CHECK(hipMalloc(&ctx->sdx_mom[PM], sizeof(double) * NM * (size_t)SDX_NODES));
    const uint4 *v = (const uint4 *)ctx->sdx_verts;
    if (PM == 1) k_sdx_node_moments<1><<<SDX_NODES, 64, 0, ctx->stream>>>(v, ctx->sdx_mom[1]);
    if (PM == 2) k_sdx_node_moments<2><<<SDX_NODES, 64, 0, ctx->stream>>>(v, ctx->sdx_mom[2]);
    if (PM == 3) k_sdx_node_moments<3><<<SDX_NODES, 64, 0, ctx->stream>>>(v, ctx->sdx_mom[3]);
    XG_LAUNCH_CHECK();
  }
  return XG_OK;
}

// one pass of the exact-rule kernel: volume side `vol_side` (NEG / POS, or
// -1 = none) and / or the interface
template <int LO, int PM>
int subdiv_exact_pass(xg_ctx *ctx, int nel, const int32_t *elems, const double *lset_nodal, const int32_t *el2v,
                      const double *coords, const double *finv, int L, int vol_side, bool with_if, double *out_vol,
                      double *out_if, int out_stride, uint8_t *dom, int32_t *stats, int stats_stride,
                      const int32_t *sub = nullptr, const int32_t *nel_dev = nullptr) {
  if (vol_side >= 0) {
    const int rc = ensure_sdx_tables(ctx, PM);
    if (rc != XG_OK) return rc;
  }
#define XG_SDX(SIDE, WHICH)                                                                                        \
  return launch_subdiv_exact<LO, PM, SIDE, WHICH>(ctx, nel, elems, lset_nodal, el2v, coords, finv, L, out_vol,    \
                                                  out_if, out_stride, dom, stats, stats_stride, sub, nel_dev)
  if (vol_side == NEG && !with_if) XG_SDX(NEG, 1);
  if (vol_side == POS && !with_if) XG_SDX(POS, 1);
  if constexpr (PM < 3) {
    if (vol_side == NEG && with_if) XG_SDX(NEG, 3);
    if (vol_side == POS && with_if) XG_SDX(POS, 3);
  }
  if (vol_side < 0 && with_if) XG_SDX(NEG, 2);
#undef XG_SDX
  return set_error(XG_ERR_ARG, "subdivision rule: bad pass");
}

// ---------------------------------------------------------------- staged path
// The one-kernel path above keeps the element's moment sums (2 x 84 FP64 for
// P3) in registers for the whole refinement walk, so the walk's latency-bound
// integer and LDS work runs at two waves per SIMD.  The staged path splits
// it into three kernels over a per-element list of uint16 node ids:
//   1. k_sdx_tree: the walk (lattice sign checks, classification) at high
//      occupancy; writes the cut leaves from the front of the element's list
//      and the uncut nodes of the kept side (table rows) from the back
//   2. k_sdx_leaves: the cut leaves' pieces and their base-rule points
//      (the register-heavy part), once per requested part (volume side or
//      interface) on the same lists; appends the complemented leaves after
//      the cut leaves
//   3. k_sdx_rows: the table rows (uncut nodes + complemented leaves),
//      added to the moments
// An element whose lists do not fit (2 cut leaves + rows > cap) is queued
// and handled by the one-kernel path.

struct SdcLds {
  uint32_t *latbits;  // sign bits (v >= 0) of the finest lattice, lattice order
  uint16_t *rowst;    // (n+1)^2: lattice index of (0, b, c)
  uint4 *nodeV;       // SDX_CH candidate vertices (level-L units)
  uint32_t *flags;    // SDX_CH
  uint16_t *cur, *nxt;
  uint32_t *latq;     // lattice offsets of levels 0..3
  double *cl;         // 20
  uint16_t *fq;       // SDX_CH: a chunk's table rows
  uint16_t *flq;      // SDX_CH: a chunk's leaves flagged cut by their vertex signs
};

__host__ __device__ inline size_t sdc_bytes(int L, SdcLds *s, char *base) {
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o = sd_align(o + bytes);
    return at;
  };
  const int n1 = (1 << L) + 1;
  const size_t a_lb = take(sizeof(uint32_t) * 2 * ((sd_nlat(L) + 63) / 64));
  const size_t a_rs = take(sizeof(uint16_t) * n1 * n1);
  const size_t a_nv = take(sizeof(uint4) * SDX_CH);
  const size_t a_fl = take(sizeof(uint32_t) * SDX_CH);
  const size_t a_cur = take(sizeof(uint16_t) * sd_maxcut(L));
  const size_t a_nxt = take(sizeof(uint16_t) * sd_maxcut(L));
  const size_t a_lq = take(sizeof(uint32_t) * SDX_LATQ);
  const size_t a_cl = take(sizeof(double) * 20);
  const size_t a_fq = take(sizeof(uint16_t) * SDX_CH);
  const size_t a_flq = take(sizeof(uint16_t) * SDX_CH);
  if (s) {
    s->fq = (uint16_t *)(base + a_fq);
    s->flq = (uint16_t *)(base + a_flq);
    s->latbits = (uint32_t *)(base + a_lb);
    s->rowst = (uint16_t *)(base + a_rs);
    s->nodeV = (uint4 *)(base + a_nv);
    s->flags = (uint32_t *)(base + a_fl);
    s->cur = (uint16_t *)(base + a_cur);
    s->nxt = (uint16_t *)(base + a_nxt);
    s->latq = (uint32_t *)(base + a_lq);
    s->cl = (double *)(base + a_cl);
  }
  return o;
}

// the lattice index of (a, b, c) through the row table (lat_index without
// the cubic terms)
__device__ __forceinline__ int lat_index_rs(const uint16_t *rowst, int a, int b, int c, int n) {
  return (int)rowst[c * (n + 1) + b] + a;
}

// q / d for 0 <= q < 2^24, d >= 1 (float estimate, one correction step)
__device__ __forceinline__ int small_div(int q, int d, float rd, int &rem) {
  int c = (int)((float)q * rd);
  int t = q - c * d;
  if (t >= d) {
    ++c;
    t -= d;
  } else if (t < 0) {
    --c;
    t += d;
  }
  rem = t;
  return c;
}

__global__ void k_zero_i32(int32_t *p) { *p = 0; }

// next element of a persistent wave: lane 0 takes it from the work counter
__device__ __forceinline__ int next_item(int32_t *queue, int lane) {
  int nx = 0;
  if (lane == 0) nx = atomicAdd(queue, 1);
  return __shfl(nx, 0, 64);
}

// SIDE: NEG or POS (the uncut nodes of that side: their table rows are
// summed here, lane-parallel over the moments, and written to out), or -1
// (cut leaves only)
template <int LO, int PM, int SIDE>
__global__ __launch_bounds__(64) void k_sdx_tree(int nel, const int32_t *__restrict__ elems,
                                                 const double *__restrict__ lset_nodal, int L,
                                                 const uint4 *__restrict__ nverts, uint16_t *__restrict__ lists,
                                                 int cap, int4 *__restrict__ cnt, uint8_t *__restrict__ dom_out,
                                                 int32_t *__restrict__ over, int32_t *__restrict__ nover,
                                                 const double *__restrict__ nmom, double *__restrict__ out,
                                                 int out_stride, int32_t *err) {
  constexpr int NL = ndof(3, LO);
  constexpr int NM = Moments<3, PM>::NM;
  extern __shared__ char smem[];
  SdcLds S;
  sdc_bytes(L, &S, smem);
  const int lane = threadIdx.x & 63;
  const int n = 1 << L;
  const int nlat = sd_nlat(L);
  const int loff = xg_lat_off_d[L];
  const int shift = SDX_DMAX - L;
  for (int k = lane; k < SDX_LATQ; k += 64) S.latq[k] = xg_lat_d[k];
  for (int k = lane; k < (n + 1) * (n + 1); k += 64) {
    const int c = k / (n + 1), b = k - c * (n + 1);
    S.rowst[k] = (uint16_t)((b + c <= n) ? lat_index(0, b, c, n) : 0);
  }
  wave_sync();
  const double inv_n = 1.0 / n;
  for (int i = blockIdx.x; i < nel; i += gridDim.x) {
    const int e = elems ? elems[i] : i;
    if (lane < NL) S.cl[lane] = lset_nodal[(size_t)e * NL + lane];
    wave_sync();
    const double *cl = S.cl;
    // signs of the level set on the finest lattice (one ballot per 64
    // points); the root's flags (its lattice is the whole lattice) follow
    uint32_t rootf = 0u;
    for (int q0 = 0; q0 < nlat; q0 += 64) {
      const int q = q0 + lane;
      bool nonneg = false;
      if (q < nlat) {
        const unsigned pk = xg_lat_d[loff + q];
        const double x[3] = {(double)(pk & 255u) * inv_n, (double)((pk >> 8) & 255u) * inv_n,
                             (double)(pk >> 16) * inv_n};
        nonneg = lagrange_eval3<LO>(cl, x) >= 0.0;
      }
      const unsigned long long b = __ballot(nonneg);
      const unsigned long long bn = __ballot(q < nlat && !nonneg);
      rootf |= (b ? 1u : 0u) | (bn ? 2u : 0u);
      if (lane == 0) {
        S.latbits[(q0 >> 5)] = (uint32_t)b;
        S.latbits[(q0 >> 5) + 1] = (uint32_t)(b >> 32);
      }
    }
    uint16_t *lst = lists + (size_t)i * cap;
    // table-row sums of the kept side: moments lane and lane + 64 (16-byte
    // loads of a few rows per instruction measured slower here: registers)
    double ta = 0.0, tb = 0.0;
    int nl = 0, nf = 0, n22 = 0;
    bool ok = true, overflow = false;
    int dom = -1;
    int ncur = 1;
    wave_sync();
    for (int k = 0; k <= L && ncur > 0 && !overflow; ++k) {
      const int r = L - k;
      const int ncand = (k == 0) ? 1 : 8 * ncur;
      const int np = xg_lat_cnt_d[r], poff = xg_lat_off_d[r];
      const float rnp = 1.0f / (float)np;
      const int koff = sdx_lvl_off(k);
      int nnxt = 0;
      for (int cbase = 0; cbase < ncand && !overflow; cbase += SDX_CH) {
        const int nc = min(SDX_CH, ncand - cbase);
        for (int c = lane; c < nc; c += 64) {
          const int gc = cbase + c;
          const int path = (k == 0) ? 0 : (int)S.cur[gc >> 3] * 8 + (gc & 7);
          S.nodeV[c] = nverts[koff + path];
          S.flags[c] = 0u;
        }
        wave_sync();
        // lattice checks above the leaves.  Up to 35 lattice points per
        // candidate (r <= 2): one lane per candidate walks them.  More: the
        // sign bits of every (candidate, lattice point) pair, OR-ed per
        // candidate with ballots (the pairs of one pass cover consecutive
        // candidates)
        if (k == 0 && r > 0) {
          if (lane == 0) S.flags[0] = rootf;
          wave_sync();
        } else if (r > 0 && r <= 2) {
          for (int c = lane; c < nc; c += 64) {
            int W[4][3];
            unpack_uint4(S.nodeV[c], shift, W);
            int E[3][3];
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
              for (int d = 0; d < 3; ++d) E[m][d] = W[m + 1][d] - W[0][d];
            uint32_t f = 0u;
            for (int t = 0; t < np; ++t) {
              const unsigned pk = S.latq[poff + t];
              const int I0 = (int)(pk & 255u), I1 = (int)((pk >> 8) & 255u), I2 = (int)(pk >> 16);
              int P[3];
#pragma unroll
              for (int d = 0; d < 3; ++d) P[d] = W[0][d] + ((I0 * E[0][d] + I1 * E[1][d] + I2 * E[2][d]) >> r);
              f |= lat_bit(S.latbits, lat_index_rs(S.rowst, P[0], P[1], P[2], n)) ? 1u : 2u;
            }
            S.flags[c] = f;
          }
          wave_sync();
        } else if (r > 0 && nc <= 8) {
          // a few candidates with many lattice points: 8 lanes per candidate
          const int cand = lane >> 3, sub = lane & 7;
          uint32_t f = 0u;
          if (cand < nc) {
            int W[4][3];
            unpack_uint4(S.nodeV[cand], shift, W);
            int E[3][3];
#pragma unroll
            for (int m = 0; m < 3; ++m)
#pragma unroll
              for (int d = 0; d < 3; ++d) E[m][d] = W[m + 1][d] - W[0][d];
            for (int t = sub; t < np; t += 8) {
              const unsigned pk = (r <= 3) ? S.latq[poff + t] : xg_lat_d[poff + t];
              const int I0 = (int)(pk & 255u), I1 = (int)((pk >> 8) & 255u), I2 = (int)(pk >> 16);
              int P[3];
#pragma unroll
              for (int d = 0; d < 3; ++d) P[d] = W[0][d] + ((I0 * E[0][d] + I1 * E[1][d] + I2 * E[2][d]) >> r);
              f |= lat_bit(S.latbits, lat_index_rs(S.rowst, P[0], P[1], P[2], n)) ? 1u : 2u;
            }
          }
          for (int cc = 0; cc < nc; ++cc) {
            const unsigned long long b1 = __ballot(cand == cc && (f & 1u));
            const unsigned long long b2 = __ballot(cand == cc && (f & 2u));
            if (lane == 0) S.flags[cc] = (b1 ? 1u : 0u) | (b2 ? 2u : 0u);
          }
          wave_sync();
        } else if (r > 0) {
          const int npairs = nc * np;
          for (int q0 = 0; q0 < npairs; q0 += 64) {
            const int q = q0 + lane;
            int c = -1;
            uint32_t bit = 0u;
            if (q < npairs) {
              int t;
              c = small_div(q, np, rnp, t);
              const unsigned pk = (r <= 3) ? S.latq[poff + t] : xg_lat_d[poff + t];
              const int I0 = (int)(pk & 255u), I1 = (int)((pk >> 8) & 255u), I2 = (int)(pk >> 16);
              int W[4][3];
              unpack_uint4(S.nodeV[c], shift, W);
              int P[3];
#pragma unroll
              for (int d = 0; d < 3; ++d)
                P[d] = W[0][d] + ((I0 * (W[1][d] - W[0][d]) + I1 * (W[2][d] - W[0][d]) + I2 * (W[3][d] - W[0][d])) >> r);
              bit = lat_bit(S.latbits, lat_index_rs(S.rowst, P[0], P[1], P[2], n)) ? 1u : 2u;
            }
            int trem;
            const int c_lo = small_div(q0, np, rnp, trem);
            const int qhi = min(q0 + 63, npairs - 1);
            const int c_hi = small_div(qhi, np, rnp, trem);
            for (int cc = c_lo; cc <= c_hi; ++cc) {
              const unsigned long long b1 = __ballot(c == cc && bit == 1u);
              const unsigned long long b2 = __ballot(c == cc && bit == 2u);
              if (lane == 0) S.flags[cc] |= (b1 ? 1u : 0u) | (b2 ? 2u : 0u);
            }
          }
          wave_sync();
        }
        // classification and the lists
        int nfc = 0;    // this chunk's table rows (S.fq)
        int nflag = 0;  // this chunk's leaves to check on their values (S.flq)
        for (int c0 = 0; c0 < nc; c0 += 64) {
          const int c = c0 + lane;
          const bool valid = c < nc;
          int W[4][3] = {};
          if (valid) unpack_uint4(S.nodeV[c], shift, W);
          uint32_t f = 0u;
          if (valid) {
            if (r > 0) f = S.flags[c];
            else
#pragma unroll
              for (int j = 0; j < 4; ++j)
                f |= lat_bit(S.latbits, lat_index_rs(S.rowst, W[j][0], W[j][1], W[j][2], n)) ? 1u : 2u;
          }
          const int cls = (f == 3u) ? IF : ((f == 1u) ? POS : NEG);
          if (k == 0) dom = __shfl(cls, 0, 64);
          // leaves whose vertex signs differ are checked on their values in a
          // second, dense pass (flq)
          const int path = (k == 0 || !valid) ? 0 : (int)S.cur[(cbase + c) >> 3] * 8 + ((cbase + c) & 7);
          const bool wantfill = SIDE >= 0 && valid && cls == SIDE;
          const bool flagged = valid && cls == IF && r == 0;
          const bool wantnext = valid && cls == IF && r > 0;
          int cf, cg, cn;
          const int sf = wave_slot(wantfill, nfc, lane, cf);
          const int sg = wave_slot(flagged, nflag, lane, cg);
          const int sn = wave_slot(wantnext, nnxt, lane, cn);
          if (wantfill) S.fq[sf] = (uint16_t)(koff + path);
          if (flagged) S.flq[sg] = (uint16_t)c;
          if (wantnext) S.nxt[sn] = (uint16_t)path;
          nfc += cf;
          nflag += cg;
          nnxt += cn;
        }
        wave_sync();
        // the flagged leaves: Simplex::CheckIfCut on the vertex values
        for (int q0 = 0; q0 < nflag && !overflow; q0 += 64) {
          const int q = q0 + lane;
          const bool valid = q < nflag;
          int lc = -2;
          bool lc22 = false;
          int path = 0;
          if (valid) {
            const int c = S.flq[q];
            path = (k == 0) ? 0 : (int)S.cur[(cbase + c) >> 3] * 8 + ((cbase + c) & 7);
            int W[4][3];
            unpack_uint4(S.nodeV[c], shift, W);
            double vv[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              const double x[3] = {(double)W[j][0] * inv_n, (double)W[j][1] * inv_n, (double)W[j][2] * inv_n};
              vv[j] = lagrange_eval3<LO>(cl, x);
            }
            lc = leaf_check(vv);
            if (lc < 0) ok = false;
            // two vertices on each side (zero counts POS): a 2-2 cut leaf
            int npos = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) npos += vv[j] >= 0.0 ? 1 : 0;
            if (lc == IF && npos == 2) lc22 = true;
          }
          const bool wantfill = SIDE >= 0 && valid && lc == SIDE;
          const bool leafcut = valid && lc == IF;
          int cf, cl2;
          const int sf = wave_slot(wantfill, nfc, lane, cf);
          const int sl = wave_slot(leafcut, nl, lane, cl2);
          // room for the cut leaves and their complements
          if (2 * (nl + cl2) > cap) {
            overflow = true;
            break;
          }
          if (wantfill) S.fq[sf] = (uint16_t)(koff + path);
          if (leafcut) lst[sl] = (uint16_t)(koff + path);
          nfc += cf;
          nl += cl2;
          n22 += __popcll(__ballot(lc22));
        }
        wave_sync();
        // the chunk's table rows, four per round trip
        if constexpr (SIDE >= 0) {
          if (!overflow) {
            int f = 0;
            for (; f + 4 <= nfc; f += 4) {
              const size_t r0 = (size_t)S.fq[f] * NM, r1 = (size_t)S.fq[f + 1] * NM;
              const size_t r2 = (size_t)S.fq[f + 2] * NM, r3 = (size_t)S.fq[f + 3] * NM;
              if (lane < NM) ta += (nmom[r0 + lane] + nmom[r1 + lane]) + (nmom[r2 + lane] + nmom[r3 + lane]);
              if (lane + 64 < NM)
                tb += (nmom[r0 + 64 + lane] + nmom[r1 + 64 + lane]) + (nmom[r2 + 64 + lane] + nmom[r3 + 64 + lane]);
            }
            for (; f < nfc; ++f) {
              const size_t r0 = (size_t)S.fq[f] * NM;
              if (lane < NM) ta += nmom[r0 + lane];
              if (lane + 64 < NM) tb += nmom[r0 + 64 + lane];
            }
          }
          wave_sync();
        }
        nf += nfc;
      }
      uint16_t *tmp = S.cur;
      S.cur = S.nxt;
      S.nxt = tmp;
      ncur = nnxt;
    }
    const bool allok = __ballot(!ok) == 0ull;
    if (lane == 0) {
      if (dom_out) dom_out[i] = (uint8_t)dom;
      if (overflow) {
        cnt[i] = make_int4(-1, 0, 0, 0);
        over[atomicAdd(nover, 1)] = i;
      } else {
        // (.y: the 2-2 cut leaves, .z: complemented leaves (the leaf
        // kernel), .w: the table rows summed here)
        cnt[i] = make_int4(nl, n22, 0, nf);
        if (!allok) atomicOr(err, 2);
      }
    }
    if constexpr (SIDE >= 0) {
      if (!overflow) {
        if (lane < NM) out[(size_t)i * out_stride + lane] = ta;
        if (lane + 64 < NM) out[(size_t)i * out_stride + 64 + lane] = tb;
      }
    }
    wave_sync();
  }
}

// the cut leaves of each listed element: pieces and base-rule points (as in
// k_subdiv_exact), moments written (not added); the complemented leaves of a
// volume pass are appended after the cut leaves (cnt.z)
template <int LO, int PM, int SIDE, int WHICH>
__global__ __launch_bounds__(64, 2) void k_sdx_leaves(int nel, const int32_t *__restrict__ elems,
                                                     const double *__restrict__ lset_nodal,
                                                     const int32_t *__restrict__ el2v,
                                                     const double *__restrict__ coords,
                                                     const double *__restrict__ finv_all, int L,
                                                     const uint4 *__restrict__ nverts,
                                                     uint16_t *__restrict__ lists, int cap,
                                                     int4 *__restrict__ cnt, double *__restrict__ out_vol,
                                                     double *__restrict__ out_if, int out_stride,
                                                     int32_t *__restrict__ stats, int stats_stride, int32_t *err,
                                                     int32_t *__restrict__ queue, int32_t *__restrict__ stats_if) {
  constexpr int NL = ndof(3, LO);
  constexpr int NM = Moments<3, PM>::NM;
  // WHICH == 3 (both parts in one launch, p = 3): lanes [0, nn) integrate the
  // volume pieces and lanes [nn, 64) the interface pieces, each lane into
  // one set of 84 sums, with one task loop (blocks of 4 points, a triangle
  // being a tet whose fourth rule coordinate is 0); nn follows the element's
  // expected point split
  static_assert(WHICH != 3 || PM == 3, "both parts in one launch: p = 3");
  // p = 3: the task loops accumulate centred affine moments (accum_aff of
  // y = lam - 1/4, 135 instead of 153 operations per point); the pool holds
  // the leaf points shifted by -1/4 (weights and normals use differences)
  // and the wave's sums are mapped to the homogeneous moments at the end
  constexpr bool AFF = PM == 3;
  constexpr double SHIFT = AFF ? 0.25 : 0.0;
  constexpr int SDX_VB = PM == 1 ? 4 : (PM == 2 ? 7 : 8), SDX_NVB = PM == 1 ? 1 : (PM == 2 ? 2 : 3);
  constexpr int SDX_IB = PM == 1 ? 3 : (PM == 2 ? 6 : 4), SDX_NIB = PM == 1 ? 1 : (PM == 2 ? 1 : 3);
  extern __shared__ char smem[];
  SdxLds S;
  const int vol_nq = xg_minrule_cnt_d[3][PM];
  const int if_nq = xg_minrule_cnt_d[2][PM];
  sdx_leaf_bytes(vol_nq, if_nq, NM, &S, smem);
  // leaf points lane-minor: (point j, coordinate d, leaf lane l)
  double *pool = &S.pool[0][0][0];
  const int lane = threadIdx.x & 63;
  const int n = 1 << L;
  const int shift = SDX_DMAX - L;
  {
    const double(*src_v)[4] = &xg_minrule_data_d[xg_minrule_off_d[3][PM]];
    const double(*src_i)[4] = &xg_minrule_data_d[xg_minrule_off_d[2][PM]];
    for (int k = lane; k < 4 * (vol_nq + if_nq); k += 64) {
      const int q = k >> 2, c = k & 3;
      if (q < vol_nq) S.rule_vol[q][c] = src_v[q][c];
      else S.rule_if[q - vol_nq][c] = src_i[q - vol_nq][c];
    }
    wave_sync();
  }
  const double inv_n = 1.0 / n;
  // elements from the work counter (persistent waves, the heavy elements do
  // not wait behind a fixed stride) or strided by the grid
  for (int i = queue ? next_item(queue, lane) : blockIdx.x; i < nel;
       i = queue ? next_item(queue, lane) : i + gridDim.x) {
    const int4 c4 = cnt[i];
    if (c4.x < 0) continue;  // queued for the one-kernel path
    const int e = elems ? elems[i] : i;
    if (lane < NL) S.cl[lane] = lset_nodal[(size_t)e * NL + lane];
    bool have_geo = false;
    if (WHICH & 2) {
      if (el2v && coords) {
        have_geo = true;
        if (lane == 0) {
          double V[4][3], F[9], Fi[9], ad;
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int v = el2v[(size_t)e * 4 + k];
#pragma unroll
            for (int d = 0; d < 3; ++d) V[k][d] = coords[(size_t)v * 3 + d];
          }
          jacobian<3>(V, F, Fi, ad);
#pragma unroll
          for (int k = 0; k < 9; ++k) S.fi[k] = Fi[k];
          S.fi[9] = ad;
        }
      } else if (finv_all) {
        have_geo = true;
        if (lane == 0) {
          double Fi[9];
#pragma unroll
          for (int k = 0; k < 9; ++k) Fi[k] = finv_all[(size_t)i * 9 + k];
          const double dinv = Fi[0] * (Fi[4] * Fi[8] - Fi[5] * Fi[7]) - Fi[1] * (Fi[3] * Fi[8] - Fi[5] * Fi[6]) +
                              Fi[2] * (Fi[3] * Fi[7] - Fi[4] * Fi[6]);
#pragma unroll
          for (int k = 0; k < 9; ++k) S.fi[k] = Fi[k];
          S.fi[9] = 1.0 / fabs(dinv);
        }
      }
    }
    wave_sync();
    const double *cl = S.cl;
    double mvol[(WHICH & 1) ? NM : 1], mif[(WHICH == 2) ? NM : 1];
#pragma unroll
    for (int k = 0; k < ((WHICH & 1) ? NM : 1); ++k) mvol[k] = 0.0;
#pragma unroll
    for (int k = 0; k < ((WHICH == 2) ? NM : 1); ++k) mif[k] = 0.0;
    const int nleaf = c4.x;
    int nn = 64;  // lanes of the volume role (WHICH == 3)
    if constexpr (WHICH == 3) {
      // points: 24 per volume tet, 12 per triangle; a 1-3 cut leaf has 1
      // tet and 1 triangle, a 2-2 cut leaf 3 tets and 2 triangles (c4.y:
      // the walk's count of 2-2 leaves)
      const int n22 = c4.y, n13 = nleaf - n22;
      const double vp = 24.0 * (n13 + 3 * n22), ip = 12.0 * (n13 + 2 * n22);
      nn = (vp + ip > 0.0) ? (int)(64.0 * vp / (vp + ip) + 0.5) : 44;
      nn = nn < 16 ? 16 : (nn > 56 ? 56 : nn);
    }
    uint16_t *lst = lists + (size_t)i * cap;
    int ncomp = 0, st_vol = 0, st_if = 0;
    bool ok = true;
    for (int l0 = 0; l0 < nleaf; l0 += 64) {
      const int l = l0 + lane;
      LeafPieces2 PC;
      PC.nvol = 0;
      PC.nif = 0;
      PC.comp = false;
      int id = 0;
      if (l < nleaf) {
        id = lst[l];
        int W[4][3];
        unpack_uint4(nverts[id], shift, W);
        double V[4][3], vv[4], cpt[4][3];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
#pragma unroll
          for (int d = 0; d < 3; ++d) V[j][d] = W[j][d] * inv_n;
          vv[j] = lagrange_eval3<LO>(cl, V[j]);
        }
        if (!leaf_pieces<SIDE>(V, vv, cpt, PC)) {
          ok = false;
          PC.nvol = 0;
          PC.nif = 0;
          PC.comp = false;
        }
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            pool[(j * 3 + d) * 64 + lane] = V[j][d] - SHIFT;
            pool[((4 + j) * 3 + d) * 64 + lane] = cpt[j][d] - SHIFT;
          }
      }
      if (!(WHICH & 1)) {
        PC.nvol = 0;
        PC.comp = false;
      }
      if (!(WHICH & 2)) PC.nif = 0;
      {
        int cc;
        const int sc = wave_slot(PC.comp, ncomp, lane, cc);
        if (PC.comp) lst[nleaf + sc] = (uint16_t)id;
        ncomp += cc;
      }
      int tot_vol, tot_if;
      const int bvol = wave_excl_scan(PC.nvol, lane, tot_vol);
      const int bif = wave_excl_scan(PC.nif, lane, tot_if);
#pragma unroll
      for (int q = 0; q < 3; ++q)
        if (q < PC.nvol) S.volrec[bvol + q] = (uint32_t)lane | (PC.vol[q] << 8) | (PC.comp ? 0x80u : 0u);
#pragma unroll
      for (int q = 0; q < 2; ++q)
        if (q < PC.nif) S.ifrec[bif + q] = (uint32_t)lane | (PC.itf[q] << 8);
      st_vol += tot_vol;
      st_if += tot_if;
      wave_sync();
      // weight factors (as in k_subdiv_exact)
      for (int q = lane; q < tot_vol; q += 64) {
        const uint32_t rec = S.volrec[q];
        const int ll = rec & 63;
        double P4[4][3];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int j = (rec >> (8 + 3 * m)) & 7;
#pragma unroll
          for (int d = 0; d < 3; ++d) P4[m][d] = pool[(j * 3 + d) * 64 + ll];
        }
        const double tf = tet_tf(P4[0], P4[1], P4[2], P4[3], true);
        S.voltf[q] = (rec & 0x80u) ? -tf : tf;
      }
      for (int q = lane; q < tot_if; q += 64) {
        const uint32_t rec = S.ifrec[q];
        const int ll = rec & 63;
        double s0[3], s1[3], s2[3], pp[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          s0[d] = pool[(((rec >> 8) & 7) * 3 + d) * 64 + ll];
          s1[d] = pool[(((rec >> 11) & 7) * 3 + d) * 64 + ll];
          s2[d] = pool[(((rec >> 14) & 7) * 3 + d) * 64 + ll];
          pp[d] = pool[(((rec >> 17) & 7) * 3 + d) * 64 + ll];
        }
        double a[3], b[3], cc[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          a[d] = s1[d] - s0[d];
          b[d] = s2[d] - s0[d];
        }
        cc[0] = a[1] * b[2] - a[2] * b[1];
        cc[1] = a[2] * b[0] - a[0] * b[2];
        cc[2] = a[0] * b[1] - a[1] * b[0];
        const double tf = norm3(cc);
        const double na = norm3(a), nb = norm3(b);
        const double rel = na > nb ? na : nb;
        double wf = 0.0;
        if (!(tf < 1e-14 * rel)) {
#pragma unroll
          for (int d = 0; d < 3; ++d) cc[d] /= tf;
          const double dd0 = pp[0] - s0[0], dd1 = pp[1] - s0[1], dd2 = pp[2] - s0[2];
          if (dd0 * cc[0] + dd1 * cc[1] + dd2 * cc[2] >= 0)
#pragma unroll
            for (int d = 0; d < 3; ++d) cc[d] *= -1.0;
          Geo geo;
          geo.Finv = have_geo ? S.fi : nullptr;
          geo.absdet = have_geo ? S.fi[9] : 1.0;
          wf = tf * geo.if_factor(cc);
        }
        S.ifw[q] = wf;
      }
      wave_sync();
      if constexpr (WHICH == 3) {
        // one task loop for both roles: (piece, block of 4 points) tasks
        const int ni = 64 - nn;
        const int nvt = tot_vol * 6, nit = tot_if * 3;
        const int rounds = max((nvt + nn - 1) / nn, (nit + ni - 1) / ni);
        const bool isv = lane < nn;
        const int rl = isv ? lane : lane - nn, rs = isv ? nn : ni;
        for (int rd = 0; rd < rounds; ++rd) {
          const int q = rd * rs + rl;
          if (isv ? (q < nvt) : (q < nit)) {
            const int pc = isv ? q / 6 : q / 3;
            const int t0 = 4 * (isv ? q - pc * 6 : q - pc * 3);
            const uint32_t rec = isv ? S.volrec[pc] : S.ifrec[pc];
            const double wf = isv ? S.voltf[pc] : S.ifw[pc];
            const double(*rule)[4] = isv ? S.rule_vol : S.rule_if;
            const int ll = rec & 63;
            double p0[3], p1[3], p2[3], p3[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              p0[d] = pool[(((rec >> 8) & 7) * 3 + d) * 64 + ll];
              p1[d] = pool[(((rec >> 11) & 7) * 3 + d) * 64 + ll];
              p2[d] = pool[(((rec >> 14) & 7) * 3 + d) * 64 + ll];
              p3[d] = pool[(((rec >> 17) & 7) * 3 + d) * 64 + ll];
            }
            for (int t = t0; t < t0 + 4; ++t) {
              const double r0 = rule[t][0], r1 = rule[t][1], r2 = rule[t][2], rw = rule[t][3];
              const double ow = 1.0 - r0 - r1 - r2;
              double y[3];
#pragma unroll
              for (int d = 0; d < 3; ++d) y[d] = ow * p0[d] + r0 * p1[d] + r1 * p2[d] + r2 * p3[d];
              Moments<3, PM>::accum_aff(y, rw * wf, mvol);
            }
          }
        }
      }
      if constexpr (WHICH == 1 && PM <= 2) {
        // p <= 2: one task per piece, its moments from its vertices
        for (int q = lane; q < tot_vol; q += 64) {
          const uint32_t rec = S.volrec[q];
          const int ll = rec & 63;
          double p0[3], p1[3], p2[3], p3[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            p0[d] = pool[(((rec >> 8) & 7) * 3 + d) * 64 + ll];
            p1[d] = pool[(((rec >> 11) & 7) * 3 + d) * 64 + ll];
            p2[d] = pool[(((rec >> 14) & 7) * 3 + d) * 64 + ll];
            p3[d] = pool[(((rec >> 17) & 7) * 3 + d) * 64 + ll];
          }
          htet<PM>(mvol, p0, p1, p2, p3, S.voltf[q]);
        }
      } else if (WHICH == 1) {
        for (int q = lane; q < tot_vol * SDX_NVB; q += 64) {
          const int pc = q / SDX_NVB, t0 = (q - pc * SDX_NVB) * SDX_VB;
          const uint32_t rec = S.volrec[pc];
          const int ll = rec & 63;
          double p0[3], p1[3], p2[3], p3[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            p0[d] = pool[(((rec >> 8) & 7) * 3 + d) * 64 + ll];
            p1[d] = pool[(((rec >> 11) & 7) * 3 + d) * 64 + ll];
            p2[d] = pool[(((rec >> 14) & 7) * 3 + d) * 64 + ll];
            p3[d] = pool[(((rec >> 17) & 7) * 3 + d) * 64 + ll];
          }
          const double tf = S.voltf[pc];
          // the next point's rule entry is read while this one accumulates
          double nr[4];
#pragma unroll
          for (int c = 0; c < 4; ++c) nr[c] = S.rule_vol[t0][c];
          for (int t = t0; t < t0 + SDX_VB; ++t) {
            const double r0 = nr[0], r1 = nr[1], r2 = nr[2], rw = nr[3];
            const int tn = t + 1 < t0 + SDX_VB ? t + 1 : t;
#pragma unroll
            for (int c = 0; c < 4; ++c) nr[c] = S.rule_vol[tn][c];
            const double ow = 1.0 - r0 - r1 - r2;
            double y[3];
#pragma unroll
            for (int d = 0; d < 3; ++d) y[d] = ow * p0[d] + r0 * p1[d] + r1 * p2[d] + r2 * p3[d];
            Moments<3, PM>::accum_aff(y, rw * tf, mvol);
          }
        }
      }
      if constexpr (WHICH == 2) {
        for (int q = lane; q < tot_if * SDX_NIB; q += 64) {
          const int pc = q / SDX_NIB, t0 = (q - pc * SDX_NIB) * SDX_IB;
          const uint32_t rec = S.ifrec[pc];
          const int ll = rec & 63;
          double s0[3], s1[3], s2[3];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            s0[d] = pool[(((rec >> 8) & 7) * 3 + d) * 64 + ll];
            s1[d] = pool[(((rec >> 11) & 7) * 3 + d) * 64 + ll];
            s2[d] = pool[(((rec >> 14) & 7) * 3 + d) * 64 + ll];
          }
          const double wf = S.ifw[pc];
          double nr[3] = {S.rule_if[t0][0], S.rule_if[t0][1], S.rule_if[t0][3]};
          for (int t = t0; t < t0 + SDX_IB; ++t) {
            const double r0 = nr[0], r1 = nr[1], rw = nr[2];
            const int tn = t + 1 < t0 + SDX_IB ? t + 1 : t;
            nr[0] = S.rule_if[tn][0];
            nr[1] = S.rule_if[tn][1];
            nr[2] = S.rule_if[tn][3];
            const double ow = 1.0 - r0 - r1;
            double lam[4];
#pragma unroll
            for (int d = 0; d < 3; ++d) lam[d] = ow * s0[d] + r0 * s1[d] + r1 * s2[d];
            if constexpr (AFF) {
              Moments<3, PM>::accum_aff(lam, rw * wf, mif);
            } else {
              lam[3] = 1.0 - lam[0] - lam[1] - lam[2];
              Moments<3, PM>::accum(lam, rw * wf, mif);
            }
          }
        }
      }
      wave_sync();
    }
    // wave reduction of the per-lane moments through LDS (the point pool is
    // free now)
    if constexpr (WHICH == 1 && PM <= 2) Moments<3, PM>::hfinish(mvol);
    if constexpr (WHICH == 3) wave_sum_lds<NM>(mvol, pool, S.red, lane, S.red + NM, nn);
    else if constexpr (WHICH == 1) wave_sum_lds<NM>(mvol, pool, S.red, lane);
    else wave_sum_lds<NM>(mif, pool, S.red + NM, lane);
    if constexpr (AFF) {
      if constexpr (WHICH == 3) aff33_to_hom<2>(S.red, lane);
      else if constexpr (WHICH == 1) aff33_to_hom<1>(S.red, lane);
      else aff33_to_hom<1>(S.red + NM, lane);
    }
    for (int k = lane; k < NM; k += 64) {
      // (the volume side adds to the walk's table-row sums)
      if ((WHICH & 1) && out_vol) out_vol[(size_t)i * out_stride + k] += S.red[k];
      if ((WHICH & 2) && out_if) out_if[(size_t)i * out_stride + k] = S.red[NM + k];
    }
    const bool allok = __ballot(!ok) == 0ull;
    if (lane == 0) {
      if (WHICH & 1) cnt[i].z = ncomp;
      if (stats) {
        stats[(size_t)i * stats_stride + 0] = (WHICH & 1) ? c4.w + ncomp : 0;
        stats[(size_t)i * stats_stride + 1] = st_vol;
        stats[(size_t)i * stats_stride + 2] = (WHICH == 3) ? 0 : st_if;
        stats[(size_t)i * stats_stride + 3] = nleaf;
      }
      if (WHICH == 3 && stats_if) {
        stats_if[(size_t)i * stats_stride + 0] = 0;
        stats_if[(size_t)i * stats_stride + 1] = 0;
        stats_if[(size_t)i * stats_stride + 2] = st_if;
        stats_if[(size_t)i * stats_stride + 3] = nleaf;
      }
      if (!allok) atomicOr(err, 2);
    }
    wave_sync();
  }
}

// table rows of each listed element (uncut nodes of the kept side from the
// back of the list, complemented leaves after the cut leaves) added to its
// moments; one wave per element.  Rows of an even length (p = 1, 3): R =
// NM / 4 lanes per row, each loading 4 moments as two 16-byte loads, so a
// wave reads 64 / R rows per instruction (3 for p = 3), four rounds of them
// in flight; the rows' partial sums meet in LDS.  p = 2 (35 moments, rows not
// 16-byte aligned): lane = moment (and moment + 64), four rows in flight.
template <int PM>
__global__ __launch_bounds__(256) void k_sdx_rows(int nel, const uint16_t *__restrict__ lists, int cap,
                                                  const int4 *__restrict__ cnt, const double *__restrict__ nmom,
                                                  double *__restrict__ out, int out_stride) {
  constexpr int NM = Moments<3, PM>::NM;
  constexpr int R = (NM + 3) / 4, G = 64 / R;  // lanes per row, rows per instruction
  __shared__ double scr[4][G][4 * R];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nw = gridDim.x * (blockDim.x >> 6);
  for (int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < nel; i += nw) {
    const int4 c4 = cnt[i];
    if (c4.x < 0) continue;
    const uint16_t *lst = lists + (size_t)i * cap;
    // the complemented leaves (.z) after the cut leaves (.x)
    const int nrow = c4.z;
    if constexpr (NM % 2 == 0) {
      const int slot = lane / R, sub = lane - slot * R, c0 = 4 * sub;
      const bool act = slot < G;
      double a[4][4] = {};
      for (int r0 = 0; r0 < nrow; r0 += 4 * G) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int rr = r0 + u * G + slot;
          if (act && rr < nrow) {
            const int id = lst[c4.x + rr];
            const double2 *src = reinterpret_cast<const double2 *>(nmom + (size_t)id * NM + c0);
            const double2 x = src[0];
            a[u][0] += x.x;
            a[u][1] += x.y;
            if (c0 + 2 < NM) {
              const double2 y = src[1];
              a[u][2] += y.x;
              a[u][3] += y.y;
            }
          }
        }
      }
      if (act) {
#pragma unroll
        for (int c = 0; c < 4; ++c) scr[w][slot][c0 + c] = (a[0][c] + a[1][c]) + (a[2][c] + a[3][c]);
      }
      wave_sync();
      for (int k = lane; k < NM; k += 64) {
        double t = 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g) t += scr[w][g][k];
        out[(size_t)i * out_stride + k] += t;
      }
      wave_sync();
    } else {
      // four rows in flight per lane (moment lane, moment lane + 64)
      double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
      for (int r0 = 0; r0 < nrow; r0 += 64) {
        // 64 row ids per round, one per lane, then broadcast
        const int rr = r0 + lane;
        int id = 0;
        if (rr < nrow) id = lst[c4.x + rr];
        const int m = min(64, nrow - r0);
        int j = 0;
        for (; j + 4 <= m; j += 4) {
          size_t ro[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) ro[u] = (size_t)__shfl(id, j + u, 64) * NM;
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            if (lane < NM) a[u] += nmom[ro[u] + lane];
            if (lane + 64 < NM) b[u] += nmom[ro[u] + 64 + lane];
          }
        }
        for (; j < m; ++j) {
          const size_t ra = (size_t)__shfl(id, j, 64) * NM;
          if (lane < NM) a[0] += nmom[ra + lane];
          if (lane + 64 < NM) b[0] += nmom[ra + 64 + lane];
        }
      }
      if (lane < NM) out[(size_t)i * out_stride + lane] += (a[0] + a[1]) + (a[2] + a[3]);
      if (lane + 64 < NM) out[(size_t)i * out_stride + 64 + lane] += (b[0] + b[1]) + (b[2] + b[3]);
    }
  }
}

constexpr int SDC_CAP = 4096;       // list entries per element (uint16)
constexpr int SDC_SUB = 1 << 19;    // elements per staged round

template <int LO, int PM>
int launch_sdx_tree(xg_ctx *ctx, int side, int nel, const int32_t *elems, const double *lset, int L,
                    uint16_t *lists, int cap, int4 *cnt, uint8_t *dom, int32_t *over, int32_t *nover, double *out,
                    int out_stride) {
  const size_t lds = sdc_bytes(L, nullptr, nullptr);
  const int blocks = nel < (1 << 22) ? nel : (1 << 22);
  KernelTimer tm(ctx, "sdx_tree");
#define XG_TREE(S)                                                                                               \
  {                                                                                                              \
    if (lds > 65536)                                                                                             \
      XG_CHECK(hipFuncSetAttribute((const void *)k_sdx_tree<LO, PM, S>,                                          \
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));                       \
    k_sdx_tree<LO, PM, S><<<blocks, 64, lds, ctx->stream>>>(nel, elems, lset, L, (const uint4 *)ctx->sdx_verts,  \
                                                             lists, cap, cnt, dom, over, nover, ctx->sdx_mom[PM], \
                                                             out, out_stride, ctx->d_err);                         \
  }
  if (side == NEG) XG_TREE(NEG)
  else if (side == POS) XG_TREE(POS)
  else XG_TREE(-1)
#undef XG_TREE
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int LO, int PM, int SIDE, int WHICH>
int launch_sdx_leaves(xg_ctx *ctx, int nel, const int32_t *elems, const double *lset, const int32_t *el2v,
                      const double *coords, const double *finv, int L, uint16_t *lists, int cap, int4 *cnt,
                      double *out_vol, double *out_if, int out_stride, int32_t *stats, int stats_stride,
                      int32_t *queue, int32_t *stats_if = nullptr) {
  constexpr int NM = Moments<3, PM>::NM;
  const void *kern = (const void *)k_sdx_leaves<LO, PM, SIDE, WHICH>;
  const size_t lds = sdx_leaf_bytes(xg_minrule_cnt[3][PM], xg_minrule_cnt[2][PM], NM, nullptr, nullptr);
  if (lds > 65536) XG_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int blocks = nel < (1 << 22) ? nel : (1 << 22);
  if (queue) {
    // as many persistent waves as fit on the device at once
    int per_cu = 0, ncu = 0;
    XG_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64, lds));
    XG_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, ctx->device));
    const long long res = (long long)(per_cu > 0 ? per_cu : 1) * (ncu > 0 ? ncu : 1);
    if (res < blocks) blocks = (int)res;
    k_zero_i32<<<1, 1, 0, ctx->stream>>>(queue);
    XG_LAUNCH_CHECK();
  }
  KernelTimer tm(ctx, WHICH == 3 ? "sdx_leaves_both" : ((WHICH & 1) ? "sdx_leaves_vol" : "sdx_leaves_if"));
  k_sdx_leaves<LO, PM, SIDE, WHICH><<<blocks, 64, lds, ctx->stream>>>(
      nel, elems, lset, el2v, coords, finv, L, (const uint4 *)ctx->sdx_verts, lists, cap, cnt, out_vol, out_if,
      out_stride, stats, stats_stride, ctx->d_err, queue, stats_if);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

template <int PM>
int launch_sdx_rows(xg_ctx *ctx, int nel, const uint16_t *lists, int cap, const int4 *cnt, double *out,
                    int out_stride) {
  const int blocks = (nel + 3) / 4 < (1 << 20) ? (nel + 3) / 4 : (1 << 20);
  KernelTimer tm(ctx, "sdx_rows");
  k_sdx_rows<PM><<<blocks, 256, 0, ctx->stream>>>(nel, lists, cap, cnt, ctx->sdx_mom[PM], out, out_stride);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

// Moments of the parts dts[0..ndt) (NEG / POS / IF) of nel elements, part u
// written at outs[u] + i * mstride (stats, nullable: stats_u[u] + i *
// sstride), through the staged kernels; the elements whose lists overflow
// go through the one-kernel path.  dom (nullable): each element's domain
// type on the lattice.
template <int LO, int PM>
int subdiv_staged(xg_ctx *ctx, int nel, const int32_t *elems, const double *lset, const int32_t *el2v,
                  const double *coords, const double *finv, int L, int ndt, const int *dts,
                  double *const *outs, int mstride, int32_t *const *stats_u, int sstride, uint8_t *dom) {
  constexpr int NM = Moments<3, PM>::NM;
  int rc = ensure_sdx_tables(ctx, PM);
  if (rc != XG_OK) return rc;
  // list capacity: SDC_CAP, or smaller through XG_SDX_CAP (16 and up; the
  // elements that do not fit go through the one-kernel path)
  int cap = SDC_CAP;
  if (const char *ev = getenv("XG_SDX_CAP")) {
    const int v = atoi(ev);
    if (v >= 16 && v < SDC_CAP) cap = v;
  }
  const int sub = nel < SDC_SUB ? nel : SDC_SUB;
  const size_t b_lists = sd_align(sizeof(uint16_t) * (size_t)sub * cap);
  const size_t b_cnt = sd_align(sizeof(int4) * (size_t)sub);
  const size_t b_over = sd_align(sizeof(int32_t) * ((size_t)sub + 2));
  // persistent waves on a work counter (XG_SDX_QUEUE=0: one block per element)
  bool use_queue = true;
  if (const char *ev = getenv("XG_SDX_QUEUE")) use_queue = atoi(ev) != 0;
  rc = ensure_buf(ctx, &ctx->sdx_ws, &ctx->sdx_ws_size, b_lists + b_cnt + b_over);
  if (rc != XG_OK) return rc;
  uint16_t *lists = (uint16_t *)ctx->sdx_ws;
  int4 *cnt = (int4 *)((char *)ctx->sdx_ws + b_lists);
  int32_t *nover = (int32_t *)((char *)ctx->sdx_ws + b_lists + b_cnt);
  int32_t *queue = use_queue ? nover + 1 : nullptr;
  int32_t *over = nover + 2;
  for (int off = 0; off < nel; off += sub) {
    const int m = nel - off < sub ? nel - off : sub;
    const int32_t *el = elems ? elems + off : nullptr;
    const double *fi = finv ? finv + (size_t)off * 9 : nullptr;
    double *o_u[3] = {nullptr, nullptr, nullptr};
    int32_t *s_u[3] = {nullptr, nullptr, nullptr};
    for (int u = 0; u < ndt; ++u) {
      o_u[u] = outs[u] + (size_t)off * mstride;
      s_u[u] = (stats_u && stats_u[u]) ? stats_u[u] + (size_t)off * sstride : nullptr;
    }
    uint8_t *dm = dom ? dom + off : nullptr;
    // volume sides in the order requested; the interface rides on the first
    // tree (its cut leaves do not depend on the side)
    bool has_if = false, if_done = false;
    for (int t = 0; t < ndt; ++t) has_if = has_if || dts[t] == IF;
    for (int t = 0; t < ndt || (has_if && !if_done && t == ndt); ++t) {
      int side;
      if (t < ndt) {
        if (dts[t] == IF) continue;
        side = dts[t];
      } else {
        side = -1;  // interface only
      }
      k_zero_i32<<<1, 1, 0, ctx->stream>>>(nover);
      XG_LAUNCH_CHECK();
      // the volume side's table-row sums go straight to its moments
      double *ovol = nullptr;
      for (int u = 0; u < ndt; ++u)
        if (dts[u] == side) ovol = o_u[u];
      rc = launch_sdx_tree<LO, PM>(ctx, side, m, el, lset, L, lists, cap, cnt, dm, over, nover, ovol, mstride);
      if (rc != XG_OK) return rc;
      // the overflowed elements (count in device memory) go through the
      // one-kernel path below; its launch strides over the count, so no
      // host round trip is needed
      // p = 3 with the volume side and the interface both requested: one
      // launch for both (lane roles), then the rows, then the overflow path
      int u_vol = -1, u_if = -1;
      for (int u = 0; u < ndt; ++u) {
        if (dts[u] == side) u_vol = u;
        if (dts[u] == IF && !if_done) u_if = u;
      }
      if constexpr (PM == 3) {
        if (u_vol >= 0 && u_if >= 0 && !getenv("XG_SDX_NOFUSE")) {
          double *ov = o_u[u_vol], *oi = o_u[u_if];
          int32_t *sv = s_u[u_vol], *si = s_u[u_if];
          if (side == NEG)
            rc = launch_sdx_leaves<LO, PM, NEG, 3>(ctx, m, el, lset, el2v, coords, fi, L, lists, cap, cnt, ov, oi,
                                                   mstride, sv, sstride, queue, si);
          else
            rc = launch_sdx_leaves<LO, PM, POS, 3>(ctx, m, el, lset, el2v, coords, fi, L, lists, cap, cnt, ov, oi,
                                                   mstride, sv, sstride, queue, si);
          if (rc != XG_OK) return rc;
          rc = launch_sdx_rows<PM>(ctx, m, lists, cap, cnt, ov, mstride);
          if (rc != XG_OK) return rc;
          rc = subdiv_exact_pass<LO, PM>(ctx, m, el, lset, el2v, coords, fi, L, side, false, ov, nullptr, mstride,
                                         nullptr, sv, sstride, over, nover);
          if (rc != XG_OK) return rc;
          rc = subdiv_exact_pass<LO, PM>(ctx, m, el, lset, el2v, coords, fi, L, -1, true, nullptr, oi, mstride,
                                         nullptr, si, sstride, over, nover);
          if (rc != XG_OK) return rc;
          if_done = true;
          if (t == ndt) break;
          continue;
        }
      }
      for (int u = 0; u < ndt; ++u) {
        const bool vol = dts[u] == side;
        const bool itf = dts[u] == IF && !if_done;
        if (!vol && !itf) continue;
        double *o = o_u[u];
        int32_t *su = s_u[u];
        if (vol) {
          if (side == NEG)
            rc = launch_sdx_leaves<LO, PM, NEG, 1>(ctx, m, el, lset, el2v, coords, fi, L, lists, cap, cnt, o, nullptr,
                                                   mstride, su, sstride, queue);
          else
            rc = launch_sdx_leaves<LO, PM, POS, 1>(ctx, m, el, lset, el2v, coords, fi, L, lists, cap, cnt, o, nullptr,
                                                   mstride, su, sstride, queue);
          if (rc != XG_OK) return rc;
          rc = launch_sdx_rows<PM>(ctx, m, lists, cap, cnt, o, mstride);
        } else {
          rc = launch_sdx_leaves<LO, PM, NEG, 2>(ctx, m, el, lset, el2v, coords, fi, L, lists, cap, cnt, nullptr, o,
                                                 mstride, su, sstride, queue);
        }
        if (rc != XG_OK) return rc;
        rc = subdiv_exact_pass<LO, PM>(ctx, m, el, lset, el2v, coords, fi, L, vol ? side : -1, !vol,
                                       vol ? o : nullptr, vol ? nullptr : o, mstride, nullptr, su, sstride, over,
                                       nover);
        if (rc != XG_OK) return rc;
        if (itf) if_done = true;
      }
      if (t == ndt) break;
    }
  }
  return XG_OK;
}

int subdiv_exact(xg_ctx *ctx, int lo, int p, int nel, const int32_t *elems, const double *lset_nodal,
                 const int32_t *el2v, const double *coords, const double *finv, int L, int vol_side, bool with_if,
                 double *out_vol, double *out_if, int out_stride, uint8_t *dom, int32_t *stats, int stats_stride) {
#define XG_SDXP(LO, PM)                                                                                    \
  if (lo == LO && p == PM)                                                                                 \
    return subdiv_exact_pass<LO, PM>(ctx, nel, elems, lset_nodal, el2v, coords, finv, L, vol_side, with_if, \
                                     out_vol, out_if, out_stride, dom, stats, stats_stride);
  XG_SDXP(1, 1) XG_SDXP(1, 2) XG_SDXP(1, 3) XG_SDXP(2, 1) XG_SDXP(2, 2) XG_SDXP(2, 3)
#undef XG_SDXP
  return set_error(XG_ERR_ARG, "subdivision rule: bad level-set order or degree");
}

// fixed-node weights of the degree-4 Lagrange lattice from degree-4 moments,
// in place: W_k = sum_alpha nodal34[k][alpha] m_alpha (one element per
// thread, its 35 moments read before any weight is written)
__global__ __launch_bounds__(128) void k_moments_to_nodal4(int nel, double *__restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  double m[35];
#pragma unroll
  for (int k = 0; k < 35; ++k) m[k] = a[(size_t)i * 35 + k];
  for (int k = 0; k < 35; ++k) {
    double w = 0.0;
#pragma unroll
    for (int q = 0; q < 35; ++q) w = fma(nodal34[k][q], m[q], w);
    a[(size_t)i * 35 + k] = w;
  }
}

// The reference-tet entry points (level-set nodal values per element, an
// optional F^{-1} per element for the interface weights) through the staged
// kernels: the NEG volume (out_vol) and / or interface (out_if) moments,
// stride nm.  Without either output the one-kernel path (domains only).
int subdiv_ref(xg_ctx *ctx, int lo, int p, int nel, const double *lset, const double *finv, int L, double *out_vol,
               double *out_if, int nm, uint8_t *dom) {
  int dts[2];
  double *outs[2];
  int ndt = 0;
  if (out_vol) {
    dts[ndt] = NEG;
    outs[ndt++] = out_vol;
  }
  if (out_if) {
    dts[ndt] = IF;
    outs[ndt++] = out_if;
  }
  if (ndt == 0)
    return subdiv_exact(ctx, lo, p, nel, nullptr, lset, nullptr, nullptr, finv, L, NEG, false, nullptr, nullptr, nm,
                        dom, nullptr, 0);
#define XG_SREF(LO, PM)                                                                                            \
  if (lo == LO && p == PM)                                                                                         \
    return subdiv_staged<LO, PM>(ctx, nel, nullptr, lset, nullptr, nullptr, finv, L, ndt, dts, outs, nm, nullptr, \
                                 0, dom);
  XG_SREF(1, 1) XG_SREF(1, 2) XG_SREF(1, 3) XG_SREF(2, 1) XG_SREF(2, 2) XG_SREF(2, 3)
#undef XG_SREF
  return set_error(XG_ERR_ARG, "subdivision rule: bad level-set order or degree");
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
  if (order >= 4) {
    // exact base rules: degree-4 moments from the exact-rule kernel, then
    // the nodal weights (the degree-4 lattice integrates exactly what the
    // moments do)
    int rc = subdiv_ref(ctx, lset_order, 2, nel, lset_nodal, finv, L, out_neg, out_if, 35, dom);
    if (rc != XG_OK) return rc;
    for (double *a : {out_neg, out_if})
      if (a) {
        k_moments_to_nodal4<<<(nel + 127) / 128, 128, 0, ctx->stream>>>(nel, a);
        XG_LAUNCH_CHECK();
      }
    return XG_OK;
  }
  if (lset_order == 2)
    return launch_subdiv<2, 2, 3, 1>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  return launch_subdiv<1, 2, 3, 1>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
}

extern "C" int xg_subdiv_moments(xg_ctx *ctx, int nel, const double *lset_nodal, int lset_order, int L, int order,
                                 int p, const double *finv, double *out_neg, double *out_if, uint8_t *dom) {
  if (!ctx || nel < 0 || L < 0 || L > XG_LAT_LMAX || order < 0 || order / 2 + 1 > XG_RULE_NMAX ||
      (lset_order != 1 && lset_order != 2) || p < 1 || p > 3)
    return set_error(XG_ERR_ARG, "xg_subdiv_moments: bad argument (L <= 5, lset_order 1|2, 1 <= p <= 3)");
  if (nel == 0) return XG_OK;
  if (!lset_nodal) return set_error(XG_ERR_ARG, "xg_subdiv_moments: null level set");
  if (order >= 2 * p) {
    // exact base rules: node-moment tables + complemented leaves
    const int nm = p == 1 ? 10 : (p == 2 ? 35 : 84);
    return subdiv_ref(ctx, lset_order, p, nel, lset_nodal, finv, L, out_neg, out_if, nm, dom);
  }
  if (lset_order == 1 && p == 1) return subdiv_moments_impl<1, 1>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  if (lset_order == 1 && p == 2) return subdiv_moments_impl<1, 2>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  if (lset_order == 1 && p == 3) return subdiv_moments_impl<1, 3>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  if (lset_order == 2 && p == 1) return subdiv_moments_impl<2, 1>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  if (lset_order == 2 && p == 2) return subdiv_moments_impl<2, 2>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  if (lset_order == 2 && p == 3) return subdiv_moments_impl<2, 3>(ctx, nel, lset_nodal, L, order, finv, out_neg, out_if, dom);
  return set_error(XG_ERR_ARG, "xg_subdiv_moments: bad argument");
}

static int subdiv_mesh_check(const char *who, int p, int L, int order, int nel, int ndt, const int *dts) {
  if (p < 1 || p > 3 || L < 0 || L > XG_LAT_LMAX || nel < 0 || ndt < 1 || ndt > 3 || !dts)
    return set_error(XG_ERR_ARG, std::string(who) + ": bad argument (1 <= p <= 3, L <= 5, 1 <= ndt <= 3)");
  if (order < 2 * p)
    return set_error(XG_ERR_UNSUPPORTED, std::string(who) + ": order >= 2p (exact base rules) only");
  for (int t = 0; t < ndt; ++t)
    if (dts[t] < 0 || dts[t] > 2) return set_error(XG_ERR_ARG, std::string(who) + ": bad domain type");
  return XG_OK;
}

static int subdiv_mesh_moments(xg_ctx *ctx, int p, int L, int nel, const int32_t *elems, const int32_t *el2v,
                               const double *coords, const double *lset_nodal, int ndt, const int *dts,
                               double *mom, int32_t *stats) {
  const int nm = p == 1 ? 10 : (p == 2 ? 35 : 84);
  double *outs[3];
  int32_t *st[3];
  for (int u = 0; u < ndt; ++u) {
    outs[u] = mom + (size_t)u * nm;
    st[u] = stats ? stats + (size_t)u * 4 : nullptr;
  }
  if (p == 1)
    return subdiv_staged<2, 1>(ctx, nel, elems, lset_nodal, el2v, coords, nullptr, L, ndt, dts, outs, ndt * nm, st,
                               ndt * 4, nullptr);
  if (p == 2)
    return subdiv_staged<2, 2>(ctx, nel, elems, lset_nodal, el2v, coords, nullptr, L, ndt, dts, outs, ndt * nm, st,
                               ndt * 4, nullptr);
  return subdiv_staged<2, 3>(ctx, nel, elems, lset_nodal, el2v, coords, nullptr, L, ndt, dts, outs, ndt * nm, st,
                             ndt * 4, nullptr);
}

extern "C" int xg_subdiv_moments_mesh(xg_ctx *ctx, int p, int L, int order, int nel, const int32_t *elems,
                                      const int32_t *el2v, const double *coords, const double *lset_nodal, int ndt,
                                      const int *dts, double *moments, int32_t *stats) {
  if (!ctx) return set_error(XG_ERR_ARG, "xg_subdiv_moments_mesh: null context");
  int rc = subdiv_mesh_check("xg_subdiv_moments_mesh", p, L, order, nel, ndt, dts);
  if (rc != XG_OK || nel == 0) return rc;
  if (!elems || !el2v || !coords || !lset_nodal || !moments)
    return set_error(XG_ERR_ARG, "xg_subdiv_moments_mesh: null pointer (elems, el2v, coords, lset_nodal, moments)");
  return subdiv_mesh_moments(ctx, p, L, nel, elems, el2v, coords, lset_nodal, ndt, dts, moments, stats);
}

extern "C" int xg_subdiv_weights_mesh(xg_ctx *ctx, int L, int order, int nel, const int32_t *elems,
                                      const int32_t *el2v, const double *coords, const double *lset_nodal, int ndt,
                                      const int *dts, double *weights, int32_t *stats) {
  if (!ctx) return set_error(XG_ERR_ARG, "xg_subdiv_weights_mesh: null context");
  int rc = subdiv_mesh_check("xg_subdiv_weights_mesh", 2, L, order, nel, ndt, dts);
  if (rc != XG_OK || nel == 0) return rc;
  if (!elems || !el2v || !coords || !lset_nodal || !weights)
    return set_error(XG_ERR_ARG, "xg_subdiv_weights_mesh: null pointer (elems, el2v, coords, lset_nodal, weights)");
  rc = subdiv_mesh_moments(ctx, 2, L, nel, elems, el2v, coords, lset_nodal, ndt, dts, weights, stats);
  if (rc != XG_OK) return rc;
  // degree-4 moments -> weights of the degree-4 lattice, each (element, dt)
  // row in place
  const long long rows = (long long)nel * ndt;
  if (rows > 0x7fffffffLL) return set_error(XG_ERR_ARG, "xg_subdiv_weights_mesh: too many elements");
  k_moments_to_nodal4<<<(unsigned)((rows + 127) / 128), 128, 0, ctx->stream>>>((int)rows, weights);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_subdiv_pipeline(xg_ctx *ctx, int p, int L, int order, int nel, const int32_t *elems,
                                  const int32_t *el2v, const double *coords, const double *lset_nodal, int ndt,
                                  const int *dts, int mask, int precision, void *out, int32_t *stats) {
  if (!ctx) return set_error(XG_ERR_ARG, "xg_subdiv_pipeline: null context");
  int rc = subdiv_mesh_check("xg_subdiv_pipeline", p, L, order, nel, ndt, dts);
  if (rc != XG_OK) return rc;
  if (mask < 1 || mask > 3) return set_error(XG_ERR_ARG, "xg_subdiv_pipeline: bad form mask");
  if (precision != XG_PREC_FP32 && precision != XG_PREC_FP64)
    return set_error(XG_ERR_UNSUPPORTED, "xg_subdiv_pipeline: FP32 or FP64 output");
  if (nel == 0) return XG_OK;
  if (!elems || !el2v || !coords || !lset_nodal || !out)
    return set_error(XG_ERR_ARG, "xg_subdiv_pipeline: null pointer (elems, el2v, coords, lset_nodal, out)");
  const int nm = p == 1 ? 10 : (p == 2 ? 35 : 84);
  const int nd = p == 1 ? 4 : (p == 2 ? 10 : 20);
  const size_t osz = precision == XG_PREC_FP64 ? sizeof(double) : sizeof(float);
  // moments of up to 2^21 elements at a time in the context workspace, then
  // the element matrices from them
  const int chunk = nel < (1 << 21) ? nel : (1 << 21);
  rc = ensure_ws(ctx, sizeof(double) * (size_t)chunk * ndt * nm);
  if (rc != XG_OK) return rc;
  double *mom = (double *)ctx->ws;
  for (int off = 0; off < nel; off += chunk) {
    const int m = nel - off < chunk ? nel - off : chunk;
    rc = subdiv_mesh_moments(ctx, p, L, m, elems + off, el2v, coords, lset_nodal, ndt, dts, mom,
                             stats ? stats + (size_t)off * ndt * 4 : nullptr);
    if (rc != XG_OK) return rc;
    rc = elmat_from_moments(ctx, p, m, ndt, elems + off, el2v, coords, mom, mask, precision == XG_PREC_FP64,
                            (char *)out + (size_t)off * ndt * nd * nd * osz);
    if (rc != XG_OK) return rc;
  }
  return XG_OK;
}
