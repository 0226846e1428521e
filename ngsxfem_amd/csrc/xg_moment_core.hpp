// xg_moment_core.hpp -- device building blocks of the moment-form element
// matrices (shared by the single-mesh kernel in xg_moment.hip and the
// batched persistent kernels in xg_batch.hip).  See xg_moment.hip.
#pragma once
#include "moments_table.h"
#include "xg_device.hpp"

namespace xg {
namespace mcore {

// a DecompEntry held in registers as six 32-bit words (byte layout of the
// struct: ns, nv, nc, ok, v[3][4], ce[4][2])
struct PackedEntry {
  uint32_t w[6];
  __device__ __forceinline__ int byte(int b) const {  // b: compile-time after unrolling
    return (int)(int8_t)((w[b >> 2] >> (8 * (b & 3))) & 255u);
  }
  __device__ __forceinline__ int ns() const { return byte(0); }
  __device__ __forceinline__ int nc() const { return byte(2); }
  __device__ __forceinline__ int ok() const { return byte(3); }
  __device__ __forceinline__ int v(int s, int k) const { return byte(4 + 4 * s + k); }
  __device__ __forceinline__ int ce(int c, int j) const { return byte(16 + 2 * c + j); }
};
static_assert(sizeof(DecompEntry) == 24, "DecompEntry layout");

// vertex `code` of a decomposition entry in reference coordinates (FP64,
// same operation order as cut_point_array)
template <int D>
__device__ __forceinline__ void piece_vertex(int code, const PackedEntry &E, double t0, double t1, double t2,
                                             double t3, double *p) {
  if (code <= D) {
#pragma unroll
    for (int d = 0; d < 3; ++d) p[d] = (d == code && code < D) ? 1.0 : 0.0;
    return;
  }
  const int c = code - (D + 1);
  int i = E.ce(0, 0), j = E.ce(0, 1);
#pragma unroll
  for (int k = 1; k < 4; ++k) {
    i = (c == k) ? E.ce(k, 0) : i;
    j = (c == k) ? E.ce(k, 1) : j;
  }
  // explicit selects (an indexed array here ends up in scratch memory)
  const double tc = (c == 0) ? t0 : ((c == 1) ? t1 : ((c == 2) ? t2 : t3));
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double pi = (d == i && i < D) ? 1.0 : 0.0;
    const double pj = (d == j && j < D) ? 1.0 : 0.0;
    p[d] = pi + tc * (pj - pi);
  }
}

// SimpleX::GetVolume (straightcutrule.cpp:66-72) on explicit vertices
__device__ __forceinline__ double piece_volume(int nv, const double (*W)[3]) {
  if (nv == 1) return 1.0;
  if (nv == 2) {
    double a[3] = {W[1][0] - W[0][0], W[1][1] - W[0][1], W[1][2] - W[0][2]};
    return norm3(a);
  }
  if (nv == 3) {
    double a[3], b[3], c[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      a[d] = W[2][d] - W[0][d];
      b[d] = W[1][d] - W[0][d];
    }
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
    return norm3(c);
  }
  double a[3], b[3], c[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a[d] = W[3][d] - W[0][d];
    b[d] = W[2][d] - W[0][d];
    c[d] = W[1][d] - W[0][d];
  }
  return fabs(det3(a, b, c));
}

// moments of one sub-simplex: base rule (NC reference coordinates per
// point) mapped by x = ow W0 + sum_q xi_q W_{q+1}
template <int D, int P, int NC>
__device__ __forceinline__ void accum_piece(const double (*rule)[4], int nq, const double (*W)[3], double wsc,
                                            double *m) {
  // the rule sits in LDS (wave-uniform addresses: broadcast reads); the next
  // point's entry is read before the current point is processed, so the
  // read latency overlaps the FP64 work
  double nx[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) nx[c] = rule[0][c];
  for (int k = 0; k < nq; ++k) {
    double r[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) r[c] = nx[c];
    if (k + 1 < nq) {
#pragma unroll
      for (int c = 0; c < 4; ++c) nx[c] = rule[k + 1][c];
    }
    double ow = 1.0;
#pragma unroll
    for (int q = 0; q < NC; ++q) ow -= r[q];
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * W[0][d];
#pragma unroll
    for (int q = 0; q < NC; ++q)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] = fma(r[q], W[q + 1][d], x[d]);
    double lam[D + 1];
    double sl = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      lam[d] = x[d];
      sl -= x[d];
    }
    lam[D] = sl;
    Moments<D, P>::accum(lam, r[3] * wsc, m);
  }
}

constexpr double inv_factorial(int n) { return n <= 1 ? 1.0 : inv_factorial(n - 1) / n; }

// Exact moments of one sub-simplex (nvs vertices W, reference coordinates)
// from its vertices (Moments::hstep, the complete homogeneous recursion):
// m_alpha += s [t^alpha] h_2P(t . l_0, .., t . l_k), k = nvs - 1, s = wsc /
// (2P + k)! (wsc = k! |T| times the interface factor), l_v the vertices'
// barycentrics; Moments::hfinish applies the alpha! factors once per domain.
// The barycentrics are scaled by s^(1/2P), so the recursion's top degree
// lands in m already weighted.  For rules exact in degree 2P this equals the
// rule's point sum to rounding (P = 2: 4 x 140 FMAs for a tetrahedron
// against 14 x 73 operations).
template <int D, int P>
__device__ __forceinline__ void hpiece(const double (*W)[3], int nvs, double wsc, double *m) {
  using MM = Moments<D, P>;
  static_assert(P <= 2, "the recursion's registers grow with the degree");
  constexpr double f_vol = inv_factorial(2 * P + D), f_if = inv_factorial(2 * P + D - 1);
  const double s = wsc * (nvs == D + 1 ? f_vol : f_if);
  double c = sqrt(s);
  if constexpr (P == 2) c = sqrt(c);
  double h[MM::NH];
#pragma unroll
  for (int k = 0; k < MM::NH; ++k) h[k] = 0.0;
#pragma unroll
  for (int v = 0; v <= D; ++v) {
    if (v >= nvs) break;
    double l[D + 1];
    double sl = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      l[d] = c * W[v][d];
      sl -= W[v][d];
    }
    l[D] = c * sl;
    MM::hstep(l, h, m);
  }
}

// Block-shared staging in LDS: the decomposition table of dimension D (all
// three domain types) and the base rules of dimensions D and D-1 for the
// rule order.  Every thread of the block must call stage_lds (it ends with a
// block barrier).  Layout: [3][16] DecompEntry, then (nq_vol + nq_if) x 4
// doubles.
struct MomLds {
  const DecompEntry *tab;  // [3][16]
  const double (*rule_vol)[4];
  const double (*rule_if)[4];
  int nq_vol, nq_if;
  bool exact;  // the rules are exact for degree 2P: moments are exact integrals
};

// Rule used by the moment form: when the requested order is at least the
// integrand degree 2P, every exact rule gives the same moments (the exact
// integrals of lam^alpha over the piece), so the minimal symmetric rule of
// degree 2P is used (tet 14 instead of 27 points for P = 2, triangle 6
// instead of 9); below 2P the rule of the requested order itself.
inline size_t mom_lds_bytes(int D, int P, int order) {  // host
  size_t nv, ni;
  if (order >= 2 * P) {
    nv = (size_t)xg_minrule_cnt[D][P];
    ni = (size_t)xg_minrule_cnt[D - 1][P];
  } else {
    const int n1 = order / 2 + 1;
    nv = (size_t)xg_rule_cnt[D][n1];
    ni = (size_t)xg_rule_cnt[D - 1][n1];
  }
  return ((sizeof(DecompEntry) * 48 + 15) & ~size_t(15)) + sizeof(double) * 4 * (nv + ni);
}

template <int D, int P>
__device__ __forceinline__ MomLds stage_lds(char *smem, int order) {
  MomLds S;
  DecompEntry *tab = reinterpret_cast<DecompEntry *>(smem);
  double(*rv)[4] = reinterpret_cast<double(*)[4]>(smem + ((sizeof(DecompEntry) * 48 + 15) & ~size_t(15)));
  const bool minimal = order >= 2 * P;
  S.nq_vol = minimal ? xg_minrule_cnt_d[D][P] : rule_cnt(D, order);
  S.nq_if = minimal ? xg_minrule_cnt_d[D - 1][P] : rule_cnt(D - 1, order);
  const double(*src_v)[4] = minimal ? &xg_minrule_data_d[xg_minrule_off_d[D][P]] : &xg_rule_data_d[rule_off(D, order)];
  const double(*src_i)[4] =
      minimal ? &xg_minrule_data_d[xg_minrule_off_d[D - 1][P]] : &xg_rule_data_d[rule_off(D - 1, order)];
  double(*ri)[4] = rv + S.nq_vol;
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int k = tid; k < 48; k += nt) tab[k] = xg_decomp_d.e[D - 1][k >> 4][k & 15];
  for (int k = tid; k < 4 * (S.nq_vol + S.nq_if); k += nt) {
    const int q = k >> 2, c = k & 3;
    if (q < S.nq_vol) rv[q][c] = src_v[q][c];
    else ri[q - S.nq_vol][c] = src_i[q - S.nq_vol][c];
  }
  __syncthreads();
  S.tab = tab;
  S.rule_vol = rv;
  S.rule_if = ri;
  S.exact = minimal;
  return S;
}

// How the moments of one (element, domain type) are obtained: the number of
// sub-simplices whose points are summed (0..3), and whether the result is
// the complement of the other side's moments.
//
// With exact rules (S.exact) the moments are the exact integrals of
// lam^alpha over the domain, so
//   * an uncut element needs no points: its moments are those of the whole
//     simplex (or zero);
//   * a volume part with D vertices of the element (the reference splits it
//     into D sub-simplices, straightcutrule.cpp:131-196) is the element minus
//     the single sub-simplex at the other vertex: one sub-simplex and the
//     exact whole-simplex moments (Moments::complement).  Only used when the
//     subtracted simplex holds at most half of the element, so the
//     subtraction loses at most one bit of relative accuracy.
// Below degree 2P the rule's own points are kept (any rule gives the
// quadrature sum of that rule).
struct Plan {
  int edt;   // table entry used: dt, or the other side when comp
  int ns;    // sub-simplices to sum
  bool cut;  // the element is cut (ns pieces from the table entry)
  bool comp; // moments = whole simplex - those of the ns pieces
  bool full; // uncut, exact: moments of the whole simplex, no points
  bool bad;  // the reference throws (straightcutrule.cpp:45-49, :148, :193)
};

template <int D>
__device__ __forceinline__ Plan make_plan(const MomLds &S, const double *l, int dt) {
  Plan pl;
  pl.edt = dt;
  pl.ns = 0;
  pl.cut = pl.comp = pl.full = pl.bad = false;
  const int eldom = check_if_straight_cut<D>(l);
  if (eldom != IF) {
    if (S.exact) pl.full = (eldom == dt);
    else pl.ns = (eldom == dt) ? 1 : 0;  // the base rule on the element itself
    return pl;
  }
  if (!cut_is_valid<D>(l)) {
    pl.bad = true;
    return pl;
  }
  const int msk = sign_mask<D>(l);
  const DecompEntry &E = S.tab[dt * 16 + msk];
  if (!E.ok) {
    pl.bad = true;
    return pl;
  }
  pl.cut = true;
  pl.ns = E.ns;
  if (S.exact && dt != IF && E.ns > 1) {
    // a valid cut has no zero values, so msk bit i = (l_i > 0); the other
    // side is a single vertex v when exactly one bit differs from dt's side
    const int npos = __builtin_popcount(msk);
    const int nother = (dt == NEG) ? npos : (D + 1 - npos);
    if (nother == 1) {
      int v = 0;
#pragma unroll
      for (int i = 0; i <= D; ++i)
        if ((((msk >> i) & 1) != 0) == (dt == NEG)) v = i;
      double lv = l[0];
#pragma unroll
      for (int i = 1; i <= D; ++i) lv = (i == v) ? l[i] : lv;
      // volume fraction of the simplex at v: prod_w lv / (lv - lw) (each
      // factor in (0, 1))
      double num = 1.0, den = 1.0;
#pragma unroll
      for (int i = 0; i <= D; ++i) {
        if (i == v) continue;
        num *= lv;
        den *= lv - l[i];
      }
      if (fabs(num) <= 0.5 * fabs(den)) {
        pl.comp = true;
        pl.edt = (dt == NEG) ? POS : NEG;
        pl.ns = 1;
      }
    }
  }
  return pl;
}

// What the domain types of one element share: the element domain, the
// sign pattern and the cut points of the sign-changing edges (their order,
// lexicographic in the edge, is the same in every table entry of that
// pattern).
struct ElemCut {
  int eldom;  // CheckIfStraightCut of the element
  int msk;    // sign pattern (bit i: l_i >= 0), valid when eldom == IF
  double tc0, tc1, tc2, tc3;
};

template <int D>
__device__ __forceinline__ ElemCut elem_cut(const MomLds &S, const double *l) {
  ElemCut c;
  c.eldom = check_if_straight_cut<D>(l);
  c.msk = 0;
  c.tc0 = c.tc1 = c.tc2 = c.tc3 = 0.0;
  if (c.eldom != IF || !cut_is_valid<D>(l)) return c;
  c.msk = sign_mask<D>(l);
  PackedEntry E;
  const uint32_t *src = reinterpret_cast<const uint32_t *>(&S.tab[IF * 16 + c.msk]);
#pragma unroll
  for (int q = 0; q < 6; ++q) E.w[q] = src[q];
  const int nc = E.nc();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    if (k >= nc) break;
    const int a = E.ce(k, 0), b = E.ce(k, 1);
    double la = l[0], lb = l[0];
#pragma unroll
    for (int j = 1; j <= D; ++j) {
      la = (j == a) ? l[j] : la;
      lb = (j == b) ? l[j] : lb;
    }
    const double t = la / (la - lb);  // straightcutrule.cpp:50
    c.tc0 = (k == 0) ? t : c.tc0;
    c.tc1 = (k == 1) ? t : c.tc1;
    c.tc2 = (k == 2) ? t : c.tc2;
    c.tc3 = (k == 3) ? t : c.tc3;
  }
  return c;
}

// The degree-2P moments of the cut rule of one (element, domain type) (rule
// topology, sub-simplex volumes, interface factor, point loop), given the
// element's shared cut data.  Returns true where the reference throws
// (straightcutrule.cpp:45-49, :148, :193).
template <int D, int P>
__device__ __forceinline__ bool moment_accumulate_ec(const MomLds &S, const double *l, const ElemCut &ec,
                                                     const double *Finv, int dt, double *m) {
  using MM = Moments<D, P>;
  const Plan pl = make_plan<D>(S, l, dt);
  const int nvs = (dt == IF) ? D : D + 1;  // vertices per sub-simplex (wave-uniform)
  PackedEntry E;
#pragma unroll
  for (int q = 0; q < 6; ++q) E.w[q] = 0u;
  if (pl.cut) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(&S.tab[pl.edt * 16 + ec.msk]);
#pragma unroll
    for (int q = 0; q < 6; ++q) E.w[q] = src[q];
  }
  const int ns = pl.ns;
  const double ifac = (dt == IF && ns > 0) ? interface_factor<D>(l, Finv) : 1.0;
  const int nq = (dt == IF) ? S.nq_if : S.nq_vol;
  const double(*rule)[4] = (dt == IF) ? S.rule_if : S.rule_vol;

#pragma unroll
  for (int k = 0; k < MM::NM; ++k) m[k] = 0.0;
#pragma unroll
  for (int s = 0; s < 3; ++s) {
    if (s >= ns) break;
    double W[D + 1][3];
    double wsc = 1.0;
    if (!pl.cut) {
      // uncut element: origin, e_0, .., e_{D-1} (identity map)
#pragma unroll
      for (int k = 0; k <= D; ++k)
#pragma unroll
        for (int d = 0; d < 3; ++d) W[k][d] = (k > 0 && d == k - 1) ? 1.0 : 0.0;
    } else {
#pragma unroll
      for (int k = 0; k <= D; ++k) {
        if (k < nvs) {
          piece_vertex<D>(E.v(s, k), E, ec.tc0, ec.tc1, ec.tc2, ec.tc3, W[k]);
        } else {
#pragma unroll
          for (int d = 0; d < 3; ++d) W[k][d] = 0.0;
        }
      }
      wsc = piece_volume(nvs, W) * ifac;
    }
    if constexpr (P <= 2) {
      // exact rules: the sub-simplex's exact moments from its vertices
      if (S.exact) {
        hpiece<D, P>(W, nvs, wsc, m);
        continue;
      }
    }
    // one code path for volume and interface pieces (a single instantiation
    // keeps the register allocation of the two from adding up): the rule
    // coordinates past the piece's dimension are exact zeros in the rule
    // table and the unused vertices are zero, so the extra terms add exact
    // zeros
    accum_piece<D, P, D>(rule, nq, W, wsc, m);
  }
  if constexpr (P <= 2) {
    if (S.exact) MM::hfinish(m);
  }
  if (pl.comp || pl.full) MM::complement(m);  // full: whole simplex - 0
  return pl.bad;
}

template <int D, int P>
__device__ __forceinline__ bool moment_accumulate(const MomLds &S, const double *l, const double *Finv, int dt,
                                                  double *m) {
  const ElemCut ec = elem_cut<D>(S, l);
  return moment_accumulate_ec<D, P>(S, l, ec, Finv, dt, m);
}

// the mass / stiffness entries of one (element, domain type) from its
// moments (generated identities); writes the ndof x ndof matrix to out
template <int D, int P, typename T>
__device__ __forceinline__ void moment_forms(const double *m, const double *Finv, double absdet, int mask, T *out) {
  using MM = Moments<D, P>;
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
  MM::template forms<T>(m, mk, Hs, sm, sk, out);
}

// One (element, domain type): rule topology, moments, identities; writes the
// ndof x ndof matrix to out.  Returns true where the reference throws
// (straightcutrule.cpp:45-49, :148, :193).
template <int D, int P, typename T>
__device__ __forceinline__ bool moment_elmat(const MomLds &S, const double *l, const double *Finv, double absdet,
                                             int dt, int mask, T *out) {
  double m[Moments<D, P>::NM];
  const bool bad = moment_accumulate<D, P>(S, l, Finv, dt, m);
  moment_forms<D, P, T>(m, Finv, absdet, mask, out);
  return bad;
}

// element vector of the source term f = 1 (SymbolicCutLinearFormIntegrator,
// xfem/symboliccutlfi.cpp:57-140): out[a] = |det F| sum_i w_i phi_a(x_i)
template <int D, int P, typename T>
__device__ __forceinline__ bool moment_elvec(const MomLds &S, const double *l, const double *Finv, double absdet,
                                             int dt, T *out) {
  using MM = Moments<D, P>;
  double m[MM::NM];
  const bool bad = moment_accumulate<D, P>(S, l, Finv, dt, m);
  MM::template vec<T>(m, absdet, out);
  return bad;
}

}  // namespace mcore
}  // namespace xg
