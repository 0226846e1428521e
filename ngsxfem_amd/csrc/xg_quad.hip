// xg_quad.hip -- straight cut rules on quadrilaterals and hexahedra
// (LevelsetCutQuadrilateral), marking and integration on such meshes.
//
// Reference: cutint/straightcutrule.cpp:74-129 (Quadrilateral volume and
// plain rule), :206-501 (LevelsetCutQuadrilateral), :503-556
// (LevelsetWrapper), :582-677 (StraightCutIntegrationRule, quad branch);
// xfem/cutinfo.cpp:61-133 (marking); cutint/python_cutint.cpp:66-98
// (IntegrateX sums).
//
// The rule of a cut unit quad / hex: the level set is the multilinear
// interpolant of the vertex values.  Along the last coordinate xi the
// element is split where an edge along xi changes sign; each piece that is
// still cut gets a tensor product rule: Gauss points xi* on its xi range
// times the cut rule of the codimension-1 slice at xi* (a cut segment in 2D,
// the 2D quad rule in 3D), weighted for the interface by |grad| / |grad in
// the slice|.  GetDimensionSwap may first transpose the level set so that
// another coordinate plays xi, or give up on the tensor product and cut the
// 2 triangles / 6 tetrahedra of the element instead (the fallback).
//
// Mapping: one lane per element (the number of points varies from element
// to element); a count pass and a fill pass produce a CSR rule like the
// simplex rule (xg_rule.hip).  The recursion depth is fixed (hex -> slice
// quad -> slice segment), so it is written as nested templates with small
// functors that map a slice point back into the element and pass it on.
#include "xg_device.hpp"
#include "xg_internal.hpp"

using namespace xg;

namespace {

constexpr double QC = 0.999;  // straightcutrule.cpp:253

// sub-simplices of the fallback (:478-479) and the edges along xi (:208-209)
__device__ const int kTri[2][3] = {{0, 1, 3}, {2, 1, 3}};
__device__ const int kTet[6][4] = {{3, 0, 1, 5}, {3, 1, 2, 5}, {3, 5, 2, 6}, {4, 5, 0, 3}, {4, 7, 5, 3}, {7, 6, 5, 3}};
__device__ const int kEdge2[2][2] = {{1, 2}, {0, 3}};
__device__ const int kEdge3[4][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}};
constexpr int ERR_DEGEN = 1;  // a simplex cut of the fallback would throw
constexpr int ERR_BOUND = 16;  // "if_scale_factor larger than bound", NaN bound

// LevelsetWrapper coefficients c[i][j][k] as c[4 i + 2 j + k]
struct QL {
  double c[8];
};

// GetCoeffsFromVals, straightcutrule.cpp:503-526
XG_DEV QL ql_quad(const double *v) {
  QL L;
#pragma unroll
  for (int k = 0; k < 8; ++k) L.c[k] = 0.0;
  L.c[0] = v[0];
  L.c[4] = v[1] - L.c[0];
  L.c[2] = v[3] - L.c[0];
  L.c[6] = v[2] - L.c[4] - L.c[2] - L.c[0];
  return L;
}

XG_DEV QL ql_hex(const double *v) {
  QL L;
  L.c[0] = v[0];
  L.c[4] = v[1] - v[0];
  L.c[2] = v[3] - v[0];
  L.c[1] = v[4] - v[0];
  L.c[6] = v[2] - L.c[4] - L.c[2] - L.c[0];
  L.c[5] = v[5] - L.c[4] - L.c[1] - L.c[0];
  L.c[3] = v[7] - L.c[2] - L.c[1] - L.c[0];
  L.c[7] = v[6] - L.c[6] - L.c[5] - L.c[3] - L.c[4] - L.c[2] - L.c[1] - L.c[0];
  return L;
}

// operator(), :542-546: terms in the loop order (i, j, k), each
// c * p0^i * p1^j * p2^k left to right (powers 0 and 1 are exact)
XG_DEV double ql_eval(const QL &L, const double *p) {
  double v = 0.0;
#pragma unroll
  for (int m = 0; m < 8; ++m) {
    double t = L.c[m];
    if (m & 4) t *= p[0];
    if (m & 2) t *= p[1];
    if (m & 1) t *= p[2];
    v += t;
  }
  return v;
}

// GetGrad, :528-534
XG_DEV void ql_grad(const QL &L, const double *p, double *g) {
  const double *c = L.c;
  g[0] = c[4] + c[6] * p[1] + c[5] * p[2] + c[7] * p[1] * p[2];
  g[1] = c[2] + c[6] * p[0] + c[3] * p[2] + c[7] * p[0] * p[2];
  g[2] = c[1] + c[3] * p[1] + c[5] * p[0] + c[7] * p[0] * p[1];
}

// an axis-parallel box [lo, hi] (Quadrilateral(bnds), straightcutrule.hpp:72-90)
struct Box {
  double lo[3], hi[3];
};

// vertex v of the box in the Quadrilateral order
XG_DEV void box_point(const Box &q, int D, int v, double *p) {
  const int b = v & 3;
  p[0] = (b == 1 || b == 2) ? q.hi[0] : q.lo[0];
  p[1] = (b >= 2) ? q.hi[1] : q.lo[1];
  p[2] = (D == 3 && v >= 4) ? q.hi[2] : q.lo[2];
}

XG_DEV Box unit_box() {
  Box q;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    q.lo[d] = 0.0;
    q.hi[d] = 1.0;
  }
  return q;
}

// CheckIfStraightCut with epsilon of the level set at the box's vertices
XG_DEV int box_check(const QL &L, int D, const Box &q, double eps) {
  bool hp = false, hn = false;
  for (int v = 0; v < (D == 2 ? 4 : 8); ++v) {
    double p[3];
    box_point(q, D, v, p);
    const double x = ql_eval(L, p);
    hp |= x > eps;
    hn |= x < -eps;
  }
  if (hp == hn) return IF;
  return hn ? NEG : POS;
}

// Quadrilateral::GetVolume / GetPlainIntegrationRule, :74-78, 102-129: the
// tensor Gauss rule (x slowest) mapped by points[0] + A ip,
// A = (p1 - p0, p3 - p0, p4 - p0)
template <class Sink>
XG_DEV void box_plain(const Box &q, int D, int order, Sink &s) {
  double p0[3], p1[3], p3[3], p4[3] = {0, 0, 0};
  box_point(q, D, 0, p0);
  box_point(q, D, 1, p1);
  box_point(q, D, 3, p3);
  if (D == 3) box_point(q, D, 4, p4);
  double a[3], b[3], tf;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    a[d] = p3[d] - p0[d];
    b[d] = p1[d] - p0[d];
  }
  if (D == 2) {
    double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
    tf = norm3(c);
  } else {
    double e4[3] = {p4[0] - p0[0], p4[1] - p0[1], p4[2] - p0[2]};
    tf = fabs(det3(e4, a, b));
  }
  double A[3][3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    A[d][0] = p1[d] - p0[d];
    A[d][1] = p3[d] - p0[d];
    A[d][2] = D == 3 ? p4[d] - p0[d] : 0.0;
  }
  if (D == 2) A[2][0] = A[2][1] = 0.0;
  const int n = rule_cnt(1, order);
  const double(*r)[4] = &xg_rule_data_d[rule_off(1, order)];
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < (D == 3 ? n : 1); ++k) {
        const double ip[3] = {r[i][0], r[j][0], D == 3 ? r[k][0] : 0.0};
        const double wq = D == 3 ? r[i][3] * r[j][3] * r[k][3] : r[i][3] * r[j][3];
        double x[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) x[d] = p0[d] + (A[d][0] * ip[0] + A[d][1] * ip[1] + A[d][2] * ip[2]);
        s(x, wq * tf);
      }
}

// base rule of dimension dim on the simplex P (SimpleX::GetPlainIntegrationRule)
template <class Sink>
XG_DEV void simplex_plain(int dim, const double (*P)[3], int order, Sink &s) {
  const double vol = simplex_measure(dim, P);
  const int nq = rule_cnt(dim, order);
  const double(*r)[4] = &xg_rule_data_d[rule_off(dim, order)];
  for (int q = 0; q < nq; ++q) {
    double ow = 1.0;
    for (int m = 0; m < dim; ++m) ow -= r[q][m];
    double x[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[d] = ow * P[0][d];
    for (int m = 0; m < dim; ++m)
#pragma unroll
      for (int d = 0; d < 3; ++d) x[d] += r[q][m] * P[m + 1][d];
    s(x, r[q][3] * vol);
  }
}

// LevelsetCutSimplex on SimpleX(ET_SEGM) = {(1,0,0), (0,0,0)} with values
// lv (CalcIFPolytopEUsingLset, :41-64, D == 1; Decompose / rule, :131-204)
template <class Sink>
XG_DEV void segment_rule(const double *lv, int dt, int order, Sink &s, int &err) {
  if (straight_cut(lv, 2) != IF) {
    err |= ERR_DEGEN;  // "You tried to cut a simplex with a plain geometry lset function"
    return;
  }
  const double t = lv[0] / (lv[0] - lv[1]);
  const double C[3] = {1.0 + t * (0.0 - 1.0), 0.0 + t * (0.0 - 0.0), 0.0 + t * (0.0 - 0.0)};
  if (dt == IF) {
    s(C, 1.0);  // ET_POINT rule, volume 1
    return;
  }
  int nr = 0, r0 = 0;
  for (int i = 0; i < 2; ++i)
    if ((dt == POS && lv[i] >= 0.0) || (dt == NEG && lv[i] < 0.0)) {
      r0 = i;
      ++nr;
    }
  if (nr != 1) {
    err |= ERR_DEGEN;  // "Cutting this part of a tetraeder is not implemented yet!"
    return;
  }
  const double P[2][3] = {{C[0], C[1], C[2]}, {r0 == 0 ? 1.0 : 0.0, 0.0, 0.0}};
  simplex_plain(1, P, order, s);
}

// GetFallbackIntegrationRule, :475-491: the 2 triangles / 6 tetrahedra of the
// unit element, each with the clamped values of update_initial_coefs
template <int D, class Sink>
XG_DEV void fallback_rule(const QL &L, int dt, int order, Sink &s, int &err) {
  const Box u = unit_box();
  for (int si = 0; si < (D == 2 ? 2 : 6); ++si) {
    double P[4][3], l[4];
    for (int k = 0; k <= D; ++k) {
      box_point(u, D, D == 2 ? kTri[si][k] : kTet[si][k], P[k]);
      const double v = ql_eval(L, P[k]);
      l[k] = fabs(v) > 1e-14 ? v : 1e-14;
    }
    const int ds = straight_cut(l, D + 1);
    if (ds != IF) {
      if (ds == dt) simplex_plain(D, P, order, s);
      continue;
    }
    double ch[3][4][3];
    bool bad = false;
    const int nch = children(D, P, l, dt, ch, bad);
    if (bad) {
      err |= ERR_DEGEN;
      continue;
    }
    const int dim = dt == IF ? D - 1 : D;
    for (int c = 0; c < nch; ++c) simplex_plain(dim, ch[c], order, s);
  }
}

template <class Sink>
XG_DEV void quad_rule(const QL &L, int dt, int order, int pol, bool consider, Sink &s, int &err);

// maps a point of the slice rule at xi* back into the box and applies the
// slice weights (GetTensorProductAlongXiIntegrationRule, :256-302)
template <int D, class Sink>
struct SliceSink {
  const QL &L;
  const Box &q;
  double xs, w1, len;
  int dt;
  Sink &s;
  int &err;
  XG_DEV void operator()(const double *p2, double w2) {
    double ip[3] = {0.0, 0.0, 0.0};
    ip[D - 1] = xs;
    ip[0] = q.lo[0] + p2[0] * (q.hi[0] - q.lo[0]);
    if (D == 3) ip[1] = q.lo[1] + p2[1] * (q.hi[1] - q.lo[1]);
    double f = 1.0;
    if (dt == IF) {
      double g[3];
      ql_grad(L, ip, g);
      if (D == 2)
        f = norm3(g) / fabs(g[0]);
      else
        f = norm3(g) / sqrt(g[0] * g[0] + g[1] * g[1]);
      if (isnan(f) || f > 1. / sqrt(1 - QC * QC)) err |= ERR_BOUND;
    }
    s(ip, w2 * w1 * len * f);
  }
};

template <int D, class Sink>
XG_DEV void tensor_rule(const QL &L, const Box &q, int dt, int order, int pol, Sink &s, int &err) {
  const double xi0 = q.lo[D - 1], xi1 = q.hi[D - 1];
  const int n = rule_cnt(1, order);
  const double(*r)[4] = &xg_rule_data_d[rule_off(1, order)];
  for (int i1 = 0; i1 < n; ++i1) {
    const double xs = xi0 + r[i1][0] * (xi1 - xi0);
    SliceSink<D, Sink> ss{L, q, xs, r[i1][3], xi1 - xi0, dt, s, err};
    if constexpr (D == 2) {
      const double pa[3] = {q.lo[0], xs, 0.0}, pb[3] = {q.hi[0], xs, 0.0};
      const double lv[2] = {ql_eval(L, pb), ql_eval(L, pa)};
      segment_rule(lv, dt, order, ss, err);
    } else {
      const double p0[3] = {q.lo[0], q.lo[1], xs}, p1[3] = {q.hi[0], q.lo[1], xs};
      const double p2[3] = {q.hi[0], q.hi[1], xs}, p3[3] = {q.lo[0], q.hi[1], xs};
      const double lv[4] = {ql_eval(L, p0), ql_eval(L, p1), ql_eval(L, p2), ql_eval(L, p3)};
      quad_rule(ql_quad(lv), dt, order, pol, true, ss, err);
    }
  }
}

// HasTopologyChangeAlongXi / Decompose / GetIntegrationRuleAlongXi,
// :206-251, 459-473, on the unit element
template <int D, class Sink>
XG_DEV void along_xi(const QL &L, int dt, int order, int pol, Sink &s, int &err) {
  const Box u = unit_box();
  const int dq = box_check(L, D, u, 0.0);
  if (dq != IF) {
    if (dq == dt) box_plain(u, D, order, s);
    return;
  }
  double xs[6] = {0.0, 1.0, 0.0, 0.0, 0.0, 0.0};
  int nx = 2;
  bool topo = false;
  for (int k = 0; k < (D == 2 ? 2 : 4); ++k) {
    const int a = D == 2 ? kEdge2[k][0] : kEdge3[k][0], b = D == 2 ? kEdge2[k][1] : kEdge3[k][1];
    double pa[3], pb[3];
    box_point(u, D, a, pa);
    box_point(u, D, b, pb);
    const double vals[2] = {ql_eval(L, pb), ql_eval(L, pa)};
    if (straight_cut(vals, 2) != IF) continue;
    topo = true;
    const double t = vals[0] / (vals[0] - vals[1]);
    const double x = 1.0 + t * (0.0 - 1.0);  // the cut point of the unit line from (1,0,0) to 0
    if (isnan(x)) continue;                   // never inserted into the reference's std::set
    bool dup = false;
    for (int m = 0; m < nx; ++m) dup |= xs[m] == x;
    if (!dup) xs[nx++] = x;
  }
  if (!topo) {
    tensor_rule<D>(L, u, dt, order, pol, s, err);
    return;
  }
  for (int i = 1; i < nx; ++i)  // ascending, as the std::set
    for (int j = i; j > 0 && xs[j] < xs[j - 1]; --j) {
      const double t = xs[j];
      xs[j] = xs[j - 1];
      xs[j - 1] = t;
    }
  for (int i = 0; i + 1 < nx; ++i) {
    Box sq = u;
    sq.lo[D - 1] = xs[i];
    sq.hi[D - 1] = xs[i + 1];
    const int ds = box_check(L, D, sq, 1e-15);
    if (ds == IF)
      tensor_rule<D>(L, sq, dt, order, pol, s, err);
    else if (ds == dt)
      box_plain(sq, D, order, s);
  }
}

enum { SW_ID, SW_XY, SW_XZ, SW_YZ, SW_NONE };

// GetExactCritsQBound2D / GetSufficientCritsQBound / GetDimensionSwap,
// :353-457 (pol: 0 FIRST_ALLOWED, 1 FIND_OPTIMAL, 2 ALWAYS_NONE)
XG_DEV int dimension_swap(const QL &L, int D, int pol, bool consider, int &err) {
  if (pol == 2) return SW_NONE;
  if (!consider) return SW_ID;
  const double *c = L.c;
  if (D == 2) {
    bool allow0 = true, allow1 = true;
    double h = -c[4] / c[6];
    if ((h > 0) && (h < 1)) allow1 = false;
    h = -c[2] / c[6];
    if ((h > 0) && (h < 1)) allow0 = false;
    double q0 = 0.0, q1 = 0.0;
    for (int k = 0; k < 4; ++k) {
      // corners (0,0), (1,0), (1,1), (0,1)
      const double cs[3] = {(double)(k == 1 || k == 2), (double)(k >= 2), 0.0};
      double g[3];
      ql_grad(L, cs, g);
      const double qy = fabs(g[1]) / norm3(g), qx = fabs(g[0]) / norm3(g);
      if (qy > q1) q1 = qy;
      if (qx > q0) q0 = qx;
    }
    if (!allow0) q0 = 2;
    if (!allow1) q1 = 2;
    if (pol == 0) return q1 < QC ? SW_ID : (q0 < QC ? SW_XY : SW_NONE);
    if ((q0 < QC) && (q1 < QC)) return q1 <= q0 ? SW_ID : SW_XY;
    if (q1 < QC) return SW_ID;
    if (q0 < QC) return SW_XY;
    return SW_NONE;
  }
  double g[8][3];
  for (int k = 0; k < 8; ++k) {
    const double p[3] = {(double)(k & 1), (double)((k >> 1) & 1), (double)((k >> 2) & 1)};
    ql_grad(L, p, g[k]);
  }
  double Vsq = 0.0;
  for (int d = 0; d < 3; ++d) {
    double mx = 0.0;
    for (int k = 0; k < 8; ++k) {
      const double v = g[k][d] * g[k][d];
      if (v > mx) mx = v;
    }
    Vsq += mx;
  }
  const double V = sqrt(Vsq);
  double qm[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) {
      const double qe = V * V / (V * V - g[k][d] * g[k][d]);
      if (qe > qm[d]) qm[d] = qe;
    }
  for (int d = 0; d < 3; ++d) qm[d] = sqrt(1 - 1 / qm[d]);
  if (isnan(qm[0]) || isnan(qm[1]) || isnan(qm[2])) {
    err |= ERR_BOUND;  // "Sufficient Criterion calculated nan Bound!"
    return SW_NONE;
  }
  if (pol == 0) return qm[2] < QC ? SW_ID : (qm[1] < QC ? SW_YZ : (qm[0] < QC ? SW_XZ : SW_NONE));
  int md = 0;  // min_element with the comparator <=: the last of equal minima
  for (int d = 1; d < 3; ++d)
    if (qm[d] <= qm[md]) md = d;
  if (qm[md] < QC) return md == 0 ? SW_XZ : (md == 1 ? SW_YZ : SW_ID);
  return SW_NONE;
}

// swaps the coordinates of the rotated element's points back (:304-351)
template <class Sink>
struct SwapSink {
  int sw;
  Sink &s;
  XG_DEV void operator()(const double *p, double w) {
    double x[3];
    if (sw == SW_XY) {
      x[0] = p[1];
      x[1] = p[0];
      x[2] = p[2];
    } else if (sw == SW_XZ) {
      x[0] = p[2];
      x[1] = p[1];
      x[2] = p[0];
    } else {
      x[0] = p[0];
      x[1] = p[2];
      x[2] = p[1];
    }
    s(x, w);
  }
};

XG_DEV QL transposed(const QL &L, int sw) {
  QL R;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        const int src = sw == SW_XY ? (j * 4 + i * 2 + k) : (sw == SW_XZ ? (k * 4 + j * 2 + i) : (i * 4 + k * 2 + j));
        R.c[i * 4 + j * 2 + k] = L.c[src];
      }
  return R;
}

// LevelsetCutQuadrilateral::GetIntegrationRule, :493-501, of a unit quad
template <class Sink>
XG_DEV void quad_rule(const QL &L, int dt, int order, int pol, bool consider, Sink &s, int &err) {
  const int sw = dimension_swap(L, 2, pol, consider, err);
  if (sw == SW_ID) {
    along_xi<2>(L, dt, order, pol, s, err);
  } else if (sw == SW_NONE) {
    fallback_rule<2>(L, dt, order, s, err);
  } else {
    SwapSink<Sink> ws{sw, s};
    along_xi<2>(transposed(L, sw), dt, order, pol, ws, err);
  }
}

// the same for a unit hex (X_Z, Y_Z: the rotated hex's rule without
// further swaps, i.e. along xi)
template <class Sink>
XG_DEV void hex_rule(const QL &L, int dt, int order, int pol, Sink &s, int &err) {
  const int sw = dimension_swap(L, 3, pol, true, err);
  if (err & ERR_BOUND) return;
  if (sw == SW_ID) {
    along_xi<3>(L, dt, order, pol, s, err);
  } else if (sw == SW_NONE) {
    fallback_rule<3>(L, dt, order, s, err);
  } else {
    SwapSink<Sink> ws{sw, s};
    along_xi<3>(transposed(L, sw), dt, order, pol, ws, err);
  }
}

// ---- element level

constexpr int MAXV = 8;

template <int D>
XG_DEV void load_values(const int32_t *el2v, const double *lset, int e, double *v) {
  constexpr int NV = 1 << D;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = lset[el2v[(size_t)e * NV + k]];
}

// multilinear element map: F = dx/dxi at xi (row-major), x(xi)
template <int D>
XG_DEV void elem_map(const double (*V)[3], const double *xi, double *F, double *x) {
  constexpr int NV = 1 << D;
#pragma unroll
  for (int k = 0; k < D * D; ++k) F[k] = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) x[d] = 0.0;
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const int b = v & 3;
    const int a[3] = {(b == 1 || b == 2), b >= 2, v >= 4};
    double N = 1.0;
#pragma unroll
    for (int m = 0; m < D; ++m) N *= a[m] ? xi[m] : 1.0 - xi[m];
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] += N * V[v][d];
#pragma unroll
    for (int c = 0; c < D; ++c) {
      double t = 1.0;
#pragma unroll
      for (int m = 0; m < D; ++m) {
        const double f = a[m] ? xi[m] : 1.0 - xi[m];
        const double df = a[m] ? 1.0 : -1.0;
        t *= (m == c) ? df : f;
      }
#pragma unroll
      for (int r = 0; r < D; ++r) F[r * D + c] += V[v][r] * t;
    }
  }
}

template <int D>
XG_DEV double det_inv(const double *F, double *Finv) {
  if (D == 2) {
    const double det = F[0] * F[3] - F[1] * F[2];
    Finv[0] = F[3] / det;
    Finv[1] = -F[1] / det;
    Finv[2] = -F[2] / det;
    Finv[3] = F[0] / det;
    return det;
  }
  const double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                     F[2] * (F[3] * F[7] - F[4] * F[6]);
  Finv[0] = (F[4] * F[8] - F[5] * F[7]) / det;
  Finv[1] = (F[2] * F[7] - F[1] * F[8]) / det;
  Finv[2] = (F[1] * F[5] - F[2] * F[4]) / det;
  Finv[3] = (F[5] * F[6] - F[3] * F[8]) / det;
  Finv[4] = (F[0] * F[8] - F[2] * F[6]) / det;
  Finv[5] = (F[2] * F[3] - F[0] * F[5]) / det;
  Finv[6] = (F[3] * F[7] - F[4] * F[6]) / det;
  Finv[7] = (F[1] * F[6] - F[0] * F[7]) / det;
  Finv[8] = (F[0] * F[4] - F[1] * F[3]) / det;
  return det;
}

// counts or writes the element's points; for dt == IF the weight gets
// |F^{-T} n_ref| at the point (TransformQuadUntrafoToIRInterface, :558-578)
template <int D>
struct ElemSink {
  const QL &L;
  const double (*V)[3];
  double4 *out;  // null: count only
  int dt;
  int n;
  double wsum;
  XG_DEV void operator()(const double *p, double w) {
    if (out) {
      if (dt == IF) {
        double g[3];
        ql_grad(L, p, g);
        const double gl = norm3(g);
#pragma unroll
        for (int d = 0; d < 3; ++d) g[d] /= gl;
        double F[D * D], Finv[D * D], x[3];
        elem_map<D>(V, p, F, x);
        det_inv<D>(F, Finv);
        double nn[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int r = 0; r < D; ++r) {
          double acc = 0.0;
#pragma unroll
          for (int c = 0; c < D; ++c) acc += Finv[c * D + r] * g[c];
          nn[r] = acc;
        }
        w = w * norm3(nn);
      }
      out[n] = make_double4(p[0], p[1], p[2], w);
    }
    wsum += w;
    ++n;
  }
};

// StraightCutIntegrationRule (:582-677) of one element on the unit element
template <int D, class Sink>
XG_DEV void element_rule(const double *vals, int dt, int order, int pol, Sink &s, int &err) {
  constexpr int NV = 1 << D;
  const int eldom = straight_cut(vals, NV);
  if (eldom != IF) {
    if (eldom == dt) box_plain(unit_box(), D, order, s);
    return;
  }
  const QL L = D == 2 ? ql_quad(vals) : ql_hex(vals);
  if constexpr (D == 2)
    quad_rule(L, dt, order, pol, true, s, err);
  else
    hex_rule(L, dt, order, pol, s, err);
}

template <int D, bool FILL>
__global__ __launch_bounds__(256) void k_quad_rule(int nel, const int32_t *__restrict__ elems,
                                                   const int32_t *__restrict__ el2v,
                                                   const double *__restrict__ coords,
                                                   const double *__restrict__ lset, int dt, int order, int pol,
                                                   int32_t *__restrict__ offsets, double4 *__restrict__ pts,
                                                   int32_t *err_flag) {
  constexpr int NV = 1 << D;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int e = elems ? elems[i] : i;
  double vals[NV];
  load_values<D>(el2v, lset, e, vals);
  double V[NV][3];
  if (FILL) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int v = el2v[(size_t)e * NV + k];
#pragma unroll
      for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
    }
  }
  const QL L = D == 2 ? ql_quad(vals) : ql_hex(vals);
  ElemSink<D> s{L, V, FILL ? pts + offsets[i] : nullptr, dt, 0, 0.0};
  int err = 0;
  element_rule<D>(vals, dt, order, pol, s, err);
  if (err) atomicOr(err_flag, err);
  if (!FILL) offsets[i] = s.n;
}

// CutInformation::Update on quads / hexes: order-0 part volumes
// (cutinfo.cpp:87-124)
template <int D>
__global__ __launch_bounds__(256) void k_quad_mark(int ne, const int32_t *__restrict__ el2v,
                                                   const double *__restrict__ lset, int pol,
                                                   uint8_t *__restrict__ dt_out, double *__restrict__ cut_ratio,
                                                   int32_t *err_flag) {
  constexpr int NV = 1 << D;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  double vals[NV];
  load_values<D>(el2v, lset, e, vals);
  const QL L = D == 2 ? ql_quad(vals) : ql_hex(vals);
  double pv[2] = {0.0, 0.0};
  int err = 0;
  for (int t = 0; t < 2; ++t) {
    const int np = t == 0 ? POS : NEG;
    ElemSink<D> s{L, nullptr, nullptr, np, 0, 0.0};
    element_rule<D>(vals, np, 0, pol, s, err);
    pv[np] = s.wsum;
  }
  if (err) atomicOr(err_flag, err);
  if (cut_ratio) cut_ratio[e] = pv[NEG] / (pv[NEG] + pv[POS]);
  const int c = pv[NEG] > 0.0 ? (pv[POS] > 0.0 ? IF : NEG) : POS;
  dt_out[e] = (uint8_t)c;
}

// IntegrateX sums on quads / hexes: one wave per element, lanes over the
// element's points, sum_q |det F(xi_q)| w_q f(x(xi_q))
template <int D>
__global__ __launch_bounds__(256) void k_quad_integrate(int nel, const int32_t *__restrict__ elems,
                                                        const int32_t *__restrict__ el2v,
                                                        const double *__restrict__ coords,
                                                        const int32_t *__restrict__ offsets,
                                                        const double4 *__restrict__ pts, Poly f,
                                                        double *__restrict__ per_elem, double *total) {
  constexpr int NV = 1 << D;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  double acc_all = 0.0;
  for (int i = wid; i < nel; i += nwaves) {
    const int e = elems ? elems[i] : i;
    double V[NV][3];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int v = el2v[(size_t)e * NV + k];
#pragma unroll
      for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
    }
    double s = 0.0;
    for (int q = offsets[i] + lane; q < offsets[i + 1]; q += 64) {
      const double4 p = pts[q];
      const double xi[3] = {p.x, p.y, p.z};
      double F[D * D], x[3];
      elem_map<D>(V, xi, F, x);
      double ad;
      if (D == 2)
        ad = fabs(F[0] * F[3] - F[1] * F[2]);
      else {
        const double c1[3] = {F[0], F[3], F[6]}, c2[3] = {F[1], F[4], F[7]}, c3[3] = {F[2], F[5], F[8]};
        ad = fabs(det3(c1, c2, c3));
      }
      double fv = 0.0;
      for (int t = 0; t < f.n; ++t) {
        double m = f.c[t];
        for (int d = 0; d < 3; ++d)
          for (int k = 0; k < f.ex[t][d]; ++k) m *= x[d];
        fv += m;
      }
      s += ad * p.w * fv;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (lane == 0 && per_elem) per_elem[i] = s;
    acc_all += s;
  }
  if (lane == 0 && total && acc_all != 0.0) atomicAdd(total, acc_all);
}

// one wave per element, at most 65536 blocks of 4 waves
int waves_grid(int nel) {
  long long blocks = ((long long)nel * 64 + 255) / 256;
  return (int)(blocks > 65536 ? 65536 : (blocks < 1 ? 1 : blocks));
}

// SymbolicCutBilinearFormIntegrator::CalcElementMatrix (xfem/symboliccutbfi.cpp:
// 76-304) for the Q1 space of quads / hexes on a materialised cut rule: one
// lane per element, its points in turn; out = sum_q |det F| w_q (sm phi_a
// phi_b + sk grad phi_a . grad phi_b), grad phi = F^{-T} grad_ref phi, the
// upper triangle accumulated in FP64 and mirrored (row = test dof)
template <int D, typename T>
__global__ __launch_bounds__(256) void k_quad_elmat(int nel, const int32_t *__restrict__ elems,
                                                    const int32_t *__restrict__ el2v,
                                                    const double *__restrict__ coords,
                                                    const int32_t *__restrict__ offsets,
                                                    const double4 *__restrict__ pts, int mask, T *__restrict__ out) {
  constexpr int NV = 1 << D;
  constexpr int NU = NV * (NV + 1) / 2;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int e = elems ? elems[i] : i;
  double V[NV][3];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int v = el2v[(size_t)e * NV + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
  }
  const double sm = (mask & 1) ? 1.0 : 0.0, sk = (mask & 2) ? 1.0 : 0.0;
  double A[NU];
#pragma unroll
  for (int k = 0; k < NU; ++k) A[k] = 0.0;
  for (int q = offsets[i]; q < offsets[i + 1]; ++q) {
    const double4 p = pts[q];
    const double xi[3] = {p.x, p.y, p.z};
    double F[D * D], Finv[D * D], x[3];
    elem_map<D>(V, xi, F, x);
    const double det = det_inv<D>(F, Finv);
    const double wq = fabs(det) * p.w;
    double phi[NV], g[NV][D];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int b = v & 3;
      const int a[3] = {(b == 1 || b == 2), b >= 2, v >= 4};
      double N = 1.0;
#pragma unroll
      for (int m = 0; m < D; ++m) N *= a[m] ? xi[m] : 1.0 - xi[m];
      phi[v] = N;
      double gr[D];
#pragma unroll
      for (int c = 0; c < D; ++c) {
        double t = 1.0;
#pragma unroll
        for (int m = 0; m < D; ++m) t *= (m == c) ? (a[m] ? 1.0 : -1.0) : (a[m] ? xi[m] : 1.0 - xi[m]);
        gr[c] = t;
      }
      // physical gradient: (F^{-T} gr)_r = sum_c Finv[c][r] gr[c]
#pragma unroll
      for (int r = 0; r < D; ++r) {
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) acc += Finv[c * D + r] * gr[c];
        g[v][r] = acc;
      }
    }
    int u = 0;
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
      for (int b = a; b < NV; ++b) {
        double gg = 0.0;
#pragma unroll
        for (int r = 0; r < D; ++r) gg += g[a][r] * g[b][r];
        A[u++] += wq * (sm * (phi[a] * phi[b]) + sk * gg);
      }
  }
  T *o = out + (size_t)i * NV * NV;
  int u = 0;
#pragma unroll
  for (int a = 0; a < NV; ++a)
#pragma unroll
    for (int b = a; b < NV; ++b) {
      o[a * NV + b] = (T)A[u];
      o[b * NV + a] = (T)A[u];
      ++u;
    }
}

// SymbolicCutLinearFormIntegrator (xfem/symboliccutlfi.cpp:57-140) with f = 1
// for the Q1 space on a materialised quad / hex rule: out_a = sum_q
// |det F| w_q phi_a(xi_q), one lane per element
template <int D>
__global__ __launch_bounds__(256) void k_quad_elvec(int nel, const int32_t *__restrict__ elems,
                                                    const int32_t *__restrict__ el2v,
                                                    const double *__restrict__ coords,
                                                    const int32_t *__restrict__ offsets,
                                                    const double4 *__restrict__ pts, double *__restrict__ out) {
  constexpr int NV = 1 << D;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nel) return;
  const int e = elems ? elems[i] : i;
  double V[NV][3];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    const int v = el2v[(size_t)e * NV + k];
#pragma unroll
    for (int d = 0; d < 3; ++d) V[k][d] = d < D ? coords[(size_t)v * D + d] : 0.0;
  }
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; ++v) acc[v] = 0.0;
  for (int q = offsets[i]; q < offsets[i + 1]; ++q) {
    const double4 p = pts[q];
    const double xi[3] = {p.x, p.y, p.z};
    double F[D * D], x[3];
    elem_map<D>(V, xi, F, x);
    double Finv[D * D];
    const double det = det_inv<D>(F, Finv);
    const double wq = fabs(det) * p.w;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int b = v & 3;
      const int a[3] = {(b == 1 || b == 2), b >= 2, v >= 4};
      double N = 1.0;
#pragma unroll
      for (int m = 0; m < D; ++m) N *= a[m] ? xi[m] : 1.0 - xi[m];
      acc[v] += wq * N;
    }
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) out[(size_t)i * NV + v] = acc[v];
}

bool quad_args_ok(int D, int dt, int order, int pol) {
  return (D == 2 || D == 3) && dt >= 0 && dt <= 2 && order >= 0 && order / 2 + 1 <= XG_RULE_NMAX && pol >= 0 &&
         pol <= 2;
}

}  // namespace

extern "C" int xg_quad_rule_count(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                  const double *lset_v, int dt, int order, int policy, int32_t *offsets) {
  if (!ctx || nel < 0 || !offsets || !quad_args_ok(D, dt, order, policy))
    return set_error(XG_ERR_ARG, "xg_quad_rule_count: bad argument");
  if (nel > 0 && (!el2v || !lset_v)) return set_error(XG_ERR_ARG, "xg_quad_rule_count: null mesh");
  if (nel == 0) {
    XG_CHECK(hipMemsetAsync(offsets, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int grid = (nel + 255) / 256;
  if (D == 2)
    k_quad_rule<2, false><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, nullptr, lset_v, dt, order, policy,
                                                         offsets, nullptr, ctx->d_err);
  else
    k_quad_rule<3, false><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, nullptr, lset_v, dt, order, policy,
                                                         offsets, nullptr, ctx->d_err);
  XG_LAUNCH_CHECK();
  return scan_exclusive_i32(ctx, offsets, nel, true);
}

extern "C" int xg_quad_rule_fill(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                 const double *coords, const double *lset_v, int dt, int order, int policy,
                                 const int32_t *offsets, double *pts) {
  if (!ctx || nel < 0 || !offsets || !quad_args_ok(D, dt, order, policy))
    return set_error(XG_ERR_ARG, "xg_quad_rule_fill: bad argument");
  if (nel == 0) return XG_OK;
  if (!el2v || !lset_v || !coords || !pts) return set_error(XG_ERR_ARG, "xg_quad_rule_fill: null pointer");
  const int grid = (nel + 255) / 256;
  double4 *o = reinterpret_cast<double4 *>(pts);
  if (D == 2)
    k_quad_rule<2, true><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, lset_v, dt, order, policy,
                                                        const_cast<int32_t *>(offsets), o, ctx->d_err);
  else
    k_quad_rule<3, true><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, lset_v, dt, order, policy,
                                                        const_cast<int32_t *>(offsets), o, ctx->d_err);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_quad_mark(xg_ctx *ctx, int D, int ne, const int32_t *el2v, const double *lset_v, int policy,
                            uint8_t *dt, double *cut_ratio, uint32_t *bits, int32_t *cut_list, int32_t *n_cut) {
  if (!ctx || (D != 2 && D != 3) || ne < 0 || !dt || !bits || policy < 0 || policy > 2 ||
      (ne > 0 && (!el2v || !lset_v)))
    return set_error(XG_ERR_ARG, "xg_quad_mark: bad argument (dt and bits required)");
  if (ne == 0) {
    if (n_cut) XG_CHECK(hipMemsetAsync(n_cut, 0, sizeof(int32_t), ctx->stream));
    return XG_OK;
  }
  const int grid = (ne + 255) / 256;
  if (D == 2)
    k_quad_mark<2><<<grid, 256, 0, ctx->stream>>>(ne, el2v, lset_v, policy, dt, cut_ratio, ctx->d_err);
  else
    k_quad_mark<3><<<grid, 256, 0, ctx->stream>>>(ne, el2v, lset_v, policy, dt, cut_ratio, ctx->d_err);
  XG_LAUNCH_CHECK();
  return bits_and_cut_list_from_dt(ctx, ne, dt, bits, cut_list, n_cut);
}

extern "C" int xg_quad_integrate(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                                 const double *coords, const int32_t *offsets, const double *pts, int nterms,
                                 const double *coef, const int32_t *expo, double *per_elem, double *total) {
  if (!ctx || (D != 2 && D != 3) || nel < 0 || nterms < 0 || nterms > 32 || (nterms > 0 && (!coef || !expo)))
    return set_error(XG_ERR_ARG, "xg_quad_integrate: bad argument");
  if (total) XG_CHECK(hipMemsetAsync(total, 0, sizeof(double), ctx->stream));
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts) return set_error(XG_ERR_ARG, "xg_quad_integrate: null pointer");
  Poly f;
  f.n = nterms;
  for (int t = 0; t < nterms; ++t) {
    f.c[t] = coef[t];
    for (int d = 0; d < 3; ++d) {
      if (expo[3 * t + d] < 0 || expo[3 * t + d] > 64)
        return set_error(XG_ERR_ARG, "xg_quad_integrate: bad exponent");
      f.ex[t][d] = expo[3 * t + d];
    }
  }
  const int grid = waves_grid(nel);
  const double4 *p = reinterpret_cast<const double4 *>(pts);
  if (D == 2)
    k_quad_integrate<2><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, f, per_elem, total);
  else
    k_quad_integrate<3><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, f, per_elem, total);
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_quad_elmat(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                             const double *coords, const int32_t *offsets, const double *pts, int mask,
                             int precision, void *out) {
  if (!ctx || (D != 2 && D != 3) || nel < 0 || mask < 1 || mask > 3 ||
      (precision != XG_PREC_FP32 && precision != XG_PREC_FP64))
    return set_error(XG_ERR_ARG, "xg_quad_elmat: bad argument");
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts || !out) return set_error(XG_ERR_ARG, "xg_quad_elmat: null pointer");
  const int grid = (nel + 255) / 256;
  const double4 *p = reinterpret_cast<const double4 *>(pts);
  if (precision == XG_PREC_FP64) {
    if (D == 2) k_quad_elmat<2, double><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, mask, (double *)out);
    else k_quad_elmat<3, double><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, mask, (double *)out);
  } else {
    if (D == 2) k_quad_elmat<2, float><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, mask, (float *)out);
    else k_quad_elmat<3, float><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, mask, (float *)out);
  }
  XG_LAUNCH_CHECK();
  return XG_OK;
}

extern "C" int xg_quad_elvec(xg_ctx *ctx, int D, int nel, const int32_t *elems, const int32_t *el2v,
                             const double *coords, const int32_t *offsets, const double *pts, double *out) {
  if (!ctx || (D != 2 && D != 3) || nel < 0) return set_error(XG_ERR_ARG, "xg_quad_elvec: bad argument");
  if (nel == 0) return XG_OK;
  if (!el2v || !coords || !offsets || !pts || !out) return set_error(XG_ERR_ARG, "xg_quad_elvec: null pointer");
  const int grid = (nel + 255) / 256;
  const double4 *p = reinterpret_cast<const double4 *>(pts);
  if (D == 2) k_quad_elvec<2><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, out);
  else k_quad_elvec<3><<<grid, 256, 0, ctx->stream>>>(nel, elems, el2v, coords, offsets, p, out);
  XG_LAUNCH_CHECK();
  return XG_OK;
}
