/*
 * xg_oracle.c -- CPU FP64 restatement of ngsxfem's cut-element hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP product path).  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  Build: oracle/Makefile (gcc -O3 -ffp-contract=off).
 *
 * Reference files restated (paths relative to the ngsxfem source tree):
 *   cutint/straightcutrule.cpp:11-204, 503-677   straight cut rule
 *   cutint/straightcutrule.cpp:74-129, 206-501    quadrilaterals / hexahedra
 *   xfem/cutinfo.cpp:61-133                       element marking
 *   xfem/symboliccutbfi.cpp:76-304                element matrix B^T D B
 *   cutint/python_cutint.cpp:66-98                IntegrateX sums
 *   cutint/xintegration.cpp:217-326, 617-1063, 1084-1388, 1626-1754
 *   cutint/xdecompose.hpp:34-146, xdecompose.cpp:16-53   subdivision rule
 */
#include "xg_oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

/* ====================================================================== */
/* Gauss-Jacobi rules on [0,1] with weight (1-t)^a, a in {0,1,2}.          */
/* Nodes: eigenvalues of the Jacobi matrix (cyclic Jacobi rotations), then  */
/* Newton-polished on the orthonormal three-term recurrence; weights by the */
/* Christoffel formula w = 1 / sum_k p_k(x)^2.                              */
/* ====================================================================== */

#define MAXN 16

static void jacobi_coeffs(int n, double alpha, double *d, double *e /* e[k], k=1..n */)
{
  const double beta = 0.0;
  for (int k = 0; k < n; ++k) {
    if (k == 0)
      d[k] = (beta - alpha) / (alpha + beta + 2.0);
    else {
      double s = 2.0 * k + alpha + beta;
      d[k] = (beta * beta - alpha * alpha) / (s * (s + 2.0));
    }
  }
  for (int k = 1; k <= n; ++k) {
    double s = 2.0 * k + alpha + beta;
    double num = 4.0 * k * (k + alpha) * (k + beta) * (k + alpha + beta);
    double den = s * s * (s + 1.0) * (s - 1.0);
    e[k] = sqrt(num / den);
  }
}

/* symmetric eigenvalues by cyclic Jacobi rotations (n <= MAXN) */
static void sym_eigvals(int n, double A[MAXN][MAXN], double *ev)
{
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) off += A[p][q] * A[p][q];
    if (off < 1e-300) break;
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        if (fabs(A[p][q]) < 1e-300) continue;
        double theta = (A[q][q] - A[p][p]) / (2.0 * A[p][q]);
        double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k][p], akq = A[k][q];
          A[k][p] = c * akp - s * akq;
          A[k][q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p][k], aqk = A[q][k];
          A[p][k] = c * apk - s * aqk;
          A[q][k] = s * apk + c * aqk;
        }
      }
  }
  for (int i = 0; i < n; ++i) ev[i] = A[i][i];
}

/* orthonormal recurrence: returns p_n-proportional value q and derivative,
   and sum of squares of p_0..p_{n-1} */
static void ortho_eval(int n, const double *d, const double *e, double mu0, double x,
                       double *q, double *dq, double *sumsq)
{
  double pm1 = 0.0, p = 1.0 / sqrt(mu0), dpm1 = 0.0, dp = 0.0, ss = 0.0;
  for (int k = 0; k < n; ++k) {
    ss += p * p;
    double pn = (x - d[k]) * p - (k > 0 ? e[k] * pm1 : 0.0);
    double dpn = p + (x - d[k]) * dp - (k > 0 ? e[k] * dpm1 : 0.0);
    if (k < n - 1) { pn /= e[k + 1]; dpn /= e[k + 1]; }
    pm1 = p; p = pn; dpm1 = dp; dp = dpn;
  }
  *q = p; *dq = dp; *sumsq = ss;
}

/* n-point rule on [0,1] for weight (1-t)^a; t ascending */
static void gauss_jacobi01(int n, int a, double *t, double *w)
{
  if (n == 1) { /* closed form, exact in FP64 for a = 0,1,2 */
    t[0] = 1.0 / (a + 2.0);
    w[0] = 1.0 / (a + 1.0);
    return;
  }
  double d[MAXN + 1], e[MAXN + 2];
  jacobi_coeffs(n, (double)a, d, e);
  double A[MAXN][MAXN];
  memset(A, 0, sizeof(A));
  for (int k = 0; k < n; ++k) {
    A[k][k] = d[k];
    if (k + 1 < n) A[k][k + 1] = A[k + 1][k] = e[k + 1];
  }
  double x[MAXN];
  sym_eigvals(n, A, x);
  /* sort ascending */
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j)
      if (x[j] < x[i]) { double tmp = x[i]; x[i] = x[j]; x[j] = tmp; }
  double mu0 = pow(2.0, a + 1.0) / (a + 1.0);
  for (int i = 0; i < n; ++i) {
    double xi = x[i], q, dq, ss;
    for (int it = 0; it < 6; ++it) {
      ortho_eval(n, d, e, mu0, xi, &q, &dq, &ss);
      if (dq == 0.0) break;
      double step = q / dq;
      xi -= step;
      if (fabs(step) < 1e-17) break;
    }
    ortho_eval(n, d, e, mu0, xi, &q, &dq, &ss);
    t[i] = (1.0 + xi) / 2.0;
    w[i] = (1.0 / ss) / pow(2.0, a + 1.0);
  }
}

int or_rule_npts(int D, int order)
{
  int n = order / 2 + 1;
  int r = 1;
  for (int i = 0; i < D; ++i) r *= n;
  return r;
}

/* Stroud conical product (collapsed coordinates):
   segm: xi0 = t
   trig: xi0 = t1, xi1 = (1-t1) t2                 t1 ~ (1-t)^1
   tet : xi0 = t1, xi1 = (1-t1) t2, xi2 = (1-t1)(1-t2) t3, t1 ~ (1-t)^2 */
int or_rule(int D, int order, double *xi, double *w)
{
  int n = order / 2 + 1;
  if (n > MAXN) return -1;
  if (D == 0) { w[0] = 1.0; return 1; }
  double t0[MAXN], w0[MAXN], t1[MAXN], w1[MAXN], t2[MAXN], w2[MAXN];
  int cnt = 0;
  if (D == 1) {
    gauss_jacobi01(n, 0, t0, w0);
    for (int i = 0; i < n; ++i) { xi[cnt] = t0[i]; w[cnt] = w0[i]; ++cnt; }
  } else if (D == 2) {
    gauss_jacobi01(n, 1, t0, w0);
    gauss_jacobi01(n, 0, t1, w1);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) {
        xi[2 * cnt + 0] = t0[i];
        xi[2 * cnt + 1] = (1.0 - t0[i]) * t1[j];
        w[cnt] = w0[i] * w1[j];
        ++cnt;
      }
  } else {
    gauss_jacobi01(n, 2, t0, w0);
    gauss_jacobi01(n, 1, t1, w1);
    gauss_jacobi01(n, 0, t2, w2);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) {
          xi[3 * cnt + 0] = t0[i];
          xi[3 * cnt + 1] = (1.0 - t0[i]) * t1[j];
          xi[3 * cnt + 2] = (1.0 - t0[i]) * (1.0 - t1[j]) * t2[k];
          w[cnt] = w0[i] * w1[j] * w2[k];
          ++cnt;
        }
  }
  return cnt;
}

/* rule cache (orders 0..30, dims 0..3) */
typedef struct { int n; double *xi; double *w; } Rule;
static Rule g_rules[4][32];
static pthread_mutex_t g_rule_mtx = PTHREAD_MUTEX_INITIALIZER;

static const Rule *get_rule(int D, int order)
{
  if (order < 0) order = 0;
  if (order > 31) order = 31;
  pthread_mutex_lock(&g_rule_mtx);
  Rule *r = &g_rules[D][order];
  if (!r->w) {
    int n = or_rule_npts(D, order);
    double *xi = (double *)malloc(sizeof(double) * (size_t)(n * (D > 0 ? D : 1)));
    double *w = (double *)malloc(sizeof(double) * (size_t)n);
    or_rule(D, order, xi, w);
    r->xi = xi;
    r->n = n;
    r->w = w;
  }
  pthread_mutex_unlock(&g_rule_mtx);
  return r;
}

/* ====================================================================== */
/* geometry                                                               */
/* ====================================================================== */

/* Determinant(col1,col2,col3) as xdecompose.hpp:91-99 */
static double det3(const double *c1, const double *c2, const double *c3)
{
  return c1[0] * (c2[1] * c3[2] - c2[2] * c3[1]) +
         c1[1] * (c2[2] * c3[0] - c2[0] * c3[2]) +
         c1[2] * (c2[0] * c3[1] - c2[1] * c3[0]);
}

static void cross3(const double *a, const double *b, double *c)
{
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

static double norm3(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

void or_jacobian(int D, const double *V, double *F, double *Finv, double *absdet)
{
  /* column c of F = V_c - V_D */
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) F[r * D + c] = V[c * D + r] - V[D * D + r];
  double det;
  if (D == 1) {
    det = F[0];
    if (Finv) Finv[0] = 1.0 / det;
  } else if (D == 2) {
    det = F[0] * F[3] - F[1] * F[2];
    if (Finv) {
      Finv[0] = F[3] / det; Finv[1] = -F[1] / det;
      Finv[2] = -F[2] / det; Finv[3] = F[0] / det;
    }
  } else {
    double a = F[0], b = F[1], c = F[2], d = F[3], e = F[4], f = F[5], g = F[6], h = F[7], i = F[8];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    det = a * A + b * B + c * C;
    if (Finv) {
      Finv[0] = A / det; Finv[1] = -(b * i - c * h) / det; Finv[2] = (b * f - c * e) / det;
      Finv[3] = B / det; Finv[4] = (a * i - c * g) / det; Finv[5] = -(a * f - c * d) / det;
      Finv[6] = C / det; Finv[7] = -(a * h - b * g) / det; Finv[8] = (a * e - b * d) / det;
    }
  }
  *absdet = fabs(det);
}

/* ====================================================================== */
/* straight cut rule                                                      */
/* ====================================================================== */

/* CheckIfStraightCut, straightcutrule.cpp:11-24 (epsilon = 0) */
int or_check_if_straight_cut(int n, const double *lset)
{
  int haspos = 0, hasneg = 0;
  for (int i = 0; i < n; ++i) {
    if (!haspos && lset[i] > 0.0) haspos = 1;
    if (!hasneg && lset[i] < 0.0) hasneg = 1;
    if (haspos && hasneg) break;
  }
  if ((hasneg && haspos) || (!hasneg && !haspos)) return OR_IF;
  return hasneg ? OR_NEG : OR_POS;
}

/* reference simplex vertices, straightcutrule.hpp:64-69 */
static void ref_vertices(int D, double P[4][3])
{
  memset(P, 0, sizeof(double) * 12);
  for (int k = 0; k < D; ++k) P[k][k] = 1.0;
}

#define MAXSUB 3
typedef struct { int np; double p[4][3]; } Simp;

/* SimpleX::CalcIFPolytopEUsingLset, straightcutrule.cpp:41-64 */
static int calc_if_polytope(int D, const double *l, double P[4][3], double C[4][3], int *nc)
{
  int n = 0;
  for (int i = 0; i <= D; ++i)
    for (int j = i + 1; j <= D; ++j)
      if ((l[i] >= 0.0) != (l[j] >= 0.0)) {
        double pair[2] = {l[i], l[j]};
        if (or_check_if_straight_cut(2, pair) != OR_IF) return OR_ERR_NOT_IMPL;
        double t = l[i] / (l[i] - l[j]);
        for (int d = 0; d < 3; ++d) C[n][d] = P[i][d] + t * (P[j][d] - P[i][d]);
        ++n;
      }
  *nc = n;
  return 0;
}

static void simp_set(Simp *s, int np, const double (*pts[])[3])
{
  s->np = np;
  for (int k = 0; k < np; ++k) memcpy(s->p[k], *pts[k], sizeof(double) * 3);
}

/* LevelsetCutSimplex::Decompose, straightcutrule.cpp:131-196, of the
   simplex with vertices P[0..D] (any simplex of dimension D in R^3) */
static int decompose_pts(int D, double P[4][3], const double *l, int dt, Simp *out, int *nsub)
{
  double C[4][3];
  if (or_check_if_straight_cut(D + 1, l) != OR_IF) return OR_ERR_NOT_IMPL;
  int nc = 0;
  int err = calc_if_polytope(D, l, P, C, &nc);
  if (err) return err;
  *nsub = 0;
  if (dt == OR_IF) {
    if (nc == D) {
      const double (*pp[4])[3] = {&C[0], &C[1], &C[2], &C[3]};
      simp_set(&out[0], nc, pp);
      *nsub = 1;
    } else if (nc == 4 && D == 3) {
      const double (*a[3])[3] = {&C[0], &C[1], &C[3]};
      const double (*b[3])[3] = {&C[0], &C[2], &C[3]};
      simp_set(&out[0], 3, a);
      simp_set(&out[1], 3, b);
      *nsub = 2;
    } else
      return OR_ERR_BAD_SCUT;
    return 0;
  }
  int R[4], nr = 0;
  for (int i = 0; i < D + 1; ++i)
    if ((dt == OR_POS && l[i] >= 0.0) || (dt == OR_NEG && l[i] < 0.0)) R[nr++] = i;
  if (nr == 1) {
    const double (*pp[4])[3];
    for (int k = 0; k < nc; ++k) pp[k] = &C[k];
    pp[nc] = &P[R[0]];
    simp_set(&out[0], nc + 1, pp);
    *nsub = 1;
  } else if (nr == 2 && D == 2) {
    const double (*a[3])[3] = {&P[R[0]], &P[R[1]], &C[1]};
    const double (*b[3])[3] = {&C[0], &C[1], &P[R[0]]};
    simp_set(&out[0], 3, a);
    simp_set(&out[1], 3, b);
    *nsub = 2;
  } else if (nr == 2 && D == 3) {
    const double (*a[4])[3] = {&P[R[1]], &C[1], &C[2], &C[3]};
    const double (*b[4])[3] = {&P[R[0]], &P[R[1]], &C[1], &C[2]};
    const double (*c[4])[3] = {&C[0], &C[1], &C[2], &P[R[0]]};
    simp_set(&out[0], 4, a);
    simp_set(&out[1], 4, b);
    simp_set(&out[2], 4, c);
    *nsub = 3;
  } else if (nr == 3 && D == 3) {
    const double (*a[4])[3] = {&C[0], &C[1], &C[2], &P[R[2]]};
    const double (*b[4])[3] = {&P[R[0]], &P[R[1]], &P[R[2]], &C[1]};
    const double (*c[4])[3] = {&C[0], &C[1], &P[R[0]], &P[R[2]]};
    simp_set(&out[0], 4, a);
    simp_set(&out[1], 4, b);
    simp_set(&out[2], 4, c);
    *nsub = 3;
  } else
    return OR_ERR_NOT_IMPL;
  return 0;
}

static int decompose(int D, const double *l, int dt, Simp *out, int *nsub)
{
  double P[4][3];
  ref_vertices(D, P);
  return decompose_pts(D, P, l, dt, out, nsub);
}

/* SimpleX::GetVolume, straightcutrule.cpp:66-72 */
static double simp_volume(const Simp *s)
{
  int Ds = s->np - 1;
  if (Ds == 0) return 1.0;
  if (Ds == 1) {
    double a[3] = {s->p[1][0] - s->p[0][0], s->p[1][1] - s->p[0][1], s->p[1][2] - s->p[0][2]};
    return norm3(a);
  }
  if (Ds == 2) {
    double a[3], b[3], c[3];
    for (int d = 0; d < 3; ++d) { a[d] = s->p[2][d] - s->p[0][d]; b[d] = s->p[1][d] - s->p[0][d]; }
    cross3(a, b, c);
    return norm3(c);
  }
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; ++d) {
    a[d] = s->p[3][d] - s->p[0][d];
    b[d] = s->p[2][d] - s->p[0][d];
    c[d] = s->p[1][d] - s->p[0][d];
  }
  return fabs(det3(a, b, c));
}

/* SimpleX::GetPlainIntegrationRule, straightcutrule.cpp:80-100 */
static int simp_rule(const Simp *s, int order, double *pts, double *w, int cap, int n0)
{
  int Ds = s->np - 1;
  double tf = simp_volume(s);
  const Rule *r = get_rule(Ds, order);
  if (n0 + r->n > cap) return OR_ERR_CAP;
  for (int q = 0; q < r->n; ++q) {
    double ow = 1.0;
    for (int m = 0; m < Ds; ++m) ow -= r->xi[q * Ds + m];
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = ow * s->p[0][d];
    for (int m = 0; m < Ds; ++m)
      for (int d = 0; d < 3; ++d) x[d] += r->xi[q * Ds + m] * s->p[m + 1][d];
    memcpy(&pts[3 * (n0 + q)], x, sizeof(x));
    w[n0 + q] = r->w[q] * tf;
  }
  return n0 + r->n;
}

/* StraightCutIntegrationRule, straightcutrule.cpp:582-677 (simplex branch) */
int or_straight_cut_rule(int D, const double *lset, int dt, int order,
                         const double *Finv, double *pts, double *w, int cap, int *nsub)
{
  *nsub = 0;
  int eldom = or_check_if_straight_cut(D + 1, lset);
  if (eldom != OR_IF) {
    if (eldom != dt) return -1;
    const Rule *r = get_rule(D, order);
    if (r->n > cap) return OR_ERR_CAP;
    for (int q = 0; q < r->n; ++q) {
      double x[3] = {0, 0, 0};
      for (int m = 0; m < D; ++m) x[m] = r->xi[q * D + m];
      memcpy(&pts[3 * q], x, sizeof(x));
      w[q] = r->w[q];
    }
    return r->n;
  }
  Simp sub[MAXSUB];
  int ns = 0;
  int err = decompose(D, lset, dt, sub, &ns);
  if (err) return err;
  *nsub = ns;
  int n = 0;
  for (int s = 0; s < ns; ++s) {
    n = simp_rule(&sub[s], order, pts, w, cap, n);
    if (n < 0) return n;
  }
  if (dt == OR_IF) {
    /* TransformQuadUntrafoToIRInterface, straightcutrule.cpp:558-578:
       w *= |F^{-T} n_ref|, n_ref = grad(lset)/|grad(lset)| (LevelsetWrapper) */
    double g[3] = {0, 0, 0};
    for (int m = 0; m < D; ++m) g[m] = lset[m] - lset[D];
    double gn = norm3(g);
    for (int m = 0; m < 3; ++m) g[m] /= gn;
    double nrm[3] = {0, 0, 0};
    for (int r2 = 0; r2 < D; ++r2) {
      double acc = 0.0;
      for (int c = 0; c < D; ++c) acc += (Finv ? Finv[c * D + r2] : (c == r2 ? 1.0 : 0.0)) * g[c];
      nrm[r2] = acc;
    }
    double len = norm3(nrm);
    for (int q = 0; q < n; ++q) w[q] = w[q] * len;
  }
  return n;
}

/* ====================================================================== */
/* multiple level sets                                                    */
/* ====================================================================== */

/* LevelsetWrapper of a P1 level set on a simplex (straightcutrule.cpp:503-
   555): coefficients from the vertex values (GetCoeffsFromVals, SEGM / TRIG /
   TET), evaluation in the loop order of operator(), and the 1e-14 clamp of
   update_initial_coefs */
static void lsw_coefs(int D, const double *vals, double c[4])
{
  c[0] = vals[D];
  c[1] = D >= 1 ? vals[0] - vals[D] : 0.0;
  c[2] = D >= 2 ? vals[1] - vals[D] : 0.0;
  c[3] = D >= 3 ? vals[2] - vals[D] : 0.0;
}

static double lsw_eval(const double c[4], const double *p)
{
  /* sum over c[i][j][k] p0^i p1^j p2^k in the order (000),(001),(010),(100);
     the bilinear coefficients are zero for simplices */
  double v = 0.0;
  v += c[0];
  v += c[3] * p[2];
  v += c[2] * p[1];
  v += c[1] * p[0];
  return v;
}

static double lsw_clamped(const double c[4], const double *p)
{
  double d = lsw_eval(c, p);
  return fabs(d) > 1e-14 ? d : 1e-14;
}

/* unit normal of the level set on the reference element (GetNormal) */
static void lsw_normal(const double c[4], double n[3])
{
  n[0] = c[1];
  n[1] = c[2];
  n[2] = c[3];
  double len = norm3(n);
  for (int d = 0; d < 3; ++d) n[d] /= len;
}

#define OR_MLSET_MAXSIMP 729

/* StraightCutsIntegrationRule, straightcutrule.cpp:679-800 (simplices).
   lsets: M x (D+1) vertex values of the level sets that cut the element,
   dts: their target domain types; F, Finv: D x D Jacobian and its inverse,
   absdet = |det F|.  Appends to pts / w from n0; returns the new count. */
static int mlset_cut_rule(int D, int M, const double *lsets, const int *dts, int order, const double *F,
                          const double *Finv, double absdet, double *pts, double *w, int cap, int n0)
{
  static __thread Simp bufA[OR_MLSET_MAXSIMP], bufB[OR_MLSET_MAXSIMP];
  Simp *cur = bufA, *nxt = bufB;
  int ncur = 0;
  double P[4][3];
  ref_vertices(D, P);
  {
    int ns = 0;
    int err = decompose_pts(D, P, &lsets[0], dts[0], cur, &ns);
    if (err) return err;
    ncur = ns;
  }
  double c[4];
  for (int i = 1; i < M; ++i) {
    lsw_coefs(D, &lsets[i * (D + 1)], c);
    int nn = 0;
    for (int s = 0; s < ncur; ++s) {
      double vals[4];
      for (int k = 0; k < cur[s].np; ++k) vals[k] = lsw_clamped(c, cur[s].p[k]);
      int dts_s = or_check_if_straight_cut(cur[s].np, vals);
      if (dts_s == OR_IF) {
        if (nn + MAXSUB > OR_MLSET_MAXSIMP) return OR_ERR_CAP;
        int ns = 0;
        int err = decompose_pts(cur[s].np - 1, cur[s].p, vals, dts[i], &nxt[nn], &ns);
        if (err) return err;
        nn += ns;
      } else if (dts_s == dts[i]) {
        if (nn + 1 > OR_MLSET_MAXSIMP) return OR_ERR_CAP;
        nxt[nn++] = cur[s];
      }
    }
    Simp *t = cur;
    cur = nxt;
    nxt = t;
    ncur = nn;
  }
  int n = n0;
  for (int s = 0; s < ncur; ++s) {
    n = simp_rule(&cur[s], order, pts, w, cap, n);
    if (n < 0) return n;
  }
  int ifi[4], nif = 0;
  for (int i = 0; i < M; ++i)
    if (dts[i] == OR_IF) {
      if (nif >= 3) return OR_ERR_NOT_IMPL;  /* "Codim possibilities available" */
      ifi[nif++] = i;
    }
  double fac = 1.0;
  if (nif == 1) {
    /* TransformQuadUntrafoToIRInterface: |F^{-T} n_ref| */
    double nr[3];
    lsw_coefs(D, &lsets[ifi[0] * (D + 1)], c);
    lsw_normal(c, nr);
    double v[3] = {0, 0, 0};
    for (int r = 0; r < D; ++r) {
      double acc = 0.0;
      for (int q = 0; q < D; ++q) acc += Finv[q * D + r] * nr[q];
      v[r] = acc;
    }
    fac = norm3(v);
  } else if (nif == 2) {
    if (D == 2) {
      fac = 1.0 / absdet;
    } else if (D == 3) {
      double n0v[3], n1v[3], cp[3];
      lsw_coefs(D, &lsets[ifi[0] * (D + 1)], c);
      lsw_normal(c, n0v);
      lsw_coefs(D, &lsets[ifi[1] * (D + 1)], c);
      lsw_normal(c, n1v);
      cross3(n0v, n1v, cp);
      double cpf = 1.0 / norm3(cp);
      double v[3];
      for (int r = 0; r < 3; ++r) {
        double acc = 0.0;
        for (int q = 0; q < 3; ++q) acc += F[r * 3 + q] * cp[q];
        v[r] = cpf * acc;
      }
      fac = norm3(v) / absdet;
    } else
      return OR_ERR_NOT_IMPL;
  } else if (nif == 3) {
    if (D != 3) return OR_ERR_NOT_IMPL;
    fac = 1.0 / absdet;
  }
  if (nif > 0)
    for (int q = n0; q < n; ++q) w[q] = w[q] * fac;
  return n;
}

/* CreateMultiLevelsetCutIntegrationRule, mlsetintegration.cpp:8-158, on one
   element: lsets M x (D+1) vertex values, dtts K x M domain tuples (the
   LevelsetIntegrationDomain's domain list).  Returns the number of points
   (0 when no tuple contributes). */
int or_mlset_rule(int D, int M, const double *lsets, int K, const int *dtts, int order, const double *F,
                  const double *Finv, double absdet, double *pts, double *w, int cap)
{
  int ldt[16];
  if (M > 16) return OR_ERR_CAP;
  for (int i = 0; i < M; ++i) ldt[i] = or_check_if_straight_cut(D + 1, &lsets[i * (D + 1)]);
  int n = 0;
  for (int k = 0; k < K; ++k) {
    const int *dts = &dtts[k * M];
    int compatible = 1, cut = 0;
    for (int i = 0; i < M; ++i) {
      if (ldt[i] != OR_IF && ldt[i] != dts[i]) compatible = 0;
      else if (ldt[i] == OR_IF) cut = 1;
    }
    if (!compatible) continue;
    if (!cut) {
      const Rule *r = get_rule(D, order);
      if (n + r->n > cap) return OR_ERR_CAP;
      for (int q = 0; q < r->n; ++q) {
        double x[3] = {0, 0, 0};
        for (int m = 0; m < D; ++m) x[m] = r->xi[q * D + m];
        memcpy(&pts[3 * (n + q)], x, sizeof(x));
        w[n + q] = r->w[q];
      }
      n += r->n;
      continue;
    }
    double cl[16 * 4];
    int cd[16], nc = 0;
    for (int i = 0; i < M; ++i)
      if (ldt[i] == OR_IF || ldt[i] != dts[i]) {
        memcpy(&cl[nc * (D + 1)], &lsets[i * (D + 1)], sizeof(double) * (D + 1));
        cd[nc++] = dts[i];
      }
    int r = mlset_cut_rule(D, nc, cl, cd, order, F, Finv, absdet, pts, w, cap, n);
    if (r < 0) return r;
    n = r;
  }
  return n;
}

/* MultiLevelsetCutInformation (cutinfo.cpp:286-415) on a mesh: bit e of
   of_type = the element's level-set domain tuple is one of dtts;
   contrib (nullable): the mlset rule of order 0 has positive weight sum. */
int or_mlset_mark(int D, int ne, const int32_t *el2v, const double *coords, int M, const double *lset_v, int nv,
                  int K, const int *dtts, uint8_t *of_type, uint8_t *contrib)
{
  double *pts = malloc(sizeof(double) * 3 * 4096), *w = malloc(sizeof(double) * 4096);
  int err = 0;
  for (int e = 0; e < ne && !err; ++e) {
    double l[16 * 4];
    int ldt[16];
    for (int i = 0; i < M; ++i) {
      for (int k = 0; k <= D; ++k) l[i * (D + 1) + k] = lset_v[(size_t)i * nv + el2v[(size_t)e * (D + 1) + k]];
      ldt[i] = or_check_if_straight_cut(D + 1, &l[i * (D + 1)]);
    }
    int match = 0;
    for (int k = 0; k < K; ++k) {
      int ok = 1;
      for (int i = 0; i < M; ++i)
        if (dtts[k * M + i] != ldt[i]) ok = 0;
      if (ok) match = 1;
    }
    of_type[e] = (uint8_t)match;
    if (contrib) {
      double V[4 * 3], F[9], Fi[9], ad;
      for (int k = 0; k <= D; ++k)
        for (int d = 0; d < D; ++d) V[k * D + d] = coords[(size_t)el2v[(size_t)e * (D + 1) + k] * D + d];
      or_jacobian(D, V, F, Fi, &ad);
      int n = or_mlset_rule(D, M, l, K, dtts, 0, F, Fi, ad, pts, w, 4096);
      if (n < 0) {
        err = n;
        break;
      }
      double sw = 0.0;
      for (int q = 0; q < n; ++q) sw += w[q];
      contrib[e] = sw > 0.0;
    }
  }
  free(pts);
  free(w);
  return err;
}

/* ====================================================================== */
/* marking, cutinfo.cpp:80-133                                             */
/* ====================================================================== */

int or_mark(int D, int ne, const int32_t *el2v, const double *lset_v,
            uint8_t *dt_out, double *cut_ratio, uint32_t *bits)
{
  int nw = (ne + 31) / 32;
  memset(bits, 0, sizeof(uint32_t) * 8 * (size_t)nw);
  double pts[3 * 16], w[16];
  for (int e = 0; e < ne; ++e) {
    double l[4];
    for (int k = 0; k <= D; ++k) l[k] = lset_v[el2v[e * (D + 1) + k]];
    double pv[2] = {0.0, 0.0};
    const int order_np[2] = {OR_POS, OR_NEG};
    for (int t = 0; t < 2; ++t) {
      int np = order_np[t], ns;
      int n = or_straight_cut_rule(D, l, np, 0, NULL, pts, w, 16, &ns);
      if (n < -1) return n;
      for (int q = 0; q < n; ++q) pv[np] += w[q];
    }
    if (cut_ratio) cut_ratio[e] = pv[OR_NEG] / (pv[OR_NEG] + pv[OR_POS]);
    int cdt;
    if (pv[OR_NEG] > 0.0)
      cdt = (pv[OR_POS] > 0.0) ? 4 /*IF*/ : 1 /*NEG*/;
    else
      cdt = 2; /*POS*/
    int dtv = cdt == 4 ? OR_IF : (cdt == 1 ? OR_NEG : OR_POS);
    if (dt_out) dt_out[e] = (uint8_t)dtv;
    uint32_t bit = 1u << (e & 31);
    int wd = e >> 5;
    for (int c = 0; c < 8; ++c)
      if ((c & cdt) || c == 7) bits[(size_t)c * nw + wd] |= bit;
  }
  /* combined arrays: index = COMBINED_DOMAIN_TYPE; bit set iff (cdt & c) != 0,
     i.e. UNCUT = NEG|POS, HASNEG = NEG|IF, HASPOS = POS|IF, ANY = all, NO = none */
  return 0;
}

/* ====================================================================== */
/* quadrilaterals and hexahedra: LevelsetCutQuadrilateral                  */
/* (straightcutrule.cpp:74-78, 102-129, 206-501, 503-556, 582-677)         */
/* The rule of a cut unit quad / hex is built as the reference does: a     */
/* tensor product along the last coordinate xi (Gauss points on [xi0,xi1]   */
/* times the cut rule of the codimension-1 slice at each point), after      */
/* splitting [0,1] at the xi where an edge along xi changes sign, with the  */
/* dimension swaps of GetDimensionSwap and the simplex fallback.            */
/* ====================================================================== */

typedef struct { double c[2][2][2]; } QL;             /* LevelsetWrapper::c */
typedef struct { int D; double p[8][3]; } QQ;          /* Quadrilateral */
typedef struct { int n, cap; double *x; double *w; } QR; /* growable rule */

static const double QC_c = 0.999;                      /* straightcutrule.cpp:253 */

static int qr_push(QR *r, const double *x, double w)
{
  if (r->n == r->cap) {
    int nc = r->cap ? 2 * r->cap : 64;
    double *nx = (double *)realloc(r->x, sizeof(double) * 3 * (size_t)nc);
    double *nw = (double *)realloc(r->w, sizeof(double) * (size_t)nc);
    if (!nx || !nw) { free(nx ? nx : r->x); free(nw ? nw : r->w); r->x = r->w = NULL; return OR_ERR_CAP; }
    r->x = nx; r->w = nw; r->cap = nc;
  }
  memcpy(&r->x[3 * r->n], x, sizeof(double) * 3);
  r->w[r->n++] = w;
  return 0;
}

static void qr_free(QR *r) { free(r->x); free(r->w); r->x = r->w = NULL; r->n = r->cap = 0; }

/* LevelsetWrapper::GetCoeffsFromVals, straightcutrule.cpp:503-526
   (et: 1 = ET_SEGM, 2 = ET_QUAD, 3 = ET_HEX) */
static void ql_from_vals(int et, const double *v, QL *L)
{
  memset(L, 0, sizeof(*L));
  double (*c)[2][2] = L->c;
  if (et == 1) {
    c[0][0][0] = v[1]; c[1][0][0] = v[0] - v[1];
  } else if (et == 2) {
    c[0][0][0] = v[0]; c[1][0][0] = v[1] - c[0][0][0]; c[0][1][0] = v[3] - c[0][0][0];
    c[1][1][0] = v[2] - c[1][0][0] - c[0][1][0] - c[0][0][0];
  } else {
    c[0][0][0] = v[0]; c[1][0][0] = v[1] - v[0]; c[0][1][0] = v[3] - v[0]; c[0][0][1] = v[4] - v[0];
    c[1][1][0] = v[2] - c[1][0][0] - c[0][1][0] - c[0][0][0];
    c[1][0][1] = v[5] - c[1][0][0] - c[0][0][1] - c[0][0][0];
    c[0][1][1] = v[7] - c[0][1][0] - c[0][0][1] - c[0][0][0];
    c[1][1][1] = v[6] - c[1][1][0] - c[1][0][1] - c[0][1][1] - c[1][0][0] - c[0][1][0] - c[0][0][1] - c[0][0][0];
  }
}

/* operator(), :542-546: sum over (i,j,k) in loop order of
   c[i][j][k] * p0^i * p1^j * p2^k (powers 0 and 1 are exact) */
static double ql_eval(const QL *L, const double *p)
{
  double v = 0.0;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k) {
        double t = L->c[i][j][k];
        if (i) t *= p[0];
        if (j) t *= p[1];
        if (k) t *= p[2];
        v += t;
      }
  return v;
}

/* GetGrad, :528-534 */
static void ql_grad(const QL *L, const double *p, double g[3])
{
  const double (*c)[2][2] = (const double (*)[2][2])L->c;
  g[0] = c[1][0][0] + c[1][1][0] * p[1] + c[1][0][1] * p[2] + c[1][1][1] * p[1] * p[2];
  g[1] = c[0][1][0] + c[1][1][0] * p[0] + c[0][1][1] * p[2] + c[1][1][1] * p[0] * p[2];
  g[2] = c[0][0][1] + c[0][1][1] * p[1] + c[1][0][1] * p[0] + c[1][1][1] * p[0] * p[1];
}

/* Quadrilateral(bnds), straightcutrule.hpp:72-90 */
static void qq_box(int D, const double *lo, const double *hi, QQ *q)
{
  memset(q, 0, sizeof(*q));
  q->D = D;
  const int nv = D == 2 ? 4 : 8;
  for (int v = 0; v < nv; ++v) {
    const int b = v & 3;
    q->p[v][0] = (b == 1 || b == 2) ? hi[0] : lo[0];
    q->p[v][1] = (b >= 2) ? hi[1] : lo[1];
    if (D == 3) q->p[v][2] = (v >= 4) ? hi[2] : lo[2];
  }
}

static void qq_unit(int D, QQ *q)
{
  const double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  qq_box(D, lo, hi, q);
}

static int qq_check(const QL *L, const QQ *q, double eps)
{
  const int nv = q->D == 2 ? 4 : 8;
  double v[8];
  for (int k = 0; k < nv; ++k) v[k] = ql_eval(L, q->p[k]);
  /* CheckIfStraightCut with epsilon, :11-24 */
  int haspos = 0, hasneg = 0;
  for (int k = 0; k < nv; ++k) {
    if (!haspos && v[k] > eps) haspos = 1;
    if (!hasneg && v[k] < -eps) hasneg = 1;
    if (haspos && hasneg) break;
  }
  if ((hasneg && haspos) || (!hasneg && !haspos)) return OR_IF;
  return hasneg ? OR_NEG : OR_POS;
}

/* Quadrilateral::GetVolume / GetPlainIntegrationRule, :74-78, 102-129:
   tensor Gauss rule of ET_QUAD / ET_HEX (x slowest), mapped by
   points[0] + A ip with A = (p1 - p0, p3 - p0, p4 - p0) */
static int qq_plain(const QQ *q, int order, QR *r)
{
  double a[3], b[3], c[3], tf;
  for (int d = 0; d < 3; ++d) { a[d] = q->p[3][d] - q->p[0][d]; b[d] = q->p[1][d] - q->p[0][d]; }
  if (q->D == 2) {
    cross3(a, b, c);
    tf = norm3(c);
  } else {
    double e4[3];
    for (int d = 0; d < 3; ++d) e4[d] = q->p[4][d] - q->p[0][d];
    tf = fabs(det3(e4, a, b));
  }
  double A[3][3] = {{0}};
  for (int d = 0; d < 3; ++d) {
    A[d][0] = q->p[1][d] - q->p[0][d];
    A[d][1] = q->p[3][d] - q->p[0][d];
    A[d][2] = q->D == 3 ? q->p[4][d] - q->p[0][d] : 0.0;
  }
  if (q->D == 2) { A[2][0] = A[2][1] = 0.0; }
  const Rule *s = get_rule(1, order);
  const int n = s->n;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < (q->D == 3 ? n : 1); ++k) {
        const double ip[3] = {s->xi[i], s->xi[j], q->D == 3 ? s->xi[k] : 0.0};
        const double wq = q->D == 3 ? s->w[i] * s->w[j] * s->w[k] : s->w[i] * s->w[j];
        double x[3];
        for (int d = 0; d < 3; ++d) x[d] = q->p[0][d] + (A[d][0] * ip[0] + A[d][1] * ip[1] + A[d][2] * ip[2]);
        int e = qr_push(r, x, wq * tf);
        if (e) return e;
      }
  return 0;
}

/* LevelsetCutSimplex(LevelsetWrapper(lv, ET_SEGM), dt, SimpleX(ET_SEGM))
   ::GetIntegrationRule, :41-64 (D == 1 branch), :131-204.  The unit
   segment's points are (1,0,0) and the origin. */
static int q_segment_rule(const double lv[2], int dt, int order, QR *r)
{
  if (or_check_if_straight_cut(2, lv) != OR_IF) return OR_ERR_PLAIN;
  const double P[2][3] = {{1, 0, 0}, {0, 0, 0}};
  const double t = lv[0] / (lv[0] - lv[1]);
  double C[3];
  for (int d = 0; d < 3; ++d) C[d] = P[0][d] + t * (P[1][d] - P[0][d]);
  if (dt == OR_IF) return qr_push(r, C, 1.0); /* ET_POINT rule, volume 1 */
  int R[2], nr = 0;
  for (int i = 0; i < 2; ++i)
    if ((dt == OR_POS && lv[i] >= 0.0) || (dt == OR_NEG && lv[i] < 0.0)) R[nr++] = i;
  if (nr != 1) return OR_ERR_NOT_IMPL;
  Simp s;
  s.np = 2;
  memcpy(s.p[0], C, sizeof(C));
  memcpy(s.p[1], P[R[0]], sizeof(C));
  double pts[3 * 16], w[16];
  int n = simp_rule(&s, order, pts, w, 16, 0);
  if (n < 0) return n;
  for (int k = 0; k < n; ++k) {
    int e = qr_push(r, &pts[3 * k], w[k]);
    if (e) return e;
  }
  return 0;
}

static int qc_rule(const QL *L, int D, int dt, int order, int pol, int consider, QR *r);

/* GetTensorProductAlongXiIntegrationRule, :256-302 */
static int qc_tensor(const QL *L, const QQ *q, int dt, int order, int pol, QR *r)
{
  const int D = q->D, xi = D - 1;
  const double xi0 = q->p[0][xi], xi1 = xi == 1 ? q->p[2][1] : q->p[4][2];
  const double Cb = 1. / sqrt(1 - QC_c * QC_c);
  const Rule *s = get_rule(1, order);
  for (int i1 = 0; i1 < s->n; ++i1) {
    const double xs = xi0 + s->xi[i1] * (xi1 - xi0);
    QR inner = {0, 0, NULL, NULL};
    int e;
    if (D == 2) {
      const double pa[3] = {q->p[0][0], xs, 0}, pb[3] = {q->p[2][0], xs, 0};
      const double lv[2] = {ql_eval(L, pb), ql_eval(L, pa)};
      e = q_segment_rule(lv, dt, order, &inner);
    } else {
      const double p0[3] = {q->p[0][0], q->p[0][1], xs}, p1[3] = {q->p[2][0], q->p[0][1], xs};
      const double p2[3] = {q->p[2][0], q->p[2][1], xs}, p3[3] = {q->p[0][0], q->p[2][1], xs};
      const double lv[4] = {ql_eval(L, p0), ql_eval(L, p1), ql_eval(L, p2), ql_eval(L, p3)};
      QL L2;
      ql_from_vals(2, lv, &L2);
      e = qc_rule(&L2, 2, dt, order, pol, 1, &inner);
    }
    if (e) { qr_free(&inner); return e; }
    for (int k = 0; k < inner.n; ++k) {
      const double *p2 = &inner.x[3 * k];
      double ip[3] = {0, 0, 0};
      ip[xi] = xs;
      ip[0] = q->p[0][0] + p2[0] * (q->p[2][0] - q->p[0][0]);
      if (D == 3) ip[1] = q->p[0][1] + p2[1] * (q->p[2][1] - q->p[0][1]);
      double f = 1;
      if (dt == OR_IF) {
        double g[3];
        ql_grad(L, ip, g);
        if (D == 2)
          f = norm3(g) / fabs(g[0]);
        else
          f = norm3(g) / sqrt(g[0] * g[0] + g[1] * g[1]);
        if (isnan(f) || f > Cb) { qr_free(&inner); return OR_ERR_BOUND; }
      }
      e = qr_push(r, ip, inner.w[k] * s->w[i1] * (xi1 - xi0) * f);
      if (e) { qr_free(&inner); return e; }
    }
    qr_free(&inner);
  }
  return 0;
}

/* HasTopologyChangeAlongXi / Decompose / GetIntegrationRuleAlongXi,
   :206-251, 459-473, on the unit quad / hex */
static int qc_along_xi(const QL *L, int D, int dt, int order, int pol, QR *r)
{
  QQ q;
  qq_unit(D, &q);
  const int dq = qq_check(L, &q, 0.0);
  if (dq != OR_IF) return dq == dt ? qq_plain(&q, order, r) : 0;
  static const int e2[2][2] = {{1, 2}, {0, 3}};
  static const int e3[4][2] = {{0, 4}, {1, 5}, {2, 6}, {3, 7}};
  const int ne = D == 2 ? 2 : 4;
  double xs[6] = {0.0, 1.0};
  int nx = 2, topo = 0;
  for (int k = 0; k < ne; ++k) {
    const int a = D == 2 ? e2[k][0] : e3[k][0], b = D == 2 ? e2[k][1] : e3[k][1];
    const double vals[2] = {ql_eval(L, q.p[b]), ql_eval(L, q.p[a])};
    if (or_check_if_straight_cut(2, vals) != OR_IF) continue;
    topo = 1;
    /* unit_line.CalcIFPolytopEUsingLset(vals).points[0][0] */
    const double t = vals[0] / (vals[0] - vals[1]);
    const double x = 1.0 + t * (0.0 - 1.0);
    if (isnan(x)) continue; /* a NaN key compares equal to the set's members: not inserted */
    int dup = 0;
    for (int m = 0; m < nx; ++m) dup |= xs[m] == x;
    if (!dup) xs[nx++] = x;
  }
  if (!topo) return qc_tensor(L, &q, dt, order, pol, r);
  for (int i = 1; i < nx; ++i) /* ascending, as the std::set iteration */
    for (int j = i; j > 0 && xs[j] < xs[j - 1]; --j) { double t = xs[j]; xs[j] = xs[j - 1]; xs[j - 1] = t; }
  for (int i = 0; i + 1 < nx; ++i) {
    QQ sq;
    double lo[3] = {q.p[0][0], 0, 0}, hi[3] = {q.p[2][0], 0, 0};
    if (D == 2) { lo[1] = xs[i]; hi[1] = xs[i + 1]; }
    else { lo[1] = q.p[0][1]; hi[1] = q.p[2][1]; lo[2] = xs[i]; hi[2] = xs[i + 1]; }
    qq_box(D, lo, hi, &sq);
    const int ds = qq_check(L, &sq, 1e-15);
    int e = 0;
    if (ds == OR_IF) e = qc_tensor(L, &sq, dt, order, pol, r);
    else if (ds == dt) e = qq_plain(&sq, order, r);
    if (e) return e;
  }
  return 0;
}

/* GetFallbackIntegrationRule, :475-491 */
static int qc_fallback(const QL *L, int D, int dt, int order, QR *r)
{
  static const int t2[2][3] = {{0, 1, 3}, {2, 1, 3}};
  static const int t3[6][4] = {{3, 0, 1, 5}, {3, 1, 2, 5}, {3, 5, 2, 6}, {4, 5, 0, 3}, {4, 7, 5, 3}, {7, 6, 5, 3}};
  QQ q;
  qq_unit(D, &q);
  const int ns = D == 2 ? 2 : 6;
  double pts[3 * 4096], w[4096];
  for (int s = 0; s < ns; ++s) {
    double P[4][3], l[4];
    for (int k = 0; k <= D; ++k) {
      const int v = D == 2 ? t2[s][k] : t3[s][k];
      memcpy(P[k], q.p[v], sizeof(P[k]));
      const double d = ql_eval(L, P[k]); /* update_initial_coefs, :548-556 */
      l[k] = fabs(d) > 1e-14 ? d : 1e-14;
    }
    const int ds = or_check_if_straight_cut(D + 1, l);
    int n = 0;
    if (ds != OR_IF && ds == dt) {
      Simp sp;
      sp.np = D + 1;
      memcpy(sp.p, P, sizeof(P));
      n = simp_rule(&sp, order, pts, w, 4096, 0);
    } else if (ds == OR_IF) {
      Simp sub[MAXSUB];
      int nsub = 0;
      int e = decompose_pts(D, P, l, dt, sub, &nsub);
      if (e) return e;
      for (int k = 0; k < nsub && n >= 0; ++k) n = simp_rule(&sub[k], order, pts, w, 4096, n);
    }
    if (n < 0) return n;
    for (int k = 0; k < n; ++k) {
      int e = qr_push(r, &pts[3 * k], w[k]);
      if (e) return e;
    }
  }
  return 0;
}

enum { QS_ID, QS_XY, QS_XZ, QS_YZ, QS_NONE };

/* GetSufficientCritsQBound / GetExactCritsQBound2D / GetDimensionSwap,
   :353-457 (pol: 0 FIRST_ALLOWED, 1 FIND_OPTIMAL, 2 ALWAYS_NONE) */
static int qc_swap(const QL *L, int D, int pol, int consider, int *sw)
{
  if (pol == 2) { *sw = QS_NONE; return 0; }
  if (!consider) { *sw = QS_ID; return 0; }
  const double (*c)[2][2] = (const double (*)[2][2])L->c;
  if (D == 2) {
    int allow[2] = {1, 1};
    double h = -c[1][0][0] / c[1][1][0];
    if ((h > 0) && (h < 1)) allow[1] = 0;
    h = -c[0][1][0] / c[1][1][0];
    if ((h > 0) && (h < 1)) allow[0] = 0;
    double qm[2] = {0, 0};
    static const double cs[4][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}};
    for (int k = 0; k < 4; ++k) {
      double g[3];
      ql_grad(L, cs[k], g);
      const double qy = fabs(g[1]) / norm3(g), qx = fabs(g[0]) / norm3(g);
      if (qy > qm[1]) qm[1] = qy;
      if (qx > qm[0]) qm[0] = qx;
    }
    for (int i = 0; i < 2; ++i) if (!allow[i]) qm[i] = 2;
    if (pol == 0) {
      *sw = qm[1] < QC_c ? QS_ID : (qm[0] < QC_c ? QS_XY : QS_NONE);
    } else {
      if ((qm[0] < QC_c) && (qm[1] < QC_c)) *sw = qm[1] <= qm[0] ? QS_ID : QS_XY;
      else if (qm[1] < QC_c) *sw = QS_ID;
      else if (qm[0] < QC_c) *sw = QS_XY;
      else *sw = QS_NONE;
    }
    return 0;
  }
  static const double cs[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {1, 1, 0},
                                  {0, 0, 1}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
  double g[8][3];
  for (int k = 0; k < 8; ++k) ql_grad(L, cs[k], g[k]);
  double Vsq = 0;
  for (int d = 0; d < 3; ++d) {
    double mx = 0;
    for (int k = 0; k < 8; ++k) {
      const double v = g[k][d] * g[k][d];
      if (v > mx) mx = v;
    }
    Vsq += mx;
  }
  const double V = sqrt(Vsq);
  double qm[3] = {0, 0, 0};
  for (int k = 0; k < 8; ++k)
    for (int d = 0; d < 3; ++d) {
      const double qe = V * V / (V * V - g[k][d] * g[k][d]);
      if (qe > qm[d]) qm[d] = qe;
    }
  for (int d = 0; d < 3; ++d) qm[d] = sqrt(1 - 1 / qm[d]);
  for (int d = 0; d < 3; ++d) if (isnan(qm[d])) return OR_ERR_BOUND;
  if (pol == 0) {
    *sw = qm[2] < QC_c ? QS_ID : (qm[1] < QC_c ? QS_YZ : (qm[0] < QC_c ? QS_XZ : QS_NONE));
  } else {
    /* min_element with comparator <=: the last of equal minima */
    int md = 0;
    for (int d = 1; d < 3; ++d) if (qm[d] <= qm[md]) md = d;
    if (qm[md] < QC_c) *sw = md == 0 ? QS_XZ : (md == 1 ? QS_YZ : QS_ID);
    else *sw = QS_NONE;
  }
  return 0;
}

/* LevelsetCutQuadrilateral::GetIntegrationRule, :493-501, with the
   permuted variants :304-351 (coefficients transposed, the rule computed
   on the unit quad / hex, coordinates swapped back) */
static int qc_rule(const QL *L, int D, int dt, int order, int pol, int consider, QR *r)
{
  int sw;
  int e = qc_swap(L, D, pol, consider, &sw);
  if (e) return e;
  if (sw == QS_ID) return qc_along_xi(L, D, dt, order, pol, r);
  if (sw == QS_NONE) return qc_fallback(L, D, dt, order, r);
  QL Lr;
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int k = 0; k < 2; ++k)
        Lr.c[i][j][k] = sw == QS_XY ? L->c[j][i][k] : (sw == QS_XZ ? L->c[k][j][i] : L->c[i][k][j]);
  QR rr = {0, 0, NULL, NULL};
  /* X_Y: GetIntegrationRuleAlongXi; X_Z, Y_Z: GetIntegrationRule of the
     rotated quad without further swaps, i.e. along xi as well */
  e = qc_along_xi(&Lr, D, dt, order, pol, &rr);
  for (int k = 0; k < rr.n && !e; ++k) {
    const double *p = &rr.x[3 * k];
    double x[3];
    if (sw == QS_XY) { x[0] = p[1]; x[1] = p[0]; x[2] = p[2]; }
    else if (sw == QS_XZ) { x[0] = p[2]; x[1] = p[1]; x[2] = p[0]; }
    else { x[0] = p[0]; x[1] = p[2]; x[2] = p[1]; }
    e = qr_push(r, x, rr.w[k]);
  }
  qr_free(&rr);
  return e;
}

/* multilinear element map of ET_QUAD / ET_HEX: F = dx/dxi at xi, |det F| */
static void q_jacobian(int D, const double *V, const double *xi, double *F, double *absdet)
{
  const int nv = D == 2 ? 4 : 8;
  for (int k = 0; k < D * D; ++k) F[k] = 0.0;
  for (int v = 0; v < nv; ++v) {
    const int b = v & 3;
    const int a[3] = {(b == 1 || b == 2), b >= 2, v >= 4};
    double dN[3];
    for (int c = 0; c < D; ++c) {
      double t = 1.0;
      for (int m = 0; m < D; ++m) {
        const double f = a[m] ? xi[m] : 1.0 - xi[m];
        const double df = a[m] ? 1.0 : -1.0;
        t *= (m == c) ? df : f;
      }
      dN[c] = t;
    }
    for (int rr = 0; rr < D; ++rr)
      for (int c = 0; c < D; ++c) F[rr * D + c] += V[v * D + rr] * dN[c];
  }
  if (D == 2) *absdet = fabs(F[0] * F[3] - F[1] * F[2]);
  else {
    const double c1[3] = {F[0], F[3], F[6]}, c2[3] = {F[1], F[4], F[7]}, c3[3] = {F[2], F[5], F[8]};
    *absdet = fabs(det3(c1, c2, c3));
  }
}

static void q_map(int D, const double *V, const double *xi, double *x)
{
  const int nv = D == 2 ? 4 : 8;
  for (int d = 0; d < 3; ++d) x[d] = 0.0;
  for (int v = 0; v < nv; ++v) {
    const int b = v & 3;
    const int a[3] = {(b == 1 || b == 2), b >= 2, v >= 4};
    double N = 1.0;
    for (int m = 0; m < D; ++m) N *= a[m] ? xi[m] : 1.0 - xi[m];
    for (int d = 0; d < D; ++d) x[d] += N * V[v * D + d];
  }
}

int or_quad_cut_rule(int D, const double *vals, int dt, int order, int pol, const double *V, double *pts,
                     double *w, int cap)
{
  const int nv = D == 2 ? 4 : 8;
  const int eldom = or_check_if_straight_cut(nv, vals);
  QR r = {0, 0, NULL, NULL};
  if (eldom != OR_IF) {
    if (eldom != dt) return -1;
    QQ q;
    qq_unit(D, &q);
    int e = qq_plain(&q, order, &r);
    if (e) { qr_free(&r); return e; }
  } else {
    QL L;
    ql_from_vals(D, vals, &L);
    int e = qc_rule(&L, D, dt, order, pol, 1, &r);
    if (e) { qr_free(&r); return e; }
    if (r.n == 0) { qr_free(&r); return -1; }
    if (dt == OR_IF) {
      /* TransformQuadUntrafoToIRInterface, :558-578 */
      for (int k = 0; k < r.n; ++k) {
        const double *p = &r.x[3 * k];
        double g[3];
        ql_grad(&L, p, g);
        const double gl = norm3(g);
        for (int d = 0; d < 3; ++d) g[d] /= gl;
        double F[9], Finv[9], ad;
        if (V) {
          q_jacobian(D, V, p, F, &ad);
          if (D == 2) {
            const double det = F[0] * F[3] - F[1] * F[2];
            Finv[0] = F[3] / det; Finv[1] = -F[1] / det; Finv[2] = -F[2] / det; Finv[3] = F[0] / det;
          } else {
            const double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                               F[2] * (F[3] * F[7] - F[4] * F[6]);
            Finv[0] = (F[4] * F[8] - F[5] * F[7]) / det; Finv[1] = (F[2] * F[7] - F[1] * F[8]) / det;
            Finv[2] = (F[1] * F[5] - F[2] * F[4]) / det; Finv[3] = (F[5] * F[6] - F[3] * F[8]) / det;
            Finv[4] = (F[0] * F[8] - F[2] * F[6]) / det; Finv[5] = (F[2] * F[3] - F[0] * F[5]) / det;
            Finv[6] = (F[3] * F[7] - F[4] * F[6]) / det; Finv[7] = (F[1] * F[6] - F[0] * F[7]) / det;
            Finv[8] = (F[0] * F[4] - F[1] * F[3]) / det;
          }
        }
        double nn[3] = {0, 0, 0};
        for (int rr = 0; rr < D; ++rr) {
          double acc = 0.0;
          for (int c = 0; c < D; ++c) acc += (V ? Finv[c * D + rr] : (c == rr ? 1.0 : 0.0)) * g[c];
          nn[rr] = acc;
        }
        r.w[k] = r.w[k] * norm3(nn);
      }
    }
  }
  if (r.n > cap) { qr_free(&r); return OR_ERR_CAP; }
  memcpy(pts, r.x, sizeof(double) * 3 * (size_t)r.n);
  memcpy(w, r.w, sizeof(double) * (size_t)r.n);
  const int n = r.n;
  qr_free(&r);
  return n;
}

/* CutInformation::Update (cutinfo.cpp:61-133) on a quad / hex mesh (el2v:
   ne x 2^D vertices in the order of the reference element) */
int or_quad_mark(int D, int ne, const int32_t *el2v, const double *lset_v, int pol, uint8_t *dt_out,
                 double *cut_ratio, uint32_t *bits)
{
  const int nv = D == 2 ? 4 : 8;
  const int nw = (ne + 31) / 32;
  memset(bits, 0, sizeof(uint32_t) * 8 * (size_t)nw);
  double pts[3 * 4096], w[4096];
  for (int e = 0; e < ne; ++e) {
    double l[8];
    for (int k = 0; k < nv; ++k) l[k] = lset_v[el2v[e * nv + k]];
    double pv[2] = {0.0, 0.0};
    const int order_np[2] = {OR_POS, OR_NEG};
    for (int t = 0; t < 2; ++t) {
      const int np = order_np[t];
      const int n = or_quad_cut_rule(D, l, np, 0, pol, NULL, pts, w, 4096);
      if (n < -1) return n;
      for (int q = 0; q < n; ++q) pv[np] += w[q];
    }
    if (cut_ratio) cut_ratio[e] = pv[OR_NEG] / (pv[OR_NEG] + pv[OR_POS]);
    const int cdt = pv[OR_NEG] > 0.0 ? (pv[OR_POS] > 0.0 ? 4 : 1) : 2;
    if (dt_out) dt_out[e] = (uint8_t)(cdt == 4 ? OR_IF : (cdt == 1 ? OR_NEG : OR_POS));
    for (int c = 0; c < 8; ++c)
      if ((c & cdt) || c == 7) bits[(size_t)c * nw + (e >> 5)] |= 1u << (e & 31);
  }
  return 0;
}

/* IntegrateX sums (python_cutint.cpp:66-98) on a quad / hex element:
   sum_q |det F(xi_q)| w_q f(x(xi_q)), f a polynomial as or_integrate_poly */
double or_quad_integrate_poly(int D, const double *V, const double *pts, const double *w, int n, int nterms,
                              const double *coef, const int32_t *expo)
{
  double s = 0.0;
  for (int q = 0; q < n; ++q) {
    double F[9], ad, x[3];
    q_jacobian(D, V, &pts[3 * q], F, &ad);
    q_map(D, V, &pts[3 * q], x);
    double fv = 0.0;
    for (int t = 0; t < nterms; ++t) {
      double m = coef[t];
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < expo[3 * t + d]; ++k) m *= x[d];
      fv += m;
    }
    s += ad * w[q] * fv;
  }
  return s;
}

/* CalcElementMatrix (symboliccutbfi.cpp:76-304) of the Q1 space on a quad /
   hex element and a rule on the unit element: mass (mask 1), stiffness
   (mask 2) or their sum, out (2^D x 2^D) overwritten */
void or_quad_elmat(int D, const double *V, const double *pts, const double *w, int n, int mask, double *out)
{
  const int nv = D == 2 ? 4 : 8;
  memset(out, 0, sizeof(double) * (size_t)(nv * nv));
  for (int q = 0; q < n; ++q) {
    const double *xi = &pts[3 * q];
    double F[9], ad, Finv[9];
    q_jacobian(D, V, xi, F, &ad);
    if (D == 2) {
      const double det = F[0] * F[3] - F[1] * F[2];
      Finv[0] = F[3] / det; Finv[1] = -F[1] / det; Finv[2] = -F[2] / det; Finv[3] = F[0] / det;
    } else {
      const double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                         F[2] * (F[3] * F[7] - F[4] * F[6]);
      Finv[0] = (F[4] * F[8] - F[5] * F[7]) / det; Finv[1] = (F[2] * F[7] - F[1] * F[8]) / det;
      Finv[2] = (F[1] * F[5] - F[2] * F[4]) / det; Finv[3] = (F[5] * F[6] - F[3] * F[8]) / det;
      Finv[4] = (F[0] * F[8] - F[2] * F[6]) / det; Finv[5] = (F[2] * F[3] - F[0] * F[5]) / det;
      Finv[6] = (F[3] * F[7] - F[4] * F[6]) / det; Finv[7] = (F[1] * F[6] - F[0] * F[7]) / det;
      Finv[8] = (F[0] * F[4] - F[1] * F[3]) / det;
    }
    double phi[8], g[8][3];
    for (int v = 0; v < nv; ++v) {
      const int b = v & 3;
      const int a[3] = {(b == 1 || b == 2), b >= 2, v >= 4};
      double N = 1.0;
      for (int m = 0; m < D; ++m) N *= a[m] ? xi[m] : 1.0 - xi[m];
      phi[v] = N;
      double gr[3];
      for (int c = 0; c < D; ++c) {
        double t = 1.0;
        for (int m = 0; m < D; ++m) t *= (m == c) ? (a[m] ? 1.0 : -1.0) : (a[m] ? xi[m] : 1.0 - xi[m]);
        gr[c] = t;
      }
      for (int r = 0; r < D; ++r) {
        double acc = 0.0;
        for (int c = 0; c < D; ++c) acc += Finv[c * D + r] * gr[c];
        g[v][r] = acc;
      }
    }
    const double wq = ad * w[q];
    for (int a = 0; a < nv; ++a)
      for (int b = 0; b < nv; ++b) {
        double gg = 0.0;
        for (int r = 0; r < D; ++r) gg += g[a][r] * g[b][r];
        out[a * nv + b] += wq * (((mask & 1) ? phi[a] * phi[b] : 0.0) + ((mask & 2) ? gg : 0.0));
      }
  }
}

/* ====================================================================== */
/* H1 Lagrange basis of degree p on the D-simplex                          */
/* dof order: vertex nodes, then edge nodes (edges (0,1),(0,2),(0,3),(1,2),*/
/* (1,3),(2,3) restricted to the simplex, nodes from the lower vertex on),  */
/* then face nodes (faces in lexicographic order), then interior nodes.    */
/* Each node is a barycentric multi-index alpha with |alpha| = p.          */
/* ====================================================================== */

static int binom(int n, int k)
{
  if (k < 0 || k > n) return 0;
  int r = 1;
  for (int i = 1; i <= k; ++i) r = r * (n - k + i) / i;
  return r;
}

int or_h1_ndof(int D, int p) { return binom(p + D, D); }

static int h1_multiindices(int D, int p, int (*al)[4]);
static int g_mi[4][9][128][4];
static int g_mi_ok[4][9];
static pthread_mutex_t g_mi_mtx = PTHREAD_MUTEX_INITIALIZER;

static const int (*mi_table(int D, int p))[4]
{
  if (!__atomic_load_n(&g_mi_ok[D][p], __ATOMIC_ACQUIRE)) {
    pthread_mutex_lock(&g_mi_mtx);
    if (!g_mi_ok[D][p]) {
      h1_multiindices(D, p, g_mi[D][p]);
      __atomic_store_n(&g_mi_ok[D][p], 1, __ATOMIC_RELEASE);
    }
    pthread_mutex_unlock(&g_mi_mtx);
  }
  return (const int (*)[4])g_mi[D][p];
}

/* enumerate multi-indices in dof order: support sets by size, then in
   lexicographic order; inside one support set, alpha descending
   lexicographically (so edge nodes run from the lower vertex on) */
static int h1_multiindices(int D, int p, int (*al)[4])
{
  /* all alpha in N^{D+1}, |alpha| = p, descending lexicographic order */
  int all[256][4], na = 0;
  int a[4] = {0, 0, 0, 0};
  for (a[0] = p; a[0] >= 0; --a[0])
    for (a[1] = (D >= 1 ? p - a[0] : 0); a[1] >= 0; --a[1])
      for (a[2] = (D >= 2 ? p - a[0] - a[1] : 0); a[2] >= 0; --a[2]) {
        int rest = p - a[0] - a[1] - a[2];
        a[3] = D >= 3 ? rest : 0;
        if (D < 3 && rest != 0) continue;
        for (int i = 0; i < 4; ++i) all[na][i] = a[i];
        ++na;
      }
  int cnt = 0;
  for (int s = 1; s <= D + 1; ++s) {
    int sup[4];
    for (int i = 0; i < s; ++i) sup[i] = i;
    while (1) {
      int mask = 0;
      for (int i = 0; i < s; ++i) mask |= 1 << sup[i];
      for (int k = 0; k < na; ++k) {
        int m = 0;
        for (int i = 0; i <= D; ++i)
          if (all[k][i] > 0) m |= 1 << i;
        if (m == mask) {
          for (int i = 0; i < 4; ++i) al[cnt][i] = all[k][i];
          ++cnt;
        }
      }
      int i = s - 1;
      while (i >= 0 && sup[i] == D + 1 - s + i) --i;
      if (i < 0) break;
      ++sup[i];
      for (int j = i + 1; j < s; ++j) sup[j] = sup[j - 1] + 1;
    }
  }
  return cnt;
}

void or_h1_node(int D, int p, int k, double *xi)
{
  const int (*al)[4] = mi_table(D, p);
  for (int m = 0; m < D; ++m) xi[m] = (double)al[k][m] / p;
}

/* L_m(t) = prod_{j<m} (p t - j)/(j+1) and derivative */
static void lag1d(int p, int m, double t, double *v, double *dv)
{
  double val = 1.0, der = 0.0;
  for (int j = 0; j < m; ++j) {
    double f = (p * t - j) / (j + 1.0);
    double df = p / (j + 1.0);
    der = der * f + val * df;
    val = val * f;
  }
  *v = val;
  *dv = der;
}

void or_h1_shape(int D, int p, const double *xi, double *phi, double *dphi)
{
  const int (*al)[4] = mi_table(D, p);
  int nd = or_h1_ndof(D, p);
  double lam[4];
  double s = 1.0;
  for (int m = 0; m < D; ++m) { lam[m] = xi[m]; s -= xi[m]; }
  lam[D] = s;
  for (int k = 0; k < nd; ++k) {
    double v[4], dv[4];
    for (int i = 0; i <= D; ++i) lag1d(p, al[k][i], lam[i], &v[i], &dv[i]);
    double val = 1.0;
    for (int i = 0; i <= D; ++i) val *= v[i];
    if (phi) phi[k] = val;
    if (dphi) {
      /* dphi/dlam_i */
      double dl[4];
      for (int i = 0; i <= D; ++i) {
        double prod = dv[i];
        for (int j = 0; j <= D; ++j)
          if (j != i) prod *= v[j];
        dl[i] = prod;
      }
      for (int m = 0; m < D; ++m) dphi[k * D + m] = dl[m] - dl[D];
    }
  }
}

/* ====================================================================== */
/* element matrix: elmat += sum_i w_i |det F| (phi phi^T + grad phi grad^T) */
/* symboliccutbfi.cpp:130-304                                              */
/* ====================================================================== */

void or_elmat_add(int D, int p, const double *V, const double *pts,
                  const double *w, int n, int mask, double *out)
{
  double F[9], Finv[9], absdet;
  or_jacobian(D, V, F, Finv, &absdet);
  int nd = or_h1_ndof(D, p);
  double phi[128], dphi[128 * 3], g[128 * 3];
  for (int q = 0; q < n; ++q) {
    double ww = w[q] * absdet;
    or_h1_shape(D, p, &pts[3 * q], phi, dphi);
    for (int k = 0; k < nd; ++k)
      for (int r = 0; r < D; ++r) {
        double acc = 0.0;
        for (int c = 0; c < D; ++c) acc += Finv[c * D + r] * dphi[k * D + c];
        g[k * D + r] = acc;
      }
    for (int a = 0; a < nd; ++a)
      for (int b = 0; b < nd; ++b) {
        double v = 0.0;
        if (mask & 1) v += phi[a] * phi[b];
        if (mask & 2)
          for (int r = 0; r < D; ++r) v += g[a * D + r] * g[b * D + r];
        out[a * nd + b] += ww * v;
      }
  }
}

double or_integrate_poly(int D, const double *V, const double *pts,
                         const double *w, int n, int nterms,
                         const double *coef, const int *expo)
{
  double F[9], Finv[9], absdet;
  or_jacobian(D, V, F, Finv, &absdet);
  double sum = 0.0;
  for (int q = 0; q < n; ++q) {
    double x[3] = {0, 0, 0};
    for (int r = 0; r < D; ++r) {
      double acc = V[D * D + r];
      for (int c = 0; c < D; ++c) acc += F[r * D + c] * pts[3 * q + c];
      x[r] = acc;
    }
    double f = 0.0;
    for (int t = 0; t < nterms; ++t) {
      double m = coef[t];
      for (int d = 0; d < 3; ++d)
        for (int e = 0; e < expo[3 * t + d]; ++e) m *= x[d];
      f += m;
    }
    sum += absdet * w[q] * f;
  }
  return sum;
}

typedef struct {
  int D, ne, dt, order, nterms;
  const int32_t *el2v; const double *coords, *lset_v, *coef; const int *expo;
  double *per_elem; double sum; int lo, hi, err;
} IntJob;

static void *int_worker(void *arg)
{
  IntJob *J = (IntJob *)arg;
  int D = J->D, cap = 4096;
  double *pts = (double *)malloc(sizeof(double) * 3 * cap), *w = (double *)malloc(sizeof(double) * cap);
  double sum = 0.0;
  for (int e = J->lo; e < J->hi; ++e) {
    double V[12], l[4], F[9], Finv[9], absdet;
    for (int k = 0; k <= D; ++k) {
      int v = J->el2v[e * (D + 1) + k];
      for (int d = 0; d < D; ++d) V[k * D + d] = J->coords[v * D + d];
      l[k] = J->lset_v[v];
    }
    or_jacobian(D, V, F, Finv, &absdet);
    int ns;
    int n = or_straight_cut_rule(D, l, J->dt, J->order, Finv, pts, w, cap, &ns);
    if (n < -1) { J->err = n; break; }
    double s = 0.0;
    if (n > 0) s = or_integrate_poly(D, V, pts, w, n, J->nterms, J->coef, J->expo);
    if (J->per_elem) J->per_elem[e] = s;
    sum += s;
  }
  J->sum = sum;
  free(pts); free(w);
  return NULL;
}

int or_integrate_mesh(int D, int ne, const int32_t *el2v, const double *coords,
                      const double *lset_v, int dt, int order, int nterms,
                      const double *coef, const int *expo, double *per_elem,
                      double *total, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  for (int d = 0; d <= D; ++d) get_rule(d, order);
  IntJob jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    IntJob *J = &jobs[t];
    J->D = D; J->ne = ne; J->dt = dt; J->order = order; J->nterms = nterms;
    J->el2v = el2v; J->coords = coords; J->lset_v = lset_v; J->coef = coef; J->expo = expo;
    J->per_elem = per_elem; J->sum = 0.0; J->err = 0;
    J->lo = (int)((long long)ne * t / nthreads);
    J->hi = (int)((long long)ne * (t + 1) / nthreads);
    if (nthreads > 1) pthread_create(&th[t], NULL, int_worker, J);
    else int_worker(J);
  }
  double s = 0.0;
  int err = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    s += jobs[t].sum;
    if (jobs[t].err) err = jobs[t].err;
  }
  *total = s;
  return err;
}

/* ====================================================================== */
/* threaded fused pipeline (the CPU baseline)                              */
/* ====================================================================== */

typedef struct {
  int D, p, order, nel, ndt, mask;
  const int32_t *elems, *el2v;
  const double *coords, *lset_v;
  const int *dts;
  double *out;
  int lo, hi, err;
} PipeJob;

static void *pipe_worker(void *arg)
{
  PipeJob *J = (PipeJob *)arg;
  int D = J->D, nd = or_h1_ndof(D, J->p);
  int cap = 4096;
  double *pts = (double *)malloc(sizeof(double) * 3 * cap);
  double *w = (double *)malloc(sizeof(double) * cap);
  for (int i = J->lo; i < J->hi; ++i) {
    int e = J->elems ? J->elems[i] : i;
    double V[12], l[4], F[9], Finv[9], absdet;
    for (int k = 0; k <= D; ++k) {
      int v = J->el2v[e * (D + 1) + k];
      for (int d = 0; d < D; ++d) V[k * D + d] = J->coords[v * D + d];
      l[k] = J->lset_v[v];
    }
    or_jacobian(D, V, F, Finv, &absdet);
    for (int t = 0; t < J->ndt; ++t) {
      double *o = J->out + ((size_t)i * J->ndt + t) * nd * nd;
      memset(o, 0, sizeof(double) * nd * nd);
      int ns;
      int n = or_straight_cut_rule(D, l, J->dts[t], J->order, Finv, pts, w, cap, &ns);
      if (n < -1) { J->err = n; goto done; }
      if (n > 0) or_elmat_add(D, J->p, V, pts, w, n, J->mask, o);
    }
  }
done:
  free(pts);
  free(w);
  return NULL;
}

int or_cut_pipeline(int D, int p, int order, int nel, const int32_t *elems,
                    const int32_t *el2v, const double *coords,
                    const double *lset_v, int ndt, const int *dts, int mask,
                    double *out, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  for (int d = 0; d <= D; ++d) get_rule(d, order);
  PipeJob jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    PipeJob *J = &jobs[t];
    J->D = D; J->p = p; J->order = order; J->nel = nel; J->ndt = ndt; J->mask = mask;
    J->elems = elems; J->el2v = el2v; J->coords = coords; J->lset_v = lset_v;
    J->dts = dts; J->out = out; J->err = 0;
    J->lo = (int)((long long)nel * t / nthreads);
    J->hi = (int)((long long)nel * (t + 1) / nthreads);
    if (nthreads > 1) pthread_create(&th[t], NULL, pipe_worker, J);
    else pipe_worker(J);
  }
  int err = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    if (jobs[t].err) err = jobs[t].err;
  }
  return err;
}

/* ====================================================================== */
/* subdivision rule for a tetrahedron (xintegration.cpp)                   */
/* ====================================================================== */

typedef struct {
  const double *c;  /* nodal values */
  int lo;           /* level set degree */
  int order;        /* rule order */
  const double *Finv;
  double absdet;    /* only a scale for the interface weight transform */
  /* materialised output */
  double *pts[3], *w[3];
  int cap, n[3];
  /* moment output */
  int q;             /* psi degree; 0 = materialise */
  double *mom[3];    /* NEG, POS, IF accumulators */
  int err;
} SubCtx;

static double lset_eval(const SubCtx *S, const double *x)
{
  double phi[10];
  or_h1_shape(3, S->lo, x, phi, NULL);
  int nd = or_h1_ndof(3, S->lo);
  double v = 0.0;
  for (int k = 0; k < nd; ++k) v += phi[k] * S->c[k];
  return v;
}

/* append a point to rule r (0 NEG, 1 POS, 2 IF); for r==IF, nrm is the
   reference-space unit normal and w the untransformed weight */
static void emit(SubCtx *S, int r, const double *x, double w, const double *nrm)
{
  if (r == OR_IF) {
    /* CutIntegrationRule, xintegration.cpp:1698-1716:
       w * | absdet * F^{-T} n | / absdet */
    double v[3];
    for (int i = 0; i < 3; ++i) {
      double acc = 0.0;
      for (int j = 0; j < 3; ++j) acc += (S->Finv ? S->Finv[j * 3 + i] : (i == j ? 1.0 : 0.0)) * nrm[j];
      v[i] = S->absdet * acc;
    }
    double len = norm3(v);
    w = w * len / S->absdet;
  }
  if (S->q > 0) {
    double psi[128];
    or_h1_shape(3, S->q, x, psi, NULL);
    int nd = or_h1_ndof(3, S->q);
    if (S->mom[r])
      for (int k = 0; k < nd; ++k) S->mom[r][k] += w * psi[k];
    return;
  }
  if (S->n[r] >= S->cap) { S->err = OR_ERR_CAP; return; }
  memcpy(&S->pts[r][3 * S->n[r]], x, sizeof(double) * 3);
  S->w[r][S->n[r]] = w;
  S->n[r]++;
}

/* FillSimplexWithRule<3>, xintegration.cpp:223-242 */
static void fill_simplex(SubCtx *S, const double *p0, const double *p1, const double *p2,
                         const double *p3, int r)
{
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; ++d) { a[d] = p1[d] - p0[d]; b[d] = p2[d] - p0[d]; c[d] = p3[d] - p0[d]; }
  const double tf = (fabs(det3(a, b, c)) / 6.0) * 6.0;
  const Rule *R = get_rule(3, S->order);
  const double *P[4] = {p0, p1, p2, p3};
  for (int k = 0; k < R->n; ++k) {
    double ow = 1.0;
    for (int m = 0; m < 3; ++m) ow -= R->xi[3 * k + m];
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = ow * p0[d];
    for (int m = 0; m < 3; ++m)
      for (int d = 0; d < 3; ++d) x[d] += R->xi[3 * k + m] * P[m + 1][d];
    emit(S, r, x, R->w[k] * tf, NULL);
  }
}

/* FillSimplexCoDim1WithRule<3>, xintegration.cpp:289-326 */
static void fill_codim1(SubCtx *S, const double *s0, const double *s1, const double *s2,
                        const double *pospoint)
{
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; ++d) { a[d] = s1[d] - s0[d]; b[d] = s2[d] - s0[d]; }
  cross3(a, b, c);
  const double tf = norm3(c);
  double na = norm3(a), nb = norm3(b);
  const double rel = na > nb ? na : nb;
  if (tf < 1e-14 * rel) return;
  for (int d = 0; d < 3; ++d) c[d] /= tf;
  double dd[3];
  for (int d = 0; d < 3; ++d) dd[d] = pospoint[d] - s0[d];
  if (dd[0] * c[0] + dd[1] * c[1] + dd[2] * c[2] >= 0)
    for (int d = 0; d < 3; ++d) c[d] *= -1.0;
  const Rule *R = get_rule(2, S->order);
  const double *P[3] = {s0, s1, s2};
  for (int k = 0; k < R->n; ++k) {
    double ow = 1.0;
    for (int m = 0; m < 2; ++m) ow -= R->xi[2 * k + m];
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = ow * s0[d];
    for (int m = 0; m < 2; ++m)
      for (int d = 0; d < 3; ++d) x[d] += R->xi[2 * k + m] * P[m + 1][d];
    emit(S, OR_IF, x, R->w[k] * tf, c);
  }
}

/* DecomposePrismIntoSimplices<3>, xdecompose.hpp:129-146 */
static void fill_prism(SubCtx *S, const double *v[6], int r)
{
  for (int i = 0; i < 3; ++i) fill_simplex(S, v[i], v[i + 1], v[i + 2], v[i + 3], r);
}

/* Simplex<3>::CheckIfCut, xdecompose.hpp:34-85 */
static int leaf_check(SubCtx *S, const double V[4][3])
{
  int haspos = 0, hasneg = 0;
  double sumneg = 0.0, sumpos = 0.0;
  for (int i = 0; i < 4; ++i) {
    double v = lset_eval(S, V[i]);
    if (v >= 0.0) { sumpos += v; haspos = 1; }
    else { sumneg -= v; hasneg = 1; }
  }
  double s = sumpos + sumneg;
  sumpos /= s;
  sumneg /= s;
  if (sumpos < 1e-14) haspos = 0;
  if (sumneg < 1e-14) hasneg = 0;
  if (haspos && hasneg) return OR_IF;
  if (haspos) return OR_POS;
  if (hasneg) return OR_NEG;
  return OR_ERR_SUBDIV;
}

/* CutSimplex<3>::MakeQuad, xintegration.cpp:1084-1388 */
static void cut_leaf(SubCtx *S, const double V[4][3])
{
  static const int edge[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
  double cut[4][3];
  const double *posp[8], *negp[8];
  int npos = 0, nneg = 0, ncut = 0;
  int posv[4], negv[4], nposv = 0, nnegv = 0;
  int v2c1[4] = {-1, -1, -1, -1}, v2c2[4] = {-1, -1, -1, -1};
  double vv[4];
  int zero[4];
  for (int j = 0; j < 4; ++j) {
    zero[j] = 0;
    vv[j] = lset_eval(S, V[j]);
    if (vv[j] > 0) { posp[npos++] = V[j]; posv[nposv++] = j; }
    else if (vv[j] < 0) { negp[nneg++] = V[j]; negv[nnegv++] = j; }
    else { posp[npos++] = V[j]; posv[nposv++] = j; zero[j] = 1; }
  }
  for (int j = 0; j < 6; ++j) {
    int lv = edge[j][0], rv = edge[j][1];
    double vl = vv[lv], vr = vv[rv];
    int hascut = (vl * vr < 0);
    if (zero[lv] && vr < 0) hascut = 1;
    if (zero[rv] && vl < 0) hascut = 1;
    if (hascut) {
      if (ncut >= 4) { S->err = OR_ERR_SUBDIV; return; }
      double cp = vl / (vl - vr);
      for (int d = 0; d < 3; ++d) cut[ncut][d] = (1 - cp) * V[lv][d] + cp * V[rv][d];
      if (v2c1[lv] == -1) v2c1[lv] = ncut; else v2c2[lv] = ncut;
      if (v2c1[rv] == -1) v2c1[rv] = ncut; else v2c2[rv] = ncut;
      ncut++;
    }
  }
  if (ncut == 3) {
    int minor_pos = nneg > npos; /* minor group is pos */
    const double **minorg = minor_pos ? posp : negp;
    int nmin = minor_pos ? npos : nneg;
    int dt_minor = minor_pos ? OR_POS : OR_NEG;
    int major_pos = nneg <= npos;
    const double **majorg = major_pos ? posp : negp;
    int nmaj = major_pos ? npos : nneg;
    int dt_major = major_pos ? OR_POS : OR_NEG;
    const int *majv = (nnegv > nposv) ? negv : posv;
    const double *g4[4];
    for (int k = 0; k < nmin; ++k) g4[k] = minorg[k];
    for (int k = 0; k < 3; ++k) g4[nmin + k] = cut[k];
    if (nmin + 3 != 4) { S->err = OR_ERR_SUBDIV; return; }
    fill_simplex(S, g4[0], g4[1], g4[2], g4[3], dt_minor);
    const double *g6[6];
    for (int k = 0; k < nmaj; ++k) g6[k] = majorg[k];
    for (int k = 0; k < 3; ++k) g6[nmaj + k] = cut[v2c1[majv[k]]];
    if (nmaj + 3 != 6) { S->err = OR_ERR_SUBDIV; return; }
    fill_prism(S, g6, dt_major);
    fill_codim1(S, cut[0], cut[1], cut[2], posp[0]);
  } else if (ncut == 4) {
    for (int side = 0; side < 2; ++side) {
      /* side 0: POS prism (uses negv[0] for the orientation switch),
         side 1: NEG prism */
      const double **own = side == 0 ? posp : negp;
      const int *ownv = side == 0 ? posv : negv;
      const int *othv = side == 0 ? negv : posv;
      const double *pr[6];
      pr[0] = own[0];
      int idxn = ownv[0];
      int c1 = v2c1[idxn], c2 = v2c2[idxn];
      pr[1] = cut[c1];
      pr[2] = cut[c2];
      pr[3] = own[1];
      int c3 = -1;
      for (int l = 0; l < 4; ++l)
        if (c1 != l && c2 != l) c3 = l;
      int c4 = 6 - c3 - c2 - c1;
      if ((v2c1[othv[0]] == c1 && v2c2[othv[0]] == c4) ||
          (v2c1[othv[0]] == c2 && v2c2[othv[0]] == c3)) {
        int tt = c3; c3 = c4; c4 = tt;
      }
      pr[4] = cut[c3];
      pr[5] = cut[c4];
      fill_prism(S, pr, side == 0 ? OR_POS : OR_NEG);
    }
    int diag1, diag2, ndiag1, ndiag2;
    if (v2c1[negv[0]] == v2c1[posv[0]] || v2c2[negv[0]] == v2c1[posv[0]]) {
      diag1 = v2c1[posv[0]]; ndiag1 = v2c2[posv[0]];
    } else {
      diag1 = v2c2[posv[0]]; ndiag1 = v2c1[posv[0]];
    }
    if (v2c1[negv[1]] == v2c1[posv[1]] || v2c2[negv[1]] == v2c1[posv[1]]) {
      diag2 = v2c1[posv[1]]; ndiag2 = v2c2[posv[1]];
    } else {
      diag2 = v2c2[posv[1]]; ndiag2 = v2c1[posv[1]];
    }
    fill_codim1(S, cut[diag1], cut[ndiag1], cut[diag2], posp[0]);
    fill_codim1(S, cut[diag2], cut[ndiag2], cut[diag1], posp[0]);
  } else {
    S->err = OR_ERR_SUBDIV;
  }
}

/* NumericalIntegrationStrategy::CheckIfCut, xintegration.cpp:617-752 */
static int lattice_check(SubCtx *S, const double V[4][3], int lvl)
{
  int haspos = 0, hasneg = 0;
  int np1 = 1 << lvl;
  double dx = 1.0 / np1;
  int I[3] = {0, 0, 0};
  while (1) {
    double pos[3];
    for (int d = 0; d < 3; ++d) pos[d] = V[0][d];
    for (int j = 0; j < 3; ++j)
      for (int d = 0; d < 3; ++d) pos[d] += I[j] * dx * (V[j + 1][d] - V[0][d]);
    double v = lset_eval(S, pos);
    if (v >= 0.0) haspos = 1; else hasneg = 1;
    if (haspos && hasneg) return OR_IF;
    I[0]++;
    int finish = 0;
    for (int cd = 0; cd < 3; ++cd) {
      int sum = I[0] + I[1] + I[2];
      if (sum >= np1 + 1) {
        if (cd == 2) { finish = 1; break; }
        I[cd] = 0;
        I[cd + 1]++;
      } else
        break;
    }
    if (finish) break;
  }
  return haspos ? OR_POS : OR_NEG;
}

/* NumericalIntegrationStrategy::MakeQuadRule, xintegration.cpp:755-1063 */
static int make_quad_rule(SubCtx *S, const double V[4][3], int lvl)
{
  static const double baryc[10][4] = {
      {0.0, 0.0, 0.0, 1.0}, {0.5, 0.0, 0.0, 0.5}, {1.0, 0.0, 0.0, 0.0}, {0.0, 0.5, 0.0, 0.5},
      {0.5, 0.5, 0.0, 0.0}, {0.0, 1.0, 0.0, 0.0}, {0.0, 0.0, 0.5, 0.5}, {0.5, 0.0, 0.5, 0.0},
      {0.0, 0.5, 0.5, 0.0}, {0.0, 0.0, 1.0, 0.0}};
  static const int tets[8][4] = {{1, 2, 4, 7}, {3, 4, 5, 8}, {6, 7, 8, 9}, {3, 4, 8, 6},
                                 {3, 4, 1, 6}, {0, 1, 3, 6}, {8, 6, 1, 7}, {8, 4, 1, 7}};
  int dt_self = lattice_check(S, V, lvl);
  if (dt_self == OR_IF) {
    if (lvl > 0) {
      for (int i = 0; i < 8 && !S->err; ++i) {
        double nv[4][3];
        for (int j = 0; j < 4; ++j) {
          for (int c = 0; c < 3; ++c) nv[j][c] = 0.0;
          for (int d = 0; d < 4; ++d)
            for (int c = 0; c < 3; ++c) nv[j][c] += baryc[tets[i][j]][d] * V[d][c];
        }
        make_quad_rule(S, nv, lvl - 1);
      }
    } else {
      int dts = leaf_check(S, V);
      if (dts < 0) { S->err = dts; return OR_IF; }
      if (dts == OR_IF)
        cut_leaf(S, V);
      else
        fill_simplex(S, V[0], V[1], V[2], V[3], dts);
    }
    return OR_IF;
  }
  /* uncut: xintegration.cpp:1019-1061 */
  double a[3], b[3], c[3];
  for (int d = 0; d < 3; ++d) { a[d] = V[1][d] - V[0][d]; b[d] = V[2][d] - V[0][d]; c[d] = V[3][d] - V[0][d]; }
  double tf = fabs(det3(a, b, c));
  const Rule *R = get_rule(3, S->order);
  for (int k = 0; k < R->n; ++k) {
    double ow = 1.0;
    for (int m = 0; m < 3; ++m) ow -= R->xi[3 * k + m];
    double x[3];
    for (int d = 0; d < 3; ++d) x[d] = ow * V[0][d];
    for (int m = 0; m < 3; ++m)
      for (int d = 0; d < 3; ++d) x[d] += R->xi[3 * k + m] * V[m + 1][d];
    double wt = 1.0 * 1.0 * R->w[k] * tf;
    emit(S, dt_self, x, wt, NULL);
  }
  return dt_self;
}

static int subdiv_run(SubCtx *S, int L)
{
  double V[4][3];
  ref_vertices(3, V);
  get_rule(3, S->order);
  get_rule(2, S->order);
  if (S->Finv) {
    /* |det F| from F^{-1} */
    const double *M = S->Finv;
    double dinv = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                  M[2] * (M[3] * M[7] - M[4] * M[6]);
    S->absdet = 1.0 / fabs(dinv);
  } else
    S->absdet = 1.0;
  int dom = make_quad_rule(S, V, L);
  if (S->err) return S->err;
  return dom;
}

int or_subdiv_rule3(const double *lset_nodal, int lset_order, int L, int order,
                    const double *Finv, double *pts[3], double *w[3], int cap, int n[3])
{
  SubCtx S;
  memset(&S, 0, sizeof(S));
  S.c = lset_nodal; S.lo = lset_order; S.order = order; S.Finv = Finv; S.cap = cap;
  for (int r = 0; r < 3; ++r) { S.pts[r] = pts[r]; S.w[r] = w[r]; }
  int dom = subdiv_run(&S, L);
  for (int r = 0; r < 3; ++r) n[r] = S.n[r];
  return dom;
}

int or_subdiv_moments3(const double *lset_nodal, int lset_order, int L, int order, int q,
                       const double *Finv, double *out_neg, double *out_if)
{
  SubCtx S;
  memset(&S, 0, sizeof(S));
  S.c = lset_nodal; S.lo = lset_order; S.order = order; S.Finv = Finv; S.q = q;
  int nd = or_h1_ndof(3, q);
  if (out_neg) memset(out_neg, 0, sizeof(double) * nd);
  if (out_if) memset(out_if, 0, sizeof(double) * nd);
  S.mom[OR_NEG] = out_neg;
  S.mom[OR_POS] = NULL;
  S.mom[OR_IF] = out_if;
  return subdiv_run(&S, L);
}

typedef struct {
  const double *ls; int lo, L, order, q, nd, nl, lo_i, hi_i, err;
  double *on, *oi;
} SubJob;

static void *sub_worker(void *arg)
{
  SubJob *J = (SubJob *)arg;
  for (int i = J->lo_i; i < J->hi_i; ++i) {
    int r = or_subdiv_moments3(J->ls + (size_t)i * J->nl, J->lo, J->L, J->order, J->q, NULL,
                               J->on ? J->on + (size_t)i * J->nd : NULL,
                               J->oi ? J->oi + (size_t)i * J->nd : NULL);
    if (r < 0) { J->err = r; break; }
  }
  return NULL;
}

int or_subdiv_moments3_batch(int nel, const double *lset_nodal, int lset_order, int L,
                             int order, int q, double *out_neg, double *out_if, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  get_rule(3, order); get_rule(2, order);
  SubJob jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    SubJob *J = &jobs[t];
    J->ls = lset_nodal; J->lo = lset_order; J->L = L; J->order = order; J->q = q;
    J->nd = or_h1_ndof(3, q); J->nl = or_h1_ndof(3, lset_order);
    J->on = out_neg; J->oi = out_if; J->err = 0;
    J->lo_i = (int)((long long)nel * t / nthreads);
    J->hi_i = (int)((long long)nel * (t + 1) / nthreads);
    if (nthreads > 1) pthread_create(&th[t], NULL, sub_worker, J);
    else sub_worker(J);
  }
  int err = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    if (jobs[t].err) err = jobs[t].err;
  }
  return err;
}

/* Threaded subdivision pipeline (the CPU baseline of P2 level sets on a mesh,
   and the checker for sampled elements): per listed tetrahedron the
   subdivision rule of level L (CutIntegrationRule with subdivlvl = L,
   xintegration.cpp:1626-1754) is materialised for every domain type of dts,
   then CalcElementMatrix (symboliccutbfi.cpp:76-304) runs on it with the
   physical element map; interface weights carry the F^{-T} n transform of
   xintegration.cpp:1698-1716. */
typedef struct {
  int p, L, order, nel, ndt, mask, lo, hi, err, cap;
  const int32_t *elems, *el2v;
  const double *coords, *lset_nodal;
  const int *dts;
  double *out;
} SubPipeJob;

static void *subpipe_worker(void *arg)
{
  SubPipeJob *J = (SubPipeJob *)arg;
  const int nd = or_h1_ndof(3, J->p);
  double *P[3], *Wt[3];
  for (int r = 0; r < 3; ++r) {
    P[r] = (double *)malloc(sizeof(double) * 3 * (size_t)J->cap);
    Wt[r] = (double *)malloc(sizeof(double) * (size_t)J->cap);
  }
  for (int i = J->lo; i < J->hi; ++i) {
    const int e = J->elems ? J->elems[i] : i;
    double V[12], F[9], Finv[9], absdet;
    for (int k = 0; k < 4; ++k) {
      const int v = J->el2v[(size_t)e * 4 + k];
      for (int d = 0; d < 3; ++d) V[k * 3 + d] = J->coords[(size_t)v * 3 + d];
    }
    or_jacobian(3, V, F, Finv, &absdet);
    int n[3];
    const int dom = or_subdiv_rule3(J->lset_nodal + (size_t)e * 10, 2, J->L, J->order, Finv, P, Wt, J->cap, n);
    if (dom < 0) { J->err = dom; break; }
    for (int t = 0; t < J->ndt; ++t) {
      double *o = J->out + ((size_t)i * J->ndt + t) * nd * nd;
      memset(o, 0, sizeof(double) * nd * nd);
      const int r = J->dts[t];
      if (n[r] > 0) or_elmat_add(3, J->p, V, P[r], Wt[r], n[r], J->mask, o);
    }
  }
  for (int r = 0; r < 3; ++r) { free(P[r]); free(Wt[r]); }
  return NULL;
}

int or_subdiv_pipeline(int p, int L, int order, int nel, const int32_t *elems, const int32_t *el2v,
                       const double *coords, const double *lset_nodal, int ndt, const int *dts, int mask,
                       double *out, int nthreads)
{
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  get_rule(3, order);
  get_rule(2, order);
  SubPipeJob jobs[256];
  pthread_t th[256];
  for (int t = 0; t < nthreads; ++t) {
    SubPipeJob *J = &jobs[t];
    J->p = p; J->L = L; J->order = order; J->nel = nel; J->ndt = ndt; J->mask = mask;
    J->elems = elems; J->el2v = el2v; J->coords = coords; J->lset_nodal = lset_nodal;
    J->dts = dts; J->out = out; J->err = 0; J->cap = 400000;
    J->lo = (int)((long long)nel * t / nthreads);
    J->hi = (int)((long long)nel * (t + 1) / nthreads);
    if (nthreads > 1) pthread_create(&th[t], NULL, subpipe_worker, J);
    else subpipe_worker(J);
  }
  int err = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (nthreads > 1) pthread_join(th[t], NULL);
    if (jobs[t].err) err = jobs[t].err;
  }
  return err;
}

/* ====================================================================== */
/* MLP (FP64 oracle for the learned rule's kernels)                        */
/* ====================================================================== */

int or_mlp_nparams(int I, int H, int O) { return H * I + H + H * H + H + O * H + O; }

static void mlp_fwd1(int I, int H, int O, const double *P, const double *x,
                     double *h1, double *h2, double *y)
{
  const double *W1 = P, *b1 = W1 + H * I, *W2 = b1 + H, *b2 = W2 + H * H, *W3 = b2 + H, *b3 = W3 + O * H;
  for (int j = 0; j < H; ++j) {
    double a = b1[j];
    for (int i = 0; i < I; ++i) a += W1[j * I + i] * x[i];
    h1[j] = tanh(a);
  }
  for (int j = 0; j < H; ++j) {
    double a = b2[j];
    for (int i = 0; i < H; ++i) a += W2[j * H + i] * h1[i];
    h2[j] = tanh(a);
  }
  for (int o = 0; o < O; ++o) {
    double a = b3[o];
    for (int i = 0; i < H; ++i) a += W3[o * H + i] * h2[i];
    y[o] = a;
  }
}

void or_mlp_forward(int I, int H, int O, const double *params, int batch,
                    const double *x, double *y)
{
  double *h1 = (double *)malloc(sizeof(double) * H), *h2 = (double *)malloc(sizeof(double) * H);
  for (int b = 0; b < batch; ++b) mlp_fwd1(I, H, O, params, x + (size_t)b * I, h1, h2, y + (size_t)b * O);
  free(h1);
  free(h2);
}

double or_mlp_loss_grad(int I, int H, int O, const double *params, int batch,
                        const double *x, const double *target, double *grad)
{
  const double *W2 = params + H * I + H, *W3 = W2 + H * H + H;
  int np = or_mlp_nparams(I, H, O);
  memset(grad, 0, sizeof(double) * np);
  double *gW1 = grad, *gb1 = gW1 + H * I, *gW2 = gb1 + H, *gb2 = gW2 + H * H, *gW3 = gb2 + H, *gb3 = gW3 + O * H;
  double *h1 = (double *)malloc(sizeof(double) * H), *h2 = (double *)malloc(sizeof(double) * H);
  double *y = (double *)malloc(sizeof(double) * O), *dy = (double *)malloc(sizeof(double) * O);
  double *d2 = (double *)malloc(sizeof(double) * H), *d1 = (double *)malloc(sizeof(double) * H);
  double loss = 0.0;
  for (int b = 0; b < batch; ++b) {
    const double *xb = x + (size_t)b * I, *tb = target + (size_t)b * O;
    mlp_fwd1(I, H, O, params, xb, h1, h2, y);
    for (int o = 0; o < O; ++o) {
      double r = y[o] - tb[o];
      loss += 0.5 * r * r;
      dy[o] = r / batch;
    }
    for (int o = 0; o < O; ++o) {
      gb3[o] += dy[o];
      for (int i = 0; i < H; ++i) gW3[o * H + i] += dy[o] * h2[i];
    }
    for (int i = 0; i < H; ++i) {
      double a = 0.0;
      for (int o = 0; o < O; ++o) a += W3[o * H + i] * dy[o];
      d2[i] = a * (1.0 - h2[i] * h2[i]);
    }
    for (int j = 0; j < H; ++j) {
      gb2[j] += d2[j];
      for (int i = 0; i < H; ++i) gW2[j * H + i] += d2[j] * h1[i];
    }
    for (int i = 0; i < H; ++i) {
      double a = 0.0;
      for (int j = 0; j < H; ++j) a += W2[j * H + i] * d2[j];
      d1[i] = a * (1.0 - h1[i] * h1[i]);
    }
    for (int j = 0; j < H; ++j) {
      gb1[j] += d1[j];
      for (int i = 0; i < I; ++i) gW1[j * I + i] += d1[j] * xb[i];
    }
  }
  free(h1); free(h2); free(y); free(dy); free(d2); free(d1);
  return loss / batch;
}

/* ---- order-0 subdivision rule on a triangle (boundary elements of a
   tetrahedral mesh with a P2 level set): the part areas that
   CutInformation::Update sums (xfem/cutinfo.cpp:80-124) from
   NumericalIntegrationStrategy<ET_TRIG>::MakeQuadRule
   (cutint/xintegration.cpp:620-755 lattice check, :797-838 four children,
   :1019-1061 uncut nodes) and CutSimplex<2>::MakeQuad (:1391-1545) on the
   cut leaves after Simplex::CheckIfCut (cutint/xdecompose.hpp:34-85).
   Reference triangle (1,0), (0,1), (0,0); nodal values in the dof order
   vertices, edges (0,1), (0,2), (1,2). */
typedef struct { const double *c; int lo; double part[2]; int err; } Sub2;

static double lset_eval2(const Sub2 *S, const double *x)
{
  double phi[6];
  or_h1_shape(2, S->lo, x, phi, NULL);
  int nd = or_h1_ndof(2, S->lo);
  double v = 0.0;
  for (int k = 0; k < nd; ++k) v += phi[k] * S->c[k];
  return v;
}

static double area2(const double *a, const double *b, const double *c)
{
  double u0 = b[0] - a[0], u1 = b[1] - a[1], w0 = c[0] - a[0], w1 = c[1] - a[1];
  return 0.5 * fabs(u0 * w1 - u1 * w0);
}

static int lattice_check2(const Sub2 *S, const double V[3][2], int lvl)
{
  int haspos = 0, hasneg = 0, np1 = 1 << lvl;
  double dx = 1.0 / np1;
  for (int i1 = 0; i1 <= np1; ++i1)
    for (int i0 = 0; i0 + i1 <= np1; ++i0) {
      int I[2] = {i0, i1};
      double pos[2] = {V[0][0], V[0][1]};
      for (int j = 0; j < 2; ++j)
        for (int d = 0; d < 2; ++d) pos[d] += I[j] * dx * (V[j + 1][d] - V[0][d]);
      double v = lset_eval2(S, pos);
      if (v >= 0.0) haspos = 1; else hasneg = 1;
      if (haspos && hasneg) return OR_IF;
    }
  return haspos ? OR_POS : OR_NEG;
}

static void cut_leaf2(Sub2 *S, const double V[3][2])
{
  static const int edge[3][2] = {{0, 1}, {0, 2}, {1, 2}};
  double vv[3], cp[2][2];
  int zero[3], pos[3], npos = 0, neg[3], nneg = 0, v2c[3] = {-1, -1, -1}, nc = 0;
  for (int j = 0; j < 3; ++j) {
    vv[j] = lset_eval2(S, V[j]);
    zero[j] = vv[j] == 0.0;
    if (vv[j] < 0.0) neg[nneg++] = j; else pos[npos++] = j;
  }
  for (int j = 0; j < 3; ++j) {
    int l = edge[j][0], r = edge[j][1];
    int cut = vv[l] * vv[r] < 0.0 || (zero[l] && vv[r] < 0.0) || (zero[r] && vv[l] < 0.0);
    if (!cut) continue;
    if (nc >= 2) { S->err = OR_ERR_SUBDIV; return; }
    double t = vv[l] / (vv[l] - vv[r]);
    for (int d = 0; d < 2; ++d) cp[nc][d] = (1 - t) * V[l][d] + t * V[r][d];
    if (v2c[l] == -1) v2c[l] = nc;
    if (v2c[r] == -1) v2c[r] = nc;
    nc++;
  }
  if (nc != 2) { S->err = OR_ERR_SUBDIV; return; }
  int minor_neg = nneg <= npos;           /* the one-vertex side */
  const int *mi = minor_neg ? neg : pos, *ma = minor_neg ? pos : neg;
  double a_min = area2(V[mi[0]], cp[0], cp[1]);
  /* DecomposePrismIntoSimplices<2> on (major0, major1, cut(major0), cut(major1)) */
  const double *q[4] = {V[ma[0]], V[ma[1]], cp[v2c[ma[0]]], cp[v2c[ma[1]]]};
  double a_maj = area2(q[0], q[1], q[2]) + area2(q[1], q[2], q[3]);
  S->part[minor_neg ? 0 : 1] += a_min;
  S->part[minor_neg ? 1 : 0] += a_maj;
}

static void make_quad_rule2(Sub2 *S, const double V[3][2], int lvl)
{
  static const double baryc[6][3] = {{0.0, 0.0, 1.0}, {0.5, 0.0, 0.5}, {1.0, 0.0, 0.0},
                                     {0.0, 0.5, 0.5}, {0.5, 0.5, 0.0}, {0.0, 1.0, 0.0}};
  static const int trigs[4][3] = {{0, 1, 3}, {1, 2, 4}, {1, 3, 4}, {3, 4, 5}};
  int dt = lattice_check2(S, V, lvl);
  if (dt == OR_IF) {
    if (lvl > 0) {
      for (int i = 0; i < 4 && !S->err; ++i) {
        double nv[3][2];
        for (int j = 0; j < 3; ++j)
          for (int c = 0; c < 2; ++c) {
            nv[j][c] = 0.0;
            for (int d = 0; d < 3; ++d) nv[j][c] += baryc[trigs[i][j]][d] * V[d][c];
          }
        make_quad_rule2(S, nv, lvl - 1);
      }
      return;
    }
    /* leaf: Simplex::CheckIfCut with its relative 1e-14 threshold */
    int haspos = 0, hasneg = 0;
    double sp = 0.0, sn = 0.0;
    for (int i = 0; i < 3; ++i) {
      double v = lset_eval2(S, V[i]);
      if (v >= 0.0) { sp += v; haspos = 1; } else { sn -= v; hasneg = 1; }
    }
    double s = sp + sn;
    sp /= s;
    sn /= s;
    if (sp < 1e-14) haspos = 0;
    if (sn < 1e-14) hasneg = 0;
    if (haspos && hasneg) { cut_leaf2(S, V); return; }
    if (!haspos && !hasneg) { S->err = OR_ERR_SUBDIV; return; }
    S->part[haspos ? 1 : 0] += area2(V[0], V[1], V[2]);
    return;
  }
  S->part[dt == OR_NEG ? 0 : 1] += area2(V[0], V[1], V[2]);
}

int or_subdiv_trig_parts(const double *lset_nodal, int lset_order, int L, double part[2])
{
  Sub2 S;
  memset(&S, 0, sizeof(S));
  S.c = lset_nodal;
  S.lo = lset_order;
  double V[3][2] = {{1.0, 0.0}, {0.0, 1.0}, {0.0, 0.0}};
  int dom = lattice_check2(&S, V, L);
  make_quad_rule2(&S, V, L);
  part[0] = S.part[0];
  part[1] = S.part[1];
  return S.err ? S.err : dom;
}
