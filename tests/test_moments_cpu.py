"""The generated moment identities (ngsxfem_amd/csrc/moments_table.h) on the
host: compiled with g++ into a throw-away library and compared with the FP64
oracle's direct quadrature sum (or_elmat_add, symboliccutbfi.cpp:229-287) on
random elements and random (inexact) point sets, for D = 1..3, p = 1..3 and
the three form masks.  The identities hold point by point, so the agreement
is to rounding (1e-12), whatever the points."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "ngsxfem_amd", "csrc")

HARNESS = r"""
#include <cmath>
#define XG_DEV inline
#define XG_MOM_NO_NODAL
#define XG_AFF_STORAGE static
#include "moments_table.h"
extern "C" int mom_elmat(int D, int P, const double *V, const double *pts, const double *w, int n, int mask,
                         double *out) {
  double F[9], Finv[9], det;
  for (int r = 0; r < D; ++r)
    for (int c = 0; c < D; ++c) F[r * D + c] = V[c * D + r] - V[D * D + r];
  if (D == 1) {
    det = F[0];
    Finv[0] = 1 / det;
  } else if (D == 2) {
    det = F[0] * F[3] - F[1] * F[2];
    Finv[0] = F[3] / det; Finv[1] = -F[1] / det; Finv[2] = -F[2] / det; Finv[3] = F[0] / det;
  } else {
    double a = F[0], b = F[1], c = F[2], d = F[3], e = F[4], f = F[5], g = F[6], h = F[7], i = F[8];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    det = a * A + b * B + c * C;
    Finv[0] = A / det; Finv[1] = -(b * i - c * h) / det; Finv[2] = (b * f - c * e) / det;
    Finv[3] = B / det; Finv[4] = (a * i - c * g) / det; Finv[5] = -(a * f - c * d) / det;
    Finv[6] = C / det; Finv[7] = -(a * h - b * g) / det; Finv[8] = (a * e - b * d) / det;
  }
  const double ad = std::fabs(det);
  // Gram matrix of the barycentric gradients (rows of F^{-1}, the last one
  // minus their sum)
  double G[4][3];
  for (int r = 0; r < D; ++r) {
    double s = 0;
    for (int i = 0; i < D; ++i) {
      G[i][r] = Finv[i * D + r];
      s += Finv[i * D + r];
    }
    G[D][r] = -s;
  }
  double Hs[10];
  int k = 0;
  for (int i = 0; i <= D; ++i)
    for (int j = i; j <= D; ++j) {
      double s = 0;
      for (int r = 0; r < D; ++r) s += G[i][r] * G[j][r];
      Hs[k++] = s;
    }
  double m[128] = {0}, mk[128] = {0};
  const double sm = (mask & 1) ? ad : 0, sk = (mask & 2) ? ad : 0;
#define CASE(DD, PP)                                          \
  if (D == DD && P == PP) {                                   \
    using M = xg::Moments<DD, PP>;                            \
    for (int q = 0; q < n; ++q) {                             \
      double l[4], s = 1;                                     \
      for (int d = 0; d < DD; ++d) {                          \
        l[d] = pts[q * 3 + d];                                \
        s -= l[d];                                            \
      }                                                       \
      l[DD] = s;                                              \
      M::accum(l, w[q], m);                                   \
    }                                                         \
    M::lower(m, mk);                                          \
    M::forms<double>(m, mk, Hs, sm, sk, out);                 \
    return 0;                                                 \
  }
  CASE(1, 1) CASE(1, 2) CASE(1, 3) CASE(2, 1) CASE(2, 2) CASE(2, 3) CASE(3, 1) CASE(3, 2) CASE(3, 3)
  return -1;
}
// Moments::complement on m (in place)
extern "C" int mom_complement(int D, int P, double *m) {
#define CCASE(DD, PP)                          \
  if (D == DD && P == PP) {                    \
    xg::Moments<DD, PP>::complement(m);        \
    return xg::Moments<DD, PP>::NM;            \
  }
  CCASE(1, 1) CCASE(1, 2) CCASE(1, 3) CCASE(2, 1) CCASE(2, 2) CCASE(2, 3) CCASE(3, 1) CCASE(3, 2) CCASE(3, 3)
  return -1;
}
// Moments::accum over a point set
extern "C" int mom_accum(int D, int P, const double *pts, const double *w, int n, double *m) {
#define ACASE(DD, PP)                                  \
  if (D == DD && P == PP) {                            \
    for (int q = 0; q < n; ++q) {                      \
      double l[4], s = 1;                              \
      for (int d = 0; d < DD; ++d) {                   \
        l[d] = pts[q * 3 + d];                         \
        s -= l[d];                                     \
      }                                                \
      l[DD] = s;                                       \
      xg::Moments<DD, PP>::accum(l, w[q], m);          \
    }                                                  \
    return xg::Moments<DD, PP>::NM;                    \
  }
  ACASE(1, 1) ACASE(1, 2) ACASE(1, 3) ACASE(2, 1) ACASE(2, 2) ACASE(2, 3) ACASE(3, 1) ACASE(3, 2) ACASE(3, 3)
  return -1;
}
// Moments::hstep over the nv vertices of one simplex (barycentrics lv,
// nv x (D+1)), every vertex scaled by s^(1/2P), then hfinish
extern "C" int mom_hsimplex(int D, int P, const double *lv, int nv, double s, double *m) {
#define HCASE(DD, PP)                                           \
  if (D == DD && P == PP) {                                     \
    using M = xg::Moments<DD, PP>;                              \
    double h[M::NH] = {0};                                      \
    const double c = std::pow(s, 1.0 / (2 * PP));               \
    for (int v = 0; v < nv; ++v) {                              \
      double l[4];                                              \
      for (int i = 0; i <= DD; ++i) l[i] = c * lv[v * (DD + 1) + i]; \
      M::hstep(l, h, m);                                        \
    }                                                           \
    M::hfinish(m);                                              \
    return M::NM;                                               \
  }
  HCASE(1, 1) HCASE(1, 2) HCASE(2, 1) HCASE(2, 2) HCASE(3, 1) HCASE(3, 2)
  return -1;
}
// Moments<3, 3>::accum_aff over a point set in centred coordinates
// y = x - 1/4, then the dense map xg_aff33 to the homogeneous moments
extern "C" int mom_accum_aff(const double *pts, const double *w, int n, double *m) {
  using M = xg::Moments<3, 3>;
  double nn[M::NM] = {0};
  for (int q = 0; q < n; ++q) {
    const double y[3] = {pts[3 * q] - 0.25, pts[3 * q + 1] - 0.25, pts[3 * q + 2] - 0.25};
    M::accum_aff(y, w[q], nn);
  }
  for (int a = 0; a < M::NM; ++a) {
    double s = 0;
    for (int r = 0; r < M::NM; ++r) s += (double)xg::xg_aff33[r][a % 64][a / 64] * nn[r];
    m[a] += s;
  }
  return M::NM;
}
"""


@pytest.fixture(scope="module")
def momlib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("mom")
    src = d / "momhost.cpp"
    src.write_text(HARNESS)
    so = d / "libmomhost.so"
    subprocess.run([gxx, "-O1", "-std=c++17", "-shared", "-fPIC", "-I" + CSRC, str(src), "-o", str(so)],
                   check=True, capture_output=True)
    lib = ctypes.CDLL(str(so))
    lib.mom_elmat.restype = ctypes.c_int
    lib.mom_hsimplex.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_void_p]
    return lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("P", [1, 2, 3])
def test_moment_identities_equal_direct_quadrature(orc, momlib, D, P):
    rng = np.random.default_rng(100 * D + P)
    for trial in range(4):
        V = rng.random((D + 1, D)) * 0.3 + np.vstack([np.eye(D), np.zeros(D)])
        V = np.ascontiguousarray(V)
        npt = 13
        pts = rng.random((npt, D))
        pts = pts / (pts.sum(1, keepdims=True) + rng.random((npt, 1)))
        w = rng.random(npt)
        pts3 = np.ascontiguousarray(np.hstack([pts, np.zeros((npt, 3 - D))]))
        for mask in (1, 2, 3):
            ref = orc.elmat(D, P, V, pts3, w, mask=mask)
            nd = ref.shape[0]
            out = np.zeros(nd * nd)
            assert momlib.mom_elmat(D, P, _p(V), _p(pts3), _p(w), npt, mask, _p(out)) == 0
            got = out.reshape(nd, nd)
            assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max(), (D, P, mask)


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("P", [1, 2])
def test_vertex_recursion_gives_exact_simplex_moments(momlib, D, P):
    """Moments::hstep / hfinish (the complete homogeneous recursion over a
    simplex's vertices) against (a) the closed form over the whole reference
    simplex (Moments::complement of zero: int lam^alpha = alpha! D! /
    (2P + D)! |T|) and (b) Grundmann-Moeller rules of degree 2P + 1 on random
    sub-simplices of dimension D and D - 1 (volume pieces and interface
    pieces), summed through Moments::accum (the product form, an independent
    code path)."""
    import math
    import sys
    sys.path.insert(0, HERE)
    from test_oracle_invariants import gm_rule
    n = 2 * P
    NM = momlib.mom_complement(D, P, _p(np.zeros(256)))
    full = np.zeros(NM)
    momlib.mom_complement(D, P, _p(full))
    got = np.zeros(NM)
    lv = np.ascontiguousarray(np.eye(D + 1)[[D] + list(range(D))])  # origin, e_1 .. e_D
    s = (1.0 / math.factorial(D)) * math.factorial(D) / math.factorial(n + D)
    momlib.mom_hsimplex(D, P, _p(lv), D + 1, s, _p(got))
    assert np.abs(got - full).max() <= 1e-15 * np.abs(full).max()
    rng = np.random.default_rng(7 * D + P)
    for k in ([D, D - 1] if D > 1 else [D]):
        for trial in range(6):
            lam = rng.random((k + 1, D + 1)) ** 2
            lam /= lam.sum(1, keepdims=True)  # vertices inside the element
            X = lam[:, :D]  # reference coordinates
            E = X[1:] - X[0]
            vol = math.sqrt(abs(np.linalg.det(E @ E.T))) / math.factorial(k)
            b, w = gm_rule(k, P)  # degree 2P + 1 on the unit k-simplex
            pts = b @ X
            pts3 = np.ascontiguousarray(np.hstack([pts, np.zeros((len(w), 3 - D))]))
            ww = np.ascontiguousarray(w * vol * math.factorial(k))
            ref = np.zeros(NM)
            momlib.mom_accum(D, P, _p(pts3), _p(ww), len(w), _p(ref))
            got = np.zeros(NM)
            s = vol * math.factorial(k) / math.factorial(n + k)
            momlib.mom_hsimplex(D, P, _p(np.ascontiguousarray(lam)), k + 1, s, _p(got))
            assert np.abs(got - ref).max() <= 1e-13 * np.abs(ref).max(), (D, P, k, trial)


def test_affine_accumulation_equals_homogeneous(momlib):
    """accum_aff (monomials of three centred affine coordinates y = lam - 1/4,
    degree <= 6, lam_3 never formed) followed by the dense map xg_aff33 gives
    the homogeneous degree-6 moments of accum: on random points with
    mixed-sign weights (complemented pieces), on points clustered at each
    vertex (where the map cancels most), and on the closed form of the whole
    reference tetrahedron (int lam^alpha = 3! alpha! / 9! |T|, a
    Grundmann-Moeller rule of degree 7)"""
    rng = np.random.default_rng(43)
    corners = np.vstack([np.eye(3), np.zeros(3)])
    for trial in range(9):
        n = 50
        pts = rng.random((n, 3))
        pts = pts / (pts.sum(1, keepdims=True) + rng.random((n, 1)))
        if trial >= 5:  # a small cluster at vertex trial - 5
            pts = corners[trial - 5] + 0.05 * (pts - corners[trial - 5])
        pts = np.ascontiguousarray(pts)
        w = np.ascontiguousarray(rng.random(n) - 0.3 * (trial % 2))
        ref = np.zeros(128)
        nm = momlib.mom_accum(3, 3, _p(pts), _p(w), n, _p(ref))
        got = np.zeros(128)
        assert momlib.mom_accum_aff(_p(pts), _p(w), n, _p(got)) == nm
        assert np.abs(got[:nm] - ref[:nm]).max() <= 1e-14 * np.abs(ref[:nm]).max(), (trial,)
    import sys
    sys.path.insert(0, HERE)
    from test_oracle_invariants import gm_rule
    b, w = gm_rule(3, 3)  # degree 7 on the unit tetrahedron (weights sum to 1/6)
    pts = np.ascontiguousarray(b[:, 1:])
    got = np.zeros(128)
    nm = momlib.mom_accum_aff(_p(pts), _p(np.ascontiguousarray(w)), len(w), _p(got))
    full = np.zeros(128)
    momlib.mom_complement(3, 3, _p(full))
    assert np.abs(got[:nm] - full[:nm]).max() <= 1e-14 * np.abs(full[:nm]).max()


def test_moment_table_is_current():
    """moments_table.h is what gen_moments.py generates"""
    import sys
    txt = subprocess.run([sys.executable, os.path.join(CSRC, "gen_moments.py")], capture_output=True, text=True,
                         check=True).stdout
    with open(os.path.join(CSRC, "moments_table.h")) as f:
        assert f.read() == txt


def test_nodal_table_maps_moments_to_fixed_node_weights(orc):
    """nodal34 (degree-4 Lagrange basis in homogeneous monomials): for random
    points, sum_i w_i psi_k(x_i) from the oracle's basis equals
    nodal34 @ (sum_i w_i lam_i^alpha)"""
    import re
    import sys
    sys.path.insert(0, CSRC)
    import gen_moments as G
    txt = open(os.path.join(CSRC, "moments_table.h")).read()
    body = txt[txt.index("nodal34[35][35] = {"):]
    body = body[:body.index("};")]
    rows = re.findall(r"\{([^{}]*)\}", body)
    A = np.array([[float(v) for v in r.split(",")] for r in rows])
    assert A.shape == (35, 35)
    mons = G.monomials(3, 4)
    rng = np.random.default_rng(7)
    pts = rng.random((20, 3))
    pts = pts / (pts.sum(1, keepdims=True) + rng.random((20, 1)))
    w = rng.random(20)
    m = np.zeros(35)
    W = np.zeros(35)
    for x, wi in zip(pts, w):
        lam = np.array([x[0], x[1], x[2], 1 - x.sum()])
        m += wi * np.array([np.prod(lam ** np.array(a)) for a in mons])
        phi, _ = orc.h1_shape(3, 4, x)
        W += wi * phi
    assert np.abs(A @ m - W).max() < 1e-12 * np.abs(W).max()


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("P", [1, 2, 3])
def test_complement_is_whole_simplex_minus(orc, momlib, D, P):
    """complement(0) = the moments of the whole reference simplex (the
    oracle's base rule of order 2P summed through accum), and for a straight
    cut the NEG moments equal complement(POS moments) (exact rules)"""
    pts, w = orc.rule(D, 2 * P)
    pts3 = np.ascontiguousarray(np.hstack([pts, np.zeros((len(w), 3 - D))]))
    full = np.zeros(128)
    nm = momlib.mom_accum(D, P, _p(pts3), _p(np.ascontiguousarray(w)), len(w), _p(full))
    got = np.zeros(128)
    assert momlib.mom_complement(D, P, _p(got)) == nm
    assert np.abs(got[:nm] - full[:nm]).max() <= 2e-15 * np.abs(full[:nm]).max()
    rng = np.random.default_rng(D * 10 + P)
    for trial in range(5):
        lset = rng.standard_normal(D + 1)
        lset[0], lset[1] = -abs(lset[0]) - 0.1, abs(lset[1]) + 0.1  # cut
        mom = {}
        for dt in (0, 1):
            r = orc.straight_cut_rule(D, lset, dt, 2 * P)
            p_, w_, _ = r
            p3 = np.ascontiguousarray(p_)
            mm = np.zeros(128)
            momlib.mom_accum(D, P, _p(p3), _p(np.ascontiguousarray(w_)), len(w_), _p(mm))
            mom[dt] = mm
        comp = mom[1].copy()
        momlib.mom_complement(D, P, _p(comp))
        assert np.abs(comp[:nm] - mom[0][:nm]).max() <= 1e-14 * np.abs(full[:nm]).max()
