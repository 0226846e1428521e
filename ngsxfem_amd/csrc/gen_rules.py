"""Generates rules_table.h: conical-product Gauss-Jacobi rules on the
reference simplices, plus degree-q Lagrange lattice data.

The reference takes its base rules from NGSolve's SelectIntegrationRule
(cutint/straightcutrule.cpp:87-90); NGSolve is not part of this build, so the
base rules are Stroud conical products of Gauss-Jacobi rules:

  segm: xi0 = t                                   t ~ Gauss-Legendre on [0,1]
  trig: xi0 = t1, xi1 = (1-t1) t2                 t1 ~ weight (1-t)
  tet : xi0 = t1, xi1 = (1-t1) t2, xi2 = (1-t1)(1-t2) t3,  t1 ~ (1-t)^2, t2 ~ (1-t)

with n = order/2 + 1 points per direction (exact to degree 2n-1).  Nodes are
eigenvalues of the Jacobi matrix, Newton-polished on the orthonormal
recurrence; weights by the Christoffel formula.  n = 1 uses the closed forms
t = 1/(a+2), w = 1/(a+1), which makes the order-0 weights (the reference
volumes 1, 1/2, 1/6) exact in FP64; marking relies on that.

Run:  python gen_rules.py > rules_table.h
"""
import numpy as np

NMAX = 8  # points per direction -> orders 0..15
LAT_LMAX = 5  # subdivision levels supported on the device


def gauss_jacobi01(n, a):
    if n == 1:
        return np.array([1.0 / (a + 2.0)]), np.array([1.0 / (a + 1.0)])
    al, be = float(a), 0.0
    d = np.zeros(n)
    e = np.zeros(n + 1)
    for k in range(n):
        if k == 0:
            d[k] = (be - al) / (al + be + 2.0)
        else:
            s = 2.0 * k + al + be
            d[k] = (be * be - al * al) / (s * (s + 2.0))
    for k in range(1, n + 1):
        s = 2.0 * k + al + be
        e[k] = np.sqrt(4.0 * k * (k + al) * (k + be) * (k + al + be) / (s * s * (s + 1.0) * (s - 1.0)))
    J = np.diag(d) + np.diag(e[1:n], 1) + np.diag(e[1:n], -1)
    x = np.sort(np.linalg.eigvalsh(J))
    mu0 = 2.0 ** (a + 1) / (a + 1)

    def ev(xx):
        pm1, p, dpm1, dp, ss = 0.0, 1.0 / np.sqrt(mu0), 0.0, 0.0, 0.0
        for k in range(n):
            ss += p * p
            pn = (xx - d[k]) * p - (e[k] * pm1 if k > 0 else 0.0)
            dpn = p + (xx - d[k]) * dp - (e[k] * dpm1 if k > 0 else 0.0)
            if k < n - 1:
                pn /= e[k + 1]
                dpn /= e[k + 1]
            pm1, p, dpm1, dp = p, pn, dp, dpn
        return p, dp, ss

    t = np.zeros(n)
    w = np.zeros(n)
    for i in range(n):
        xi = float(x[i])
        for _ in range(6):
            q, dq, _ss = ev(xi)
            if dq == 0.0:
                break
            step = q / dq
            xi -= step
            if abs(step) < 1e-17:
                break
        q, dq, ss = ev(xi)
        t[i] = (1.0 + xi) / 2.0
        w[i] = (1.0 / ss) / 2.0 ** (a + 1)
    return t, w


def conical(D, n):
    pts, ws = [], []
    if D == 0:
        return [(0.0, 0.0, 0.0)], [1.0]
    if D == 1:
        t, w = gauss_jacobi01(n, 0)
        for i in range(n):
            pts.append((t[i], 0.0, 0.0))
            ws.append(w[i])
    elif D == 2:
        t0, w0 = gauss_jacobi01(n, 1)
        t1, w1 = gauss_jacobi01(n, 0)
        for i in range(n):
            for j in range(n):
                pts.append((t0[i], (1.0 - t0[i]) * t1[j], 0.0))
                ws.append(w0[i] * w1[j])
    else:
        t0, w0 = gauss_jacobi01(n, 2)
        t1, w1 = gauss_jacobi01(n, 1)
        t2, w2 = gauss_jacobi01(n, 0)
        for i in range(n):
            for j in range(n):
                for k in range(n):
                    pts.append((t0[i], (1.0 - t0[i]) * t1[j], (1.0 - t0[i]) * (1.0 - t1[j]) * t2[k]))
                    ws.append(w0[i] * w1[j] * w2[k])
    return pts, ws


# Minimal symmetric rules on the triangle and the tetrahedron, exact for
# polynomials of degree 2, 4, 6 (3 / 6 / 12 and 4 / 14 / 24 points, positive
# weights, interior points).  Orbit parameters solved once from the
# moment equations (least squares to 1e-17 residual) and kept here as exact
# FP64 values; weights normalised to the simplex volume 1/D!.  The moment form
# of the element matrices uses them when the requested rule order is at
# least the integrand degree 2p: every exact rule gives the same moments.
MINRULES = {
    (2, 2): (['S21'], ['0x1.5555555555555p-3', '0x1.5555555555555p-3']),
    (2, 4): (['S21', 'S21'], ['0x1.c8a6b8a0bd0d9p-2', '0x1.c97c4971907c9p-4', '0x1.77189ea1db0c9p-4',
                              '0x1.c25cc272345c2p-5']),
    (2, 6): (['S21', 'S21', 'S111'], ['0x1.fe8a0c8eaec18p-3', '0x1.de5b492ddd0d9p-5', '0x1.0269a05fa5610p-4',
                                      '0x1.a0857f40e737fp-6', '0x1.b35d3f60e374ep-5', '0x1.3dcd086db1b8ep-2',
                                      '0x1.535ba6438255ep-5']),
    (3, 2): (['S31'], ['0x1.1b06d1d200913p-3', '0x1.5555555555555p-5']),
    (3, 4): (['S31', 'S31', 'S22'], ['0x1.7bd7f53458af5p-4', '0x1.915eb8b8ec6f6p-7', '0x1.3e58e0e01751bp-2',
                                     '0x1.33b691d134718p-6', '0x1.74c42f2220f9bp-5', '0x1.d0b74bf80003fp-8']),
    (3, 6): (['S31', 'S31', 'S31', 'S211'], ['0x1.b781b5cd7c6dap-3', '0x1.b41019c99bffbp-8', '0x1.4d3379368f7d7p-5',
                                             '0x1.b847b29c95d64p-10', '0x1.4a12f1a1b0de0p-2', '0x1.2e52f30703380p-7',
                                             '0x1.04c16611573a7p-4', '0x1.1424fbd0ff873p-2', '0x1.0750750750763p-7']),
}
ORBIT_NPAR = {'S21': 1, 'S111': 2, 'S31': 1, 'S22': 1, 'S211': 2}


def orbit_points(D, kind, par):
    """distinct permutations of the orbit's barycentric tuple, sorted"""
    import itertools
    if kind == 'S21':
        t = (par[0], par[0], 1.0 - 2.0 * par[0])
    elif kind == 'S111':
        t = (par[0], par[1], 1.0 - par[0] - par[1])
    elif kind == 'S31':
        t = (par[0], par[0], par[0], 1.0 - 3.0 * par[0])
    elif kind == 'S22':
        t = (par[0], par[0], 0.5 - par[0], 0.5 - par[0])
    else:  # S211
        t = (par[0], par[0], par[1], 1.0 - 2.0 * par[0] - par[1])
    return sorted(set(itertools.permutations(t)))


def minrule(D, deg):
    """points (reference coordinates = the first D barycentrics) and weights"""
    if D <= 1:
        return conical(D, deg // 2 + 1)
    kinds, xs = MINRULES[(D, deg)]
    x = [float.fromhex(v) for v in xs]
    pts, ws = [], []
    k = 0
    for kind in kinds:
        npar = ORBIT_NPAR[kind]
        par, w = x[k:k + npar], x[k + npar]
        k += npar + 1
        for b in orbit_points(D, kind, par):
            pts.append(tuple(b[:D]) + (0.0,) * (3 - D))
            ws.append(w)
    return pts, ws


def main():
    out = []
    out.append("// Generated by gen_rules.py -- do not edit.")
    out.append("// Conical-product Gauss-Jacobi rules: for dimension D in 0..3 and n in 1..%d" % NMAX)
    out.append("// points per direction (order = 2n-1 exactness), entries (xi0, xi1, xi2, w).")
    out.append("#pragma once")
    out.append("#define XG_RULE_NMAX %d" % NMAX)
    offs = {}
    rows = []
    for D in range(4):
        for n in range(1, NMAX + 1):
            pts, ws = conical(D, n)
            offs[(D, n)] = (len(rows), len(ws))
            for p, w in zip(pts, ws):
                rows.append((p[0], p[1], p[2], w))
    out.append("#define XG_RULE_OFF_INIT { \\")
    for D in range(4):
        out.append("  {0, " + ", ".join(str(offs[(D, n)][0]) for n in range(1, NMAX + 1)) + "}, \\")
    out.append("}")
    out.append("#define XG_RULE_CNT_INIT { \\")
    for D in range(4):
        out.append("  {0, " + ", ".join(str(offs[(D, n)][1]) for n in range(1, NMAX + 1)) + "}, \\")
    out.append("}")
    out.append("#define XG_RULE_ROWS %d" % len(rows))
    out.append("#define XG_RULE_DATA_INIT { \\")
    for r in rows:
        out.append("  {" + ", ".join(float(v).hex() for v in r) + "}, \\")
    out.append("}")
    out.append("static const int xg_rule_off[4][%d] = XG_RULE_OFF_INIT;" % (NMAX + 1))
    out.append("static const int xg_rule_cnt[4][%d] = XG_RULE_CNT_INIT;" % (NMAX + 1))
    out.append("static const double xg_rule_data[XG_RULE_ROWS][4] = XG_RULE_DATA_INIT;")
    out.append("#if defined(__HIP__)")
    out.append("static __device__ const int xg_rule_off_d[4][%d] = XG_RULE_OFF_INIT;" % (NMAX + 1))
    out.append("static __device__ const int xg_rule_cnt_d[4][%d] = XG_RULE_CNT_INIT;" % (NMAX + 1))
    out.append("static __device__ const double xg_rule_data_d[XG_RULE_ROWS][4] = XG_RULE_DATA_INIT;")
    out.append("#endif")
    # minimal exact rules of degree 2, 4, 6 ([D][deg / 2], deg / 2 = 1..3)
    moff, mcnt, mrows = {}, {}, []
    for D in range(4):
        for h in (1, 2, 3):
            pts, ws = minrule(D, 2 * h)
            moff[(D, h)] = len(mrows)
            mcnt[(D, h)] = len(ws)
            for p, w in zip(pts, ws):
                mrows.append((p[0], p[1], p[2], w))
    out.append("// minimal symmetric rules exact for degree 2h, h = 1..3 (index [D][h])")
    out.append("#define XG_MINRULE_OFF_INIT { " + ", ".join(
        "{0, " + ", ".join(str(moff[(D, h)]) for h in (1, 2, 3)) + "}" for D in range(4)) + " }")
    out.append("#define XG_MINRULE_CNT_INIT { " + ", ".join(
        "{0, " + ", ".join(str(mcnt[(D, h)]) for h in (1, 2, 3)) + "}" for D in range(4)) + " }")
    out.append("#define XG_MINRULE_ROWS %d" % len(mrows))
    out.append("#define XG_MINRULE_DATA_INIT { \\")
    for r in mrows:
        out.append("  {" + ", ".join(float(v).hex() for v in r) + "}, \\")
    out.append("}")
    out.append("static const int xg_minrule_off[4][4] = XG_MINRULE_OFF_INIT;")
    out.append("static const int xg_minrule_cnt[4][4] = XG_MINRULE_CNT_INIT;")
    out.append("static const double xg_minrule_data[XG_MINRULE_ROWS][4] = XG_MINRULE_DATA_INIT;")
    out.append("#if defined(__HIP__)")
    out.append("static __device__ const int xg_minrule_off_d[4][4] = XG_MINRULE_OFF_INIT;")
    out.append("static __device__ const int xg_minrule_cnt_d[4][4] = XG_MINRULE_CNT_INIT;")
    out.append("static __device__ const double xg_minrule_data_d[XG_MINRULE_ROWS][4] = XG_MINRULE_DATA_INIT;")
    out.append("#endif")
    # lattice points I (I0 + I1 + I2 <= 2^lvl) for the subdivision rule's cut
    # check (xintegration.cpp:617-752), packed I0 | I1 << 8 | I2 << 16
    lat_off, lat_cnt, lat = [], [], []
    for lvl in range(LAT_LMAX + 1):
        n = 1 << lvl
        lat_off.append(len(lat))
        for i2 in range(n + 1):
            for i1 in range(n + 1 - i2):
                for i0 in range(n + 1 - i2 - i1):
                    lat.append(i0 | (i1 << 8) | (i2 << 16))
        lat_cnt.append(len(lat) - lat_off[-1])
    out.append("#define XG_LAT_LMAX %d" % LAT_LMAX)
    out.append("#define XG_LAT_ROWS %d" % len(lat))
    out.append("#if defined(__HIP__)")
    out.append("static __device__ const int xg_lat_off_d[%d] = {%s};" % (LAT_LMAX + 1, ", ".join(map(str, lat_off))))
    out.append("static __device__ const int xg_lat_cnt_d[%d] = {%s};" % (LAT_LMAX + 1, ", ".join(map(str, lat_cnt))))
    out.append("static __device__ const unsigned xg_lat_d[XG_LAT_ROWS] = {")
    for k in range(0, len(lat), 16):
        out.append("  " + ", ".join(str(v) for v in lat[k:k + 16]) + ",")
    out.append("};")
    out.append("#endif")
    print("\n".join(out))


if __name__ == "__main__":
    main()
