"""Generates moments_table.h: the moment form of the H1 element matrices.

For a quadrature rule {(x_i, w_i)} on a simplex T with barycentric
coordinates lam(x) (lam_k = x_k for k < D, lam_D = 1 - sum x), and the
degree-P Lagrange basis phi_a of T,

    M_ab = sum_i w_i phi_a(x_i) phi_b(x_i)
    K_ab = sum_i w_i grad phi_a(x_i) . grad phi_b(x_i)

are linear in the homogeneous barycentric moments of the rule,

    m_alpha = sum_i w_i lam(x_i)^alpha,   |alpha| = 2P.

phi_a is written as a homogeneous polynomial of degree P in lam (each
constant c of the Lagrange factors (P lam_k - c)/j is replaced by c * S,
S = sum_k lam_k = 1), so phi_a phi_b = sum_alpha C_ab,alpha lam^alpha
holds at every point, and M = sum_alpha C_alpha m_alpha is the same
quadrature sum reassociated (an identity for any rule, not only an exact
one).  The reference gradient d phi / d xi_d = d phi/d lam_d - d phi/d lam_D
is homogeneous of degree P-1; with H = F^{-1} F^{-T},

    K_ab = sum_{d,e<D} H_de sum_i w_i G_ad(x_i) G_be(x_i),

whose degree-(2P-2) moments follow from the degree-2P ones through
lam^beta = lam^beta S^2.

The generated code (all FP64):
  accum(lam, w, m)   m_alpha += w lam^alpha (per quadrature point)
  complement(m)      m_alpha <- (exact moment of the whole simplex) - m_alpha
  lower(m, mk)       degree-(2P-2) moments
  forms(...)         M and K entries, written to the element matrix
The dof order is the one of xg_device.hpp (Lattice<D,P>).
"""
import itertools
import math
import sys
from fractions import Fraction


def lattice(D, P):
    """multi-indices of the degree-P Lagrange nodes in dof order (as Lattice<D,P>)"""
    allm = []
    for a0 in range(P, -1, -1):
        for a1 in (range(P - a0, -1, -1) if D >= 1 else [0]):
            for a2 in (range(P - a0 - a1, -1, -1) if D >= 2 else [0]):
                rest = P - a0 - a1 - a2
                if D < 3 and rest != 0:
                    continue
                allm.append((a0, a1, a2, rest if D >= 3 else 0))
    out = []
    for s in range(1, D + 2):
        for sup in itertools.combinations(range(D + 1), s):
            mask = sum(1 << i for i in sup)
            for a in allm:
                m = sum(1 << i for i in range(D + 1) if a[i] > 0)
                if m == mask:
                    out.append(a[:D + 1])
    assert len(out) == len(allm)
    return out


def monomials(D, n):
    """exponent tuples of length D+1 with |alpha| = n (descending lex)"""
    res = []

    def rec(prefix, left, k):
        if k == D:
            res.append(tuple(prefix + [left]))
            return
        for v in range(left, -1, -1):
            rec(prefix + [v], left - v, k + 1)
    rec([], n, 0)
    return res


def pmul(p, q):
    r = {}
    for a, ca in p.items():
        for b, cb in q.items():
            k = tuple(x + y for x, y in zip(a, b))
            r[k] = r.get(k, 0) + ca * cb
    return {k: v for k, v in r.items() if v != 0}


def padd(p, q, s=1):
    r = dict(p)
    for k, v in q.items():
        r[k] = r.get(k, 0) + s * v
    return {k: v for k, v in r.items() if v != 0}


def pdiff(p, i):
    r = {}
    for a, c in p.items():
        if a[i] > 0:
            b = list(a)
            b[i] -= 1
            r[tuple(b)] = r.get(tuple(b), 0) + c * a[i]
    return r


def basis(D, P):
    """homogeneous degree-P Lagrange basis in barycentric monomials"""
    nv = D + 1
    one = tuple([0] * nv)
    phis = []
    for a in lattice(D, P):
        p = {one: Fraction(1)}
        for k in range(nv):
            for j in range(1, a[k] + 1):
                lin = {}
                for i in range(nv):
                    c = Fraction(P * (i == k) - (j - 1), j)
                    if c != 0:
                        e = [0] * nv
                        e[i] = 1
                        lin[tuple(e)] = c
                p = pmul(p, lin)
        phis.append(p)
    return phis


def basis_affine(D, P):
    """the same Lagrange basis with its constants kept (not homogenised); its
    partial derivatives in the barycentrics are sparse (a vertex function
    depends on its own barycentric only)"""
    nv = D + 1
    one = tuple([0] * nv)
    phis = []
    for a in lattice(D, P):
        p = {one: Fraction(1)}
        for k in range(nv):
            for j in range(1, a[k] + 1):
                lin = {}
                e = [0] * nv
                e[k] = 1
                lin[tuple(e)] = Fraction(P, j)
                if j > 1:
                    lin[one] = Fraction(-(j - 1), j)
                p = pmul(p, lin)
        phis.append(p)
    return phis


def homogenise(p, D, n):
    """multiply each term of degree d < n by S^(n-d), S = sum lam (= 1)"""
    nv = D + 1
    S1 = {tuple(1 if i == k else 0 for i in range(nv)): Fraction(1) for k in range(nv)}
    out = {}
    for a, c in p.items():
        t = {a: c}
        for _ in range(n - sum(a)):
            t = pmul(t, S1)
        out = padd(out, t)
    return out


def fmt(c):
    v = float(c)
    if v == int(v) and abs(v) < 1e15:
        return "%.1f" % v
    return repr(v)


def lin_expr(terms, var):
    """sum of c * var[idx] as a C expression (terms: list of (idx, coef))"""
    parts = []
    for idx, c in terms:
        if c == 1:
            parts.append("%s[%d]" % (var, idx))
        elif c == -1:
            parts.append("-%s[%d]" % (var, idx))
        else:
            parts.append("%s * %s[%d]" % (fmt(c), var, idx))
    if not parts:
        return "0.0"
    s = parts[0]
    for t in parts[1:]:
        s += (" - " + t[1:]) if t.startswith("-") else (" + " + t)
    return s


def gen_hstep(D, P, L):
    """Exact moments of a simplex from its vertices (rules exact for degree
    2P): for a k-simplex T with vertices whose barycentric coordinates in the
    element are l_0 .. l_k, and n = 2P,

        int_T (t . lam)^n = |T| n! k! / (n + k)! h_n(t . l_0, .., t . l_k),

    h_n the complete homogeneous symmetric polynomial (the integral of a
    power of an affine function over a simplex).  Comparing the coefficients
    of t^alpha (the left side is sum_alpha n!/alpha! t^alpha m_alpha):

        m_alpha = |T| k! alpha! / (n + k)! [t^alpha] h_n.

    [t^alpha] h_n follows from the recursion over the vertices
    H_j <- H_j + (t . l_v) H_{j-1} (j = 1..n ascending, H_0 = 1), whose terms
    are all products of nonnegative barycentrics (no cancellation).  hstep
    runs one vertex, H_1 .. H_{n-1} in h (zero at the start of a simplex), H_n
    accumulated into m; scaling every l_v by s^(1/n) multiplies the simplex's
    contribution by s, so m collects the pieces of a domain directly;
    hfinish applies the alpha! factors once at the end."""
    nv = D + 1
    n = 2 * P
    mons = [monomials(D, k) for k in range(n + 1)]
    idx = [{a: i for i, a in enumerate(mk)} for mk in mons]
    off = {}
    o = 0
    for k in range(1, n):
        off[k] = o
        o += len(mons[k])
    L.append("  static constexpr int NH = %d;  // h: coefficients of H_1 .. H_%d" % (max(o, 1), n - 1))
    L.append("  // one vertex of the complete homogeneous recursion (l: barycentrics scaled by s^(1/%d))" % n)
    L.append("  XG_DEV static void hstep(const double *l, double *h, double *m) {")
    for k in range(1, n + 1):
        for g in mons[k]:
            tgt = ("m[%d]" % idx[n][g]) if k == n else ("h[%d]" % (off[k] + idx[k][g]))
            expr = tgt
            for i in range(nv):
                if g[i] == 0:
                    continue
                b = list(g)
                b[i] -= 1
                b = tuple(b)
                if k == 1:
                    expr = "%s + l[%d]" % (expr, i)
                else:
                    expr = "fma(l[%d], h[%d], %s)" % (i, off[k - 1] + idx[k - 1][b], expr)
            L.append("    %s = %s;" % (tgt, expr))
    L.append("  }")
    L.append("  XG_DEV static void hfinish(double *m) {")
    for g in mons[n]:
        f = 1
        for x in g:
            f *= math.factorial(x)
        if f != 1:
            L.append("    m[%d] *= %.1f;" % (idx[n][g], f))
    L.append("  }")


def gen_forms_bernstein(D, P, L):
    """forms through the Bernstein basis B_c = P!/c! lam^c (|c| = P): the
    products of Bernstein functions and of their barycentric partials are
    single monomials,

        M^B_cd = P!/c! P!/d! m_{c+d}
        K^B_cd = P!/c! P!/d! sum_{i,j} c_i d_j Gs_ij mk_{c+d-e_i-e_j},

    and the Lagrange matrices follow by the change of basis phi_a =
    sum_c T_ac B_c:  L = T (sm M^B + sk K^B) T^T, evaluated one column of
    X = A T^T at a time.  Same quadrature sum, reassociated by exact
    identities, in fewer operations than the expanded coefficients."""
    nv = D + 1
    monP = monomials(D, P)
    mon2 = monomials(D, 2 * P)
    idx2 = {a: i for i, a in enumerate(mon2)}
    monK = monomials(D, 2 * P - 2)
    idxK = {a: i for i, a in enumerate(monK)}
    phis = basis(D, P)
    N = len(phis)
    NB = len(monP)
    assert N == NB
    Pf = math.factorial(P)

    def mult(c):
        f = Pf
        for x in c:
            f //= math.factorial(x)
        return Fraction(f)

    T = [[Fraction(0)] * NB for _ in range(N)]
    for a, p in enumerate(phis):
        for mono, coef in p.items():
            T[a][monP.index(mono)] = coef / mult(mono)
    gpairs = [(i, j) for i in range(nv) for j in range(i, nv)]
    gidx = {pr: k for k, pr in enumerate(gpairs)}
    nterms = 0
    # entries of A: (m index, its factor, {(g, mk index): factor})
    ents = {}
    for ci in range(NB):
        for di in range(ci, NB):
            c, d = monP[ci], monP[di]
            kc = mult(c) * mult(d)
            mi = idx2[tuple(x + y for x, y in zip(c, d))]
            kt = {}
            for i in range(nv):
                if c[i] == 0:
                    continue
                for j in range(nv):
                    if d[j] == 0:
                        continue
                    b = [x + y for x, y in zip(c, d)]
                    b[i] -= 1
                    b[j] -= 1
                    key = (gidx[(min(i, j), max(i, j))], idxK[tuple(b)])
                    kt[key] = kt.get(key, 0) + kc * c[i] * d[j]
            ents[(ci, di)] = (mi, kc, {k: v for k, v in kt.items() if v != 0})
    # common factors go into scaled copies of sm / sk; the remaining term
    # factors are 1 or 2 (2 Gs precomputed), so every term is one FMA
    fm = sorted({e[1] for e in ents.values()})
    fk = {}
    for key, (mi, kc, kt) in ents.items():
        if kt:
            q = min(abs(v) for v in kt.values())
            assert all(v / q in (1, 2) for v in kt.values()), (D, P, kt)
            fk[key] = q
    def scal(name, f):
        return name if f == 1 else "%s%s_" % (name, str(f).replace("/", "o"))
    for f in fm:
        if f != 1:
            L.append("    const double %s = %s * sm;" % (scal("sm", f), fmt(f)))
    for f in sorted(set(fk.values())):
        if f != 1:
            L.append("    const double %s = %s * sk;" % (scal("sk", f), fmt(f)))
    need2 = sorted({g for key, (mi, kc, kt) in ents.items() for (g, b), v in kt.items() if v / fk[key] == 2})
    for g in need2:
        L.append("    const double Gd%d = 2.0 * Gs[%d];" % (g, g))
    L.append("    // A = sm M^B + sk K^B (Bernstein basis, upper triangle)")
    for (ci, di), (mi, kc, kt) in ents.items():
        nterms += 1 + len(kt)
        if kt:
            q = fk[(ci, di)]
            parts = []
            for (g, b), v in sorted(kt.items()):
                parts.append("%s * mk[%d]" % ("Gs[%d]" % g if v / q == 1 else "Gd%d" % g, b))
            L.append("    const double A%d_%d = %s * m[%d] + %s * (%s);" % (ci, di, scal("sm", kc), mi, scal("sk", q),
                                                                            " + ".join(parts)))
        else:
            L.append("    const double A%d_%d = %s * m[%d];" % (ci, di, scal("sm", kc), mi))

    def A(c, d):
        return "A%d_%d" % (min(c, d), max(c, d))

    def chain(terms):
        """sum of coefficient * name (coefficients exact small numbers)"""
        parts = []
        for name, v in terms:
            if v == 1:
                parts.append(name)
            elif v == -1:
                parts.append("-" + name)
            else:
                parts.append("%s * %s" % (fmt(v), name))
        e = parts[0]
        for t in parts[1:]:
            e += (" - " + t[1:]) if t.startswith("-") else (" + " + t)
        return e

    L.append("    // L = T A T^T: a column b of X = A T^T where T's row b has several")
    L.append("    // entries, otherwise the entries of A directly")
    for b in range(N):
        L.append("    {")
        tb = [(d, T[b][d]) for d in range(NB) if T[b][d] != 0]
        direct = len(tb) == 1
        if not direct:
            for c in range(NB):
                nterms += len(tb)
                L.append("      const double X%d = %s;" % (c, chain([(A(c, d), v) for d, v in tb])))
        for a in range(b + 1):
            ta = [(c, T[a][c]) for c in range(NB) if T[a][c] != 0]
            if direct:
                d0, v0 = tb[0]
                terms = [(A(c, d0), v * v0) for c, v in ta]
            else:
                terms = [("X%d" % c, v) for c, v in ta]
            nterms += len(terms)
            L.append("      const T v%d_ = (T)(%s);" % (a, chain(terms)))
            L.append("      out[%d] = v%d_;" % (a * N + b, a))
            if a != b:
                L.append("      out[%d] = v%d_;" % (b * N + a, a))
        L.append("    }")
    return nterms


def gen_affine(D, P, L, idx2):
    """accum_aff: the affine moments n_beta = sum_i w_i y_i^beta, |beta| <= 2P,
    in D affine coordinates y (lam_D = 1 - sum is not formed): the monomials
    of degree 2..P take 16 products instead of the homogeneous ones' 30
    (D = 3, P = 3), and the constant's weighted monomial is w itself.
    n_beta is kept in the slot of the homogeneous moment (beta, 2P - |beta|).
    With y = lam - 1/4 (centred coordinates) the homogeneous moments follow
    by the dense map of gen_affine_table."""
    n2 = 2 * P
    zero = tuple([0] * D)

    def slot(b):
        return idx2[tuple(b) + (n2 - sum(b),)]
    L.append("  // n_beta += w y^beta, |beta| <= 2P, in the slot of the homogeneous (beta, 2P - |beta|)")
    L.append("  XG_DEV static void accum_aff(const double *x, double w, double *m) {")
    have = {tuple(1 if i == k else 0 for i in range(D)): "x[%d]" % k for k in range(D)}
    cnt = 0
    for deg in range(2, P + 1):
        for a in monomials(D - 1, deg):
            k = next(i for i in range(D) if a[i] > 0)
            b = list(a)
            b[k] -= 1
            b = tuple(b)
            nm = "t%d" % cnt
            cnt += 1
            L.append("    const double %s = %s * x[%d];" % (nm, have[b], k))
            have[a] = nm
    groups = {}
    for a in affine_monomials(D, n2):
        beta = []
        left = P
        for i in range(D):
            t = min(a[i], left)
            beta.append(t)
            left -= t
        beta = tuple(beta)
        gamma = tuple(x - y for x, y in zip(a, beta))
        groups.setdefault(beta, []).append((a, gamma))
    for b in affine_monomials(D, P):
        if b not in groups:
            continue
        L.append("    {")
        L.append("      const double wb = %s;" % ("w" if b == zero else "w * " + have[b]))
        for a, gamma in groups[b]:
            s = slot(a)
            if gamma == zero:
                L.append("      m[%d] += wb;" % s)
            else:
                L.append("      m[%d] = fma(wb, %s, m[%d]);" % (s, have[gamma], s))
        L.append("    }")
    L.append("  }")


def affine_monomials(D, n):
    """exponent tuples of length D with |beta| <= n, ascending degree"""
    res = []
    for deg in range(n + 1):
        res.extend(monomials(D - 1, deg))
    return res


def gen_affine_table(D, P):
    """m_alpha = sum_beta C[beta][alpha] n_beta for the centred affine
    moments n (y_k = lam_k - 1/4, k < D; lam_D = 1/4 - sum y): C from the
    expansion of prod_k (1/4 + y_k)^alpha_k (1/4 - sum y)^alpha_D.  Centring
    keeps the map well conditioned (sum |C n| / max |m| about 20 on the
    whole simplex, against 2,815 for y = lam), and its entries (multinomials
    times powers of 1/4) are exact in FP32.  Layout [slot of beta][lane][2]:
    the columns alpha = lane and alpha = lane + 64 (zero past NM), so a wave
    reads one 8-byte pair per lane and row."""
    n2 = 2 * P
    mon2 = monomials(D, n2)
    idx2 = {a: i for i, a in enumerate(mon2)}
    NM = len(mon2)
    q = Fraction(1, 4)
    cz = Fraction(1) - D * q
    C = [[Fraction(0)] * NM for _ in range(NM)]
    for ai, a in enumerate(mon2):
        p = {tuple([0] * D): Fraction(1)}
        for k in range(D):
            lk = {tuple([0] * D): q, tuple(1 if i == k else 0 for i in range(D)): Fraction(1)}
            for _ in range(a[k]):
                p = pmul(p, lk)
        ld = {tuple([0] * D): cz}
        for k in range(D):
            ld[tuple(1 if i == k else 0 for i in range(D))] = Fraction(-1)
        for _ in range(a[D]):
            p = pmul(p, ld)
        for beta, c in p.items():
            C[idx2[tuple(beta) + (n2 - sum(beta),)]][ai] = c
    L = []
    L.append("// centred affine moments (y = lam - 1/4) -> homogeneous degree-%d moments, D = %d:" % (n2, D))
    L.append("// m[alpha] = sum_slot xg_aff%d%d[slot][alpha %% 64][alpha / 64] n[slot]" % (D, P))
    L.append("#ifndef XG_AFF_STORAGE")
    L.append("#define XG_AFF_STORAGE static __device__")
    L.append("#endif")
    L.append("XG_AFF_STORAGE const float xg_aff%d%d[%d][64][2] __attribute__((aligned(16))) = {" % (D, P, NM))
    for r in range(NM):
        vals = []
        for lane in range(64):
            for h in range(2):
                al = lane + 64 * h
                c = C[r][al] if al < NM else Fraction(0)
                f = float(c)
                import struct
                assert Fraction(struct.unpack("f", struct.pack("f", f))[0]) == c, "not exact in FP32"
                vals.append("0" if c == 0 else repr(f))
        L.append("    {" + ", ".join("{%s, %s}" % (vals[2 * j], vals[2 * j + 1]) for j in range(64)) + "},")
    L.append("};")
    return "\n".join(L)


def gen(D, P):
    nv = D + 1
    mon2 = monomials(D, 2 * P)
    idx2 = {a: i for i, a in enumerate(mon2)}
    monP = monomials(D, P)
    idxP = {a: i for i, a in enumerate(monP)}
    monK = monomials(D, 2 * P - 2)
    idxK = {a: i for i, a in enumerate(monK)}
    N = len(lattice(D, P))
    phis = basis(D, P)
    L = []
    L.append("template <>")
    L.append("struct Moments<%d, %d> {" % (D, P))
    L.append("  static constexpr int NM = %d;  // degree-%d moments" % (len(mon2), 2 * P))
    L.append("  static constexpr int NK = %d;  // degree-%d moments" % (len(monK), 2 * P - 2))
    L.append("  static constexpr int NQ = %d;  // degree-%d monomials per point" % (len(monP), P))
    # ---- accum: q_beta = lam^beta (|beta| = P), m_alpha += (w q_beta) q_gamma
    L.append("  XG_DEV static void accum(const double *l, double w, double *m) {")
    # build monomials of degree 1..P incrementally
    have = {tuple(1 if i == k else 0 for i in range(nv)): "l[%d]" % k for k in range(nv)}
    names = {}
    cnt = 0
    for deg in range(2, P + 1):
        for a in monomials(D, deg):
            # a = b + e_k for the first k with a_k > 0
            k = next(i for i in range(nv) if a[i] > 0)
            b = list(a)
            b[k] -= 1
            b = tuple(b)
            nm = "t%d" % cnt
            cnt += 1
            L.append("    const double %s = %s * l[%d];" % (nm, have[b], k))
            have[a] = nm
    # split alpha = beta + gamma, |beta| = |gamma| = P: take beta greedily
    # from the front; the updates are grouped by beta and each weighted
    # monomial w q_beta is formed right before its group, so only one of them
    # is live at a time (register pressure of the degree-6 moments)
    groups = {}
    for a in mon2:
        beta = []
        left = P
        for i in range(nv):
            t = min(a[i], left)
            beta.append(t)
            left -= t
        beta = tuple(beta)
        gamma = tuple(x - y for x, y in zip(a, beta))
        groups.setdefault(beta, []).append((a, gamma))
    for b in monP:
        if b not in groups:
            continue
        L.append("    {")
        L.append("      const double wb = w * %s;" % have[b])
        for a, gamma in groups[b]:
            L.append("      m[%d] = fma(wb, %s, m[%d]);" % (idx2[a], have[gamma], idx2[a]))
        L.append("    }")
    L.append("  }")
    if D == 3:
        gen_affine(D, P, L, idx2)
    # ---- complement: m <- (moments of the whole reference simplex) - m.
    # int_T lam^alpha = |T| D! alpha! / (|alpha| + D)! with |T| = 1/D! (the
    # rule weights of the reference simplex sum to its volume)
    L.append("  // m_alpha <- int_T lam^alpha - m_alpha over the whole reference simplex T")
    L.append("  XG_DEV static void complement(double *m) {")
    for a in mon2:
        num = 1
        for x in a:
            num *= math.factorial(x)
        full = Fraction(num, math.factorial(2 * P + D))
        L.append("    m[%d] = %s - m[%d];" % (idx2[a], repr(float(full)), idx2[a]))
    L.append("  }")
    # ---- lower: mk_beta = sum_{|g|=2} (2!/g!) m_{beta+g}
    L.append("  XG_DEV static void lower(const double *m, double *mk) {")
    g2 = monomials(D, 2)
    for b in monK:
        terms = []
        for g in g2:
            mult = 2
            for x in g:
                if x == 2:
                    mult = 1
            a = tuple(x + y for x, y in zip(b, g))
            terms.append((idx2[a], mult))
        L.append("    mk[%d] = %s;" % (idxK[b], lin_expr(terms, "m")))
    L.append("  }")
    if P <= 2:
        gen_hstep(D, P, L)
    # ---- forms
    # stiffness through the barycentric gradients: grad phi = sum_i
    # (d phi / d lam_i) grad lam_i for any extension of phi off the plane
    # sum lam = 1 (the grad lam_i sum to zero); with the affine form of the
    # basis the partials are sparse.  K_ab = sum_{i<=j} Gs_ij <...>, Gs_ij =
    # grad lam_i . grad lam_j, the products homogenised to degree 2P-2.
    aff = basis_affine(D, P)
    dG = [[pdiff(p, i) for i in range(nv)] for p in aff]
    gpairs = [(i, j) for i in range(nv) for j in range(i, nv)]
    L.append("  // out[a*N+b] = sm * M_ab + sk * K_ab; Gs = grad lam_i . grad lam_j, i <= j (row-major upper)")
    L.append("  static constexpr int NG = %d;" % len(gpairs))
    L.append("  template <typename T>")
    L.append("  XG_DEV static void forms(const double *m, const double *mk, const double *Gs, double sm, double sk,")
    L.append("                           T *out) {")
    nterms = 0
    if P == 2:
        nterms = gen_forms_bernstein(D, P, L)
    for a in (range(N) if P != 2 else []):
        for b in range(a, N):
            mp = pmul(phis[a], phis[b])
            mterms = sorted((idx2[k], v) for k, v in mp.items())
            nterms += len(mterms)
            kparts = []
            for pi, (i, j) in enumerate(gpairs):
                t = pmul(dG[a][i], dG[b][j])
                if i != j:
                    t = padd(t, pmul(dG[a][j], dG[b][i]))
                if not t:
                    continue
                t = homogenise(t, D, 2 * P - 2)
                kt = sorted((idxK[k], v) for k, v in t.items())
                nterms += len(kt)
                if kt:
                    kparts.append("Gs[%d] * (%s)" % (pi, lin_expr(kt, "mk")))
            kexpr = " + ".join(kparts) if kparts else "0.0"
            L.append("    {")
            L.append("      const double M_ = %s;" % lin_expr(mterms, "m"))
            L.append("      const double K_ = %s;" % kexpr)
            L.append("      const T v_ = (T)(sm * M_ + sk * K_);")
            L.append("      out[%d] = v_;" % (a * N + b))
            if a != b:
                L.append("      out[%d] = v_;" % (b * N + a))
            L.append("    }")
    L.append("  }")
    # ---- vec: F_a = sum_i w_i phi_a(x_i) = sum_alpha (phi_a S^P)_alpha m_alpha
    S1 = {tuple(1 if i == k else 0 for i in range(nv)): Fraction(1) for k in range(nv)}
    SP = {tuple([0] * nv): Fraction(1)}
    for _ in range(P):
        SP = pmul(SP, S1)
    L.append("  // out[a] = s * sum_i w_i phi_a(x_i)  (source term of a linear form)")
    L.append("  template <typename T>")
    L.append("  XG_DEV static void vec(const double *m, double s, T *out) {")
    for a in range(N):
        t = sorted((idx2[k], v) for k, v in pmul(phis[a], SP).items())
        L.append("    out[%d] = (T)(s * (%s));" % (a, lin_expr(t, "m")))
    L.append("  }")
    L.append("};  // %d coefficient terms" % nterms)
    return "\n".join(L)


def gen_nodal(D, Q):
    """psi_k = sum_alpha A[k][alpha] lam^alpha: the degree-Q Lagrange basis in
    homogeneous degree-Q monomials (monomial order of Moments<D, Q/2>), so
    that fixed-node weights W_k = sum_i w_i psi_k(x_i) = sum_alpha A m_alpha"""
    mon = monomials(D, Q)
    idx = {a: i for i, a in enumerate(mon)}
    phis = basis(D, Q)
    L = []
    L.append("// degree-%d Lagrange basis of the reference simplex (D = %d) in homogeneous" % (Q, D))
    L.append("// monomials: W_k = sum_alpha nodal%d%d[k][alpha] m_alpha" % (D, Q))
    L.append("static __device__ const double nodal%d%d[%d][%d] = {" % (D, Q, len(phis), len(mon)))
    for p in phis:
        row = [0.0] * len(mon)
        for a, c in p.items():
            row[idx[a]] = float(c)
        L.append("    {" + ", ".join(repr(v) for v in row) + "},")
    L.append("};")
    return "\n".join(L)


def main():
    out = []
    out.append("// moments_table.h -- GENERATED by gen_moments.py; do not edit.")
    out.append("// Moment form of the H1 mass and stiffness element matrices (see gen_moments.py).")
    out.append("#pragma once")
    out.append("#ifndef XG_DEV  // host builds (tests) define XG_DEV themselves")
    out.append("#include \"xg_device.hpp\"")
    out.append("#endif")
    out.append("#include <math.h>")
    out.append("namespace xg {")
    out.append("template <int D, int P>")
    out.append("struct Moments;")
    out.append("#pragma clang fp contract(fast)")
    for D in (1, 2, 3):
        for P in (1, 2, 3):
            out.append(gen(D, P))
    out.append(gen_affine_table(3, 3))
    out.append("#ifndef XG_MOM_NO_NODAL  // host builds of the identities leave the device table out")
    out.append(gen_nodal(3, 4))
    out.append("#endif")
    out.append("}  // namespace xg")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
