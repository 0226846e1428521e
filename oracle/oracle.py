"""ctypes binding of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module; the product package never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

NEG, POS, IF = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_up = ctypes.POINTER(ctypes.c_uint32)
_bp = ctypes.POINTER(ctypes.c_uint8)


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def _load():
    if not os.path.exists(LIB):
        build()
    lib = ctypes.CDLL(LIB)
    lib.or_rule_npts.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.or_rule.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.or_jacobian.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
    lib.or_check_if_straight_cut.argtypes = [ctypes.c_int, _dp]
    lib.or_straight_cut_rule.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                         _dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.or_mark.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _bp, _dp, _up]
    lib.or_quad_cut_rule.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                     ctypes.c_int]
    lib.or_quad_mark.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, _bp, _dp, _up]
    lib.or_quad_integrate_poly.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp,
                                           ctypes.POINTER(ctypes.c_int)]
    lib.or_quad_integrate_poly.restype = ctypes.c_double
    lib.or_quad_elmat.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp]
    lib.or_quad_elmat.restype = None
    lib.or_mlset_rule.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _ip, ctypes.c_int, _dp, _dp,
                                  ctypes.c_double, _dp, _dp, ctypes.c_int]
    lib.or_mlset_mark.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, _dp, ctypes.c_int,
                                  ctypes.c_int, _ip, _bp, _bp]
    lib.or_h1_ndof.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.or_h1_shape.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    lib.or_h1_node.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    lib.or_elmat_add.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                 ctypes.c_int, _dp]
    lib.or_integrate_poly.argtypes = [ctypes.c_int, _dp, _dp, _dp, ctypes.c_int, ctypes.c_int,
                                      _dp, ctypes.POINTER(ctypes.c_int)]
    lib.or_integrate_poly.restype = ctypes.c_double
    lib.or_integrate_mesh.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, ctypes.c_int,
                                      ctypes.c_int, ctypes.c_int, _dp,
                                      ctypes.POINTER(ctypes.c_int), _dp, _dp, ctypes.c_int]
    lib.or_cut_pipeline.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip,
                                    _ip, _dp, _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                    ctypes.c_int, _dp, ctypes.c_int]
    lib.or_subdiv_pipeline.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp,
                                       _dp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int, _dp,
                                       ctypes.c_int]
    lib.or_subdiv_trig_parts.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
    lib.or_subdiv_rule3.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp,
                                    ctypes.POINTER(_dp), ctypes.POINTER(_dp), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int)]
    lib.or_subdiv_moments3.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_int, _dp, _dp, _dp]
    lib.or_subdiv_moments3_batch.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int]
    lib.or_mlp_nparams.argtypes = [ctypes.c_int] * 3
    lib.or_mlp_forward.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_int, _dp, _dp]
    lib.or_mlp_loss_grad.argtypes = [ctypes.c_int] * 3 + [_dp, ctypes.c_int, _dp, _dp, _dp]
    lib.or_mlp_loss_grad.restype = ctypes.c_double
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load()
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def rule(D, order):
    n = lib().or_rule_npts(D, order)
    xi = np.zeros((n, max(D, 1)))
    w = np.zeros(n)
    lib().or_rule(D, order, _d(xi), _d(w))
    return xi[:, :D], w


def jacobian(V):
    V = _f64(V)
    D = V.shape[1]
    F = np.zeros(D * D)
    Fi = np.zeros(D * D)
    ad = np.zeros(1)
    lib().or_jacobian(D, _d(V), _d(F), _d(Fi), _d(ad))
    return F.reshape(D, D), Fi.reshape(D, D), ad[0]


def check_if_straight_cut(lset):
    l = _f64(lset)
    return lib().or_check_if_straight_cut(len(l), _d(l))


def straight_cut_rule(D, lset, dt, order, Finv=None, cap=4096):
    """Returns (pts[n,3], w[n], nsub), or None for 'no rule'. Raises on error."""
    l = _f64(lset)
    pts = np.zeros((cap, 3))
    w = np.zeros(cap)
    ns = ctypes.c_int(0)
    Fi = _f64(Finv).reshape(-1) if Finv is not None else None
    n = lib().or_straight_cut_rule(D, _d(l), dt, order, _d(Fi) if Fi is not None else None,
                                   _d(pts), _d(w), cap, ctypes.byref(ns))
    if n == -1:
        return None
    if n < -1:
        raise RuntimeError("oracle straight_cut_rule error %d" % n)
    return pts[:n].copy(), w[:n].copy(), ns.value


def mark(D, el2v, lset_v):
    el2v = np.ascontiguousarray(el2v, dtype=np.int32)
    l = _f64(lset_v)
    ne = el2v.shape[0]
    nw = (ne + 31) // 32
    dt = np.zeros(ne, np.uint8)
    cr = np.zeros(ne)
    bits = np.zeros((8, nw), np.uint32)
    r = lib().or_mark(D, ne, el2v.ctypes.data_as(_ip), _d(l), dt.ctypes.data_as(_bp), _d(cr),
                      bits.ctypes.data_as(_up))
    if r:
        raise RuntimeError("oracle mark error %d" % r)
    return dt, cr, bits


QUAD_POLICY = {"FIRST": 0, "OPTIMAL": 1, "FALLBACK": 2}


def quad_cut_rule(D, vals, dt, order, pol=1, V=None, cap=40000):
    """Straight cut rule of a quad (D=2) / hex (D=3) element
    (LevelsetCutQuadrilateral): vals = 2^D vertex values in the reference
    element's order, pol 0 FIRST / 1 OPTIMAL / 2 FALLBACK, V (2^D, D)
    physical vertices for the interface weights.  Returns (pts[n,3], w[n]) or
    None for 'no rule'."""
    v = _f64(vals)
    pts = np.zeros((cap, 3))
    w = np.zeros(cap)
    Vf = _f64(V).reshape(-1) if V is not None else None
    n = lib().or_quad_cut_rule(D, _d(v), dt, order, pol, _d(Vf) if Vf is not None else None, _d(pts), _d(w), cap)
    if n == -1:
        return None
    if n < -1:
        raise RuntimeError("oracle quad_cut_rule error %d" % n)
    return pts[:n].copy(), w[:n].copy()


def quad_mark(D, el2v, lset_v, pol=1):
    el2v = np.ascontiguousarray(el2v, dtype=np.int32)
    l = _f64(lset_v)
    ne = el2v.shape[0]
    nw = (ne + 31) // 32
    dt = np.zeros(ne, np.uint8)
    cr = np.zeros(ne)
    bits = np.zeros((8, nw), np.uint32)
    r = lib().or_quad_mark(D, ne, el2v.ctypes.data_as(_ip), _d(l), pol, dt.ctypes.data_as(_bp), _d(cr),
                           bits.ctypes.data_as(_up))
    if r:
        raise RuntimeError("oracle quad_mark error %d" % r)
    return dt, cr, bits


def quad_integrate_poly(D, V, pts, w, terms):
    coef = _f64([t[0] for t in terms])
    ex = np.ascontiguousarray([e for t in terms for e in t[1]], dtype=np.int32)
    return lib().or_quad_integrate_poly(D, _d(_f64(V)), _d(_f64(pts)), _d(_f64(w)), len(w), len(terms), _d(coef),
                                        ex.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))


def quad_elmat(D, V, pts, w, mask=3):
    """Q1 element matrix on a quad / hex element and a unit-element rule"""
    nv = 2 ** D
    out = np.zeros((nv, nv))
    if len(w):
        lib().or_quad_elmat(D, _d(_f64(V)), _d(_f64(pts)), _d(_f64(w)), len(w), mask, _d(out))
    return out


def quad_integrate_mesh(coords, el2v, lset_v, dt, order, terms=((1.0, (0, 0, 0)),), pol=1, element_wise=False):
    """IntegrateX on a quad / hex mesh with the oracle's rules (element loop
    in Python: small meshes only)."""
    coords = _f64(coords)
    D = coords.shape[1]
    out = np.zeros(len(el2v))
    for e, vs in enumerate(np.asarray(el2v)):
        r = quad_cut_rule(D, np.asarray(lset_v)[vs], dt, order, pol, V=coords[vs])
        if r is not None:
            out[e] = quad_integrate_poly(D, coords[vs], r[0], r[1], terms)
    return out if element_wise else float(out.sum())


def mlset_rule(D, lsets, dtts, order, V=None, cap=20000):
    """mlset rule of one element (CreateMultiLevelsetCutIntegrationRule):
    lsets (M, D+1) vertex values, dtts (K, M) domain tuples, V (D+1, D)
    physical vertices (None: the reference element).  Returns (pts[n,3],
    w[n]) with reference weights (physical = w * |det F|)."""
    l = _f64(lsets)
    M = l.shape[0]
    dt = np.ascontiguousarray(np.asarray(dtts, dtype=np.int32).reshape(-1, M))
    if V is None:
        V = np.vstack([np.eye(D), np.zeros(D)])
    F, Fi, ad = jacobian(V)
    pts = np.zeros((cap, 3))
    w = np.zeros(cap)
    n = lib().or_mlset_rule(D, M, _d(l), dt.shape[0], dt.ctypes.data_as(_ip), order, _d(_f64(F)), _d(_f64(Fi)),
                            ad, _d(pts), _d(w), cap)
    if n < 0:
        raise RuntimeError("oracle mlset rule error %d" % n)
    return pts[:n], w[:n]


def mlset_mark(D, el2v, coords, lset_v, dtts, contrib=True):
    """(of_type, contrib) per element (MultiLevelsetCutInformation)"""
    lv = _f64(lset_v)
    M, nv = lv.shape
    dt = np.ascontiguousarray(np.asarray(dtts, dtype=np.int32).reshape(-1, M))
    e = np.ascontiguousarray(el2v, dtype=np.int32)
    ne = e.shape[0]
    ot = np.zeros(ne, np.uint8)
    co = np.zeros(ne, np.uint8) if contrib else None
    r = lib().or_mlset_mark(D, ne, e.ctypes.data_as(_ip), _d(_f64(coords)), M, _d(lv), nv, dt.shape[0],
                            dt.ctypes.data_as(_ip), ot.ctypes.data_as(_bp),
                            co.ctypes.data_as(_bp) if co is not None else None)
    if r < 0:
        raise RuntimeError("oracle mlset mark error %d" % r)
    return ot.astype(bool), (co.astype(bool) if co is not None else None)


def h1_ndof(D, p):
    return lib().or_h1_ndof(D, p)


def h1_shape(D, p, xi):
    nd = h1_ndof(D, p)
    x = _f64(list(xi) + [0.0] * (3 - len(xi)))
    phi = np.zeros(nd)
    dphi = np.zeros(nd * D)
    lib().or_h1_shape(D, p, _d(x), _d(phi), _d(dphi))
    return phi, dphi.reshape(nd, D)


def h1_nodes(D, p):
    nd = h1_ndof(D, p)
    out = np.zeros((nd, D))
    x = np.zeros(3)
    for k in range(nd):
        lib().or_h1_node(D, p, k, _d(x))
        out[k] = x[:D]
    return out


def elmat(D, p, V, pts, w, mask=3):
    nd = h1_ndof(D, p)
    out = np.zeros(nd * nd)
    V = _f64(V)
    pts = _f64(pts)
    w = _f64(w)
    lib().or_elmat_add(D, p, _d(V), _d(pts), _d(w), len(w), mask, _d(out))
    return out.reshape(nd, nd)


def integrate_poly(D, V, pts, w, terms):
    """terms: list of (coef, (a, b, c)) for coef * x^a y^b z^c."""
    coef = _f64([t[0] for t in terms])
    ex = np.ascontiguousarray([e for t in terms for e in t[1]], dtype=np.int32)
    V = _f64(V)
    pts = _f64(pts)
    w = _f64(w)
    return lib().or_integrate_poly(D, _d(V), _d(pts), _d(w), len(w), len(terms), _d(coef),
                                   ex.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))


def integrate_mesh(mesh_or_arrays, lset_v, dt, order, terms=((1.0, (0, 0, 0)),),
                   element_wise=False, nthreads=4):
    if isinstance(mesh_or_arrays, tuple):
        coords, el2v = mesh_or_arrays
    else:
        coords, el2v = mesh_or_arrays.coords, mesh_or_arrays.el2v
    coords = _f64(coords)
    el2v = np.ascontiguousarray(el2v, dtype=np.int32)
    D = coords.shape[1]
    ne = el2v.shape[0]
    coef = _f64([t[0] for t in terms])
    ex = np.ascontiguousarray([e for t in terms for e in t[1]], dtype=np.int32)
    l = _f64(lset_v)
    pe = np.zeros(ne) if element_wise else None
    tot = np.zeros(1)
    r = lib().or_integrate_mesh(D, ne, el2v.ctypes.data_as(_ip), _d(coords), _d(l), dt, order,
                                len(terms), _d(coef), ex.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                _d(pe) if pe is not None else None, _d(tot), nthreads)
    if r:
        raise RuntimeError("oracle integrate error %d" % r)
    return pe if element_wise else tot[0]


def cut_pipeline(D, p, order, elems, el2v, coords, lset_v, dts, mask=3, nthreads=1):
    nd = h1_ndof(D, p)
    el2v = np.ascontiguousarray(el2v, dtype=np.int32)
    coords = _f64(coords)
    l = _f64(lset_v)
    if elems is None:
        nel = el2v.shape[0]
        ep = None
    else:
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        nel = len(elems)
        ep = elems.ctypes.data_as(_ip)
    dtsa = (ctypes.c_int * len(dts))(*dts)
    out = np.zeros((nel, len(dts), nd, nd))
    r = lib().or_cut_pipeline(D, p, order, nel, ep, el2v.ctypes.data_as(_ip), _d(coords), _d(l),
                              len(dts), dtsa, mask, _d(out), nthreads)
    if r:
        raise RuntimeError("oracle pipeline error %d" % r)
    return out


def subdiv_rule3(lset_nodal, lset_order, L, order, Finv=None, cap=200000):
    c = _f64(lset_nodal)
    P = [np.zeros((cap, 3)) for _ in range(3)]
    W = [np.zeros(cap) for _ in range(3)]
    pp = (_dp * 3)(*[_d(a) for a in P])
    ww = (_dp * 3)(*[_d(a) for a in W])
    n = (ctypes.c_int * 3)()
    Fi = _f64(Finv).reshape(-1) if Finv is not None else None
    dom = lib().or_subdiv_rule3(_d(c), lset_order, L, order, _d(Fi) if Fi is not None else None,
                                pp, ww, cap, n)
    if dom < 0:
        raise RuntimeError("oracle subdiv error %d" % dom)
    return dom, [(P[r][:n[r]].copy(), W[r][:n[r]].copy()) for r in range(3)]


def subdiv_pipeline(p, L, order, elems, el2v, coords, lset_nodal, dts, mask=3, nthreads=1):
    """or_subdiv_pipeline: level-L subdivision rule + degree-p element
    matrices per listed element; returns nel x ndt x ndof x ndof"""
    nd = h1_ndof(3, p)
    el2v = np.ascontiguousarray(el2v, dtype=np.int32)
    coords = _f64(coords)
    ln = _f64(lset_nodal)
    if elems is None:
        nel, ep = el2v.shape[0], None
    else:
        elems = np.ascontiguousarray(elems, dtype=np.int32)
        nel, ep = len(elems), elems.ctypes.data_as(_ip)
    dtsa = (ctypes.c_int * len(dts))(*dts)
    out = np.zeros((nel, len(dts), nd, nd))
    r = lib().or_subdiv_pipeline(p, L, order, nel, ep, el2v.ctypes.data_as(_ip), _d(coords), _d(ln), len(dts),
                                 dtsa, mask, _d(out), nthreads)
    if r:
        raise RuntimeError("oracle subdivision pipeline error %d" % r)
    return out


def subdiv_trig_parts(lset_nodal, lset_order, L):
    """(lattice domain, NEG area, POS area) of the order-0 subdivision rule on
    the reference triangle (or_subdiv_trig_parts)"""
    c = _f64(lset_nodal)
    part = np.zeros(2)
    dom = lib().or_subdiv_trig_parts(_d(c), lset_order, L, _d(part))
    if dom < 0:
        raise RuntimeError("oracle subdiv error %d" % dom)
    return dom, part[0], part[1]


def subdiv_moments3(lset_nodal, lset_order, L, order, q, Finv=None):
    nd = h1_ndof(3, q)
    c = _f64(lset_nodal)
    on = np.zeros(nd)
    oi = np.zeros(nd)
    Fi = _f64(Finv).reshape(-1) if Finv is not None else None
    dom = lib().or_subdiv_moments3(_d(c), lset_order, L, order, q,
                                   _d(Fi) if Fi is not None else None, _d(on), _d(oi))
    if dom < 0:
        raise RuntimeError("oracle subdiv error %d" % dom)
    return dom, on, oi


def subdiv_moments3_batch(lset_nodal, lset_order, L, order, q, nthreads=1):
    nd = h1_ndof(3, q)
    c = _f64(lset_nodal)
    nel = c.shape[0]
    on = np.zeros((nel, nd))
    oi = np.zeros((nel, nd))
    r = lib().or_subdiv_moments3_batch(nel, _d(c), lset_order, L, order, q, _d(on), _d(oi),
                                       nthreads)
    if r < 0:
        raise RuntimeError("oracle subdiv batch error %d" % r)
    return on, oi


def mlp_nparams(I, H, O):
    return lib().or_mlp_nparams(I, H, O)


def mlp_forward(I, H, O, params, x):
    params = _f64(params)
    x = _f64(x)
    B = x.shape[0]
    y = np.zeros((B, O))
    lib().or_mlp_forward(I, H, O, _d(params), B, _d(x), _d(y))
    return y


def mlp_loss_grad(I, H, O, params, x, target):
    params = _f64(params)
    x = _f64(x)
    t = _f64(target)
    g = np.zeros_like(params)
    loss = lib().or_mlp_loss_grad(I, H, O, _d(params), x.shape[0], _d(x), _d(t), _d(g))
    return loss, g
