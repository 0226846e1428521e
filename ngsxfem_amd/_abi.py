"""ctypes binding of libxg.so, the C ABI declared in include/xg_abi.h.

Device buffers are torch CUDA tensors (PyTorch is used here only for device
memory and streams); every call passes raw device pointers to the HIP
kernels.  There is no CPU fallback: loading fails loudly when libxg.so is
missing, and every call raises XgError on a non-zero status.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libxg.so")

XG_OK = 0
XG_NEG, XG_POS, XG_IF = 0, 1, 2
XG_PREC_FP64, XG_PREC_FP32, XG_PREC_BF16X3, XG_PREC_BF16 = 0, 1, 2, 3
XG_METHOD_MOMENTS = 16  # OR-ed into the precision of xg_cut_pipeline
XG_FORM_MASS, XG_FORM_LAPLACE = 1, 2

_vp = ctypes.c_void_p
_i = ctypes.c_int


class XgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("xg error %d: %s" % (code, msg))
        self.code = code


# signatures: name -> argtypes (all return int)
class _XgCf(ctypes.Structure):
    """struct xg_cf (include/xg_abi.h): a coefficient's postfix program"""
    _fields_ = [("ncode", ctypes.c_int32), ("code", _vp), ("nconst", ctypes.c_int32), ("consts", _vp),
                ("ngf", ctypes.c_int32), ("gf", _vp)]

    @staticmethod
    def make(code, consts, gfs=()):
        c = (ctypes.c_int32 * max(len(code), 1))(*code)
        k = (ctypes.c_double * max(len(consts), 1))(*consts)
        g = (_vp * max(len(gfs), 1))(*[ptr(t) for t in gfs])
        s = _XgCf(len(code), ctypes.cast(c, _vp), len(consts), ctypes.cast(k, _vp), len(gfs), ctypes.cast(g, _vp))
        s._keep = (c, k, g)
        return s


_SIGS = {
    "xg_ctx_create": [_i, ctypes.POINTER(_vp)],
    "xg_ctx_destroy": [_vp],
    "xg_ctx_set_stream": [_vp, _vp],
    "xg_sync": [_vp],
    "xg_rule_npts": [_i, _i],
    "xg_p1_perturb": [_vp, _i, _vp, ctypes.c_double],
    "xg_mark": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "xg_mark_p2": [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp],
    "xg_mark_p2_trig": [_vp, _i, _vp, _i, _vp, _vp, _vp, _vp, _vp],
    "xg_facets_with_neighbor_types": [_vp, _i, _vp, _vp, _vp, _i, _i, _i, _vp],
    "xg_mark_from_items": [_vp, _i, _i, _vp, _vp, _i, _vp],
    "xg_mlset_rule_count": [_vp, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _i, _vp],
    "xg_mlset_rule_fill": [_vp, _i, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _i, _vp, _vp],
    "xg_mlset_mark": [_vp, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp],
    "xg_elements_with_neighbor_facets": [_vp, _i, _vp, _vp, _i, _vp],
    "xg_elements_with_shared_vertex": [_vp, _i, _i, _vp, _vp, _i, _vp],
    "xg_node_domains": [_vp, _i, _i, _vp, _vp, _vp, _i, _vp, _vp],
    "xg_cut_rule_count": [_vp, _i, _i, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "xg_cut_rule_fill": [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp, _vp],
    "xg_integrate": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "xg_quad_rule_count": [_vp, _i, _i, _vp, _vp, _vp, _i, _i, _i, _vp],
    "xg_quad_rule_fill": [_vp, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _i, _vp, _vp],
    "xg_quad_mark": [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp],
    "xg_quad_integrate": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "xg_quad_elmat": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "xg_quad_elvec": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "xg_integrate_nodes": [_vp, _i, _vp, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, _vp],
    "xg_elmat": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "xg_rule_scale_cf": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "xg_elvec": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "xg_facet_rule_to_element": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "xg_p1_interpolate_cf": [_vp, _i, _i, _vp, _vp, ctypes.c_double, _vp],
    "xg_cut_pipeline": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _vp],
    "xg_subdiv_weights": [_vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp],
    "xg_subdiv_moments": [_vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp],
    "xg_subdiv_pipeline": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _i, _vp, _vp],
    "xg_mark_noalloc": [_vp, _i, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "xg_learned_moments": [_vp, _i, _i, _vp, _vp, _vp, ctypes.c_double, _vp, _vp, _vp],
    "xg_subdiv_weights_mesh": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    "xg_assemble_csr_scaled": [_vp, _i, _i, _vp, _vp, _i, ctypes.c_double, _vp, _vp, _vp],
    "xg_timing_enable": [_vp, _i],
    "xg_timing_read": [_vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)],
    "xg_csr_pattern": [_vp, _i, _i, _vp, _i, _vp, _vp, ctypes.c_longlong, _vp],
    "xg_subdiv_moments_mesh": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _vp, _vp],
    "xg_elmat_moments": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "xg_cut_elvec": [_vp, _i, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _vp, _i, _vp],
    "xg_assemble_csr": [_vp, _i, _i, _vp, _vp, _i, _vp, _vp, _vp],
    "xg_assemble_vec": [_vp, _i, _i, _vp, _vp, _i, _vp],
    "xg_elmat_weights": [_vp, _i, _i, _i, _vp, _vp, _vp, _vp, _i, _i, _vp],
    "xg_mlp_nparams": [_i, _i, _i],
    "xg_mlp_forward": [_vp, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _i],
    "xg_mlp_train_step": [_vp, _i, _i, _i, _vp, _vp, _vp, _i, ctypes.c_float, _i, _vp, _vp, _vp, _vp, _vp],
    "xg_moment_residual": [_vp, _i, _i, _vp, ctypes.c_double, _vp, ctypes.c_double, _vp, _vp, _vp],
    "xg_batch_create": [_vp, _i, _i, _vp, _vp, _i, _vp, _i, _vp, _vp, _vp, ctypes.POINTER(_vp)],
    "xg_batch_destroy": [_vp],
    "xg_batch_tune": [_vp, _i, _i, _i, _i],
    "xg_batch_method": [_vp, _i, _i],
    "xg_batch_plan": [_vp, _vp],
    "xg_batch_info": [_vp, _vp, _vp, _vp],
    "xg_batch_run": [_vp, _vp, _vp, ctypes.c_longlong, _i, _i],
}


def declared_symbols():
    """Entry points declared in include/xg_abi.h (parsed from the header)."""
    import re
    hdr = os.path.join(HERE, "..", "include", "xg_abi.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|const int32_t \*)\s*(xg_\w+)\s*\(", txt, re.M)))


_lib = None


def _nmom(D, p):
    """number of degree-2p homogeneous monomials in D+1 barycentrics"""
    from math import comb
    return comb(2 * p + D, D)


def load_hip_runtime():
    """Make the process use a single HIP runtime.  PyTorch-ROCm wheels bundle
    their own libamdhip64; when libxg.so (linked against the system one) is
    loaded first, two runtimes end up in the process and the second to
    initialise sees no device.  Importing torch first lets libxg.so bind to
    the runtime torch already loaded (same soname)."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    global _lib
    if _lib is None:
        load_hip_runtime()
        if not os.path.exists(LIBPATH):
            raise ImportError("libxg.so not built (%s); run __graft_entry__.build() or "
                              "python -m ngsxfem_amd.build_ext" % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _i
        L.xg_batch_cut_list.argtypes = [_vp]
        L.xg_batch_cut_list.restype = _vp
        L.xg_last_error.argtypes = []
        L.xg_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def check(code):
    if code != XG_OK:
        raise XgError(code, lib().xg_last_error().decode(errors="replace"))
    return code


def ptr(t):
    """device pointer of a torch tensor (or None)"""
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


class Context:
    """xg_ctx bound to a device and a (torch) stream."""

    def __init__(self, device=0, stream=None):
        self.h = ctypes.c_void_p()
        check(lib().xg_ctx_create(device, ctypes.byref(self.h)))
        self.device = device
        if stream is not None:
            self.set_stream(stream)

    def set_stream(self, stream):
        s = stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)
        check(lib().xg_ctx_set_stream(self.h, ctypes.c_void_p(s)))

    def sync(self):
        check(lib().xg_sync(self.h))

    def close(self):
        if self.h:
            lib().xg_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --- thin wrappers (device tensors in, device tensors out) -------------
    def p1_perturb(self, v, eps=1e-14):
        check(lib().xg_p1_perturb(self.h, v.numel(), ptr(v), eps))

    def mark(self, D, el2v, lset, dt=None, cut_ratio=None, bits=None, cut_list=None, n_cut=None):
        ne = el2v.shape[0]
        check(lib().xg_mark(self.h, D, ne, ptr(el2v), ptr(lset), ptr(dt), ptr(cut_ratio), ptr(bits),
                            ptr(cut_list), ptr(n_cut)))

    def timing_enable(self, on=True):
        """per-kernel HIP event timing of the subdivision kernels (cleared when enabled)"""
        check(lib().xg_timing_enable(self.h, 1 if on else 0))

    def timing_read(self, name):
        """(summed ms, launches) recorded under a kernel name"""
        ms, n = ctypes.c_double(0.0), ctypes.c_int(0)
        check(lib().xg_timing_read(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n)))
        return ms.value, n.value

    def mark_noalloc(self, D, el2v, lset, dt=None, cut_ratio=None, bits=None, cut_list=None, n_cut=None):
        """xg_mark without allocation (fixed device scratch)"""
        ne = el2v.shape[0]
        check(lib().xg_mark_noalloc(self.h, D, ne, ptr(el2v), ptr(lset), ptr(dt), ptr(cut_ratio), ptr(bits),
                                    ptr(cut_list), ptr(n_cut)))

    def mark_p2(self, lset_nodal, L, dt, bits, cut_list=None, n_cut=None):
        ne = lset_nodal.shape[0]
        check(lib().xg_mark_p2(self.h, ne, ptr(lset_nodal), L, ptr(dt), ptr(bits), ptr(cut_list), ptr(n_cut)))

    def mark_p2_trig(self, lset_nodal, L, dt, cut_ratio, bits, cut_list=None, n_cut=None):
        ne = lset_nodal.shape[0]
        check(lib().xg_mark_p2_trig(self.h, ne, ptr(lset_nodal), L, ptr(dt), ptr(cut_ratio), ptr(bits),
                                    ptr(cut_list), ptr(n_cut)))

    def facets_with_neighbor_types(self, f2el, a, b, bnd_val_a, bnd_val_b, use_and, out):
        check(lib().xg_facets_with_neighbor_types(self.h, f2el.shape[0], ptr(f2el), ptr(a), ptr(b),
                                                  int(bnd_val_a), int(bnd_val_b), int(use_and), ptr(out)))

    def mark_from_items(self, item2node, mark, nn, out):
        """out bit d for each node d of a marked item (mark None: all items)"""
        n, k = item2node.shape
        check(lib().xg_mark_from_items(self.h, n, k, ptr(item2node), ptr(mark), nn, ptr(out)))

    def elements_with_neighbor_facets(self, f2el, facet_mark, ne, out):
        check(lib().xg_elements_with_neighbor_facets(self.h, f2el.shape[0], ptr(f2el), ptr(facet_mark), ne,
                                                     ptr(out)))

    def elements_with_shared_vertex(self, D, el2v, elem_mark, nv, out):
        """D: element dimension (the vertex count per element is el2v's width)"""
        check(lib().xg_elements_with_shared_vertex(self.h, el2v.shape[1], el2v.shape[0], ptr(el2v), ptr(elem_mark), nv,
                                                   ptr(out)))

    def node_domains(self, el2node, dt, cut_ratio, nn, cutnb=None, dom=None):
        """el2node None: the element level (nn == number of elements)"""
        ne = dt.shape[0]
        k = el2node.shape[1] if el2node is not None else 0
        check(lib().xg_node_domains(self.h, ne, k, ptr(el2node), ptr(dt), ptr(cut_ratio), nn, ptr(cutnb),
                                    ptr(dom)))

    @staticmethod
    def _mlset_args(lsets, dtts):
        M = len(lsets)
        arr = (ctypes.c_void_p * M)(*[t.data_ptr() for t in lsets])
        ptrs = ctypes.cast(arr, ctypes.c_void_p)
        ptrs._keep = arr
        import numpy as np
        dt = np.ascontiguousarray(np.asarray(dtts, dtype=np.int8).reshape(-1, M))
        return M, ptrs, dt.shape[0], dt

    def mlset_rule_count(self, D, el2v, lsets, dtts, order, offsets, elems=None):
        """lsets: list of device tensors (vertex values), dtts: (K, M) domain tuples"""
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        M, ptrs, K, dt = self._mlset_args(lsets, dtts)
        check(lib().xg_mlset_rule_count(self.h, D, nel, ptr(elems), ptr(el2v), M, ptrs, K,
                                        dt.ctypes.data_as(ctypes.c_void_p), order, ptr(offsets)))

    def mlset_rule_fill(self, D, el2v, coords, lsets, dtts, order, offsets, pts, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        M, ptrs, K, dt = self._mlset_args(lsets, dtts)
        check(lib().xg_mlset_rule_fill(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), M, ptrs, K,
                                       dt.ctypes.data_as(ctypes.c_void_p), order, ptr(offsets), ptr(pts)))

    def mlset_mark(self, D, el2v, coords, lsets, dtts, of_type=None, contrib=None):
        M, ptrs, K, dt = self._mlset_args(lsets, dtts)
        check(lib().xg_mlset_mark(self.h, D, el2v.shape[0], ptr(el2v), ptr(coords), M, ptrs, K,
                                  dt.ctypes.data_as(ctypes.c_void_p), ptr(of_type), ptr(contrib)))

    def cut_rule_count(self, D, el2v, lset, dt, order, offsets, elems=None, nsub=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_cut_rule_count(self.h, D, nel, ptr(elems), ptr(el2v), ptr(lset), dt, order,
                                      ptr(offsets), ptr(nsub)))

    def cut_rule_fill(self, D, el2v, coords, lset, dt, order, offsets, pts, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_cut_rule_fill(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(lset), dt,
                                     order, ptr(offsets), ptr(pts)))

    def integrate(self, D, el2v, coords, offsets, pts, terms, total, per_elem=None, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        n = len(terms)
        coef = (ctypes.c_double * max(n, 1))(*[t[0] for t in terms])
        ex = (ctypes.c_int32 * max(3 * n, 1))(*[e for t in terms for e in t[1]])
        check(lib().xg_integrate(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(offsets),
                                 ptr(pts), n, ctypes.cast(coef, _vp), ctypes.cast(ex, _vp),
                                 ptr(per_elem), ptr(total)))

    # quadrilaterals / hexahedra (el2v: 2^D vertices per element); policy
    # 0 FIRST_ALLOWED, 1 FIND_OPTIMAL, 2 ALWAYS_NONE
    def quad_rule_count(self, D, el2v, lset, dt, order, policy, offsets, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_quad_rule_count(self.h, D, nel, ptr(elems), ptr(el2v), ptr(lset), dt, order, policy,
                                       ptr(offsets)))

    def quad_rule_fill(self, D, el2v, coords, lset, dt, order, policy, offsets, pts, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_quad_rule_fill(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(lset), dt, order,
                                      policy, ptr(offsets), ptr(pts)))

    def quad_mark(self, D, el2v, lset, policy, dt, bits, cut_ratio=None, cut_list=None, n_cut=None):
        check(lib().xg_quad_mark(self.h, D, el2v.shape[0], ptr(el2v), ptr(lset), policy, ptr(dt), ptr(cut_ratio),
                                 ptr(bits), ptr(cut_list), ptr(n_cut)))

    def quad_integrate(self, D, el2v, coords, offsets, pts, terms, total, per_elem=None, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        n = len(terms)
        coef = (ctypes.c_double * max(n, 1))(*[t[0] for t in terms])
        ex = (ctypes.c_int32 * max(3 * n, 1))(*[e for t in terms for e in t[1]])
        check(lib().xg_quad_integrate(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(offsets),
                                      ptr(pts), n, ctypes.cast(coef, _vp), ctypes.cast(ex, _vp),
                                      ptr(per_elem), ptr(total)))

    def quad_elmat(self, D, el2v, coords, offsets, pts, mask, precision, out, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_quad_elmat(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(offsets), ptr(pts), mask,
                                  precision, ptr(out)))

    def quad_elvec(self, D, el2v, coords, offsets, pts, out, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_quad_elvec(self.h, D, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(offsets), ptr(pts),
                                  ptr(out)))

    def rule_scale_cf(self, D, el2v, coords, offsets, pts, code, consts, gfs=(), elems=None, A=None):
        """pts[q].w *= cf(x_q); code / consts: the coefficient's postfix
        program, gfs: device FP64 vertex values of its P1 grid functions;
        coords: A (default D) values per vertex"""
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        cf = _XgCf.make(code, consts, gfs)
        check(lib().xg_rule_scale_cf(self.h, D, D if A is None else A, nel, ptr(elems), ptr(el2v), ptr(coords),
                                     ptr(offsets), ptr(pts), ctypes.byref(cf)))

    def elvec(self, D, p, el2v, coords, offsets, pts, out, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_elvec(self.h, D, p, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(offsets), ptr(pts),
                             ptr(out)))

    def elmat(self, D, p, el2v, coords, offsets, pts, mask, precision, out, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        check(lib().xg_elmat(self.h, D, p, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(offsets),
                             ptr(pts), mask, precision, ptr(out)))

    def subdiv_weights(self, lset_nodal, lset_order, L, order, q, out_neg=None, out_if=None, dom=None,
                       finv=None):
        nel = lset_nodal.shape[0]
        check(lib().xg_subdiv_weights(self.h, nel, ptr(lset_nodal), lset_order, L, order, q, ptr(finv),
                                      ptr(out_neg), ptr(out_if), ptr(dom)))

    def subdiv_moments(self, lset_nodal, lset_order, L, order, p, out_neg=None, out_if=None, dom=None,
                       finv=None):
        nel = lset_nodal.shape[0]
        check(lib().xg_subdiv_moments(self.h, nel, ptr(lset_nodal), lset_order, L, order, p, ptr(finv),
                                      ptr(out_neg), ptr(out_if), ptr(dom)))

    def subdiv_pipeline(self, p, L, order, elems, el2v, coords, lset_nodal, dts, mask, precision, out,
                        stats=None):
        """level-L subdivision rule + degree-p element matrices of the listed
        elements (xg_subdiv_pipeline); out: nel x ndt x ndof^2"""
        arr = (ctypes.c_int * len(dts))(*dts)
        check(lib().xg_subdiv_pipeline(self.h, p, L, order, elems.shape[0], ptr(elems), ptr(el2v), ptr(coords),
                                       ptr(lset_nodal), len(dts), ctypes.cast(arr, _vp), mask, precision, ptr(out),
                                       ptr(stats)))

    def subdiv_moments_mesh(self, p, L, order, elems, el2v, coords, lset_nodal, dts, moments, stats=None):
        """degree-2p moments of the level-L subdivision rule of the listed
        elements (xg_subdiv_moments_mesh); moments: nel x ndt x NM FP64"""
        arr = (ctypes.c_int * len(dts))(*dts)
        check(lib().xg_subdiv_moments_mesh(self.h, p, L, order, elems.shape[0], ptr(elems), ptr(el2v), ptr(coords),
                                           ptr(lset_nodal), len(dts), ctypes.cast(arr, _vp), ptr(moments),
                                           ptr(stats)))

    def subdiv_weights_mesh(self, L, order, elems, el2v, coords, lset_nodal, dts, weights, stats=None):
        """degree-4 fixed-node weights of the level-L subdivision rule of the
        listed elements (xg_subdiv_weights_mesh); weights: nel x ndt x 35"""
        arr = (ctypes.c_int * len(dts))(*dts)
        check(lib().xg_subdiv_weights_mesh(self.h, L, order, elems.shape[0], ptr(elems), ptr(el2v), ptr(coords),
                                           ptr(lset_nodal), len(dts), ctypes.cast(arr, _vp), ptr(weights),
                                           ptr(stats)))

    def csr_pattern(self, dofs, n, row_ptr, col_idx, nnz):
        """matrix graph of the element dof table on the device (xg_csr_pattern)"""
        nel, ndof = dofs.shape
        check(lib().xg_csr_pattern(self.h, nel, ndof, ptr(dofs), n, ptr(row_ptr), ptr(col_idx), col_idx.numel(),
                                   ptr(nnz)))

    def cut_elvec(self, D, p, order, el2v, coords, lset, dts, precision, out, elems=None):
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        arr = (ctypes.c_int * len(dts))(*dts)
        check(lib().xg_cut_elvec(self.h, D, p, order, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(lset),
                                 len(dts), ctypes.cast(arr, _vp), precision, ptr(out)))

    def assemble_csr(self, dofs, elmat, precision, row_ptr, col_idx, values):
        nel, ndof = dofs.shape
        check(lib().xg_assemble_csr(self.h, nel, ndof, ptr(dofs), ptr(elmat), precision, ptr(row_ptr),
                                    ptr(col_idx), ptr(values)))

    def assemble_vec(self, dofs, elvec, precision, out):
        nel, ndof = dofs.shape
        check(lib().xg_assemble_vec(self.h, nel, ndof, ptr(dofs), ptr(elvec), precision, ptr(out)))

    def elmat_moments(self, D, p, moments, mask, precision, out, el2v=None, coords=None, elems=None, nel=None):
        if nel is None:
            nel = elems.shape[0] if elems is not None else moments.numel() // _nmom(D, p)
        check(lib().xg_elmat_moments(self.h, D, p, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(moments), mask,
                                     precision, ptr(out)))

    def cut_pipeline(self, D, p, order, el2v, coords, lset, dts, mask, precision, out, elems=None,
                     nel=None):
        if nel is None:
            nel = elems.shape[0] if elems is not None else el2v.shape[0]
        arr = (ctypes.c_int * len(dts))(*dts)
        check(lib().xg_cut_pipeline(self.h, D, p, order, nel, ptr(elems), ptr(el2v), ptr(coords),
                                    ptr(lset), len(dts), ctypes.cast(arr, _vp), mask, precision, ptr(out)))

    def prepare_cut_pipeline(self, D, p, order, el2v, coords, lset, dts, mask, precision, out, elems=None):
        """the same call with its ctypes arguments built once: returns a
        function that enqueues one launch (for loops that repeat it)"""
        nel = elems.shape[0] if elems is not None else el2v.shape[0]
        arr = (ctypes.c_int * len(dts))(*dts)
        args = (self.h, D, p, order, nel, ptr(elems), ptr(el2v), ptr(coords), ptr(lset), len(dts),
                ctypes.cast(arr, _vp), mask, precision, ptr(out))
        f = lib().xg_cut_pipeline
        keep = (arr, elems, el2v, coords, lset, out)

        def launch():
            rc = f(*args)
            if rc != XG_OK:
                check(rc)
            return keep
        return launch


class Batch:
    """xg_batch: many meshes, one captured launch sequence (see xg_abi.h)."""

    def __init__(self, ctx, coords, el2v, lset, elem_off, mesh_p, dts=(XG_NEG, XG_IF), mask=3):
        import numpy as np
        self.ctx = ctx
        self.nmesh = len(mesh_p)
        self._keep = (coords, el2v, lset)  # device arrays must outlive the batch
        off = (ctypes.c_int32 * (self.nmesh + 1))(*[int(v) for v in elem_off])
        pp = (ctypes.c_int32 * self.nmesh)(*[int(v) for v in mesh_p])
        d = (ctypes.c_int * len(dts))(*dts)
        self.h = ctypes.c_void_p()
        check(lib().xg_batch_create(ctx.h, 3, self.nmesh, ctypes.cast(off, _vp), ctypes.cast(pp, _vp), len(dts),
                                    ctypes.cast(d, _vp), mask, ptr(el2v), ptr(coords), ptr(lset),
                                    ctypes.byref(self.h)))
        self.mesh_p = np.asarray(mesh_p)
        self.ndt = len(dts)

    def tune(self, p, waves, blocks_per_cu, grab):
        check(lib().xg_batch_tune(self.h, p, waves, blocks_per_cu, grab))

    def method(self, method, p=0):
        check(lib().xg_batch_method(self.h, p, method))

    def plan(self):
        check(lib().xg_batch_plan(self.ctx.h, self.h))
        return self.info()

    def info(self):
        import numpy as np
        cnt = np.zeros(self.nmesh, np.int32)
        off = np.zeros(self.nmesh + 1, np.int64)
        check(lib().xg_batch_info(self.ctx.h, self.h, cnt.ctypes.data_as(_vp), off.ctypes.data_as(_vp)))
        return cnt, off

    def cut_list_ptr(self):
        return lib().xg_batch_cut_list(self.h)

    def cut_list_host(self, n):
        """the first n cut element ids (after plan/run), copied to the host"""
        import numpy as np
        out = np.zeros(n, np.int32)
        if n:
            hip = ctypes.CDLL("libamdhip64.so")
            hip.hipDeviceSynchronize()
            rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.cut_list_ptr()),
                               ctypes.c_size_t(4 * n), 2)  # hipMemcpyDeviceToHost
            if rc != 0:
                raise XgError(-2, "hipMemcpy of the cut list failed (%d)" % rc)
        return out

    def run(self, out, use_graph=True, only_p=0):
        check(lib().xg_batch_run(self.ctx.h, self.h, ptr(out), out.numel(), int(use_graph), int(only_p)))

    def close(self):
        if self.h:
            lib().xg_batch_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
