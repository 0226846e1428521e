"""ngsxfem's Python surface for the cut-element path, on the MI355X build.

Mirror of the reference package ``xfem`` (python/__init__.py): the compiled
module ``xfem.xfem`` (here ``ngsxfem_amd.xfem.xfem``, pybind11 over the HIP
C ABI) provides CutInfo, BitArray, GetFacetsWithNeighborTypes,
GetElementsWithNeighborFacets, GetElementsWithSharedVertex,
GetDofsOfElements, GetDofsOfFacets, IntegrateX and the enums; this file adds HAS, Integrate, dCut, SymbolicCutBFI,
InterpolateToP1 and the small NGSolve stand-ins the path needs (meshes, P1
grid functions, polynomial coefficient functions, u*v / grad(u)*grad(v)
forms, BilinearForm.Assemble).

    from ngsxfem_amd.xfem import *
    mesh = MakeStructured2DMesh(nx=30, ny=30)
    lsetp1 = GridFunction(H1(mesh, order=1))
    InterpolateToP1(sqrt(x*x + y*y) - 0.5, lsetp1)
    ci = CutInfo(mesh, lsetp1)
    a = BilinearForm(H1(mesh, order=2))
    a += (grad(u)*grad(v) + u*v) * dCut(lsetp1, NEG, definedonelements=ci.GetElementsOfType(HASNEG))
    a.Assemble()
    Integrate({"levelset": lsetp1, "domain_type": IF}, cf=1, mesh=mesh, order=2)
"""
import math

import numpy as np

from .._abi import load_hip_runtime
from ..meshes import Make1DMesh, MakeStructured2DMesh, MakeStructured3DMesh, Mesh  # noqa: F401

load_hip_runtime()  # before the host module loads libxg.so: one HIP runtime per process

from .xfem import (ANY, BND, CDOM_IF, CDOM_NEG, CDOM_POS, FALLBACK, FIRST, HASNEG, HASPOS, IF, NEG, NO,  # noqa: F401
                   OPTIMAL, POS, UNCUT, VOL, AssembleCSR, AssembleCutFormsDevice, AssembleVector, BitArray,
                   COMBINED_DOMAIN_TYPE, CutElementMatrices, CutElementMatricesCF, CutElementVectors,
                   CutElementVectorsCF, CutInfo, ElementBoundaryMatrices, IntegrateCF, InterpolateCFToP1, DOMAIN_TYPE, GetDofsOfElements,
                   GetDofsOfFacets, GetElementsWithNeighborFacets, GetElementsWithSharedVertex,
                   GetFacetsWithNeighborTypes, IntegrateX, MlsetElementMatrices, MultiLevelsetCutInfo, NODE_TYPE,
                   QuadElementMatrices, SubdivElementMatrices,
                   NT_CELL, NT_EDGE, NT_ELEMENT, NT_FACE, NT_FACET, NT_VERTEX, PerturbP1, QUAD_DIRECTION_POLICY,
                   VorB)

PREC_FP64, PREC_FP32 = 0, 1


class _Globals:
    """ngsxfemglobals (utils/python_utils.cpp:232-280): the epsilons of the path"""

    def __init__(self):
        self.SetDefaults()

    def SetDefaults(self):
        self.eps_P1_perturbation = 1e-14
        self.simd_eval = True
        self.elmat_precision = PREC_FP32

    def MultiplyAllEps(self, fac):
        self.eps_P1_perturbation *= fac

    def Output(self):
        print("eps_P1_perturbation =", self.eps_P1_perturbation)


ngsxfemglobals = _Globals()


def HAS(domain_type):
    """python/__init__.py:30-42"""
    if domain_type == NEG:
        return HASNEG
    elif domain_type == POS:
        return HASPOS
    elif domain_type == IF:
        return IF
    else:
        raise Exception("invalid domain type")


# ------------------------------------------------------------------ coefficient functions
class CoefficientFunction:
    """Scalar coefficient function of the physical coordinates.  Polynomials
    (built from x, y, z, constants, +, -, *, integer powers) are kept as
    monomial term lists and integrated by the GPU reduction kernel; every
    other expression keeps its tree (constants, coordinates, + - * /, powers,
    sqrt sin cos tan exp log atan tanh abs, IfPos, max / min, P1 grid
    functions) and is compiled into the postfix program that the GPU
    evaluates at the rule points (``program``, include/xg_abi.h xg_cf).
    ``evaluate`` computes the same values with numpy (for checks)."""

    def __init__(self, val=None, terms=None, expr=None):
        if terms is not None:
            self._terms = {k: v for k, v in terms.items() if v != 0.0}
            self._expr = None
        elif expr is not None:
            self._terms = None
            self._expr = expr
        else:
            self._terms = {(0, 0, 0): float(val)} if float(val) != 0.0 else {}
            self._expr = None

    @staticmethod
    def _lift(o):
        if isinstance(o, CoefficientFunction):
            return o
        if isinstance(o, GridFunction):
            if isinstance(o.space, L2):
                return CoefficientFunction(expr=("el", o))
            if o.order != 1:
                raise NotImplementedError("grid functions in coefficients: P1 and L2 order 0")
            return CoefficientFunction(expr=("gf", o))
        return CoefficientFunction(float(o))

    def is_poly(self):
        return self._terms is not None

    def poly_terms(self):
        if not self.is_poly():
            raise TypeError("not a polynomial")
        t = [(c, e) for e, c in sorted(self._terms.items())]
        return t if t else [(0.0, (0, 0, 0))]

    def _node(self):
        return ("poly", self._terms) if self.is_poly() else self._expr

    def evaluate(self, X):
        X = np.asarray(X, dtype=np.float64)
        return np.asarray(_eval_node(self._node(), X), dtype=np.float64) * np.ones(X.shape[0])

    def __call__(self, *p):
        X = np.zeros((1, 3))
        X[0, :len(p)] = p
        return float(self.evaluate(X)[0])

    def program(self):
        """(code, consts, grid functions) of the postfix program"""
        code, consts, gfs = [], [], []
        _compile_node(self._node(), code, consts, gfs)
        return code, consts, gfs

    def _binop(self, o, op):
        o = self._lift(o)
        if self.is_poly() and o.is_poly():
            if op == "+" or op == "-":
                s = -1.0 if op == "-" else 1.0
                t = dict(self._terms)
                for e, c in o._terms.items():
                    t[e] = t.get(e, 0.0) + s * c
                return CoefficientFunction(terms=t)
            if op == "*":
                t = {}
                for e1, c1 in self._terms.items():
                    for e2, c2 in o._terms.items():
                        e = (e1[0] + e2[0], e1[1] + e2[1], e1[2] + e2[2])
                        t[e] = t.get(e, 0.0) + c1 * c2
                return CoefficientFunction(terms=t)
        return CoefficientFunction(expr=("bin", op, self._node(), o._node()))

    def __add__(self, o):
        return self._binop(o, "+")

    def __radd__(self, o):
        return self._lift(o)._binop(self, "+")

    def __sub__(self, o):
        return self._binop(o, "-")

    def __rsub__(self, o):
        return self._lift(o)._binop(self, "-")

    def __mul__(self, o):
        if isinstance(o, (_Form, _LinForm, _Proxy, _CutDifferentialSymbol, _Measure)):
            return NotImplemented
        return self._binop(o, "*")

    def __rmul__(self, o):
        return self._lift(o)._binop(self, "*")

    def __truediv__(self, o):
        o = self._lift(o)
        if o.is_poly() and list(o._terms.keys()) in ([(0, 0, 0)],):
            return self * (1.0 / o._terms[(0, 0, 0)])
        return self._binop(o, "/")

    def __rtruediv__(self, o):
        return self._lift(o).__truediv__(self)

    def __neg__(self):
        return self * (-1.0)

    def __abs__(self):
        return CoefficientFunction(expr=("un", "abs", self._node()))

    def __pow__(self, k):
        if isinstance(k, int) and k >= 0 and self.is_poly():
            r = CoefficientFunction(1.0)
            for _ in range(k):
                r = r * self
            return r
        if isinstance(k, int):
            return CoefficientFunction(expr=("powi", self._node(), k))
        return CoefficientFunction(expr=("bin", "pow", self._node(), self._lift(k)._node()))


# postfix opcodes of the device coefficient programs (include/xg_abi.h)
_OP = {"const": 1, "coord": 2, "+": 3, "-": 4, "*": 5, "/": 6, "neg": 7, "powi": 8, "pow": 9, "sqrt": 10,
       "sin": 11, "cos": 12, "tan": 13, "exp": 14, "log": 15, "atan": 16, "tanh": 17, "abs": 18, "gf": 19,
       "ifpos": 20, "max": 21, "min": 22, "el": 23}
_NP = {"sqrt": np.sqrt, "sin": np.sin, "cos": np.cos, "tan": np.tan, "exp": np.exp, "log": np.log,
       "atan": np.arctan, "tanh": np.tanh, "abs": np.abs}


def _eval_node(n, X):
    k = n[0]
    if k == "poly":
        out = np.zeros(X.shape[0])
        for (a, b, c), co in n[1].items():
            out += co * X[:, 0] ** a * X[:, 1] ** b * X[:, 2] ** c
        return out
    if k == "bin":
        a, b = _eval_node(n[2], X), _eval_node(n[3], X)
        return {"+": lambda: a + b, "-": lambda: a - b, "*": lambda: a * b, "/": lambda: a / b,
                "pow": lambda: np.power(a, b), "max": lambda: np.maximum(a, b),
                "min": lambda: np.minimum(a, b)}[n[1]]()
    if k == "un":
        return _NP[n[1]](_eval_node(n[2], X))
    if k == "powi":
        return _eval_node(n[1], X) ** float(n[2])
    if k == "ifpos":
        return np.where(_eval_node(n[1], X) > 0, _eval_node(n[2], X), _eval_node(n[3], X))
    if k in ("gf", "el"):
        raise NotImplementedError("grid functions are evaluated per element on the GPU (no host point location)")
    raise ValueError(k)


def _compile_node(n, code, consts, gfs):
    def const(v):
        v = float(v)
        if v not in consts:
            consts.append(v)
        code.extend([_OP["const"], consts.index(v)])

    k = n[0]
    if k == "poly":
        terms = sorted(n[1].items()) or [((0, 0, 0), 0.0)]
        for t, ((a, b, c), co) in enumerate(terms):
            const(co)
            for d, e in enumerate((a, b, c)):
                if e:
                    code.extend([_OP["coord"], d])
                    if e > 1:
                        code.extend([_OP["powi"], e])
                    code.append(_OP["*"])
            if t:
                code.append(_OP["+"])
    elif k == "bin":
        _compile_node(n[2], code, consts, gfs)
        _compile_node(n[3], code, consts, gfs)
        code.append(_OP[n[1]])
    elif k == "un":
        _compile_node(n[2], code, consts, gfs)
        code.append(_OP[n[1]])
    elif k == "powi":
        _compile_node(n[1], code, consts, gfs)
        code.extend([_OP["powi"], int(n[2])])
    elif k == "ifpos":
        for c in n[1:]:
            _compile_node(c, code, consts, gfs)
        code.append(_OP["ifpos"])
    elif k in ("gf", "el"):
        if n[1] not in gfs:
            gfs.append(n[1])
        code.extend([_OP[k], gfs.index(n[1])])
    else:
        raise ValueError(k)


x = CoefficientFunction(terms={(1, 0, 0): 1.0})
y = CoefficientFunction(terms={(0, 1, 0): 1.0})
z = CoefficientFunction(terms={(0, 0, 1): 1.0})


def _unary(name):
    f = _NP[name]

    def g(cf):
        if isinstance(cf, (int, float)):
            return float(f(cf))
        return CoefficientFunction(expr=("un", name, CoefficientFunction._lift(cf)._node()))
    return g


sqrt = _unary("sqrt")
sin = _unary("sin")
cos = _unary("cos")
tan = _unary("tan")
exp = _unary("exp")
log = _unary("log")
atan = _unary("atan")
tanh = _unary("tanh")


def IfPos(cf, cf_then, cf_else):
    """NGSolve IfPos: cf_then where cf > 0, else cf_else"""
    L = CoefficientFunction._lift
    return CoefficientFunction(expr=("ifpos", L(cf)._node(), L(cf_then)._node(), L(cf_else)._node()))


def Max(a, b):
    L = CoefficientFunction._lift
    return CoefficientFunction(expr=("bin", "max", L(a)._node(), L(b)._node()))


def Min(a, b):
    L = CoefficientFunction._lift
    return CoefficientFunction(expr=("bin", "min", L(a)._node(), L(b)._node()))


# ------------------------------------------------------------------ spaces, grid functions
class H1:
    """Lagrange space of degree `order` (P1 vertex dofs are what the cut path
    reads; higher orders define the element matrices' basis)."""

    def __init__(self, mesh, order=1, **kwargs):
        self.mesh = mesh
        self.order = order

    @property
    def ndof(self):
        return self.mesh.dofs(self.order)[1]

    def dofs(self):
        """element dof table (ne x ndof_el) and the number of global dofs"""
        return self.mesh.dofs(self.order)

    def TrialFunction(self):
        return _Proxy("u", self)

    def TestFunction(self):
        return _Proxy("v", self)

    def TnT(self):
        return self.TrialFunction(), self.TestFunction()


class L2:
    """Piecewise constant space (order 0): one value per element, in the
    element numbering.  Its grid functions enter coefficient functions
    (CutRatioGF, IndicatorCF)."""
    is_l2 = True

    def __init__(self, mesh, order=0, **kwargs):
        if order != 0:
            raise NotImplementedError("L2: order 0")
        self.mesh = mesh
        self.order = 0

    @property
    def ndof(self):
        return self.mesh.ne


class GridFunction:
    def __init__(self, space, name="gf"):
        self.space = space
        self.order = space.order
        self.vec = np.zeros(space.ndof)

    # a grid function is a coefficient function (P1: vertex values times
    # barycentrics, L2 order 0: the element's value)
    def _cf(self):
        return CoefficientFunction._lift(self)

    def __add__(self, o):
        return self._cf().__add__(o)

    def __radd__(self, o):
        return self._cf().__radd__(o)

    def __sub__(self, o):
        return self._cf().__sub__(o)

    def __rsub__(self, o):
        return self._cf().__rsub__(o)

    def __mul__(self, o):
        return self._cf().__mul__(o)

    def __rmul__(self, o):
        return self._cf().__rmul__(o)

    def __truediv__(self, o):
        return self._cf().__truediv__(o)

    def __rtruediv__(self, o):
        return self._cf().__rtruediv__(o)

    def __neg__(self):
        return self._cf().__neg__()

    def __pow__(self, k):
        return self._cf().__pow__(k)

    def Set(self, cf):
        """nodal interpolation: at the vertices (P1), at the vertices and edge
        midpoints (P2, this build's dof numbering)"""
        if isinstance(self.space, L2):
            raise NotImplementedError("GridFunction(L2).Set: assign .vec")
        if self.order == 1:
            if isinstance(cf, (CoefficientFunction, GridFunction)):
                # the coefficient's program at the vertices on the GPU (no perturbation)
                self.vec = InterpolateCFToP1(self.space.mesh, CoefficientFunction._lift(cf), 0.0)
                return
            self.vec = _eval_at_vertices(cf, self.space.mesh)
        elif self.order == 2:
            mesh = self.space.mesh
            e2v, _ = mesh.edges()
            X = np.vstack([mesh.coords, 0.5 * (mesh.coords[e2v[:, 0]] + mesh.coords[e2v[:, 1]])])
            self.vec = _eval_at_points(cf, X)
        else:
            raise NotImplementedError("GridFunction.Set: orders 1 and 2")


def _eval_at_points(cf, P):
    X = np.zeros((P.shape[0], 3))
    X[:, :P.shape[1]] = P
    if isinstance(cf, CoefficientFunction):
        return np.asarray(cf.evaluate(X), dtype=np.float64)
    if callable(cf):
        return np.array([cf(*p) for p in P], dtype=np.float64)
    return np.full(P.shape[0], float(cf))


def _eval_at_vertices(cf, mesh):
    return _eval_at_points(cf, mesh.coords)


def InterpolateToP1(gf_or_cf, gf_p1, eps_perturbation=None, heapsize=1000000):
    """utils/p1interpol.cpp:27-113: vertex values, |v| < eps -> +eps (GPU)"""
    eps = ngsxfemglobals.eps_P1_perturbation if eps_perturbation is None else eps_perturbation
    if isinstance(gf_or_cf, CoefficientFunction):
        # the coefficient's program at the vertices and the perturbation, on the GPU
        gf_p1.vec = InterpolateCFToP1(gf_p1.space.mesh, gf_or_cf, eps)
        return
    if isinstance(gf_or_cf, GridFunction):
        vals = np.asarray(gf_or_cf.vec, dtype=np.float64).copy()
    else:
        vals = _eval_at_vertices(gf_or_cf, gf_p1.space.mesh)
    gf_p1.vec = PerturbP1(vals, eps)


def IndicatorCF(mesh, ba, facets=False):
    """python/__init__.py:441-453: 1 on the elements where the BitArray is
    set, 0 elsewhere (a piecewise constant grid function's coefficient)"""
    if facets:
        raise NotImplementedError("IndicatorCF on facets")
    g = GridFunction(L2(mesh, order=0))
    g.vec = np.asarray(ba.numpy() if isinstance(ba, BitArray) else ba, dtype=np.float64).copy()
    return CoefficientFunction._lift(g)


def BitArrayCF(ba, mesh=None):
    """NGSolve's BitArrayCF (elements only): the mesh comes from the caller"""
    if mesh is None:
        raise NotImplementedError("BitArrayCF: pass mesh= (no global mesh in this build)")
    return IndicatorCF(mesh, ba)


def CutRatioGF(cutinfo):
    """python/__init__.py:455-462: the cut ratios NEG / (NEG + POS) as a
    piecewise constant grid function"""
    g = GridFunction(L2(cutinfo.Mesh(), order=0))
    g.vec = np.asarray(cutinfo.GetCutRatios(VOL), dtype=np.float64).copy()
    return g


def kappa(mesh, lset_approx, subdivlvl=0):
    """python/__init__.py:464-478 (deprecated there): the NEG and POS volume
    fractions of every element as piecewise constant grid functions"""
    print("kappa-function is deprecated - use CutRatioGF instead")
    k1 = CutRatioGF(CutInfo(mesh, lset_approx, subdivlvl=subdivlvl))
    k2 = GridFunction(L2(mesh, order=0))
    k2.vec = 1.0 - k1.vec
    return k1, k2


def IsCut(mesh, lset_approx, subdivlvl=0):
    """python/__init__.py:480-510 (deprecated there): 1 on elements with a
    volume fraction strictly between 1e-16 and 1 - 1e-16"""
    print("IsCut-function is deprecated - use CutInfo-functionality instead")
    k1, _ = kappa(mesh, lset_approx, subdivlvl)
    g = GridFunction(L2(mesh, order=0))
    g.vec = ((k1.vec > 1e-16) & (k1.vec < 1.0 - 1e-16)).astype(np.float64)
    return g


# ------------------------------------------------------------------ Integrate
def Integrate_X_special_args(levelset_domain={}, cf=None, mesh=None, VOL_or_BND=VOL, order=5, time_order=-1,
                             region_wise=False, element_wise=False, heapsize=1000000, deformation=None,
                             ip_container=None):
    """python/__init__.py:339-358"""
    lsd = dict(levelset_domain)
    if "order" not in lsd or lsd["order"] == -1:
        lsd["order"] = order
    if "time_order" not in lsd or lsd["time_order"] == -1:
        lsd["time_order"] = time_order
    if isinstance(cf, (int, float)):
        cf = CoefficientFunction(float(cf))
    if VOL_or_BND == BND:
        if deformation is not None or ip_container is not None:
            raise NotImplementedError("Integrate on the boundary: no deformation, no ip_container")
        return _integrate_bnd(mesh, lsd["levelset"], lsd["domain_type"], lsd["order"], cf, element_wise)
    return IntegrateX(levelset_domain=lsd, mesh=mesh, cf=cf, deformation=deformation,
                      ip_container=ip_container, element_wise=element_wise, heapsize=heapsize)


def _bnd_submesh(mesh):
    """The boundary elements as a mesh of dimension D - 1: every boundary
    simplex with its own vertex copies in an isometric frame of its own
    (v0 at the origin, v1 on the first axis).  Cut rules in reference
    coordinates depend only on the level set values, the interface weights'
    |F^-T n| and the measures are invariant under isometries, and so are the
    trace space's element matrices, so the (D-1)-dimensional kernels compute
    the boundary integrals; coefficients are evaluated at the ambient
    coordinates.  Returns (submesh, vmap = mesh vertex of each copy,
    ambient = their coordinates); cached on the mesh."""
    c = getattr(mesh, "_bnd_sub", None)
    if c is not None:
        return c
    D = mesh.dim
    if D not in (2, 3) or getattr(mesh, "is_quad", False):
        raise NotImplementedError("boundary integrals: triangle and tetrahedral meshes")
    mesh._bnd_sub = _isometric_submesh(mesh, np.ascontiguousarray(mesh.bnd2v, dtype=np.int32))
    return mesh._bnd_sub


def _isometric_submesh(mesh, b):
    """simplices of dimension D - 1 given by their vertices b (nb x D) as a
    mesh of their own: vertex copies in an isometric frame per simplex;
    returns (submesh, vmap, ambient coordinates)"""
    D = mesh.dim
    nb = b.shape[0]
    X = mesh.coords[b]  # nb x D x D
    e1 = X[:, 1] - X[:, 0]
    l1 = np.sqrt(np.sum(e1 * e1, axis=1))
    if D == 2:
        loc = np.zeros((nb, 2, 1))
        loc[:, 1, 0] = l1
    else:
        u = e1 / l1[:, None]
        e2 = X[:, 2] - X[:, 0]
        a = np.sum(e2 * u, axis=1)
        h = np.sqrt(np.maximum(np.sum(e2 * e2, axis=1) - a * a, 0.0))
        loc = np.zeros((nb, 3, 2))
        loc[:, 1, 0] = l1
        loc[:, 2, 0] = a
        loc[:, 2, 1] = h
    from ..meshes import Mesh
    sub = Mesh(loc.reshape(nb * D, D - 1), np.arange(nb * D, dtype=np.int32).reshape(nb, D))
    vmap = b.reshape(-1).copy()
    amb = np.ascontiguousarray(mesh.coords[vmap])
    return sub, vmap, amb


def _facet_submesh(mesh, els):
    """the D + 1 facets of the elements els (facet k of element i: its
    vertices without vertex k, ascending; submesh element i (D + 1) + k)"""
    D = mesh.dim
    E = mesh.el2v[els]
    fv = np.stack([E[:, [j for j in range(D + 1) if j != k]] for k in range(D + 1)], axis=1)
    return _isometric_submesh(mesh, np.ascontiguousarray(fv.reshape(-1, D), dtype=np.int32))


def _bnd_lset(lset, vmap):
    if getattr(lset, "order", 1) != 1 or isinstance(lset, (tuple, list)):
        raise NotImplementedError("boundary integrals: one P1 level set")
    v = np.asarray(lset.vec if hasattr(lset, "vec") else lset, dtype=np.float64)
    return np.ascontiguousarray(v[vmap])


def _integrate_bnd(mesh, lset, dt, order, cf, element_wise):
    """Integrate over the boundary elements (VOL_or_BND=BND): the cut rule of
    the level set's trace on each boundary simplex, cf on the GPU"""
    sub, vmap, amb = _bnd_submesh(mesh)
    return IntegrateCF(sub, _bnd_lset(lset, vmap), int(dt), int(order), CoefficientFunction._lift(cf),
                       bool(element_wise), amb, vmap)


def Integrate(levelset_domain=None, *args, **kwargs):
    """python/__init__.py:361-439: level-set integration when a dict is given;
    otherwise `cf * dCut(...)` integrals"""
    if levelset_domain is not None and type(levelset_domain) == dict:
        return Integrate_X_special_args(levelset_domain, *args, **kwargs)
    if isinstance(levelset_domain, _SumOfIntegrals):
        mesh = args[0] if args else kwargs["mesh"]
        return sum(i.integrate(mesh, kwargs.get("element_wise", False)) for i in levelset_domain.parts)
    raise NotImplementedError("plain (non level-set) Integrate is NGSolve's and not part of this build")


# ------------------------------------------------------------------ forms
def _is_coef(o):
    return isinstance(o, (int, float, CoefficientFunction, GridFunction))


def _cmul(a, b):
    """product of two form coefficients (numbers or coefficient functions)"""
    if isinstance(a, (int, float)) and isinstance(b, (int, float)):
        return float(a) * float(b)
    return CoefficientFunction._lift(a) * CoefficientFunction._lift(b)


def _as_const(c):
    """a number for constant coefficients, None for coefficient functions"""
    if isinstance(c, (int, float)):
        return float(c)
    if isinstance(c, CoefficientFunction) and c.is_poly() and set(c._terms) <= {(0, 0, 0)}:
        return float(c._terms.get((0, 0, 0), 0.0))
    return None


class _Proxy:
    """trial / test function (or its gradient), times a coefficient"""

    def __init__(self, kind, space, deriv=0, coef=1.0):
        self.kind, self.space, self.deriv, self.coef = kind, space, deriv, coef

    def __mul__(self, o):
        if isinstance(o, _Proxy):
            if {self.kind, o.kind} != {"u", "v"} or self.deriv != o.deriv:
                raise NotImplementedError("forms: u*v and grad(u)*grad(v) only")
            return _Form({("mass" if self.deriv == 0 else "laplace"): _cmul(self.coef, o.coef)}, self.space)
        if _is_coef(o):
            if self.kind == "v" and self.deriv == 0:
                return _LinForm(_cmul(self.coef, o), self.space)
            return _Proxy(self.kind, self.space, self.deriv, _cmul(self.coef, o))
        return NotImplemented

    def __rmul__(self, o):
        return self.__mul__(o) if _is_coef(o) else NotImplemented


class _LinForm:
    """f * v with a number or a coefficient function f
    (SymbolicCutLinearFormIntegrator's source term)"""

    def __init__(self, f, space):
        self.f, self.space = f, space

    def __mul__(self, o):
        if isinstance(o, (_CutDifferentialSymbol, _Measure)):
            return _SumOfIntegrals([_CutIntegral(self, o)])
        if _is_coef(o):
            return _LinForm(_cmul(self.f, o), self.space)
        return NotImplemented

    __rmul__ = lambda self, o: self.__mul__(o) if _is_coef(o) else NotImplemented


def grad(p):
    if isinstance(p, _Proxy):
        return _Proxy(p.kind, p.space, 1, p.coef)
    raise NotImplementedError("grad of a proxy function only")


class _Form:
    def __init__(self, parts, space):
        self.parts, self.space = dict(parts), space

    def __add__(self, o):
        d = dict(self.parts)
        for k, c in o.parts.items():
            d[k] = (d[k] + c) if k in d else c
        return _Form(d, self.space)

    def __mul__(self, o):
        if isinstance(o, (_CutDifferentialSymbol, _Measure)):
            return _SumOfIntegrals([_CutIntegral(self, o)])
        if _is_coef(o):
            return _Form({k: _cmul(c, o) for k, c in self.parts.items()}, self.space)
        return NotImplemented

    __rmul__ = lambda self, o: self.__mul__(o) if _is_coef(o) else NotImplemented


class _Measure:
    """NGSolve dx(definedonelements=...): the uncut base rule on the listed
    elements (a constant negative level set selects it)."""

    def __init__(self, definedonelements=None, order=None):
        self.definedonelements, self.order = definedonelements, order


def dx(definedonelements=None, bonus_intorder=0, **kwargs):
    return _Measure(definedonelements)


class _CutDifferentialSymbol:
    """CutDifferentialSymbol (cutint/cutintegral.hpp:55-82)"""

    def __init__(self, lsetdom, definedon=None, vb=VOL, element_boundary=False, element_vb=VOL, skeleton=False,
                 deformation=None, definedonelements=None):
        if element_vb == BND:
            element_boundary = True
        if deformation is not None or skeleton or vb not in (VOL, BND) or (element_boundary and vb != VOL):
            raise NotImplementedError("dCut: element (VOL), boundary (BND) and element-boundary integrals without "
                                      "deformation or skeleton in this build")
        self.lsetdom = lsetdom
        self.definedonelements = definedonelements
        self.vb = vb
        self.element_boundary = bool(element_boundary)

    def __rmul__(self, o):
        if isinstance(o, (_Form, _LinForm)):
            return o.__mul__(self)
        if isinstance(o, (CoefficientFunction, GridFunction, int, float)):
            return _SumOfIntegrals([_CutIntegral(CoefficientFunction._lift(o), self)])
        return NotImplemented


def dCut(levelset, domain_type, order=None, subdivlvl=None, time_order=-1, levelset_domain=None, **kwargs):
    """python/__init__.py:781-864"""
    if levelset_domain is not None and type(levelset_domain) == dict:
        lsetdom = levelset_domain
    else:
        lsetdom = {"levelset": levelset, "domain_type": domain_type}
    if order is not None and "order" not in lsetdom:
        lsetdom["order"] = order
    if subdivlvl is not None and "subdivlvl" not in lsetdom:
        lsetdom["subdivlvl"] = subdivlvl
    if time_order > -1 and "time_order" not in lsetdom:
        lsetdom["time_order"] = time_order
    if "tref" in kwargs:
        raise NotImplementedError("tref (space-time) is outside this build")
    return _CutDifferentialSymbol(lsetdom, **kwargs)


class _CutIntegral:
    """CutIntegral (cutint/cutintegral.hpp:21-51)"""

    def __init__(self, integrand, measure):
        self.integrand, self.measure = integrand, measure

    def _rule_args(self, mesh, p):
        m = self.measure
        if isinstance(m, _Measure):
            lset = np.full(mesh.nv, -1.0)
            return lset, NEG, (m.order if m.order is not None else 2 * p), m.definedonelements
        d = m.lsetdom
        if isinstance(d["levelset"], (tuple, list)):
            return tuple(d["levelset"]), d["domain_type"], d.get("order", 2 * p), m.definedonelements
        return d["levelset"], int(d["domain_type"]), d.get("order", 2 * p), m.definedonelements

    def is_bnd(self):
        return isinstance(self.measure, _CutDifferentialSymbol) and self.measure.vb == BND

    def is_eb(self):
        return isinstance(self.measure, _CutDifferentialSymbol) and self.measure.element_boundary

    def _element_matrices_eb(self, mesh, p, precision):
        """forms on the element boundaries (dCut(..., element_boundary=True),
        symboliccutbfi.cpp:488-640): the cut rules of each element's facets
        with the volume basis; rule order the sum of the trial and test
        orders (2p) unless dCut sets it"""
        lset, dt, order, els = self._rule_args(mesh, p)
        if isinstance(lset, (tuple, list)) or getattr(lset, "order", 1) != 1 or \
                self.measure.lsetdom.get("subdivlvl", 0) or getattr(mesh, "is_quad", False):
            raise NotImplementedError("element-boundary integrals: one P1 level set on simplices")
        if els is None:
            el = np.arange(mesh.ne, dtype=np.int32)
        elif isinstance(els, BitArray):
            el = np.nonzero(np.asarray(els.numpy(), dtype=bool))[0].astype(np.int32)
        else:
            el = np.asarray(els, dtype=np.int32)
        fm, vmap, _ = _facet_submesh(mesh, el)
        fl = _bnd_lset(lset, vmap)
        order = int(self.measure.lsetdom.get("order", 2 * p))
        e, tot = None, None
        for part, c in self.integrand.parts.items():
            k = _as_const(c)
            e, m = ElementBoundaryMatrices(mesh, fm, fl, int(dt), order, p, 1 if part == "mass" else 2, el,
                                           precision, None if k is not None else CoefficientFunction._lift(c))
            m = k * m if k is not None else m
            tot = m if tot is None else tot + m
        return e, tot

    def dof_table(self, space):
        """global dofs of the elements this integral runs over"""
        return space.mesh.bnd_dofs(space.order)[0] if self.is_bnd() else space.dofs()[0]

    def _bnd_args(self, mesh, p):
        lset, dt, order, els = self._rule_args(mesh, p)
        if self.measure.lsetdom.get("subdivlvl", 0):
            raise NotImplementedError("boundary integrals: straight cut rules (subdivlvl 0)")
        sub, vmap, amb = _bnd_submesh(mesh)
        return sub, vmap, amb, _bnd_lset(lset, vmap), int(dt), els

    def _cf_order(self, p):
        d = self.measure.lsetdom if isinstance(self.measure, _CutDifferentialSymbol) else {}
        order = d.get("order") if isinstance(self.measure, _CutDifferentialSymbol) else self.measure.order
        if order is None:
            order = 2 * p if "mass" in self.integrand.parts else max(2 * p - 2, 0)
        return int(order)

    def _element_matrices_bnd(self, mesh, p, precision):
        """forms on the boundary elements (dCut(..., vb=BND)): the trace
        space's element matrices on the boundary simplices' cut rules"""
        sub, vmap, amb, lsub, dt, els = self._bnd_args(mesh, p)
        parts = self.integrand.parts
        if all(_as_const(c) is not None for c in parts.values()):
            order = self.measure.lsetdom.get("order", 2 * p)
            e, tot = None, None
            for part, c in parts.items():
                e, m = CutElementMatrices(sub, lsub, dt, int(order), p, 1 if part == "mass" else 2, els, precision)
                tot = _as_const(c) * m if tot is None else tot + _as_const(c) * m
            return e, tot
        order = self._cf_order(p)
        e, tot = None, None
        for part, c in parts.items():
            e, m = CutElementMatricesCF(sub, lsub, dt, order, p, 1 if part == "mass" else 2, els, precision,
                                        CoefficientFunction._lift(c), amb, vmap)
            tot = m if tot is None else tot + m
        return e, tot

    def element_matrices(self, mesh, p, precision):
        if self.is_bnd():
            return self._element_matrices_bnd(mesh, p, precision)
        if self.is_eb():
            return self._element_matrices_eb(mesh, p, precision)
        lset, dt, order, els = self._rule_args(mesh, p)
        mask = (1 if "mass" in self.integrand.parts else 0) | (2 if "laplace" in self.integrand.parts else 0)
        if any(_as_const(c) is None for c in self.integrand.parts.values()):
            return self._element_matrices_cf(mesh, p, precision)
        coefs = {k: _as_const(c) for k, c in self.integrand.parts.items()}
        L = self.measure.lsetdom.get("subdivlvl", 0) if isinstance(self.measure, _CutDifferentialSymbol) else 0
        if getattr(mesh, "is_quad", False):
            # quadrilaterals / hexahedra: the Q1 space on the quad / hex cut rule
            if p != 1 or isinstance(lset, (tuple, list)) or L > 0:
                raise NotImplementedError("quad / hex meshes: order-1 spaces, one P1 level set, subdivlvl 0")
            pol = self.measure.lsetdom.get("quad_dir_policy", OPTIMAL) if isinstance(
                self.measure, _CutDifferentialSymbol) else OPTIMAL

            def kernel(msk):
                return QuadElementMatrices(mesh, lset, dt, order if order is not None else 2, int(pol), msk, els,
                                           precision)
        elif isinstance(lset, (tuple, list)):
            # several level sets: dt holds the domain tuples
            def kernel(msk):
                return MlsetElementMatrices(mesh, tuple(lset), dt, order, p, msk, els, precision)
        elif L > 0 or getattr(lset, "order", 1) == 2:
            # subdivision rule (P2 level sets, subdivlvl > 0)
            def kernel(msk):
                return SubdivElementMatrices(mesh, lset, dt, L, order, p, msk, els, precision)
        else:
            def kernel(msk):
                return CutElementMatrices(mesh, lset, dt, order, p, msk, els, precision)
        if mask == 3 and coefs["mass"] != coefs["laplace"]:
            e1, m1 = kernel(1)
            e2, m2 = kernel(2)
            return e1, coefs["mass"] * m1 + coefs["laplace"] * m2
        e, m = kernel(mask)
        c = coefs.get("mass", coefs.get("laplace", 1.0))
        return e, c * m

    def _cf_rule_args(self, mesh, p, what):
        lset, dt, order, els = self._rule_args(mesh, p)
        L = self.measure.lsetdom.get("subdivlvl", 0) if isinstance(self.measure, _CutDifferentialSymbol) else 0
        if what == "element matrices" and isinstance(lset, (tuple, list)) and not getattr(mesh, "is_quad", False):
            return lset, dt, els  # several level sets: MlsetElementMatrices with cf
        if getattr(mesh, "is_quad", False) or isinstance(lset, (tuple, list)) or L > 0 or \
                getattr(lset, "order", 1) == 2:
            raise NotImplementedError(what + " with coefficient functions: straight cut rules of one P1 level "
                                      "set on simplices")
        return lset, dt, els

    def _element_matrices_cf(self, mesh, p, precision):
        """c u v / c grad u . grad v with coefficient functions: the rule's
        weights scaled by c on the GPU, then the element matrices.  Rule order
        as the reference's SymbolicCutBFI (symboliccutbfi.cpp:107-118): the
        sum of the trial and test orders minus their derivative orders
        (2p, or 2p - 2 for a form with gradients only), unless dCut sets it."""
        lset, dt, els = self._cf_rule_args(mesh, p, "element matrices")
        parts = self.integrand.parts
        order = self._cf_order(p)
        e, tot = None, None
        for part, c in parts.items():
            msk = 1 if part == "mass" else 2
            if isinstance(lset, (tuple, list)):
                e, m = MlsetElementMatrices(mesh, tuple(lset), dt, int(order), p, msk, els, precision,
                                            CoefficientFunction._lift(c))
            else:
                e, m = CutElementMatricesCF(mesh, lset, dt, int(order), p, msk, els, precision,
                                            CoefficientFunction._lift(c))
            tot = m if tot is None else tot + m
        return e, tot

    def element_vectors(self, mesh, p):
        if self.is_eb():
            raise NotImplementedError("element-boundary linear forms (the reference's linear form integrator "
                                      "has no element-boundary variant)")
        f = self.integrand.f
        if self.is_bnd():
            sub, vmap, amb, lsub, dt, els = self._bnd_args(mesh, p)
            order = int(self.measure.lsetdom.get("order", 2 * p))
            if _as_const(f) is None:
                return CutElementVectorsCF(sub, lsub, dt, order, p, els, CoefficientFunction._lift(f), amb, vmap)
            e, v = CutElementVectors(sub, lsub, dt, order, p, els)
            return e, _as_const(f) * v
        if _as_const(f) is None:
            # f v with a coefficient function: rule order 2p
            # (symboliccutlfi.cpp:93-94), f evaluated on the GPU
            lset, dt, els = self._cf_rule_args(mesh, p, "element vectors")
            order = self.measure.lsetdom.get("order") if isinstance(self.measure, _CutDifferentialSymbol) else \
                self.measure.order
            return CutElementVectorsCF(mesh, lset, dt, int(order if order is not None else 2 * p), p, els,
                                       CoefficientFunction._lift(f))
        lset, dt, order, els = self._rule_args(mesh, p)
        e, v = CutElementVectors(mesh, lset, dt, order, p, els)
        return e, _as_const(f) * v

    def integrate(self, mesh, element_wise=False):
        if self.is_eb():
            raise NotImplementedError("Integrate over element boundaries")
        if self.is_bnd():
            sub, vmap, amb, lsub, dt, els = self._bnd_args(mesh, 0)
            order = self.measure.lsetdom.get("order", 5)
            r = IntegrateCF(sub, lsub, dt, int(order if order is not None else 5),
                            CoefficientFunction._lift(self.integrand), True, amb, vmap)
            if els is not None:
                r = r * np.asarray(els.numpy(), dtype=np.float64)
            return r if element_wise else float(np.sum(r))
        lset, dt, order, els = self._rule_args(mesh, 0)
        lsd = {"levelset": lset, "domain_type": dt, "order": order if order is not None else 5}
        if isinstance(self.measure, _CutDifferentialSymbol) and self.measure.lsetdom.get("subdivlvl", 0):
            lsd["subdivlvl"] = self.measure.lsetdom["subdivlvl"]
        r = IntegrateX(levelset_domain=lsd, mesh=mesh, cf=self.integrand, element_wise=True)
        if els is not None:
            r = r * np.asarray(els.numpy(), dtype=np.float64)
        return r if element_wise else float(np.sum(r))


class _SumOfIntegrals:
    def __init__(self, parts):
        self.parts = list(parts)

    def __add__(self, o):
        return _SumOfIntegrals(self.parts + o.parts)


def SymbolicCutBFI(levelset_domain, form, VOL_or_BND=VOL, element_boundary=False, skeleton=False, definedon=None,
                   definedonelements=None, deformation=None):
    """xfem/python_xfem.cpp:855-931"""
    return _SumOfIntegrals([_CutIntegral(form, _CutDifferentialSymbol(
        levelset_domain, vb=VOL_or_BND, definedonelements=definedonelements, deformation=deformation,
        element_boundary=element_boundary, skeleton=skeleton))])


def SymbolicCutLFI(levelset_domain, form, VOL_or_BND=VOL, element_boundary=False, skeleton=False, definedon=None,
                   definedonelements=None, deformation=None):
    """xfem/python_xfem.cpp:1015-1065: the linear form integrator on a level
    set domain (f * v with a number or a coefficient function f)"""
    if element_boundary or skeleton:
        raise Exception("No Facet LFI with Symbolic cuts..")
    return _SumOfIntegrals([_CutIntegral(form, _CutDifferentialSymbol(
        levelset_domain, vb=VOL_or_BND, definedonelements=definedonelements, deformation=deformation))])


def _wrapper_lsetdom(levelset_domain, kwargs):
    """the level-set dictionary with the wrapper's order / time_order
    keywords moved into it (python/__init__.py:228-237)"""
    d = dict(levelset_domain)
    if "order" in kwargs:
        if "order" not in d:
            d["order"] = kwargs["order"]
        del kwargs["order"]
    if "time_order" in kwargs:
        if "time_order" not in d or d["time_order"] == -1:
            d["time_order"] = kwargs["time_order"]
        del kwargs["time_order"]
    return d


def _plain_integral(form, kwargs):
    """NGSolve's SymbolicBFI / SymbolicLFI on uncut elements: the form on
    dx(definedonelements=...)"""
    if kwargs.get("VOL_or_BND", VOL) != VOL or kwargs.get("element_boundary") or kwargs.get("skeleton"):
        raise NotImplementedError("plain (non level-set) integrators: volume elements only")
    return _SumOfIntegrals([_CutIntegral(form, _Measure(kwargs.get("definedonelements")))])


def SymbolicBFIWrapper(levelset_domain=None, *args, **kwargs):
    """python/__init__.py:156-246: SymbolicCutBFI when a level-set
    dictionary is given, otherwise the plain integrator"""
    if levelset_domain is not None and type(levelset_domain) == dict:
        return SymbolicCutBFI(_wrapper_lsetdom(levelset_domain, kwargs), *args, **kwargs)
    form = levelset_domain if levelset_domain is not None else (args[0] if args else kwargs.pop("form"))
    return _plain_integral(form, kwargs)


def SymbolicLFIWrapper(levelset_domain=None, *args, **kwargs):
    """python/__init__.py:248-337: SymbolicCutLFI when a level-set
    dictionary is given, otherwise the plain integrator"""
    if levelset_domain is not None and type(levelset_domain) == dict:
        return SymbolicCutLFI(_wrapper_lsetdom(levelset_domain, kwargs), *args, **kwargs)
    form = levelset_domain if levelset_domain is not None else (args[0] if args else kwargs.pop("form"))
    return _plain_integral(form, kwargs)


SymbolicBFI = SymbolicBFIWrapper
SymbolicLFI = SymbolicLFIWrapper


def _csr_pattern(dofs, n):
    """sorted CSR sparsity pattern of the element dof couplings (host)"""
    nd = dofs.shape[1]
    rows = np.repeat(dofs, nd, axis=1).ravel().astype(np.int64)
    cols = np.tile(dofs, (1, nd)).ravel().astype(np.int64)
    key = np.unique(rows * n + cols)
    r, c = key // n, key % n
    row_ptr = np.zeros(n + 1, np.int32)
    np.add.at(row_ptr, r + 1, 1)
    return np.cumsum(row_ptr).astype(np.int32), c.astype(np.int32)


class BilinearForm:
    """Collects cut integrals; Assemble() builds the CSR matrix (scipy) of the
    P1 or P2 space.  On simplex meshes with one level set the whole assembly
    stays on the GPU (AssembleCutFormsDevice: the matrix graph of the dof
    table, the element matrices of every integral and their scatter); other
    meshes and several level sets compute the element matrices on the GPU and
    scatter them there from host copies.  `elmats` gives the element matrices
    of each integral on request (computed again on the GPU)."""

    def __init__(self, space, symmetric=False, **kwargs):
        self.space = space
        self.integrals = []
        self._elmats = None
        self._prec = None
        self.mat = None

    def __iadd__(self, o):
        if isinstance(o, _SumOfIntegrals):
            self.integrals += o.parts
        else:
            raise TypeError("BilinearForm += expects cut integrals")
        return self

    @property
    def elmats(self):
        if self._elmats is None:
            mesh = self.space.mesh
            self._elmats = [ci.element_matrices(mesh, self.space.order, self._prec) for ci in self.integrals]
        return self._elmats

    def _pattern_elements(self):
        """elements whose dof couplings make the matrix graph (None: all)"""
        return None

    def _device_terms(self, mesh, p):
        """(lset, dt, order, p, mask, elements, coef, subdivlvl) per integral and
        form part, or None when an integral needs the host path"""
        if getattr(mesh, "is_quad", False):
            return None
        terms = []
        for ci in self.integrals:
            if ci.is_bnd() or ci.is_eb():
                return None  # boundary / element-boundary forms: element matrices on the GPU, host scatter
            lset, dt, order, els = ci._rule_args(mesh, p)
            if isinstance(lset, (tuple, list)):
                return None
            L = ci.measure.lsetdom.get("subdivlvl", 0) if isinstance(ci.measure, _CutDifferentialSymbol) else 0
            if (L > 0 or getattr(lset, "order", 1) == 2) and order < 2 * p:
                return None
            parts = {k: _as_const(c) for k, c in ci.integrand.parts.items()}
            if any(c is None for c in parts.values()):
                return None  # coefficient functions: element matrices on the GPU, host scatter
            mass, lap = "mass" in parts, "laplace" in parts
            if mass and lap and parts["mass"] == parts["laplace"]:
                terms.append((lset, int(dt), int(order), p, 3, els, float(parts["mass"]), int(L)))
            else:
                if mass:
                    terms.append((lset, int(dt), int(order), p, 1, els, float(parts["mass"]), int(L)))
                if lap:
                    terms.append((lset, int(dt), int(order), p, 2, els, float(parts["laplace"]), int(L)))
        return terms

    def Assemble(self, precision=None):
        from scipy.sparse import csr_matrix
        prec = ngsxfemglobals.elmat_precision if precision is None else precision
        self._prec = prec
        self._elmats = None
        mesh = self.space.mesh
        p = self.space.order
        if p in (1, 2):
            dofs, n = self.space.dofs()
            terms = self._device_terms(mesh, p)
            pat = self._pattern_elements()
            if terms is not None:
                row_ptr, col_idx, values = AssembleCutFormsDevice(mesh, np.ascontiguousarray(dofs, dtype=np.int32),
                                                                  int(n), terms, int(prec), pat)
            else:
                row_ptr, col_idx = _csr_pattern(dofs if pat is None else dofs[pat], n)
                values = np.zeros(len(col_idx))
                for ci, (els, mats) in zip(self.integrals, self.elmats):
                    tab = ci.dof_table(self.space)
                    values += AssembleCSR(np.ascontiguousarray(tab[els]), np.asarray(mats, dtype=np.float64),
                                          row_ptr, col_idx)
            self.mat = csr_matrix((values, col_idx, row_ptr), shape=(n, n))
        return self


class RestrictedBilinearForm(BilinearForm):
    """A bilinear form whose matrix graph holds only the couplings of the
    active elements (python/__init__.py:973-1037, utils/restrictedblf.cpp:
    8-64,145-164): rows and columns of dofs outside them stay empty.  The
    graph is built on the GPU from the active elements' dofs; an integral
    with an element outside them has entries outside the graph, which is an
    error here (XG_ERR_CAPACITY from the scatter).  facet_restriction only
    concerns DG couplings, which the H1 spaces of this build do not have."""

    def __init__(self, space=None, name="blf", element_restriction=None, facet_restriction=None, trialspace=None,
                 testspace=None, **kwargs):
        if trialspace is not None and testspace is not None:
            if trialspace is not testspace:
                raise NotImplementedError("RestrictedBilinearForm: one space for trial and test functions")
            space = trialspace
        if space is None:
            raise Exception("No space given for RestrictedBilinearForm!")
        super().__init__(space, **kwargs)
        self.name = name
        self.element_restriction = element_restriction
        self.facet_restriction = facet_restriction

    def _pattern_elements(self):
        er = self.element_restriction
        if er is None:
            return None
        if isinstance(er, BitArray):
            return np.nonzero(np.asarray(er.numpy(), dtype=bool))[0].astype(np.int32)
        return np.asarray(er, dtype=np.int32)


class LinearForm:
    """Collects cut source terms f*v*dCut(...); Assemble() computes the element
    vectors on the GPU (xg_cut_elvec, SymbolicCutLinearFormIntegrator) and
    scatters them on the GPU (xg_assemble_vec)."""

    def __init__(self, space, **kwargs):
        self.space = space
        self.integrals = []
        self.vec = None

    def __iadd__(self, o):
        if isinstance(o, _SumOfIntegrals):
            self.integrals += o.parts
        else:
            raise TypeError("LinearForm += expects cut integrals")
        return self

    def Assemble(self):
        mesh = self.space.mesh
        p = self.space.order
        dofs, n = self.space.dofs()
        out = np.zeros(n)
        for ci in self.integrals:
            els, vecs = ci.element_vectors(mesh, p)
            tab = ci.dof_table(self.space)
            out += AssembleVector(np.ascontiguousarray(tab[els]), np.asarray(vecs, dtype=np.float64), n)
        self.vec = out
        return self
