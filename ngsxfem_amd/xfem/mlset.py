"""Domains of several level sets (the reference's ``xfem.mlset``,
python/mlset.py): DomainTypeArray, TensorUnion, TensorIntersection.

A domain is a set of tuples with one domain type (NEG, POS, IF) per level
set; all tuples share the co-dimension (the number of IF entries).  ANY
entries expand into NEG and POS.  The rules and element marks of such
domains run on the GPU (xg_mlset_*: IntegrateX / Integrate with a tuple of
level sets, MultiLevelsetCutInfo, dCut).
"""
import itertools

from . import ANY, IF, NEG, POS


def _norm(dt):
    """enum value -> plain int (NEG 0, POS 1, IF 2; ANY stays ANY)"""
    return int(dt)


class DomainTypeArray:
    """Container of domain tuples (python/mlset.py:10-449).

    dtlist: a tuple of domain types or a list of such tuples; lsets: level
    sets to compress against (regions of zero measure are dropped);
    persistent_compress: keep compressing the results of operations."""

    def __init__(self, dtlist, lsets=None, persistent_compress=False, do_safety_checks=True):
        if isinstance(dtlist, tuple):
            dtlist = [dtlist]
        if not isinstance(dtlist, list) or len(dtlist) == 0 or not isinstance(dtlist[0], tuple):
            raise TypeError("Invalid input: dtlist must be a tuple or a list of tuples")
        n = len(dtlist[0])
        self.codim = sum(1 for dt in dtlist[0] if _norm(dt) == _norm(IF))
        if do_safety_checks:
            for dtt in dtlist:
                if len(dtt) != n:
                    raise Exception("DomainTypeArray initialised with tuples of different length!")
                for dt in dtt:
                    if _norm(dt) not in (_norm(NEG), _norm(POS), _norm(IF), _norm(ANY)):
                        raise Exception("Initialised with something other than NEG, POS, IF, ANY")
                if sum(1 for dt in dtt if _norm(dt) == _norm(IF)) != self.codim:
                    raise Exception("DomainTypeArray initialised with tuples of different co-dimension!")
            if lsets is not None and len(lsets) != n:
                raise Exception("The number of level sets does not match the length of the arrays domain tuples!")
        out = set()
        for dtt in dtlist:
            dtt = tuple(_norm(d) for d in dtt)
            anys = [i for i, d in enumerate(dtt) if d == _norm(ANY)]
            if not anys:
                out.add(dtt)
                continue
            for choice in itertools.product((_norm(POS), _norm(NEG)), repeat=len(anys)):
                t = list(dtt)
                for i, c in zip(anys, choice):
                    t[i] = c
                out.add(tuple(t))
        self.as_list = list(out)
        self.lsets = lsets
        self.persistent_compress = persistent_compress
        if self.lsets:
            self.Compress(self.lsets, self.persistent_compress)
        if not self.persistent_compress:
            self.lsets = None

    def __len__(self):
        return len(self.as_list)

    def __iter__(self):
        return iter(self.as_list)

    def __contains__(self, dtt):
        return tuple(_norm(d) for d in dtt) in self.as_list

    def __eq__(self, other):
        return (isinstance(other, DomainTypeArray) and self.codim == other.codim
                and sorted(self.as_list) == sorted(other.as_list))

    def __ne__(self, other):
        return not self.__eq__(other)

    def __str__(self):
        return "DomainTypeArray:\n  Co-dim : %d\n  Domains: [%s]" % (
            self.codim, ", ".join(str(t) for t in self.as_list))

    def _combined_compression(self, other):
        if self.persistent_compress != other.persistent_compress:
            if self.persistent_compress or other.persistent_compress:
                print("WARNING: Combined domains with non matching persistent compression status: \n"
                      " The returned domain will not be compressed.")
            return None, False
        if not self.persistent_compress:
            return None, False
        for a, b in zip(self.lsets, other.lsets):
            if a is not b:
                raise Exception("Operator only possible for domains with the same level sets")
        return self.lsets, True

    def __or__(self, other):
        if self.codim != other.codim:
            raise Exception("Co-dims don't match: Union not possible!")
        ls, pers = self._combined_compression(other)
        return DomainTypeArray(self.as_list + other.as_list, ls, pers)

    def __and__(self, other):
        ls, pers = self._combined_compression(other)
        common = [t for t in self.as_list if t in other.as_list]
        if not common:
            # regions meet where, level set by level set, the domain types
            # agree, or one of them is the interface: then the interface
            for a, b in itertools.product(self.as_list, other.as_list):
                t = []
                for x, y in zip(a, b):
                    if x == y:
                        t.append(x)
                    elif x in (_norm(NEG), _norm(POS)) and y in (_norm(NEG), _norm(POS)):
                        break
                    else:
                        t.append(_norm(IF))
                else:
                    common.append(tuple(t))
        return DomainTypeArray(common, ls, pers)

    def __invert__(self):
        n = len(self.as_list[0])
        if self.codim == 0:
            allt = set(itertools.product((_norm(POS), _norm(NEG)), repeat=n))
        else:
            allt = {t for t in itertools.product((_norm(POS), _norm(NEG), _norm(IF)), repeat=n)
                    if t.count(_norm(IF)) == self.codim}
        return DomainTypeArray(list(allt - set(self.as_list)), self.lsets, self.persistent_compress)

    def __ior__(self, other):
        return self.__or__(other)

    def __iand__(self, other):
        return self.__and__(other)

    def _measure(self, lsets, dtt):
        from . import Integrate
        mesh = lsets[0].space.mesh
        return Integrate({"levelset": tuple(lsets), "domain_type": dtt}, cf=1, mesh=mesh, order=0)

    def Compress(self, lsets, persistent=False):
        """drop the tuples whose region has zero measure for lsets"""
        if len(lsets) != len(self.as_list[0]):
            raise Exception("The number of level sets does not match the length of the arrays domain tuples!")
        self.as_list = [t for t in self.as_list if abs(self._measure(lsets, t)) > 1e-12]
        if persistent:
            self.persistent_compress = True
            self.lsets = lsets
        return None

    def Boundary(self):
        """the boundary of the region: each non-IF entry whose flip is not in
        the region becomes IF (python/mlset.py:280-322)"""
        if self.codim >= 3:
            raise Exception("Boundary does not make sense for codim >=3")
        out = []
        for t in self.as_list:
            for i, dt in enumerate(t):
                if dt == _norm(IF):
                    continue
                if (t[:i] + (_norm(POS),) + t[i + 1:] in self.as_list
                        and t[:i] + (_norm(NEG),) + t[i + 1:] in self.as_list):
                    continue
                b = t[:i] + (_norm(IF),) + t[i + 1:]
                if self.lsets and abs(self._measure(self.lsets, b)) < 1e-12:
                    continue
                out.append(b)
        if self.persistent_compress:
            return DomainTypeArray(out, self.lsets, True)
        return DomainTypeArray(out)

    def Indicator(self, lsets):
        raise NotImplementedError("DomainTypeArray.Indicator needs NGSolve coefficient functions")

    def IndicatorSmoothed(self, lsets, eps=0.03):
        raise NotImplementedError("DomainTypeArray.IndicatorSmoothed needs NGSolve coefficient functions")

    def GetOuterNormals(self, lsetsp1):
        raise NotImplementedError("DomainTypeArray.GetOuterNormals needs NGSolve coefficient functions")


def TensorUnion(*args):
    """union of regions that depend on disjoint sets of level sets
    (python/mlset.py:451-490): the tuples are concatenated level set block by
    level set block, every other block ANY"""
    codim = args[0].codim
    for a in args:
        if not isinstance(a, DomainTypeArray):
            raise TypeError("TensorUnion only possible for DomainTypeArrays")
        if a.codim != codim:
            raise Exception("Cannot form TensorUnion for arrays of different codimension")
    widths = [len(a.as_list[0]) for a in args]
    total = sum(widths)
    out = []
    start = 0
    for a, w in zip(args, widths):
        for t in a.as_list:
            full = [_norm(ANY)] * total
            full[start:start + w] = t
            out.append(tuple(full))
        start += w
    return DomainTypeArray(out)


def TensorIntersection(*args):
    """intersection of regions that depend on disjoint sets of level sets
    (python/mlset.py:493-519): every combination of one tuple per region,
    concatenated"""
    for a in args:
        if not isinstance(a, DomainTypeArray):
            raise TypeError("TensorIntersection only possible for DomainTypeArrays")
    out = [tuple(itertools.chain(*combo)) for combo in itertools.product(*[a.as_list for a in args])]
    return DomainTypeArray(out)
