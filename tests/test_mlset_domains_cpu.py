"""DomainTypeArray / TensorUnion / TensorIntersection (ngsxfem_amd.xfem.mlset)
against the known answers of the reference's
tests/pytests/test_mlset_module_functionality.py (pure set logic, no GPU)."""
import pytest

from ngsxfem_amd.xfem import ANY, IF, NEG, POS
from ngsxfem_amd.xfem.mlset import DomainTypeArray, TensorIntersection, TensorUnion


def test_triangle():
    triangle = DomainTypeArray([(NEG, NEG, NEG)])
    assert ~triangle == DomainTypeArray([(NEG, NEG, POS), (NEG, POS, NEG), (NEG, POS, POS), (POS, NEG, NEG),
                                         (POS, NEG, POS), (POS, POS, NEG), (POS, POS, POS)])
    bnd = triangle.Boundary()
    target = DomainTypeArray([(IF, NEG, NEG), (NEG, IF, NEG), (NEG, NEG, IF)])
    assert bnd == target and bnd.codim == 1
    assert (target | bnd) == bnd
    assert (target & bnd) == bnd
    bb = bnd.Boundary()
    assert bb == DomainTypeArray([(IF, IF, NEG), (NEG, IF, IF), (IF, NEG, IF)]) and bb.codim == 2


def test_operators():
    dta1 = DomainTypeArray((POS, NEG))
    dta2 = DomainTypeArray((POS, POS))
    dta3 = dta1 | dta2
    assert (POS, NEG) in dta3 and (POS, POS) in dta3 and len(dta3) == 2
    dta3 |= DomainTypeArray([(NEG, POS)])
    assert dta3 == DomainTypeArray([(POS, NEG), (POS, POS), (NEG, POS)])
    assert (dta3 & dta1) == DomainTypeArray([(POS, NEG)])
    dta3 &= dta1 | dta2
    assert dta3 == DomainTypeArray([(POS, NEG), (POS, POS)])
    dta5 = DomainTypeArray([(IF, NEG)]) & DomainTypeArray([(POS, NEG)])
    assert dta5 == DomainTypeArray([(IF, NEG)])
    dta5 &= DomainTypeArray([(POS, IF)])
    assert dta5 == DomainTypeArray([(IF, IF)])
    assert (~dta3) == DomainTypeArray([(NEG, POS), (NEG, NEG)])
    dta6 = DomainTypeArray([(IF, NEG, POS)])
    dta6 &= DomainTypeArray([(POS, IF, POS)])
    assert (~dta6) == DomainTypeArray([(IF, NEG, IF), (POS, IF, IF), (IF, POS, IF), (NEG, IF, IF), (IF, IF, NEG)])
    assert DomainTypeArray([(NEG, NEG, NEG)]).Boundary() == DomainTypeArray([(NEG, IF, NEG), (NEG, NEG, IF),
                                                                            (IF, NEG, NEG)])
    assert DomainTypeArray([(POS, IF, NEG)]).Boundary() == DomainTypeArray([(POS, IF, IF), (IF, IF, NEG)])


def test_any_expansion():
    assert DomainTypeArray((NEG, ANY, ANY)) == DomainTypeArray([(NEG, NEG, NEG), (NEG, NEG, POS), (NEG, POS, NEG),
                                                                (NEG, POS, POS)])
    assert DomainTypeArray([(NEG, ANY, ANY), (POS, ANY, ANY)]) == DomainTypeArray([(ANY, ANY, ANY)])
    assert DomainTypeArray([(IF, ANY, ANY), (POS, IF, ANY)]) == DomainTypeArray(
        [(IF, NEG, NEG), (IF, NEG, POS), (IF, POS, NEG), (IF, POS, POS), (POS, IF, NEG), (POS, IF, POS)])


def test_tensor_operators():
    dta1 = DomainTypeArray((POS,))
    dta2 = DomainTypeArray([(POS, NEG), (POS, POS)])
    dta3 = DomainTypeArray((NEG, NEG, IF))
    assert TensorUnion(dta1, dta2) == DomainTypeArray([(POS, ANY, ANY), (ANY, POS, NEG), (ANY, POS, POS)])
    with pytest.raises(Exception):
        TensorUnion(dta1, dta2, dta3)
    assert TensorIntersection(dta1, dta2) == DomainTypeArray([(POS, POS, NEG), (POS, POS, POS)])
    assert TensorIntersection(dta1, dta2, dta3) == DomainTypeArray([(POS, POS, NEG, NEG, NEG, IF),
                                                                    (POS, POS, POS, NEG, NEG, IF)])


def test_safety_checks():
    with pytest.raises(Exception):
        DomainTypeArray([(NEG, IF), (NEG, NEG)])  # different co-dimension
    with pytest.raises(Exception):
        DomainTypeArray([(NEG, NEG), (NEG,)])  # different length
    with pytest.raises(TypeError):
        DomainTypeArray("NEG")
