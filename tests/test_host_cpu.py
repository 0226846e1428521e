"""CPU checks of the host layer: the pybind11 module loads and its
GPU-independent pieces (enums, BitArray, polynomial coefficient functions,
HAS) behave like the reference's."""
import numpy as np


def test_module_loads_and_enums():
    from ngsxfem_amd import xfem as X
    assert int(X.NEG) == 0 and int(X.POS) == 1 and int(X.IF) == 2
    assert [int(getattr(X, n)) for n in ("NO", "CDOM_NEG", "CDOM_POS", "UNCUT", "CDOM_IF", "HASNEG", "HASPOS", "ANY")] \
        == list(range(8))
    assert X.HAS(X.NEG) == X.HASNEG and X.HAS(X.POS) == X.HASPOS and X.HAS(X.IF) == X.IF


def test_bitarray_layout_and_ops():
    from ngsxfem_amd import xfem as X
    a = X.BitArray(70)
    a[0] = True
    a[33] = True
    a[69] = True
    assert a.NumSet() == 3 and len(a) == 70
    w = a.words()
    assert w[0] == 1 and w[1] == 2 and w[2] == 32
    b = ~a
    assert b.NumSet() == 67
    assert (a | b).NumSet() == 70 and (a & b).NumSet() == 0
    assert list(np.nonzero(a.numpy())[0]) == [0, 33, 69]


def test_polynomial_cf_terms():
    from ngsxfem_amd import xfem as X
    cf = (1 - 2 * X.x - 2 * X.y) ** 2 + X.z * 3
    X3 = np.array([[0.1, 0.2, 0.3]])
    assert abs(cf.evaluate(X3)[0] - ((1 - 0.6) ** 2 + 0.9)) < 1e-14
    assert cf.is_poly() and not X.sqrt(cf).is_poly()
