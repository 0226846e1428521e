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


def test_isometric_submeshes():
    """boundary and facet submeshes (the (D-1)-dimensional copies the
    boundary and element-boundary integrals run on): every copy keeps its
    simplex's edge lengths (so measures, interface factors and the trace
    basis' matrices are unchanged), maps back to the right mesh vertices,
    and facet k of element i is its vertices without vertex k, ascending"""
    from ngsxfem_amd import xfem as X
    for mesh in (X.MakeStructured2DMesh(nx=3, ny=2), X.MakeStructured3DMesh(nx=2, ny=3, nz=2)):
        D = mesh.dim
        sub, vmap, amb = X._bnd_submesh(mesh)
        assert sub.el2v.shape == mesh.bnd2v.shape and sub.coords.shape[1] == D - 1
        assert np.array_equal(vmap, mesh.bnd2v.reshape(-1)) and np.array_equal(amb, mesh.coords[vmap])
        for s2, b in ((sub, mesh.bnd2v), X._facet_submesh(mesh, np.arange(mesh.ne))[:1] + (None,)):
            if b is None:
                E = mesh.el2v
                b = np.stack([E[:, [j for j in range(D + 1) if j != k]] for k in range(D + 1)], 1).reshape(-1, D)
            P, Q = s2.coords[s2.el2v], mesh.coords[b]
            for i in range(D):
                for j in range(i + 1, D):
                    np.testing.assert_allclose(np.linalg.norm(P[:, i] - P[:, j], axis=1),
                                               np.linalg.norm(Q[:, i] - Q[:, j], axis=1), rtol=1e-14, atol=1e-15)
        fm, fvm, _ = X._facet_submesh(mesh, np.array([1, 0]))
        assert fm.el2v.shape == (2 * (D + 1), D)
        assert np.array_equal(fvm.reshape(2, D + 1, D)[0, 0], mesh.el2v[1][1:])
        assert np.array_equal(fvm.reshape(2, D + 1, D)[1, D], mesh.el2v[0][:D])
