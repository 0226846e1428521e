"""CPU-side checks of the C ABI library: it loads and exports every entry
point include/xg_abi.h declares (no compute calls without a GPU)."""
import ctypes

from ngsxfem_amd import _abi


def test_library_exports_declared_symbols():
    L = _abi.lib()
    declared = _abi.declared_symbols()
    assert len(declared) >= 10
    missing = [s for s in declared if not hasattr(L, s)]
    assert not missing, missing


def test_rule_npts_matches_tables():
    # pure host function: conical product sizes n^D, n = order/2 + 1
    for D in (1, 2, 3):
        for order in range(0, 15):
            n = order // 2 + 1
            assert _abi.lib().xg_rule_npts(D, order) == n ** D
    assert _abi.lib().xg_rule_npts(3, 100) < 0
