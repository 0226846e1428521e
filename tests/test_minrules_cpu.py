"""The minimal symmetric rules of gen_rules.py (used by the moment form when
the rule order is at least the integrand degree): exact for every monomial
of degree <= 2h in the barycentric coordinates, against the closed form
int lam^a = a! / (|a| + D)! on the reference simplex; positive weights,
interior points; and the generated header holds exactly these rows."""
import itertools
import os
import sys
from math import factorial

import numpy as np
import pytest

CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "ngsxfem_amd", "csrc")
sys.path.insert(0, CSRC)
import gen_rules as G  # noqa: E402


@pytest.mark.parametrize("D", [1, 2, 3])
@pytest.mark.parametrize("h", [1, 2, 3])
def test_minimal_rule_exact(D, h):
    pts, ws = G.minrule(D, 2 * h)
    P = np.array(pts)[:, :D]
    W = np.array(ws)
    lam = np.hstack([P, 1.0 - P.sum(axis=1, keepdims=True)])
    assert W.min() > 0 and lam.min() > 0
    for a in itertools.product(range(2 * h + 1), repeat=D + 1):
        if sum(a) > 2 * h:
            continue
        exact = np.prod([factorial(x) for x in a]) / factorial(sum(a) + D)
        got = float(np.sum(W * np.prod(lam ** np.array(a), axis=1)))
        assert abs(got - exact) <= 1e-15, (D, h, a)
    # one degree higher is not exact for D >= 2 (the rule is really minimal-order)
    if D >= 2 and h < 3:
        worst = 0.0
        for a in itertools.product(range(2 * h + 3), repeat=D + 1):
            if sum(a) != 2 * h + 2:
                continue
            exact = np.prod([factorial(x) for x in a]) / factorial(sum(a) + D)
            worst = max(worst, abs(float(np.sum(W * np.prod(lam ** np.array(a), axis=1))) - exact))
        assert worst > 1e-8


def test_minimal_rule_counts():
    counts = {(D, h): len(G.minrule(D, 2 * h)[1]) for D in (2, 3) for h in (1, 2, 3)}
    assert counts == {(2, 1): 3, (2, 2): 6, (2, 3): 12, (3, 1): 4, (3, 2): 14, (3, 3): 24}


def test_rules_table_is_current():
    import subprocess
    txt = subprocess.run([sys.executable, os.path.join(CSRC, "gen_rules.py")], capture_output=True, text=True,
                         check=True).stdout
    with open(os.path.join(CSRC, "rules_table.h")) as f:
        assert f.read().rstrip("\n") == txt.rstrip("\n")
