"""Coefficient functions compiled to the device programs (xg_cf.hpp), on the
host: the interpreter compiled with g++ into a throw-away library and run on
random points against numpy evaluations of the same expression trees
(CoefficientFunction.evaluate), the program checks (stack depth, operands),
and the P1 grid-function opcode against the barycentric sum it stands for."""
import ctypes
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "ngsxfem_amd", "csrc")

HARNESS = r"""
#include <cmath>
#include <cstdio>
#define XG_CF_FN inline
#include "xg_cf.hpp"
using namespace xg::cf;
static Prog prog;
// 0 or -1 with the message copied to msg
extern "C" int cf_load(int ncode, const int *code, int nconst, const double *consts, int ngf, const double *const *gf,
                       char *msg, int nmsg) {
  const char *m = validate(ncode, code, nconst, consts, ngf, gf, prog);
  if (!m) return 0;
  snprintf(msg, nmsg, "%s", m);
  return -1;
}
// values at n points x (n x 3), barycentrics lam (n x 4) of a tetrahedron
// with vertex numbers vv
extern "C" void cf_eval(int n, const double *x, const double *lam, const int *vv, double *out) {
  for (int i = 0; i < n; ++i) out[i] = eval<3>(prog, x + 3 * i, lam + 4 * i, vv, 5);
}
"""


@pytest.fixture(scope="module")
def cflib(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("cf")
    src = d / "cfhost.cpp"
    src.write_text(HARNESS)
    so = d / "libcfhost.so"
    subprocess.run([gxx, "-O1", "-std=c++17", "-shared", "-fPIC", "-I" + CSRC, str(src), "-o", str(so)],
                   check=True, capture_output=True)
    return ctypes.CDLL(str(so))


@pytest.fixture(scope="module")
def X():
    from ngsxfem_amd import xfem
    return xfem


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def load(lib, code, consts, gfs=()):
    c = np.ascontiguousarray(code, dtype=np.int32)
    k = np.ascontiguousarray(consts if len(consts) else [0.0], dtype=np.float64)
    g = (ctypes.c_void_p * max(len(gfs), 1))(*[_p(a) for a in gfs])
    msg = ctypes.create_string_buffer(200)
    r = lib.cf_load(len(code), _p(c), len(consts), _p(k), len(gfs), g, msg, 200)
    return r, msg.value.decode()


def run(lib, n, x, lam=None, vv=(0, 1, 2, 3)):
    lam = np.zeros((n, 4)) if lam is None else lam
    out = np.zeros(n)
    lib.cf_eval(n, _p(np.ascontiguousarray(x)), _p(np.ascontiguousarray(lam)),
                _p(np.ascontiguousarray(vv, dtype=np.int32)), _p(out))
    return out


def expressions(X):
    x, y, z = X.x, X.y, X.z
    return [
        X.sin(x) * X.exp(y) + X.sqrt(x * x + 1),
        X.IfPos(x - 0.5, x * y, 2.0 + z),
        (1 + x) ** -2 + z ** 3,
        X.atan(y) / (1 + x * x),
        X.tanh(3 * z - 1) - X.log(2 + x),
        abs(x - 0.5) * X.cos(2 * y),
        X.Max(x, y) - X.Min(y, z) + X.tan(0.3 * z),
        (x + 1) ** 1.5 * (2.0 - y),
        3.0 / (1.0 + x + y + z),
        x * y * z + 2 * x ** 2 - 0.5,
    ]


def test_programs_match_numpy(cflib, X):
    rng = np.random.default_rng(3)
    P = rng.random((200, 3))
    for k, cf in enumerate(expressions(X)):
        code, consts, gfs = cf.program()
        r, msg = load(cflib, code, consts)
        assert r == 0, (k, msg)
        got = run(cflib, len(P), P)
        ref = cf.evaluate(P)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-15, err_msg=str(k))


def test_program_checks(cflib, X):
    x = X.x
    cf = X.sin(x)
    for _ in range(20):  # right-nested sum: 21 values on the stack at once
        cf = X.sin(x) + cf
    code, consts, _ = cf.program()
    r, msg = load(cflib, code, consts)
    assert r == -1 and "stack" in msg
    assert load(cflib, [1, 5], [1.0])[0] == -1          # constant index out of range
    assert load(cflib, [2, 3], [])[0] == -1             # coordinate 3
    assert load(cflib, [3], [])[0] == -1                # ADD on an empty stack
    assert load(cflib, [2, 0, 2, 1], [])[0] == -1       # two values left
    assert load(cflib, [99], [])[0] == -1               # unknown opcode
    assert load(cflib, [19, 0], [])[0] == -1            # grid function slot without grid functions
    assert load(cflib, [2, 0, 8], [])[0] == -1          # POWI without its exponent
    assert load(cflib, [2, 0, 8, -3], [])[0] == 0       # x^-3
    got = run(cflib, 1, np.array([[2.0, 0, 0]]))
    assert abs(got[0] - 0.125) < 1e-16


def test_grid_function_opcode(cflib):
    """GF1 s = sum_k lam_k u_s[v_k] for the element's vertex numbers"""
    rng = np.random.default_rng(5)
    u = rng.random(10)
    vv = np.array([7, 2, 9, 4], dtype=np.int32)
    lam = rng.random((50, 4))
    lam /= lam.sum(1, keepdims=True)
    # u * x + 2: GF1 0, COORD 0, MUL, CONST 0, ADD
    r, msg = load(cflib, [19, 0, 2, 0, 5, 1, 0, 3], [2.0], [u])
    assert r == 0, msg
    P = rng.random((50, 3))
    got = run(cflib, 50, P, lam, vv)
    ref = (lam * u[vv]).sum(1) * P[:, 0] + 2.0
    np.testing.assert_allclose(got, ref, rtol=1e-14)
