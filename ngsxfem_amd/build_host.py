"""Builds the pybind11 host module ngsxfem_amd/xfem/xfem*.so (C++17, g++),
linked against libxg.so (rpath $ORIGIN/..) and the HIP runtime."""
import os
import subprocess
import sysconfig

import pybind11

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "host", "python_xfem.cpp")
OUTDIR = os.path.join(HERE, "xfem")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def target():
    suffix = sysconfig.get_config_var("EXT_SUFFIX") or ".so"
    return os.path.join(OUTDIR, "xfem" + suffix)


def build(verbose=False):
    out = target()
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-D__HIP_PLATFORM_AMD__",
           "-I" + pybind11.get_include(), "-I" + sysconfig.get_paths()["include"], "-I" + ROCM + "/include",
           SRC, "-o", out + ".tmp", "-L" + HERE, "-l:libxg.so", "-L" + ROCM + "/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath," + ROCM + "/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("host module build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-6000:]))
    os.replace(out + ".tmp", out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)
