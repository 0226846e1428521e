"""Builds libxg.so (HIP kernels + C ABI for gfx950) in-tree.

    python -m ngsxfem_amd.build_ext          # or __graft_entry__.build()

Every .hip file under csrc/ is compiled with hipcc for gfx950 and linked into
ngsxfem_amd/libxg.so.  Flags:
  -ffp-contract=off      FP64 geometry in IEEE source order (bit-exact marking);
                         hot FP32 loops opt back in with a local pragma
  -packed-fp32-ops off   no v_pk_{add,mul,fma}_f32: some of their forms have given
                         wrong results on this hardware now and then
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxg.so")
ARCH = os.environ.get("XG_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH, "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-I" + os.path.join(HERE, "..", "include"),
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CXXFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-8000:]))
    return obj


def regen_tables():
    """rules_table.h and moments_table.h are generated; regenerate when the
    generator is newer."""
    for gen, out in (("gen_rules.py", "rules_table.h"), ("gen_moments.py", "moments_table.h")):
        gen = os.path.join(CSRC, gen)
        out = os.path.join(CSRC, out)
        if not os.path.exists(out) or os.path.getmtime(gen) > os.path.getmtime(out):
            txt = subprocess.run([sys.executable, gen], capture_output=True, text=True, check=True).stdout
            with open(out, "w") as f:
                f.write(txt)


def build(verbose=False, jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    regen_tables()
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=min(jobs, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-8000:]))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
