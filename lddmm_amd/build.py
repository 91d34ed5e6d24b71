"""Builds liblddmm_amd.so (gfx950) in-tree with hipcc.

`python -m lddmm_amd.build` or `lddmm_amd.build.build()`.  Objects go to
lddmm_amd/_build/, the library to lddmm_amd/liblddmm_amd.so.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "liblddmm_amd.so")
ARCH = os.environ.get("LDDMM_OFFLOAD_ARCH", "gfx950")

SOURCES = ["pair_kernels.hip", "varifold_kernels.hip", "vector_kernels.hip", "batch.hip", "mlp.hip",
           "noalloc.hip", "rates.hip", "engine.cpp", "capi.cpp"]

# -packed-fp32-ops: keep the compiler from forming v_pk_{add,mul,fma}_f32
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src, verbose):
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + ".o")
    cmd = [_hipcc(), "-c", os.path.join(CSRC, src), "-o", obj, "--offload-arch=" + ARCH] + COMMON
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msg = "\n".join(line for line in (r.stderr or "").splitlines()
                    if "packed-fp32-ops" not in line and "ignoring feature" not in line)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, r.stderr))
    return obj, msg.strip()


def build(verbose=False):
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    for _, msg in results:
        if msg and verbose:
            print(msg, file=sys.stderr)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [_hipcc(), "-shared", "-o", tmp] + objs + ["--offload-arch=" + ARCH, "-fPIC",
                                                     "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
