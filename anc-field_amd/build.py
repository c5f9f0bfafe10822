"""Build the afx library (libafx.so) and the REPL executable (afx-field) for gfx950.

Run as ``python anc-field_amd/build.py``; ``__graft_entry__.build()`` calls
``build()``.  Outputs land in ``anc-field_amd/lib`` and ``anc-field_amd/bin``
(in-tree, so they travel with the source tree).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
BIN_DIR = os.path.join(HERE, "bin")
OBJ_DIR = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AFX_ARCH", "gfx950")

# IEEE arithmetic in the reference's operation order: no contraction, exact
# division, denormals kept.  Packed-FP32 instructions are disabled.
COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    "-ffp-contract=off",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # the first 16 kernel-argument dwords arrive in SGPRs at wave start
    "-mllvm", "-amdgpu-kernarg-preload-count=16",
]

SOURCES = ["afx_kernels.hip", "afx_host.cpp", "afx_kerkythea.cpp", "afx_mlp.hip"]


def _run(cmd):
    p = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in p.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)
    return out


def _compile(src):
    obj = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"]
    _run([HIPCC] + COMMON + DEVICE + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def lib_path():
    return os.path.join(LIB_DIR, "libafx.so")


def repl_path():
    return os.path.join(BIN_DIR, "afx-field")


def build(verbose=False):
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    os.makedirs(BIN_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = lib_path() + ".tmp"
    _run([HIPCC] + DEVICE + ["-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, lib_path())
    repl_src = os.path.join(CSRC, "afx_repl.cpp")
    if os.path.exists(repl_src):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-o", repl_path(), repl_src,
              "-L" + LIB_DIR, "-lafx", "-Wl,-rpath,$ORIGIN/../lib"])
    ab_src = os.path.join(CSRC, "afx_interior_ab.hip")
    if os.path.exists(ab_src):
        _run([HIPCC] + COMMON + DEVICE + ["-o", os.path.join(BIN_DIR, "afx-interior-ab"), ab_src])
    if verbose:
        print("built", lib_path())
    return lib_path()


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
