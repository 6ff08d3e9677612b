"""Build recipe for the HIP library (gfx950) and the CPU oracle.

`build_library()` compiles every HIP/C++ source under gaia_amd/csrc into
gaia_amd/libgaia_vbd.so with hipcc; `build_oracle()` runs oracle/Makefile.
Both always recompile (this is the "does it build" check).
"""
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
LIB_PATH = os.path.join(PKG_DIR, "libgaia_vbd.so")
ORACLE_DIR = os.path.join(REPO_DIR, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

ARCH = "gfx950"


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def build_library(verbose=False):
    tmp = LIB_PATH + ".tmp"
    cmd = [
        _hipcc(),
        "--offload-arch=" + ARCH,
        "-O3",
        "-std=c++17",
        "-fPIC",
        "-shared",
        # export only the C ABI (GVBD_API in include/gaia_vbd.h), not the C++ internals
        "-fvisibility=hidden",
        "-fvisibility-inlines-hidden",
        # template instantiations of the standard library keep default visibility: the version
        # script makes the gvbd_* entry points the only exported symbols
        "-Wl,--version-script=" + os.path.join(CSRC, "exports.map"),
        "-Wall",
        "-Wno-unused-function",
        "-Wno-unused-result",
        # no packed-FP32 instructions in device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", os.path.join(REPO_DIR, "include"),
        "-o", tmp,
    ] + sources()
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("hipcc build of libgaia_vbd.so failed")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def build_oracle(verbose=False):
    res = subprocess.run(["make", "-B", "-C", ORACLE_DIR], capture_output=True, text=True)
    if verbose:
        print(res.stdout, flush=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("oracle build failed")
    return ORACLE_LIB


if __name__ == "__main__":
    print(build_library(verbose=True))
    print(build_oracle(verbose=True))
