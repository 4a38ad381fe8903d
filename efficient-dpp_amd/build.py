"""Build recipe for libedpp.so: every HIP source compiled for gfx950 with hipcc.

Run as ``python efficient-dpp_amd/build.py`` or through ``__graft_entry__.build()``.
The shared library is written next to this file so that it travels with the
source tree; object files go to ``_build/`` at the repository root.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
INCLUDE = os.path.join(REPO, "include")
BUILD = os.path.join(REPO, "_build")
LIB = os.path.join(PKG_DIR, "libedpp.so")
ARCH = "gfx950"

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

# -packed-fp32-ops: no v_pk_*_f32 instructions are generated
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I", INCLUDE, "-I", CSRC,
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, "-x", "hip", *COMMON, "-c", src, "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    msgs = [ln for ln in r.stderr.splitlines()
            if "warning" in ln and "packed-fp32-ops" not in ln and "not a recognized feature" not in ln]
    if msgs and verbose:
        print("\n".join(msgs))
    return obj


def build(verbose=False, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    # the C++ driver of the reference-shaped API (tests/cpp), linked against the library
    drv_src = os.path.join(REPO, "tests", "cpp", "test_primitives.cpp")
    if os.path.exists(drv_src):
        drv = os.path.join(REPO, "tests", "cpp", "test_primitives")
        obj = os.path.join(BUILD, "test_primitives.o")
        for cmd in ([HIPCC, "-x", "hip", "-O2", "-std=c++17", "-Wno-unused-result", f"--offload-arch={ARCH}",
                     "-I", INCLUDE, "-c", drv_src, "-o", obj],
                    [HIPCC, f"--offload-arch={ARCH}", obj, "-L", PKG_DIR, "-ledpp",
                     "-Wl,-rpath,$ORIGIN/../../efficient-dpp_amd", "-o", drv]):
            if verbose:
                print(" ".join(cmd), flush=True)
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError(f"driver build failed:\n{r.stdout}\n{r.stderr}")
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
