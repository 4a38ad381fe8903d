"""Build recipe for the checker libraries under oracle/ (test infrastructure).

* ``oracle/liboracle.so``   — oracle.c, the sequential restatement (gcc -O2).
* ``oracle/libbaseline.so`` — baseline_omp.c, the OpenMP CPU baseline, built
  with the reference's flags ``-O3 -fopenmp -mavx2`` (openmp/CMakeLists.txt:5).
* ``oracle/_ref/libedpp_ref_openmp.so`` — the reference's own OpenMP sources
  compiled from where they lie under /root/reference (only when that tree is
  present): openmp/test_gather_scatter_CPU.cpp, openmp/test_bandwidth_CPU.cpp
  and openmp/util/{utility,log}.cpp, each ``main`` renamed by a -D flag, plus
  ref_shim.cpp for C entry points.  openmp/test_scan_CPU.cpp is not built: it
  needs oneTBB's removed ``tbb/task_scheduler_init.h``, which this image lacks.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference"
REF_OUT = os.path.join(HERE, "_ref")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_oracle(verbose=False):
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", os.path.join(HERE, "oracle.c"), "-lm",
          "-o", os.path.join(HERE, "liboracle.so")], verbose)
    _run(["gcc", "-O3", "-fopenmp", "-mavx2", "-fPIC", "-shared", os.path.join(HERE, "baseline_omp.c"),
          "-o", os.path.join(HERE, "libbaseline.so")], verbose)


def build_ref(verbose=False):
    """Compile the reference's OpenMP path; returns False when the tree is absent."""
    src = os.path.join(REF, "openmp")
    if not os.path.isdir(src):
        return False
    os.makedirs(REF_OUT, exist_ok=True)
    flags = ["g++", "-O3", "-fopenmp", "-mavx2", "-fPIC", "-w", "-std=c++11",
             "-I", src, "-I", os.path.join(src, "util")]
    units = [
        (os.path.join(src, "test_gather_scatter_CPU.cpp"), ["-Dmain=edpp_ref_unused_main_gs"]),
        (os.path.join(src, "test_bandwidth_CPU.cpp"), ["-Dmain=edpp_ref_unused_main_bw"]),
        (os.path.join(src, "util", "utility.cpp"), []),
        (os.path.join(src, "util", "log.cpp"), []),
        (os.path.join(HERE, "ref_shim.cpp"), []),
    ]
    objs = []
    for path, extra in units:
        obj = os.path.join(REF_OUT, os.path.basename(path) + ".o")
        _run(flags + extra + ["-c", path, "-o", obj], verbose)
        objs.append(obj)
    _run(["g++", "-shared", "-fopenmp", "-o", os.path.join(REF_OUT, "libedpp_ref_openmp.so")] + objs, verbose)
    for o in objs:
        os.remove(o)
    return True


if __name__ == "__main__":
    v = "-v" in sys.argv
    build_oracle(v)
    print("reference OpenMP build:", build_ref(v))
