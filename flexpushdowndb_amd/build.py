"""Builds libfpdb_amd.so (HIP kernels + C ABI, gfx950) and the C++ operator
layer in-tree.  Invoked by __graft_entry__.build() and usable directly:

    python -m flexpushdowndb_amd.build
"""
import concurrent.futures
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
POP = os.path.join(PKG, "pop")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FPDB_OFFLOAD_ARCH", "gfx950")

HIP_SOURCES = ["context.hip", "bloom.hip", "hashjoin.hip", "groupby.hip", "starjoin.hip", "serve.hip", "shuffle.hip",
               "segcache.hip", "wire.hip", "strings.hip"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # never emit packed-FP32 VALU ops (v_pk_*_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
    f"-I{ROOT}/include",
]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", src, "-o", out])
    return out


def build_library(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    target = os.path.join(LIB, "libfpdb_amd.so")
    tmp = target + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, target)
    return target


def build_pop(jobs=8):
    """C++ drop-in operator layer (host code over the C ABI)."""
    srcs = sorted(os.path.join(POP, f) for f in os.listdir(POP) if f.endswith(".cpp"))
    if not srcs:
        return None
    os.makedirs(OBJ, exist_ok=True)
    flags = ["-O2", "-fPIC", "-std=c++17", "-Wall", f"-I{ROOT}/include", f"-I{POP}"]

    def cc(src):
        out = os.path.join(OBJ, "pop_" + os.path.basename(src) + ".o")
        _run(["g++", *flags, "-c", src, "-o", out])
        return out

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(cc, srcs))
    target = os.path.join(LIB, "libfpdb_pop.so")
    tmp = target + ".tmp"
    _run(["g++", "-shared", "-fPIC", "-o", tmp, *objs, f"-L{LIB}", "-lfpdb_amd",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, target)
    return target


def build_pop_tests():
    """The operator-level C++ test program (run by tests/test_pop_gpu.py)."""
    src = os.path.join(ROOT, "tests", "cpp", "pop_test.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(ROOT, "tests", "cpp", "pop_test")
    _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{ROOT}/include", f"-I{POP}", src, "-o", out + ".tmp",
          f"-L{LIB}", "-lfpdb_pop", "-lfpdb_amd", f"-Wl,-rpath,{LIB}", "-Wl,-rpath,$ORIGIN/../../flexpushdowndb_amd/lib"])
    os.replace(out + ".tmp", out)
    # the Arrow C Data Interface hooks tests/test_arrow_c_data.py loads with ctypes
    hook = os.path.join(ROOT, "tests", "cpp", "libcdata_roundtrip.so")
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-Wall", f"-I{ROOT}/include", f"-I{POP}",
          os.path.join(ROOT, "tests", "cpp", "cdata_roundtrip.cpp"), "-o", hook + ".tmp",
          f"-L{LIB}", "-lfpdb_pop", "-lfpdb_amd", "-Wl,-rpath,$ORIGIN/../../flexpushdowndb_amd/lib"])
    os.replace(hook + ".tmp", hook)
    return out


def build_all():
    lib = build_library()
    pop = build_pop()
    test = build_pop_tests()
    return lib, pop, test


if __name__ == "__main__":
    for p in build_all():
        if p:
            print(p)
    sys.exit(0)
