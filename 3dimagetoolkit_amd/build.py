"""Build libtk.so (HIP, gfx950) in-tree, and the oracle library for the tests.

Usage: python 3dimagetoolkit_amd/build.py [--no-oracle] [-j N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libtk.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TK_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: the FP64 path reproduces the reference's un-fused arithmetic.
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: " + " ".join(cmd) + "\n" + r.stdout)
    return r.stdout


def build_lib(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted((f for f in os.listdir(CSRC) if f.endswith(".hip") or f.endswith(".cpp")),
                  key=lambda f: -os.path.getsize(os.path.join(CSRC, f)))
    cmds = []
    objs = []
    for s in srcs:
        obj = os.path.join(BUILD, s + ".o")
        objs.append(obj)
        src = os.path.join(CSRC, s)
        if s.endswith(".hip"):
            cmds.append([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        else:
            cmds.append(["g++"] + CXX_FLAGS + ["-c", src, "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        outs = list(ex.map(_run, cmds))
    for o in outs:
        # the host pass does not know the GPU-only feature flag; that notice is expected
        lines = [ln for ln in o.splitlines()
                 if "packed-fp32-ops' is not a recognized feature" not in ln
                 and "warnings generated when compiling for host" not in ln
                 and "warning generated when compiling for host" not in ln]
        if any(ln.strip() for ln in lines):
            sys.stderr.write("\n".join(lines) + "\n")
    tmp = LIB + ".tmp"
    # -z defs: an instance a translation unit expects from another one must exist
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-C", odir, "-j4"])
    return os.path.join(odir, "liboracle.so")


CPP_TEST_SRC = os.path.join(ROOT, "tests", "cpp", "test_class_api.cpp")
CPP_TEST_LIB = os.path.join(ROOT, "tests", "cpp", "libtk_cpptest.so")


def build_cpp_tests():
    """The C++ class-layer test library: include/tk/*.hxx over libtk.so, checked against the
    oracle library (test infrastructure; loaded by tests/test_gpu_cpp_api.py)."""
    odir = os.path.join(ROOT, "oracle")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC", "-o", CPP_TEST_LIB, CPP_TEST_SRC,
          "-L", PKG, "-ltk", "-L", odir, "-loracle", "-fopenmp",
          "-Wl,-rpath,$ORIGIN/../../3dimagetoolkit_amd", "-Wl,-rpath,$ORIGIN/../../oracle"])
    return CPP_TEST_LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build_lib(a.j))
    if not a.no_oracle and shutil.which("make"):
        print(build_oracle())
        print(build_cpp_tests())


if __name__ == "__main__":
    main()
