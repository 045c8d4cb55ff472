"""The C++ class layer (include/tk/*.hxx: im3d::image3d, conv::filtering,
lapl::unsteady_poisson_functor, segm::rsfe_splitbregman) on the device, each case checked
against the CPU oracle inside tests/cpp/test_class_api.cpp (built by
3dimagetoolkit_amd/build.py as tests/cpp/libtk_cpptest.so and called in-process)."""
import ctypes as C
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "cpp", "libtk_cpptest.so")
SRC = os.path.join(HERE, "cpp", "test_class_api.cpp")


def test_class_layer_headers_compile():
    """CPU: the headers and the test program compile as C++17 without warnings."""
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        SRC], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_class_layer_vs_oracle(tk):
    assert os.path.exists(LIB), "tests/cpp/libtk_cpptest.so is not built (build.py)"
    lib = C.CDLL(LIB)
    lib.tk_cpp_class_api_test.restype = C.c_int
    buf = C.create_string_buffer(1 << 16)
    rc = lib.tk_cpp_class_api_test(buf, len(buf))
    log = buf.value.decode()
    print(log, flush=True)
    assert rc == 0 and log.rstrip().endswith("ALL OK"), log
