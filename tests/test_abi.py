"""C-ABI library checks that need no GPU: libtk.so loads, exports every entry point
include/tk.h declares, the host-side generators are deterministic, and a context
request without a device fails with a status code instead of crashing."""
import ctypes
import os
import re

import numpy as np

from conftest import ROOT, has_gpu


def declared_functions():
    src = open(os.path.join(ROOT, "include", "tk.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b(tk_[a-z0-9_]+)\s*\(", src)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(tk):
    lib = ctypes.CDLL(tk.LIB_PATH)
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_binding_covers_every_declared_symbol(tk):
    # every C entry point is reachable from the Python binding with a prototype
    for n in declared_functions():
        fn = getattr(tk.lib, n)
        assert fn.argtypes is not None, n


def test_params_struct_matches_header(tk):
    src = open(os.path.join(ROOT, "include", "tk.h")).read()
    body = re.search(r"typedef struct tk_rsfe_params \{(.*?)\} tk_rsfe_params;", src, re.S).group(1)
    body = re.sub(r"/\*.*?\*/", "", body, flags=re.S)
    fields = re.findall(r"\b([a-z_][a-z_0-9]*)\s*(?:\[\d+\])?\s*[,;]", body)
    assert [f for f, _ in tk.RsfeParams._fields_] == fields
    p = tk.default_params()
    # the connected-component and dump defaults of the same constructor
    assert (p.auto_extract_conn_comp, p.auto_extract_conn_comp_freq, list(p.cc_init_pixel),
            p.cc_binary_output, p.cc_init_variables) == (0, 0, [-1, -1, -1], 1, 1)
    assert (p.dumpfrequency, p.outputname, p.cc_threshold) == (0, b"result", 0.5)
    # rsfe_splitbregman_imp.hxx:25-41
    assert (p.sigma, p.lamda, p.lamda1, p.lamda2, p.beta, p.dt) == (3, 5e-2, 1e-5, 1e-5, 100, 1)
    assert (p.ls_steps, p.epsilon, p.a0, p.b0, p.nu, p.tol, p.ls_tol) == (5, 1e-3, 0, 1, 1, 1e-5, 0)
    assert (p.maxiter, p.bc_dir, p.gaussian_pixel_approach) == (100, 0, 0)


def test_generator_deterministic_and_shaped(tk):
    a, pa = tk.gen_standard((64, 64, 64), 11, 1, 2)
    b, pb = tk.gen_standard((64, 64, 64), 11, 1, 2)
    c, _ = tk.gen_standard((64, 64, 64), 12, 1, 2)
    assert np.array_equal(a, b) and np.array_equal(pa, pb) and not np.array_equal(a, c)
    # sphere of radius 20 about the centre
    i, j, k = np.meshgrid(*[np.arange(64)] * 3, indexing="ij")
    r2 = (i - 31.5) ** 2 + (j - 31.5) ** 2 + (k - 31.5) ** 2
    assert np.array_equal(pa, (r2 <= 400).astype(float))
    # background 0.2 times the x bias ramp 0.7 -> 1.3, noise 0.05 twice
    low = a < 0.45
    assert abs(a[low].mean() - 0.2) < 0.01
    assert abs(a[0][a[0] < 0.45].mean() - 0.2 * 0.7) < 0.01
    assert abs(a[63][a[63] < 0.45].mean() - 0.2 * 1.3) < 0.01
    # two ellipsoids of intensity 0.8 (times bias): roughly 4/3 pi (12*9*7 + 8*8*10) voxels
    hi = (a > 0.5).sum()
    assert 0.7 * 5850 < hi < 1.3 * 5850


def test_batch_dims_generator(tk):
    d, ne = tk.gen_batch_dims(5, 64)
    assert d.shape == (64, 3)
    assert set(np.unique(d)) <= {48, 64, 80, 96, 128, 160}
    assert ne.min() >= 1 and ne.max() <= 4


def test_ct_generator_small(tk):
    img, seed = tk.gen_ct((64, 64, 48), 3, 4)
    assert img.min() >= 0 and img.max() <= 1.2 + 1e-12
    assert all(10 <= s for s in seed)
    # the seed lies inside vessel 0: brighter than the tissue slab
    assert img[seed] > 0.6


def test_no_device_is_a_status_not_a_crash(tk):
    if has_gpu():
        return
    h = ctypes.c_void_p()
    rc = tk.lib.tk_ctx_create(0, ctypes.byref(h))
    assert rc in (tk.TK_ERR_HIP, tk.TK_ERR_ARG)
    assert tk.lib.tk_last_error() != b""
