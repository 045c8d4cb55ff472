"""tk_norm_l2_diff_noalloc (the entry point that allocates nothing) against tk_norm_l2_diff
and tk_norm_l2, the entry points it varies: two calls in a row on one GPU, the second with a
smaller input, in FP64 and FP32.  The two reduce in different orders (the scratch-array
variant sums blocks of 8192 consecutive voxels), so FP64 results agree to 1e-13 relative."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TK_ERR_ARG = -1


def reference(tk, ctx, va, vb):
    d, n = ctypes.c_double(), ctypes.c_double()
    tk._check(tk.lib.tk_norm_l2_diff(ctx.h, va.h, vb.h, ctypes.byref(d)))
    tk._check(tk.lib.tk_norm_l2(ctx.h, vb.h, ctypes.byref(n)))
    return d.value, n.value


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_noalloc_matches_norm_l2_diff(tk, ctx, dt):
    rng = np.random.default_rng(7)
    out = tk.Volume(ctx, (2, 1, 1), np.float64)
    h = (0.5, 1.25, 2.0)
    for shape in ((96, 80, 64), (20, 30, 41)):  # the second call is smaller
        a = rng.normal(size=shape).astype(dt)
        b = rng.normal(size=shape).astype(dt)
        va, vb = ctx.from_numpy(a, h=h), ctx.from_numpy(b, h=h)
        ref = reference(tk, ctx, va, vb)
        tk._check(tk.lib.tk_norm_l2_diff_noalloc(
            tk.lib.tk_ctx_stream(ctx.h), tk.lib.tk_vol_device_ptr(va.h),
            tk.lib.tk_vol_device_ptr(vb.h), a.size, tk._NP2TK[np.dtype(dt)], *h,
            tk.lib.tk_vol_device_ptr(out.h)))
        ctx.sync()
        got = out.numpy().reshape(-1)
        print(f"{np.dtype(dt).name} {shape}: noalloc {got.tolist()} vs {ref}", flush=True)
        assert got[0] == pytest.approx(ref[0], rel=1e-13)
        assert got[1] == pytest.approx(ref[1], rel=1e-13)
        # without b: ||a||_L2 and 0
        tk._check(tk.lib.tk_norm_l2_diff_noalloc(
            tk.lib.tk_ctx_stream(ctx.h), tk.lib.tk_vol_device_ptr(va.h), None, a.size,
            tk._NP2TK[np.dtype(dt)], *h, tk.lib.tk_vol_device_ptr(out.h)))
        ctx.sync()
        na = ctypes.c_double()
        tk._check(tk.lib.tk_norm_l2(ctx.h, va.h, ctypes.byref(na)))
        got = out.numpy().reshape(-1)
        assert got[0] == pytest.approx(na.value, rel=1e-13) and got[1] == 0.0
        va.free()
        vb.free()


def test_noalloc_rejects_sizes_beyond_its_scratch(tk, ctx):
    """checked on the host before anything is enqueued"""
    out = tk.Volume(ctx, (2, 1, 1), np.float64)
    p = tk.lib.tk_vol_device_ptr(out.h)
    rc = tk.lib.tk_norm_l2_diff_noalloc(tk.lib.tk_ctx_stream(ctx.h), p, None, 134217729,
                                        tk.TK_F64, 1.0, 1.0, 1.0, p)
    assert rc == TK_ERR_ARG
    assert b"134217728" in tk.lib.tk_last_error()
