"""The image3d operator API on the device against the reference's semantics.

im3d::image3d<T> compound operators with a scalar or an image and the binary operators
built on them (image3d_imp.hxx:155-378), scalarprod (385-407), norm1 (1387-1398), norminf
with its quirk (1427-1452), vector_abs (892-943), element access (image3d.hxx:252, 262); the
block restriction and trilinear prolongation of the learned coarse-grid correction; the
verbose iteration record of apply (rsfe_splitbregman_imp.hxx:1786-1793, 1835-1844).

Elementwise results are compared bit for bit with numpy in the volume's own type (IEEE
single or double arithmetic, as the reference's loops); reductions against the CPU oracle
(test infrastructure) in FP64, whose summation order differs, to 1e-12 relative; maxima
exactly.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPS = [(0, np.add), (1, np.subtract), (2, np.multiply), (3, np.divide)]
TK_ERR_SHAPE = -4


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("op", OPS, ids=["add", "sub", "mul", "div"])
def test_compound_with_scalar(tk, ctx, rng, dt, op):
    a = (rng.normal(size=(5, 6, 7)) + 3.0).astype(dt)
    s = 1.7000000000000002  # static_cast<T>(scalar) first: FP32 rounds it
    v = ctx.from_numpy(a)
    tk._check(tk.lib.tk_vol_op_scalar(ctx.h, v.h, op[0], s))
    assert np.array_equal(v.numpy(), op[1](a, dt(s)))


@pytest.mark.parametrize("dt,db", [(np.float32, np.float32), (np.float64, np.float64),
                                   (np.float32, np.float64), (np.float64, np.float32)])
@pytest.mark.parametrize("op", OPS, ids=["add", "sub", "mul", "div"])
def test_compound_with_image(tk, ctx, rng, dt, db, op):
    a = (rng.normal(size=(9, 4, 13)) + 2.0).astype(dt)
    b = (rng.normal(size=(9, 4, 13)) + 4.0).astype(db)
    va, vb = ctx.from_numpy(a), ctx.from_numpy(b)
    tk._check(tk.lib.tk_vol_op_vol(ctx.h, va.h, op[0], vb.h))
    # static_cast<T>(addend(i, j, k)), then T arithmetic
    assert np.array_equal(va.numpy(), op[1](a, b.astype(dt)))


def test_compound_with_image_of_other_dims_leaves_it_unchanged(tk, ctx, rng):
    """image3d_imp.hxx:220-223: dims that disagree -> WARNING, nothing done"""
    a = rng.normal(size=(4, 5, 6))
    va, vb = ctx.from_numpy(a), ctx.from_numpy(rng.normal(size=(4, 5, 7)))
    assert tk.lib.tk_vol_op_vol(ctx.h, va.h, 0, vb.h) == TK_ERR_SHAPE
    assert np.array_equal(va.numpy(), a)


def test_testcxx_operator_sequence(tk, orc, ctx):
    """examples/test.cxx:109-201 on 3^3 images: image1 = image2 = 1.5, image3 = 3; then
    += -= *= /= with an image and a scalar, the binary operators and scalarprod, each value
    as the test prints it"""
    one = lambda x: np.full((3, 3, 3), x)
    i1, i2, i3 = ctx.from_numpy(one(1.5)), ctx.from_numpy(one(1.5)), ctx.from_numpy(one(3.0))
    steps = [(0, 1.5, 3.0, 4.5), (1, 1.5, 1.5, 3.0), (2, 1.5, 2.25, 4.5), (3, 1.5, 1.5, 3.0)]
    for op, s, e1, e3 in steps:
        tk._check(tk.lib.tk_vol_op_vol(ctx.h, i1.h, op, i2.h))
        tk._check(tk.lib.tk_vol_op_scalar(ctx.h, i3.h, op, s))
        assert np.all(i1.numpy() == e1) and np.all(i3.numpy() == e3)
        assert np.all(i2.numpy() == 1.5)
    # binary operators: copy of the left operand, then the compound operator
    for op, e in ((0, 3.0), (1, 0.0), (2, 2.25), (3, 1.0)):
        tk._check(tk.lib.tk_vol_copy(i3.h, i1.h))
        tk._check(tk.lib.tk_vol_op_vol(ctx.h, i3.h, op, i2.h))
        assert np.all(i3.numpy() == e)
        tk._check(tk.lib.tk_vol_copy(i3.h, i1.h))
        tk._check(tk.lib.tk_vol_op_scalar(ctx.h, i3.h, op, 1.5))
        assert np.all(i3.numpy() == e)
    i3.fill(1.0)
    import ctypes
    r = ctypes.c_double()
    tk._check(tk.lib.tk_scalarprod(ctx.h, i1.h, i3.h, ctypes.byref(r)))
    assert r.value == orc.scalarprod(i1.numpy(), i3.numpy()) == 40.5


def test_reductions_against_oracle(tk, orc, ctx, rng):
    import ctypes
    d = ctypes.c_double
    for dt in (np.float64, np.float32):
        a = (rng.normal(size=(33, 17, 20)) * 3).astype(dt)
        b = rng.normal(size=(33, 17, 20)).astype(dt)
        a[0, 0, 0] = -7.5  # the running value of norminf starts at |a(0,0,0)|
        a[5, 5, 5] = -1e3  # a large negative entry: ignored by the quirk
        va, vb = ctx.from_numpy(a), ctx.from_numpy(b)
        sp, n1, ninf, nd, nb = d(), d(), d(), d(), d()
        tk._check(tk.lib.tk_scalarprod(ctx.h, va.h, vb.h, ctypes.byref(sp)))
        tk._check(tk.lib.tk_norm1(ctx.h, va.h, ctypes.byref(n1)))
        tk._check(tk.lib.tk_norminf(ctx.h, va.h, ctypes.byref(ninf)))
        tk._check(tk.lib.tk_norminf_diff(ctx.h, va.h, vb.h, ctypes.byref(nd), ctypes.byref(nb)))
        a64, b64 = a.astype(np.float64), b.astype(np.float64)
        # the products of a FP32 scalarprod are formed in FP32
        prod = (a * b).astype(np.float64)
        assert sp.value == pytest.approx(float(np.sum(prod)), rel=1e-12)
        if dt == np.float64:
            assert sp.value == pytest.approx(orc.scalarprod(a64, b64), rel=1e-12)
        assert n1.value == pytest.approx(orc.norm1(a64), rel=1e-12)
        assert ninf.value == orc.norminf(a64) == max(7.5, float(a64.max()))
        assert ninf.value < 1e3
        diff = (a - b).astype(np.float64)  # the temporary of the reference, in T
        assert nd.value == orc.norminf(diff) and nb.value == orc.norminf(b64)


def test_vector_abs(tk, ctx, rng):
    for dt in (np.float32, np.float64):
        f = [rng.normal(size=(6, 7, 8)).astype(dt) for _ in range(3)]
        v = [ctx.from_numpy(x) for x in f]
        res = tk.Volume(ctx, (6, 7, 8), dt)
        tk._check(tk.lib.tk_vector_abs(ctx.h, v[0].h, v[1].h, v[2].h, res.h))
        assert np.array_equal(res.numpy(), np.sqrt((f[0] * f[0] + f[1] * f[1]) + f[2] * f[2]))
        tk._check(tk.lib.tk_vector_abs(ctx.h, v[0].h, v[1].h, None, res.h))
        assert np.array_equal(res.numpy(), np.sqrt(f[0] * f[0] + f[1] * f[1]))


def test_element_access(tk, ctx, rng):
    import ctypes
    for dt in (np.float32, np.float64):
        a = rng.normal(size=(4, 5, 6)).astype(dt)
        v = ctx.from_numpy(a)
        tk._check(tk.lib.tk_vol_set(v.h, 3, 1, 5, 2.25))
        a[3, 1, 5] = 2.25
        assert np.array_equal(v.numpy(), a)
        x = ctypes.c_double()
        tk._check(tk.lib.tk_vol_get(v.h, 2, 4, 0, ctypes.byref(x)))
        assert x.value == float(a[2, 4, 0])
        assert tk.lib.tk_vol_get(v.h, 4, 0, 0, ctypes.byref(x)) == -1  # outside


def prolong_np(c, f):
    """fine = trilinear interpolation of the cell-centred coarse values, constant beyond the
    outermost centres"""
    out = c
    for ax in range(3):
        n = c.shape[ax]
        q = np.arange(n * f)
        x = np.clip((q + 0.5) / f - 0.5, 0, n - 1)
        i0 = np.minimum(np.floor(x).astype(int), max(n - 2, 0))
        w = x - i0 if n > 1 else np.zeros_like(x)
        i1 = np.minimum(i0 + 1, n - 1)
        sh = [1, 1, 1]
        sh[ax] = -1
        w = w.reshape(sh)
        out = np.take(out, i0, axis=ax) * (1 - w) + np.take(out, i1, axis=ax) * w
    return out


@pytest.mark.parametrize("f,cdims", [(4, (8, 6, 5)), (2, (7, 9, 4)), (3, (5, 2, 6))])
def test_restrict_and_prolong(tk, ctx, rng, f, cdims):
    fine = rng.normal(size=tuple(f * c for c in cdims))
    coarse = rng.normal(size=cdims)
    vf, vc = ctx.from_numpy(fine), tk.Volume(ctx, cdims, np.float64)
    tk._check(tk.lib.tk_vol_restrict(ctx.h, vf.h, vc.h))
    ref = fine.reshape(cdims[0], f, cdims[1], f, cdims[2], f).mean(axis=(1, 3, 5))
    assert np.allclose(vc.numpy(), ref, rtol=1e-14, atol=1e-15)
    vc.upload(coarse)
    tk._check(tk.lib.tk_vol_prolong_add(ctx.h, vc.h, vf.h, 2.5))
    assert np.allclose(vf.numpy(), fine + 2.5 * prolong_np(coarse, f), rtol=1e-13, atol=1e-13)


def test_verbose_iteration_record(tk, orc, ctx):
    """verbosity: ls_err of the last of ls_steps sweeps (ls_tol 0) and the norminf ratio, as
    the reference prints them (rsfe_splitbregman_imp.hxx:1786-1793, 1835-1844), against the
    oracle's pieces in FP64 (lexicographic sweeps are bitwise the oracle's); phi is the same
    bit for bit with and without the record"""
    shape = (24, 20, 28)
    u0, phi0 = tk.gen_standard(shape, 5, 1, 2)
    u255 = orc.change_range(u0, 255.0, 0.0)
    phis, stats = [], []
    for verb in (1, 0):
        r = tk.Rsfe(ctx, ctx.from_numpy(u255), ctx.from_numpy(phi0),
                    tk.default_params(sigma=2.0, maxiter=5, tol=0.0, ls_steps=5, ls_tol=0.0,
                                      gs_order=tk.TK_GS_LEX, verbosity=verb))
        st = []
        ph = []
        for _ in range(2):
            st += r.iterate(1)
            ph.append(r.phi())
        r.close()
        phis.append(ph)
        stats.append(st)
    assert all(np.array_equal(a, b) for a, b in zip(phis[0], phis[1]))
    assert stats[1][0]["ls_err"] == -1 and stats[1][0]["tol_inf"] == -1
    # iteration 1: rhs from phi0 (divdb 0); 4 sweeps, aux, 1 sweep
    p = orc.default_params(sigma=2.0, ls_steps=5)
    rhs = orc.rsfe_forcing(u255, phi0, np.zeros(shape), p)
    aux, _, _ = orc.poisson_loop(phi0, rhs, 4, 0.0)
    last = orc.gs_sweep(aux, rhs)
    ls_err = orc.norm_l2(last - aux) / orc.norm_l2(aux)
    tol_inf = [orc.norminf(phis[0][0] - phi0) / orc.norminf(phi0),
               orc.norminf(phis[0][1] - phis[0][0]) / orc.norminf(phis[0][0])]
    say(f"verbose record: ls_err {stats[0][0]['ls_err']:.6e} (oracle {ls_err:.6e}), norm_inf "
        f"{[s['tol_inf'] for s in stats[0]]} (oracle {tol_inf})")
    assert stats[0][0]["ls_err"] == pytest.approx(ls_err, rel=1e-12)
    for s, t in zip(stats[0], tol_inf):
        assert s["tol_inf"] == pytest.approx(t, rel=1e-14)
