"""GPU parity of the whole RSFE Split-Bregman loop (rsfe_splitbregman::apply) against
the CPU FP64 oracle: live at small sizes, and against the committed 64^3 golden
fixtures (tests/golden/make_golden.py).  FP64 plain mode (lexicographic Gauss-Seidel):
phi within 1e-9 max-abs.  FP32 fast mode (red-black): error and Dice stated."""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def say(*a):
    print(*a, flush=True)


def dice(a, b):
    s = a.sum() + b.sum()
    return 1.0 if s == 0 else 2.0 * (a & b).sum() / s


def run_gpu(tk, ctx, u0, phi0, dtype, n_iter, **kw):
    """the harness of examples/rsfe_splitbregman.cxx: image rescaled to [0, 255] on the
    device (image3d::change_range_of_intensity), then apply(myim, init)."""
    vu = ctx.from_numpy(u0.astype(dtype))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0.astype(dtype))
    p = tk.default_params(**kw)
    r = tk.Rsfe(ctx, vu, vi, p)
    st = r.iterate(n_iter)
    phi = r.phi()
    r.close()
    return phi, st


@pytest.mark.parametrize("shape,ls_tol", [((24, 20, 28), 0.0), ((32, 30, 20), 1e-7)])
def test_loop_f64_matches_oracle_small(tk, orc, ctx, shape, ls_tol):
    say("loop f64", shape, ls_tol)
    u0, phi0 = tk.gen_standard(shape, 5, 1, 2)
    p = orc.default_params(sigma=2.0, maxiter=6, tol=0.0, ls_steps=5, ls_tol=ls_tol)
    ref, sw, tl, _ = orc.rsfe_apply(orc.change_range(u0, 255.0, 0.0), phi0, p)
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float64, 6, sigma=2.0, maxiter=6, tol=0.0,
                      ls_steps=5, ls_tol=ls_tol, gs_order=tk.TK_GS_LEX)
    assert [s["sweeps"] for s in st] == list(sw)
    err = np.max(np.abs(phi - ref))
    say(f"loop {shape} ls_tol {ls_tol}: max|phi-ref| = {err:.3e}")
    # FP64 plain-mode bar: 1e-9 max-abs (FFT rounding differs from the oracle's FFT)
    assert err < 1e-9
    assert np.allclose([s["tollerance"] for s in st], tl, rtol=1e-9)


@pytest.mark.parametrize("dtype,sigma,acyclic", [(np.float64, 2.0, 0), (np.float64, 3.0, 1),
                                                 (np.float32, 3.0, 0), (np.float32, 6.0, 0)])
def test_forcing_term_vs_oracle(tk, orc, ctx, dtype, sigma, acyclic):
    """R6 + R7 on their own: the forcing term of the first outer iteration (Heaviside,
    the convolutions of H u0 and H, the region fit, the second pair of convolutions and the
    forcing, rsfe_splitbregman_imp.hxx:1700-1759, with divdb = 0), read back with
    tk_rsfe_rhs, against the oracle's or_rsfe_forcing"""
    shape = (32, 28, 36)
    say("forcing", np.dtype(dtype).name, sigma, "acyclic" if acyclic else "cyclic")
    u0, phi0 = tk.gen_standard(shape, 11, 1, 2)
    u255 = orc.change_range(u0, 255.0, 0.0)
    p = orc.default_params(sigma=sigma, acyclic=acyclic)
    ref = orc.rsfe_forcing(u255, orc.change_range(phi0, 1.0, 0.0), np.zeros(shape), p)
    r = tk.Rsfe(ctx, ctx.from_numpy(u255.astype(dtype)), ctx.from_numpy(phi0.astype(dtype)),
                tk.default_params(sigma=sigma, maxiter=1, tol=0.0, ls_steps=1, ls_tol=0.0,
                                  gs_order=tk.TK_GS_LEX, acyclic=acyclic))
    r.iterate(1)
    rhs = r.rhs()
    r.close()
    err = np.max(np.abs(rhs - ref)) / np.max(np.abs(ref))
    say(f"forcing: max |rhs - ref| / max |ref| = {err:.3e}")
    # FP64: convolution rounding only; FP32: single-precision sums over 2 x 3 axes of taps
    assert err < (1e-11 if dtype == np.float64 else 2e-5)


@pytest.mark.parametrize("shape,sigma2", [((24, 20, 28), 1.6), ((32, 32, 32), 3.0)])
def test_onthego_f64_matches_oracle(tk, orc, ctx, shape, sigma2, tmp_path):
    """tk_rsfe_onthego, the on-the-go re-read (rsfe_splitbregman_imp.hxx:545-578, called at
    1816-1829) after iteration 2: sigma (the filter, kones and kmyim again), lamda1,
    ls_steps, dt and lamda change; FP64 lexicographic against the oracle's switch; then
    save_current and end_now; a structural change is refused"""
    say("onthego f64", shape, sigma2)
    u0, phi0 = tk.gen_standard(shape, 5, 1, 2)
    p1 = orc.default_params(sigma=2.0, maxiter=5, tol=0.0, ls_steps=4, ls_tol=0.0)
    p2 = orc.default_params(sigma=sigma2, maxiter=5, tol=0.0, ls_steps=6, ls_tol=0.0,
                            lamda1=1.4e-5, dt=0.9, lamda=0.08)
    ref, sw, tl, _ = orc.rsfe_apply_onthego(orc.change_range(u0, 255.0, 0.0), phi0, p1, 2, p2)
    vu = ctx.from_numpy(u0)
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0)
    q1 = tk.default_params(sigma=2.0, maxiter=5, tol=0.0, ls_steps=4, ls_tol=0.0,
                           gs_order=tk.TK_GS_LEX)
    q2 = tk.default_params(sigma=sigma2, maxiter=5, tol=0.0, ls_steps=6, ls_tol=0.0,
                           lamda1=1.4e-5, dt=0.9, lamda=0.08, gs_order=tk.TK_GS_LEX)
    r = tk.Rsfe(ctx, vu, vi, q1)
    st = r.iterate(2)
    r.onthego(q2)
    st += r.iterate(5)
    phi = r.phi()
    assert [s["sweeps"] for s in st] == list(sw) == [4, 4, 6, 6, 6]
    err = np.max(np.abs(phi - ref))
    say(f"onthego {shape}: max|phi-ref| = {err:.3e}")
    assert err < 1e-9
    assert np.allclose([s["tollerance"] for s in st], tl, rtol=1e-9)
    r.close()
    # save_current writes <outputname>_2; end_now ends the run
    out = str(tmp_path / "og")
    r = tk.Rsfe(ctx, vu, vi, q1)
    r.iterate(2)
    q3 = tk.default_params(sigma=2.0, maxiter=5, tol=0.0, ls_steps=4, ls_tol=0.0,
                           gs_order=tk.TK_GS_LEX, outputname=out)
    r.onthego(q3, save_current=True, end_now=True)
    assert r.iterate(3) == []
    saved = ctx.read_mhd(out + "_2.mhd").numpy()
    assert np.array_equal(saved, r.phi())
    with pytest.raises(tk.TkError):
        r.onthego(tk.default_params(sigma=2.0, maxiter=5, tol=0.0, ls_steps=4, ls_tol=0.0,
                                    gs_order=tk.TK_GS_REDBLACK))
    r.close()


@pytest.mark.parametrize("shape,sigma", [((24, 20, 28), 2.0), ((32, 32, 32), 3.0),
                                         ((32, 28, 36), 3.0)])
def test_loop_acyclic_f64_matches_oracle(tk, orc, ctx, shape, sigma):
    """set_cv(acyclic_fftw_convolution), the reference's default cv: every convolution of
    the loop (kones, kmyim, the region fit, the forcing) with zero outside the volume.
    At sigma 3 (19 taps) on dims that are multiples of 4 the FP64 register window runs with
    zero halos; 24x20x28 at sigma 2 runs the generic taps pass."""
    say("loop acyclic f64", shape, sigma)
    u0, phi0 = tk.gen_standard(shape, 7, 1, 3)
    gs = tk.TK_GS_LEX
    p = orc.default_params(sigma=sigma, maxiter=5, tol=0.0, ls_steps=5, ls_tol=0.0, acyclic=1)
    ref, sw, tl, _ = orc.rsfe_apply(orc.change_range(u0, 255.0, 0.0), phi0, p)
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float64, 5, sigma=sigma, maxiter=5, tol=0.0,
                      ls_steps=5, ls_tol=0.0, gs_order=gs, acyclic=1)
    err = np.max(np.abs(phi - ref))
    say(f"loop acyclic {shape}: max|phi-ref| = {err:.3e}")
    assert [s["sweeps"] for s in st] == list(sw)
    assert err < 1e-9
    assert np.allclose([s["tollerance"] for s in st], tl, rtol=1e-9)
    # and it is not the cyclic loop
    phic, _ = run_gpu(tk, ctx, u0, phi0, np.float64, 5, sigma=sigma, maxiter=5, tol=0.0,
                      ls_steps=5, ls_tol=0.0, gs_order=gs)
    assert np.max(np.abs(phic - ref)) > 1e-6


@pytest.mark.parametrize("shape,ls_tol", [((24, 20, 28), 0.0), ((32, 30, 20), 0.02)])
def test_loop_neumann_f64_matches_oracle(tk, orc, ctx, shape, ls_tol):
    """set_bc(NEU): onestep_poisson = NeuGaussSeidel with derivative 0 (imp:1294-1322).
    The pure Neumann problem is singular and the loop's right-hand side has a non-zero mean,
    so phi drifts while the relative update falls like 1/sweeps.  A tight gate lets it drift
    out of [a0, b0] for good (ls_tol 1e-3: ~1000 sweeps, then the cut leaves phi = 0 and the
    next right-hand side is rounding noise, whose sweep count no two implementations share);
    the gate here, 0.02, keeps a mask (44-45 sweeps per iteration)."""
    say("loop neumann f64", shape, ls_tol)
    u0, phi0 = tk.gen_standard(shape, 5, 1, 2)
    p = orc.default_params(sigma=2.0, maxiter=6, tol=0.0, ls_steps=5, ls_tol=ls_tol, neumann=1)
    ref, sw, tl, _ = orc.rsfe_apply(orc.change_range(u0, 255.0, 0.0), phi0, p)
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float64, 6, sigma=2.0, maxiter=6, tol=0.0,
                      ls_steps=5, ls_tol=ls_tol, gs_order=tk.TK_GS_LEX, neumann=1)
    assert [s["sweeps"] for s in st] == list(sw)
    assert (ref >= 0.5).sum() > 500
    err = np.max(np.abs(phi - ref))
    say(f"loop neumann {shape} ls_tol {ls_tol}: max|phi-ref| = {err:.3e}, sweeps {list(sw)}")
    assert err < 1e-9
    assert np.allclose([s["tollerance"] for s in st], tl, rtol=1e-9)
    pd = orc.default_params(sigma=2.0, maxiter=6, tol=0.0, ls_steps=5, ls_tol=ls_tol)
    assert np.max(np.abs(orc.rsfe_apply(orc.change_range(u0, 255.0, 0.0), phi0, pd)[0] - ref)) > 1e-6


@pytest.mark.parametrize("acyclic", [0, 1])
def test_loop_blurred_edge_detector_f64_vs_oracle(tk, orc, ctx, acyclic):
    """edge_detector_sigma != 0 (imp:214-227): g from the image convolved by cv with a
    Gaussian of that sigma; FP64 against the oracle, FP32 mask Dice"""
    shape = (24, 20, 28)
    u0, phi0 = tk.gen_standard(shape, 9, 1, 2)
    kw = dict(sigma=2.0, maxiter=5, tol=0.0, ls_steps=5, ls_tol=0.0, acyclic=acyclic,
              edge_detector_sigma=1.5)
    ref, sw, tl, _ = orc.rsfe_apply(orc.change_range(u0, 255.0, 0.0), phi0,
                                    orc.default_params(**kw))
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float64, 5, gs_order=tk.TK_GS_LEX, **kw)
    err = np.max(np.abs(phi - ref))
    say(f"blurred edge detector acyclic {acyclic}: max|phi-ref| = {err:.3e}")
    assert err < 1e-9
    assert np.allclose([s["tollerance"] for s in st], tl, rtol=1e-9)
    plain, _ = run_gpu(tk, ctx, u0, phi0, np.float64, 5, gs_order=tk.TK_GS_LEX,
                       **dict(kw, edge_detector_sigma=0.0))
    assert np.max(np.abs(plain - ref)) > 1e-6
    p32, _ = run_gpu(tk, ctx, u0, phi0, np.float32, 5, gs_order=tk.TK_GS_LEX, **kw)
    d = dice(p32 >= 0.5, ref >= 0.5)
    say(f"  FP32 Dice {d:.6f}")
    assert d > 0.999


def test_loop_neumann_refusals(tk, ctx):
    u0, phi0 = tk.gen_standard((16, 16, 16), 5, 1, 2)
    vu, vi = ctx.from_numpy(u0), ctx.from_numpy(phi0)
    for kw in (dict(gs_order=tk.TK_GS_REDBLACK), dict(ls_res_tol=1e-6)):
        with pytest.raises(tk.TkError):
            tk.Rsfe(ctx, vu, vi, tk.default_params(neumann=1, **kw))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("init_shape", ["same", (15, 16, 11), (16, 15, 12), None, (10, 16, 12)])
def test_init_phi_cases_vs_oracle(tk, orc, ctx, rng, init_shape, dtype):
    """initialize_phi_with_init / _with_cube (rsfe_splitbregman_imp.hxx:48-170) through
    tk_rsfe_create: same dims, one voxel short (faces extended), no init, other dims."""
    shape = (16, 16, 12)
    say("init phi", init_shape, np.dtype(dtype).name)
    u0 = rng.random(shape)
    init = None
    if init_shape is not None:
        init = rng.random(shape if init_shape == "same" else init_shape)
    vu = ctx.from_numpy(u0.astype(dtype))
    vi = ctx.from_numpy(init.astype(dtype)) if init is not None else None
    r = tk.Rsfe(ctx, vu, vi, tk.default_params(a0=0.0, b0=1.0))
    phi = r.phi()
    r.close()
    src = init.astype(dtype).astype(np.float64) if init is not None else None
    ref, case = orc.init_phi(src, shape, a0=0.0, b0=1.0)
    tol = 1e-15 if dtype == np.float64 else 5e-7
    assert np.max(np.abs(phi.astype(np.float64) - ref)) <= tol
    assert case == (0 if init_shape in (None, (10, 16, 12)) else 1)


def load_golden(name):
    d = np.load(os.path.join(GOLDEN, name))
    return d["phi"], d["sweeps"], d["tols"], json.loads(str(d["meta"]))


def golden_case(tk, name):
    phi_ref, sweeps, tols, meta = load_golden(name)
    u0, phi0 = tk.gen_standard(tuple(meta["shape"]), meta["seed"], 1, 2)
    return u0, phi0, phi_ref, sweeps, tols, meta


def test_c0_s5_f64_vs_golden(tk, ctx):
    u0, phi0, ref, sweeps, tols, meta = golden_case(tk, "s5_64.npz")
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float64, meta["maxiter"], sigma=meta["sigma"],
                      maxiter=meta["maxiter"], tol=0.0, ls_steps=5, ls_tol=0.0,
                      gs_order=tk.TK_GS_LEX)
    err = np.max(np.abs(phi - ref))
    nd = int(((phi >= 0.5) != (ref >= 0.5)).sum())
    say(f"s5 64^3 FP64: max|phi-ref| = {err:.3e}, mask voxels differing = {nd}")
    assert err < 1e-9
    assert nd == 0


def test_c0_gated_f64_vs_golden(tk, ctx):
    path = os.path.join(GOLDEN, "c0_64.npz")
    if not os.path.exists(path):
        pytest.skip("c0 golden not generated")
    u0, phi0, ref, sweeps, tols, meta = golden_case(tk, "c0_64.npz")
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float64, meta["maxiter"], sigma=meta["sigma"],
                      maxiter=meta["maxiter"], tol=0.0, ls_steps=5, ls_tol=meta["ls_tol"],
                      gs_order=tk.TK_GS_LEX)
    err = np.max(np.abs(phi - ref))
    nd = int(((phi >= 0.5) != (ref >= 0.5)).sum())
    say(f"c0 64^3 FP64 gated: max|phi-ref| = {err:.3e}, mask voxels differing = {nd}, "
        f"sweeps {[s['sweeps'] for s in st][:5]}... vs {list(sweeps[:5])}")
    assert [s["sweeps"] for s in st] == list(sweeps)
    assert err < 1e-9
    assert nd == 0


def test_c0_s5_f32_lex_vs_golden(tk, ctx):
    """FP32 with FP64 reductions and the reference's lexicographic ordering: the only
    difference from the oracle is precision."""
    u0, phi0, ref, sweeps, tols, meta = golden_case(tk, "s5_64.npz")
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float32, meta["maxiter"], sigma=meta["sigma"],
                      maxiter=meta["maxiter"], tol=0.0, ls_steps=5, ls_tol=0.0,
                      gs_order=tk.TK_GS_LEX)
    err = np.max(np.abs(phi - ref))
    d = dice(phi >= 0.5, ref >= 0.5)
    nd = int(((phi >= 0.5) != (ref >= 0.5)).sum())
    say(f"s5 64^3 FP32 lexicographic: max|phi-ref| = {err:.3e}, Dice = {d:.6f}, "
        f"mask voxels differing = {nd}")
    assert d >= 0.999


def test_c0_s5_f32_redblack_vs_golden(tk, ctx):
    """FP32 red-black: with only 5 sweeps per outer iteration the Poisson solve is far
    from converged, so the sweep ordering itself changes the trajectory; the deviation
    from the lexicographic FP64 oracle is stated here (Dice >= 0.98)."""
    u0, phi0, ref, sweeps, tols, meta = golden_case(tk, "s5_64.npz")
    phi, st = run_gpu(tk, ctx, u0, phi0, np.float32, meta["maxiter"], sigma=meta["sigma"],
                      maxiter=meta["maxiter"], tol=0.0, ls_steps=5, ls_tol=0.0,
                      gs_order=tk.TK_GS_REDBLACK)
    err = np.max(np.abs(phi - ref))
    d = dice(phi >= 0.5, ref >= 0.5)
    say(f"s5 64^3 FP32 red-black: max|phi-ref| = {err:.3e}, Dice = {d:.5f}")
    assert d >= 0.98


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("shape,h", [((24, 20, 28), 1.0), ((33, 21, 64), 1.0), ((9, 70, 68), 0.7),
                                     ((96, 40, 136), 1.0), ((3, 20, 28), 1.0), ((4, 18, 24), 1.0),
                                     ((5, 26, 32), 1.0)])
def test_fused_update_and_div_equal_two_passes(tk, ctx, dtype, shape, h):
    """the update and div in one pass (tiles with recomputed rings, the first and last planes
    of volumes of 3, 4, 5 and more planes, rows and columns, general spacing) against the
    two passes: phi and rhs (which holds the
    divergence of the iteration before) bitwise equal after several iterations; the
    tolerance, whose sums are added in another order, to rounding"""
    say("fused update", shape, h, dtype.__name__)
    u0, phi0 = tk.gen_standard(shape, 7, 1, 2)
    out = []
    for fused in (False, True):
        ctx.set_fused_update(fused)
        vu = ctx.from_numpy(u0.astype(dtype), h=(h, 1.0, h))
        tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
        vi = ctx.from_numpy(phi0.astype(dtype), h=(h, 1.0, h))
        p = tk.default_params(sigma=2.0, maxiter=6, tol=0.0, ls_steps=5, ls_tol=0.0,
                              gs_order=tk.TK_GS_LEX)
        r = tk.Rsfe(ctx, vu, vi, p)
        st = r.iterate(6)
        out.append((r.phi(), r.rhs(), [x["tollerance"] for x in st]))
        r.close()
    ctx.set_fused_update(True)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.allclose(out[0][2], out[1][2], rtol=1e-9 if dtype == np.float32 else 1e-12, atol=0)


@pytest.mark.parametrize("dtype,order", [(np.float32, "redblack"), (np.float64, "lex")])
def test_iterate_ahead_equals_one_call_per_iteration(tk, ctx, dtype, order):
    """tk_rsfe_iterate(n) enqueues each next iteration's convolution passes before it waits
    for the tollerance (tol <= 0, another iteration due in the same call); n calls of one
    iteration each never do.  Both give the same phi, bit for bit, and the same stats."""
    shape = (36, 28, 32)
    u0, phi0 = tk.gen_standard(shape, 5, 1, 2)
    gs = tk.TK_GS_REDBLACK if order == "redblack" else tk.TK_GS_LEX
    p = tk.default_params(sigma=2.0, maxiter=50, tol=0.0, ls_steps=5, ls_tol=0.0, gs_order=gs)
    runs = []
    for _ in range(2):
        vu = ctx.from_numpy(u0.astype(dtype))
        tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
        vi = ctx.from_numpy(phi0.astype(dtype))
        runs.append((tk.Rsfe(ctx, vu, vi, p), vu, vi))
    a, b = runs[0][0], runs[1][0]
    st_a = a.iterate(6)
    st_b = [b.iterate(1)[0] for _ in range(6)]
    phi_a, phi_b = a.phi(), b.phi()
    a.close()
    b.close()
    say(f"ahead {order}: tollerances {[s['tollerance'] for s in st_a]}")
    assert np.array_equal(phi_a, phi_b)
    assert [s["sweeps"] for s in st_a] == [s["sweeps"] for s in st_b]
    assert [s["tollerance"] for s in st_a] == [s["tollerance"] for s in st_b]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("level", [0.25, 0.0])
def test_iterate_ahead_loop_stops_on_its_tollerance(tk, orc, ctx, dtype, level):
    """a loop that stops while the next iteration's convolution passes are already enqueued
    (tol 0, iterate(5)).  a0 == b0 == level: initialize_phi_with_init gives phi == level
    everywhere (x /= max / (b0 - a0) divides by +inf) and cut_phi clamps every Poisson step
    back to it, so phi is exactly stationary.  The tollerance ||phi - phiold|| / ||phiold|| is
    0 for level 0.25 and 0 / 0 = NaN for level 0; either fails `tollerance > tol`
    (rsfe_splitbregman_imp.hxx:1659), so the reference runs one iteration.  Same count, phi
    and tollerance as the oracle; a later call runs nothing."""
    shape = (24, 20, 28)
    u0, phi0 = tk.gen_standard(shape, 5, 1, 2)
    u255 = orc.change_range(u0, 255.0, 0.0)
    p = tk.default_params(sigma=2.0, maxiter=50, tol=0.0, ls_steps=5, ls_tol=0.0, a0=level,
                          b0=level, gs_order=tk.TK_GS_LEX)
    vu = ctx.from_numpy(u255.astype(dtype))
    vi = ctx.from_numpy(phi0.astype(dtype))
    r = tk.Rsfe(ctx, vu, vi, p)
    st = r.iterate(5)
    phi = r.phi()
    again = r.iterate(3)
    phi_again = r.phi()
    r.close()
    ref, sweeps, tols, ost = orc.rsfe_apply(u255, phi0, orc.default_params(
        sigma=2.0, maxiter=50, tol=0.0, ls_steps=5, ls_tol=0.0, a0=level, b0=level))
    say(f"stationary phi = {level} ({np.dtype(dtype).name}): iterations {len(st)} (oracle "
        f"{ost.iters}), tollerance {st[0]['tollerance']} (oracle {tols[0]})")
    assert len(st) == ost.iters == 1
    assert np.all(phi == level) and np.array_equal(phi, ref)
    if level == 0.0:
        assert np.isnan(st[0]["tollerance"]) and np.isnan(tols[0])
    else:
        assert st[0]["tollerance"] == 0.0 == tols[0]
    assert st[0]["sweeps"] == sweeps[0] == 5
    assert again == [] and np.array_equal(phi_again, phi)
