"""The RSFE Split-Bregman path at the full sizes of the benchmark configurations.

C1: 256^3, the ellipsoid generator scaled x4, sigma 6, 50 outer iterations, ls_steps 5,
    ls_tol 0.  FP64 plain mode (lexicographic Gauss-Seidel) against the CPU FP64 oracle's
    golden (tests/golden/c1_256.npz: mask of every voxel, phi at 65536 seeded samples, FP64
    sum and sum of squares, per-iteration tollerance); FP32 (red-black) error and Dice
    stated against the same golden.
C2: CT-like 512x512x384 FP32 (tk_gen_ct: background -1000 +- 20, tissue slab, 12 vessels on
    Catmull-Rom splines, rescaled to [0,1], z bias 0.8 -> 1.2), sigma 4, 50 iterations.  The
    converged mask's seeded 6-connected component (seed = first control point of vessel 0)
    must be bit-exact against the oracle's connected_component on the same phi; the FP64
    plain mode is checked against the CPU oracle's own C2 run (tests/golden/c2_ct.npz) at
    iterations 5 and 10, within the spread the oracle shows under a 1e-12 input change; the
    FP32 mask is stated against both.  The image is rescaled to [0, 255] as the reference's
    driver does and phi starts from the reference's cube (apply(myim)).
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
C1_SEED = 20240602
C2_SEED = 20240603


def say(*a):
    print(*a, flush=True)


def dice(a, b):
    s = int(a.sum()) + int(b.sum())
    return 1.0 if s == 0 else 2.0 * int((a & b).sum()) / s


def run(tk, ctx, u0, phi0, dtype, n_iter, **kw):
    vu = ctx.from_numpy(u0.astype(dtype))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0.astype(dtype))
    del u0, phi0
    r = tk.Rsfe(ctx, vu, vi, tk.default_params(**kw))
    vi.free()
    st = r.iterate(n_iter)
    phi = r.phi()
    r.close()
    vu.free()
    return phi, st


@pytest.fixture(scope="module")
def c1_golden():
    path = os.path.join(GOLDEN, "c1_256.npz")
    d = np.load(path)
    return {k: d[k] for k in d.files}


@pytest.fixture(scope="module")
def c1_input(tk):
    return tk.gen_standard((256, 256, 256), C1_SEED, 4, 2)


def test_c1_256_f64_vs_golden(tk, ctx, c1_golden, c1_input):
    g = c1_golden
    meta = json.loads(str(g["meta"]))
    say("C1 256^3 FP64 lexicographic,", meta["maxiter"], "iterations")
    u0, phi0 = c1_input
    phi, st = run(tk, ctx, u0, phi0, np.float64, 50, sigma=6.0, maxiter=50, tol=0.0,
                  ls_steps=5, ls_tol=0.0, gs_order=tk.TK_GS_LEX)
    flat = phi.reshape(-1)
    err = float(np.max(np.abs(flat[g["sample_idx"]] - g["phi_samples"])))
    mask = flat >= 0.5
    ref_mask = np.unpackbits(g["mask_bits"])[:flat.size].astype(bool)
    nd = int(np.count_nonzero(mask != ref_mask))
    rs = abs(float(flat.sum()) - float(g["phi_sum"])) / abs(float(g["phi_sum"]))
    rq = abs(float((flat * flat).sum()) - float(g["phi_sumsq"])) / float(g["phi_sumsq"])
    tols = np.array([s["tollerance"] for s in st])
    say(f"C1 FP64: max|phi-ref| over {g['sample_idx'].size} samples = {err:.3e}, "
        f"mask voxels differing = {nd} of {flat.size}, sum rel {rs:.2e}, sumsq rel {rq:.2e}, "
        f"last tollerance {tols[-1]:.6e} vs {g['tols'][-1]:.6e}")
    assert err < 1e-9
    assert nd == 0
    assert rs < 1e-12 and rq < 1e-12
    assert np.allclose(tols, g["tols"], rtol=1e-9)


def test_c1_256_f32_redblack_vs_golden(tk, ctx, c1_golden, c1_input):
    g = c1_golden
    u0, phi0 = c1_input
    phi, st = run(tk, ctx, u0, phi0, np.float32, 50, sigma=6.0, maxiter=50, tol=0.0,
                  ls_steps=5, ls_tol=0.0, gs_order=tk.TK_GS_REDBLACK)
    flat = phi.reshape(-1).astype(np.float64)
    err = float(np.max(np.abs(flat[g["sample_idx"]] - g["phi_samples"])))
    mask = flat >= 0.5
    ref_mask = np.unpackbits(g["mask_bits"])[:flat.size].astype(bool)
    d = dice(mask, ref_mask)
    nd = int(np.count_nonzero(mask != ref_mask))
    say(f"C1 FP32 red-black: max|phi-ref| (samples) = {err:.3e}, Dice = {d:.6f}, "
        f"mask voxels differing = {nd}")
    # red-black iterates differ from the lexicographic ones the oracle runs (same fixed
    # point); the bar is the segmentation, stated as Dice
    assert d >= 0.98


def test_c1_256_f32_lex_vs_golden(tk, ctx, c1_golden, c1_input):
    g = c1_golden
    u0, phi0 = c1_input
    phi, st = run(tk, ctx, u0, phi0, np.float32, 50, sigma=6.0, maxiter=50, tol=0.0,
                  ls_steps=5, ls_tol=0.0, gs_order=tk.TK_GS_LEX)
    flat = phi.reshape(-1).astype(np.float64)
    err = float(np.max(np.abs(flat[g["sample_idx"]] - g["phi_samples"])))
    mask = flat >= 0.5
    ref_mask = np.unpackbits(g["mask_bits"])[:flat.size].astype(bool)
    d = dice(mask, ref_mask)
    say(f"C1 FP32 lexicographic: max|phi-ref| (samples) = {err:.3e}, Dice = {d:.6f}, "
        f"mask voxels differing = {int(np.count_nonzero(mask != ref_mask))}")
    assert d >= 0.999


@pytest.fixture(scope="module")
def c2_input(tk):
    img, seed = tk.gen_ct((512, 512, 384), C2_SEED, 12)
    return img.astype(np.float32), seed


def c2_run(tk, ctx, img32, dtype, order):
    """apply(myim) (cube initial contour) on the image rescaled to [0, 255] as the driver
    does; returns the run (caller closes it)."""
    vu = ctx.from_numpy(img32.astype(dtype))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    r = tk.Rsfe(ctx, vu, None, tk.default_params(sigma=4.0, maxiter=50, tol=0.0, ls_steps=5,
                                                  ls_tol=0.0, gs_order=order))
    vu.free()
    r.iterate(50)
    return r


def bw_float32(phi32, thr=0.5):
    """im_to_black_and_white for T = float (image3d_imp.hxx:1604): pivot formed as
    min + (max - min) * threshold with the product in double, stored in T."""
    mn, mx = np.float32(phi32.min()), np.float32(phi32.max())
    pivot = np.float32(np.float64(mn) + np.float64(np.float32(mx - mn)) * thr)
    return (phi32 >= pivot).astype(np.uint8)


def test_c2_ct_f32_mask_and_cc_bit_exact(tk, orc, ctx, c2_input):
    img, seed = c2_input
    say("C2 512x512x384 FP32 red-black, sigma 4, 50 iterations; CC seed", seed)
    r = c2_run(tk, ctx, img, np.float32, tk.TK_GS_REDBLACK)
    vphi = tk.Volume(ctx, img.shape, np.float32)
    tk._check(tk.lib.tk_vol_copy(vphi.h, tk.lib.tk_rsfe_phi(r.h)))
    r.close()
    phi = vphi.numpy()
    # threshold on the device equals the float semantics of the reference
    bw = tk.Volume(ctx, img.shape, np.uint8)
    tk._check(tk.lib.tk_im_to_black_and_white(ctx.h, vphi.h, 0.5, 0, bw.h))
    mask = bw.numpy()
    ref_bw = bw_float32(phi)
    assert np.array_equal(mask, ref_bw)
    # seeded 6-connected component on the device vs the oracle on the same mask
    cc = tk.Volume(ctx, img.shape, np.uint8)
    passes = tk.cc_seeded(ctx, vphi, seed, 0.5, False, cc)
    got = cc.numpy()
    ref, sweeps = orc.connected_component(ref_bw.astype(np.float64), seed, 0.5, False)
    nd = int(np.count_nonzero(got != ref.astype(np.uint8)))
    say(f"C2: mask voxels {int(mask.sum())}, component voxels {int(got.sum())}, "
        f"device passes {passes}, oracle sweeps {sweeps}, voxels differing {nd}")
    assert nd == 0
    assert got.sum() > 0


def c2_masks(tk, ctx, img, dt, order, checkpoints, perturb=0.0):
    x = img.astype(np.float64)
    if perturb:
        # every 7th voxel raised by `perturb` (the min-max rescale keeps it: a scaling
        # of the whole image would cancel)
        x.reshape(-1)[::7] += perturb
    vu = ctx.from_numpy(x.astype(dt))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    r = tk.Rsfe(ctx, vu, None, tk.default_params(sigma=4.0, maxiter=50, tol=0.0, ls_steps=5,
                                                  ls_tol=0.0, gs_order=order))
    vu.free()
    out, done = [], 0
    for n in checkpoints:
        r.iterate(n - done)
        done = n
        out.append(r.phi().astype(np.float64).reshape(-1))
    r.close()
    return out


def test_c2_ct_f32_vs_f64_plain(tk, ctx, c2_input):
    """FP32 against the FP64 plain mode.  The C2 trajectory is unstable: two FP64 runs
    whose inputs differ by 1e-12 at every 7th voxel (of values in [0, 1.2]) separate by
    ~1e4 every five iterations and end far apart at iteration 50 (printed below), so the
    FP32/FP64 agreement is asserted at iteration 10 and only stated at 50."""
    img, seed = c2_input
    cps = [10, 50]
    ref = c2_masks(tk, ctx, img, np.float64, tk.TK_GS_LEX, cps)
    pert = c2_masks(tk, ctx, img, np.float64, tk.TK_GS_LEX, cps, perturb=1e-12)
    lex = c2_masks(tk, ctx, img, np.float32, tk.TK_GS_LEX, cps)
    rb = c2_masks(tk, ctx, img, np.float32, tk.TK_GS_REDBLACK, cps)
    res = {}
    for name, got in (("FP64, input + 1e-12 at every 7th voxel", pert), ("FP32 lexicographic", lex),
                      ("FP32 red-black", rb)):
        for q, n in enumerate(cps):
            e = float(np.max(np.abs(got[q] - ref[q])))
            d = dice(got[q] >= 0.5, ref[q] >= 0.5)
            res[(name, n)] = (e, d)
            say(f"C2 {name} vs FP64 at iteration {n}: max|dphi| = {e:.3e}, Dice = {d:.6f}")
    e, d = res[("FP32 lexicographic", 10)]
    assert e < 5e-3 and d >= 0.9999
    assert res[("FP32 red-black", 10)][1] >= 0.99


@pytest.fixture(scope="module")
def c2_golden():
    d = np.load(os.path.join(GOLDEN, "c2_ct.npz"))
    return {k: d[k] for k in d.files}


def test_c2_ct_f64_lex_vs_oracle_golden(tk, ctx, c2_input, c2_golden):
    """The FP64 plain mode against the CPU oracle's own run of C2 (tests/golden/c2_ct.npz)
    at iterations 5 and 10.  The oracle's second run, whose input differs by 1e-12 at every
    7th voxel, moved phi by max|dphi| = pert[0] at iteration 5 and pert[1] at 10 with both
    masks unchanged; the GPU run (whose FFT convolution differs from the oracle's FFT in
    the last bits) must stay within 100x of that and reproduce both masks voxel for voxel.
    The oracle's later checkpoints show the same perturbation growing to Dice 0.94 at
    iteration 30, so the comparison stops at 10."""
    g = c2_golden
    meta = json.loads(str(g["meta"]))
    pert = meta["pert_max_abs"]
    say(f"C2 oracle sensitivity (input + 1e-12 at every 7th voxel): max|dphi| {pert[0]:.3e} "
        f"at 5, {pert[1]:.3e} at 10; Dice at {meta['checkpoints']}: {meta['pert_dice']}")
    img, seed = c2_input
    idx = g["sample_idx"]
    cps = [5, 10]
    f64 = c2_masks(tk, ctx, img, np.float64, tk.TK_GS_LEX, cps)
    f32 = c2_masks(tk, ctx, img, np.float32, tk.TK_GS_LEX, cps)
    n = f64[0].size
    for q, it in enumerate(cps):
        ref_mask = np.unpackbits(g[f"mask_bits_{it}"])[:n].astype(bool)
        e = float(np.max(np.abs(f64[q][idx] - g[f"phi_samples_{it}"])))
        s = float(f64[q].sum())
        ss = float((f64[q] * f64[q]).sum())
        nd = int(np.count_nonzero((f64[q] >= 0.5) != ref_mask))
        d32 = dice(f32[q] >= 0.5, ref_mask)
        e32 = float(np.max(np.abs(f32[q][idx] - g[f"phi_samples_{it}"])))
        say(f"C2 FP64 lexicographic vs oracle at iteration {it}: max|dphi| (samples) {e:.3e}, "
            f"sum rel {abs(s - g[f'phi_sum_{it}']) / abs(g[f'phi_sum_{it}']):.2e}, sumsq rel "
            f"{abs(ss - g[f'phi_sumsq_{it}']) / g[f'phi_sumsq_{it}']:.2e}, mask voxels "
            f"differing {nd} of {int(ref_mask.sum())}; FP32 lexicographic max|dphi| {e32:.3e}, "
            f"Dice {d32:.6f}")
        assert e <= 100 * pert[q] + 1e-11
        assert nd == 0
        assert abs(s - g[f"phi_sum_{it}"]) <= 1e-9 * abs(g[f"phi_sum_{it}"])
        assert abs(ss - g[f"phi_sumsq_{it}"]) <= 1e-9 * g[f"phi_sumsq_{it}"]
        assert d32 >= 0.9999


C3_SEED = 20240604


def rescale255(u0):
    """image3d::change_range_of_intensity(255, 0) (image3d_imp.hxx:1627-1635) in FP64."""
    mn, mx = u0.min(), u0.max()
    return (u0 - mn) / ((mx - mn) / 255.0) + 0.0


def c3_batch(tk, ctx, dims, nell, vols, dtype=np.float32, gs_order=None):
    if gs_order is None:
        gs_order = tk.TK_GS_REDBLACK if dtype == np.float32 else tk.TK_GS_LEX
    p = tk.default_params(sigma=3.0, maxiter=30, tol=0.0, ls_steps=5, ls_tol=0.0,
                          gs_order=gs_order)
    b = tk.Batch(ctx, dims[vols], p, dtype)
    for q, v in enumerate(vols):
        u0, phi0 = tk.gen_standard(tuple(int(x) for x in dims[v]), C3_SEED + v, 1, int(nell[v]))
        b.set_volume(q, rescale255(u0), phi0)
    return b


def test_c3_graph_replays_match_eager(tk, ctx):
    """the captured two-iteration graph, replayed, gives bitwise the eager iterations"""
    dims, nell = tk.gen_batch_dims(C3_SEED, 64)
    vols = list(range(8))
    ga = c3_batch(tk, ctx, dims, nell, vols)
    ea = c3_batch(tk, ctx, dims, nell, vols)
    ga.iterate(2, use_graph=True)
    ea.iterate(2, use_graph=False)
    for q in range(len(vols)):  # host work between replays, then more replays
        assert np.array_equal(ga.phi(q), ea.phi(q))
    ga.iterate(5, use_graph=True)  # two replays + one eager iteration
    ea.iterate(5, use_graph=False)
    info = ga.info()
    say("C3 graph:", info)
    assert info["graph_nodes"] > 0 and info["iterations"] == 7
    for q in range(len(vols)):
        assert np.array_equal(ga.phi(q), ea.phi(q))
    # after an odd count the buffers have swapped roles: the graph is captured again
    ga.iterate(2, use_graph=True)
    ea.iterate(2, use_graph=False)
    assert ga.info()["iterations"] == 9
    for q in range(len(vols)):
        assert np.array_equal(ga.phi(q), ea.phi(q))
    ga.close()
    ea.close()


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_c3_batch_one_launch_sweeps_equal_single_runs(tk, ctx, dtype):
    """the staged batch (each red-black sweep of every volume in one launch) gives bitwise
    the values of each volume run on its own (its own sweep launches and tile choice), in
    eager mode and through the captured graph"""
    dims, nell = tk.gen_batch_dims(C3_SEED, 64)
    vols = list(range(12))
    b = c3_batch(tk, ctx, dims, nell, vols, dtype=dtype, gs_order=tk.TK_GS_REDBLACK)
    g = c3_batch(tk, ctx, dims, nell, vols, dtype=dtype, gs_order=tk.TK_GS_REDBLACK)
    b.iterate(3, use_graph=False)
    g.iterate(3, use_graph=True)
    for q, v in enumerate(vols):
        shape = tuple(int(x) for x in dims[v])
        u0, phi0 = tk.gen_standard(shape, C3_SEED + v, 1, int(nell[v]))
        vu, vi = ctx.from_numpy(rescale255(u0).astype(dtype)), ctx.from_numpy(phi0.astype(dtype))
        r = tk.Rsfe(ctx, vu, vi, tk.default_params(sigma=3.0, maxiter=30, tol=0.0, ls_steps=5,
                                                   ls_tol=0.0, gs_order=tk.TK_GS_REDBLACK))
        r.iterate(3)
        single = r.phi(dtype)
        r.close()
        assert np.array_equal(b.phi(q).astype(dtype), single), v
        assert np.array_equal(g.phi(q).astype(dtype), single), v
    say("C3 staged batch:", g.info())
    b.close()
    g.close()


def test_c3_batch_f32_lex_10_iterations_vs_golden(tk, ctx):
    """every one of the 64 volumes after 10 outer iterations, before rounding-level
    differences can grow along an unstable trajectory: the FP32 batch with the reference's
    (lexicographic) sweeps against the FP64 oracle stopped at the same iteration"""
    d = np.load(os.path.join(GOLDEN, "c3_batch_it10.npz"))
    dims, nell = d["dims"], d["nell"]
    vols = list(range(64))
    b = c3_batch(tk, ctx, dims, nell, vols, gs_order=tk.TK_GS_LEX)
    b.iterate(10, use_graph=True)
    dices = []
    for v in vols:
        n = int(np.prod(dims[v]))
        ref = np.unpackbits(d[f"mask_{v}"])[:n].astype(bool)
        dices.append(dice((b.phi(v) >= 0.5).reshape(-1), ref))
    b.close()
    dices = np.array(dices)
    say(f"C3 64 volumes FP32 lex, 10 iterations, vs FP64 oracle: Dice min {dices.min():.6f} "
        f"mean {dices.mean():.6f}")
    assert dices.min() >= 0.999


def c3_masks(tk, ctx, dims, nell, dtype, order, n_iter=30):
    b = c3_batch(tk, ctx, dims, nell, list(range(64)), dtype=dtype, gs_order=order)
    b.iterate(n_iter, use_graph=True)
    out = [(b.phi(v) >= 0.5).reshape(-1) for v in range(64)]
    b.close()
    return out


def c3_dices(a, b):
    return np.array([dice(x, y) for x, y in zip(a, b)])


def c3_golden_masks(d):
    return [np.unpackbits(d[f"mask_{v}"])[:int(np.prod(d["dims"][v]))].astype(bool)
            for v in range(64)]


def say_dices(name, dc):
    low = [(int(v), round(float(dc[v]), 5)) for v in np.argsort(dc)[:4]]
    say(f"C3 {name}: Dice min {dc.min():.6f} mean {dc.mean():.6f}, lowest (volume, Dice) {low}")


def test_c3_batch_f64_lex_vs_golden(tk, ctx):
    """the 64 uneven volumes in FP64 with the reference's (lexicographic) sweeps after 30
    outer iterations: every mask equals the FP64 oracle's"""
    d = np.load(os.path.join(GOLDEN, "c3_batch.npz"))
    dc = c3_dices(c3_masks(tk, ctx, d["dims"], d["nell"], np.float64, tk.TK_GS_LEX),
                  c3_golden_masks(d))
    say_dices("FP64 lex vs FP64 oracle", dc)
    assert dc.min() >= 0.99999


@pytest.mark.parametrize("order", ["lex", "redblack"])
def test_c3_batch_f32_vs_golden(tk, ctx, order):
    """per-volume Dice of the FP32 batch against the FP64 oracle, split into its two causes.
    Precision: the FP32 batch against the FP64 batch of the same sweep order; every volume
    but one agrees to >= 0.995, and volume 14 (160x96x128, 3 ellipsoids) to ~0.954 in either
    order.  The FP64 oracle itself does not move on that volume when its input is rounded to
    FP32 or raised by 1e-7 at every 7th voxel (tests/golden/c3_sensitivity.json), so the
    deviation comes from FP32 arithmetic inside the loop, not from an unstable trajectory.
    Sweep order: lexicographic FP64 equals the oracle (test above); red-black sweeps (5 per
    outer iteration, far from converged) follow another iteration, in FP64 as well: volume 41
    ends at Dice 0.814 and volume 60 at 0.905 whatever the precision."""
    d = np.load(os.path.join(GOLDEN, "c3_batch.npz"))
    dims, nell = d["dims"], d["nell"]
    gs = tk.TK_GS_LEX if order == "lex" else tk.TK_GS_REDBLACK
    m32 = c3_masks(tk, ctx, dims, nell, np.float32, gs)
    m64 = c3_masks(tk, ctx, dims, nell, np.float64, gs)
    ref = c3_golden_masks(d)
    prec, vs_ref, order_part = c3_dices(m32, m64), c3_dices(m32, ref), c3_dices(m64, ref)
    say_dices(f"FP32 {order} vs FP64 {order} (precision)", prec)
    say_dices(f"FP64 {order} vs FP64 oracle (sweep order)", order_part)
    say_dices(f"FP32 {order} vs FP64 oracle", vs_ref)
    sens = json.load(open(os.path.join(GOLDEN, "c3_sensitivity.json")))
    assert min(min(v for k, v in r.items() if k.startswith(("input", "every"))) for r in
               sens.values()) == 1.0
    assert prec.min() >= 0.95 and int((prec < 0.995).sum()) <= 1 and prec.mean() >= 0.999
    if order == "lex":
        assert order_part.min() >= 0.99999
        assert vs_ref.min() >= 0.95 and int((vs_ref < 0.995).sum()) <= 1
    else:
        assert order_part.min() >= 0.80 and order_part.mean() >= 0.975
        assert vs_ref.min() >= 0.80 and vs_ref.mean() >= 0.975
