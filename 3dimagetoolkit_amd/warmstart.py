"""The learned Poisson warm start: training pairs from the solver, training, evaluation.

Host orchestration over libtk.so (every array operation runs on the device):

- collect_pairs: Split-Bregman runs (FP64, red-black) on seeded ellipsoid phantoms; at
  chosen outer iterations the Poisson right-hand side the solver built is logged together
  with its converged solution phi* (the Dirichlet problem solved from zero to a relative
  residual of 1e-6).  These are the (rhs, solution) pairs the solver itself emits.
- train: Adam steps of tk_unet_train_step on random pairs; out_scale is the mean ratio
  rms(phi*) / rms(rhs) of the training pairs.
- evaluate: the same Split-Bregman runs with the Poisson step gated on the residual,
  once with the reference's start (previous iterate) and once with the learned guess
  (kept only when, after 16 sweeps each, its residual is lower).

The runs are FP64: a relative residual of 1e-6 is below what FP32 storage of phi allows
(phi is ~100x rhs at 64^3, and its FP32 rounding alone leaves a residual of ~5e-6).  The
network itself computes in BF16 with FP32 accumulation either way.
"""
import ctypes as C
import importlib
import time

import numpy as np

SEED_TRAIN = 30000
SEED_EVAL = 50000


def _tk():
    return importlib.import_module("3dimagetoolkit_amd")


def _run(tk, ctx, seed, shape, dtype, order, **kw):
    u0, phi0 = tk.gen_standard(shape, seed, max(1, shape[0] // 64), 1 + seed % 4)
    vu = ctx.from_numpy(u0.astype(dtype))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0.astype(dtype))
    p = tk.default_params(gs_order=order, **kw)
    r = tk.Rsfe(ctx, vu, vi, p)
    vu.free()
    vi.free()
    return r


def collect_pairs(ctx, n_seeds, shape=(64, 64, 64), sigma=3.0, log_iters=(1, 3, 6, 10, 15),
                  res_tol=1e-6, seed0=SEED_TRAIN, max_sweeps=200000,
                  residual_snapshots=(16, 128, 512, 1024, 2048), log=None):
    """returns lists of float32 arrays (input, target) and the sweeps each solve took: per
    logged iteration the pair (rhs, phi*) and, for each snapshot of the solve from zero after
    k sweeps, the pair (residual r_k, correction phi* - phi_k) of the same linear problem"""
    tk = _tk()
    rhs_l, sol_l, sweeps_l = [], [], []
    last = max(log_iters)
    for q in range(n_seeds):
        r = _run(tk, ctx, seed0 + q, shape, np.float64, tk.TK_GS_REDBLACK, sigma=sigma,
                 maxiter=1 << 20, tol=0.0, ls_steps=5, ls_tol=0.0)
        for it in range(1, last + 1):
            r.iterate(1)
            if it not in log_iters:
                continue
            rhs = tk.Volume(ctx, shape, np.float64)
            tk._check(tk.lib.tk_vol_copy(rhs.h, tk.lib.tk_rsfe_rhs(r.h)))
            phi = tk.Volume(ctx, shape, np.float64)
            res_v = tk.Volume(ctx, shape, np.float64)
            # residuals along the solve from zero, for (residual, correction) pairs
            snaps, done = [], 0
            for k in residual_snapshots:
                tk._check(tk.lib.tk_poisson_dir_sweeps(ctx.h, phi.h, rhs.h, 1.0, 0.0,
                                                        tk.TK_GS_REDBLACK, k - done))
                done = k
                tk._check(tk.lib.tk_poisson_residual_field(ctx.h, phi.h, rhs.h, res_v.h))
                snaps.append((res_v.numpy(), phi.numpy()))
            sw, res = C.c_int(), C.c_double()
            tk._check(tk.lib.tk_poisson_solve_residual(ctx.h, phi.h, rhs.h, 1.0, 0.0,
                                                        tk.TK_GS_REDBLACK, res_tol, 64,
                                                        max_sweeps, C.byref(sw), C.byref(res)))
            sol = phi.numpy()
            rhs_l.append(rhs.numpy().astype(np.float32))
            sol_l.append(sol.astype(np.float32))
            for rk, pk in snaps:
                rhs_l.append(rk.astype(np.float32))
                sol_l.append((sol - pk).astype(np.float32))
            sweeps_l.append(sw.value + done)
            rhs.free()
            phi.free()
            res_v.free()
        r.close()
        if log and (q + 1) % 10 == 0:
            log(f"pairs: {q + 1}/{n_seeds} seeds, {len(rhs_l)} pairs, "
                f"sweeps from zero (mean) {np.mean(sweeps_l):.0f}")
    return rhs_l, sol_l, sweeps_l


def train(ctx, net, rhs_l, sol_l, steps, lr=1e-3, seed=7, log=None):
    """Adam on random pairs with a cosine learning-rate decay; returns the loss history"""
    tk = _tk()
    shape = net.dims
    ratios = [np.sqrt(np.mean(s.astype(np.float64) ** 2)) / np.sqrt(np.mean(r.astype(np.float64) ** 2))
              for r, s in zip(rhs_l, sol_l)]
    net.out_scale = float(np.mean(ratios))
    vr = [ctx.from_numpy(r.astype(np.float32)) for r in rhs_l]
    vs = [ctx.from_numpy(s.astype(np.float32)) for s in sol_l]
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for k in range(steps):
        i = int(rng.integers(len(vr)))
        a = lr * 0.5 * (1 + np.cos(np.pi * k / steps))
        hist.append(net.train_step(vr[i], vs[i], a))
        if log and (k + 1) % max(1, steps // 10) == 0:
            log(f"train: step {k + 1}/{steps}, loss (last 50) {np.mean(hist[-50:]):.4e}, "
                f"{time.perf_counter() - t0:.1f} s")
    for v in vr + vs:
        v.free()
    return hist


def guess_quality(ctx, net, rhs_l, sol_l):
    """relative L2 error of the network's guess against phi*, per pair"""
    tk = _tk()
    out = []
    for r, s in zip(rhs_l, sol_l):
        vr = ctx.from_numpy(r.astype(np.float64))
        vg = tk.Volume(ctx, r.shape, np.float64)
        net.predict(vr, vg)
        g = vg.numpy().astype(np.float64)
        out.append(np.linalg.norm(g - s) / np.linalg.norm(s))
        vr.free()
        vg.free()
    return np.array(out)


def evaluate(ctx, net, n_seeds, shape=(64, 64, 64), sigma=3.0, n_iter=15, res_tol=1e-6,
             seed0=SEED_EVAL, rounds=8, presweeps=512, tight_tol=None, log=None):
    """per seed: sweeps and seconds of the residual-gated loop with the plain start and with
    the learned start, the Dice of their masks, and how often the guess was kept.  tight_tol:
    also a plain run gated at that tighter residual, and the Dice of each mask against its
    mask (how close each run comes to a more accurate solution of the same loop)"""
    tk = _tk()
    rows = []
    modes = ("plain", "learned") + (("tight",) if tight_tol else ())
    for q in range(n_seeds):
        res = {}
        for mode in modes:
            r = _run(tk, ctx, seed0 + q, shape, np.float64, tk.TK_GS_REDBLACK, sigma=sigma,
                     maxiter=n_iter, tol=0.0,
                     ls_res_tol=tight_tol if mode == "tight" else res_tol, res_check_every=64,
                     max_sweeps=400000, warm_start=rounds if mode == "learned" else 0,
                     ws_presweeps=presweeps)
            if mode == "learned":
                tk._check(tk.lib.tk_rsfe_set_warm_start(r.h, net.h))
            ctx.sync()
            t0 = time.perf_counter()
            st = r.iterate(n_iter)
            ctx.sync()
            dt = time.perf_counter() - t0
            res[mode] = dict(sweeps=sum(s["sweeps"] for s in st), seconds=dt,
                             used=sum(s["guess_used"] for s in st),
                             mask=r.phi() >= 0.5)
            r.close()
        def _dice(a, b):
            tot = int(a.sum()) + int(b.sum())
            return 1.0 if tot == 0 else 2 * int((a & b).sum()) / tot

        dice = _dice(res["plain"]["mask"], res["learned"]["mask"])
        rows.append(dict(seed=seed0 + q, sweeps_plain=res["plain"]["sweeps"],
                         sweeps_learned=res["learned"]["sweeps"],
                         s_plain=res["plain"]["seconds"], s_learned=res["learned"]["seconds"],
                         guess_used=res["learned"]["used"], dice=float(dice)))
        if tight_tol:
            rows[-1]["dice_plain_tight"] = _dice(res["plain"]["mask"], res["tight"]["mask"])
            rows[-1]["dice_learned_tight"] = _dice(res["learned"]["mask"], res["tight"]["mask"])
        if log:
            log(f"eval seed {seed0 + q}: sweeps {rows[-1]['sweeps_plain']} -> "
                f"{rows[-1]['sweeps_learned']}, {rows[-1]['s_plain']:.3f} s -> "
                f"{rows[-1]['s_learned']:.3f} s, guess kept {rows[-1]['guess_used']}/{n_iter}, "
                f"Dice {dice:.6f}")
    return rows


# ---- 256^3: the network as a learned coarse-grid correction --------------------------
SEED_TRAIN_256 = 70000
SEED_EVAL_256 = 90000


def _vol_copy(tk, ctx, src_h, shape, dtype=np.float64, h=(1.0, 1.0, 1.0)):
    v = tk.Volume(ctx, shape, dtype, h=h)
    tk._check(tk.lib.tk_vol_copy(v.h, src_h))
    return v


def coarse_cycles(tk, ctx, phi, rhs, f, pre, max_cycles, res_tol=1e-6, target_tol=1e-4,
                  pairs=None, net=None):
    """two-grid cycles on the fine problem -lap phi = rhs (Dirichlet 0 faces): `pre` red-black
    sweeps, the residual averaged over f^3 blocks, the coarse problem (spacing f) solved by
    red-black sweeps to a relative residual target_tol from zero, its interpolation added,
    `pre` sweeps.  With pairs a list, (restricted residual, coarse solution) is appended for
    every cycle (FP32 device volumes of the coarse dims): the pairs the coarse solve emits.
    With net (a network on the coarse dims), the cycles add the network's correction instead
    of the coarse solution (the pairs then follow the states the network itself produces).
    Returns the fine sweeps run and the final relative residual."""
    shape = tuple(s // f for s in (phi.shape))
    H = (float(f),) * 3
    resid = tk.Volume(ctx, phi.shape, np.float64)
    cr = tk.Volume(ctx, shape, np.float64, h=H)
    ce = tk.Volume(ctx, shape, np.float64, h=H)
    cg = tk.Volume(ctx, shape, np.float64, h=H) if net is not None else None
    total, res = 0, 1.0
    for _ in range(max_cycles):
        tk._check(tk.lib.tk_poisson_dir_sweeps(ctx.h, phi.h, rhs.h, 1.0, 0.0, tk.TK_GS_REDBLACK,
                                                pre))
        tk._check(tk.lib.tk_poisson_residual_field(ctx.h, phi.h, rhs.h, resid.h))
        tk._check(tk.lib.tk_vol_restrict(ctx.h, resid.h, cr.h))
        ce.fill(0.0)
        sw, rr = C.c_int(), C.c_double()
        # dt 1e6: the sweep's fixed point does not depend on dt, and a large one relaxes with
        # omega ~ 1 instead of 6/7 (at spacing 4, 1/dt would dominate 2 sum 1/h^2)
        tk._check(tk.lib.tk_poisson_solve_residual(ctx.h, ce.h, cr.h, 1e6, 0.0,
                                                    tk.TK_GS_REDBLACK, target_tol, 64, 200000,
                                                    C.byref(sw), C.byref(rr)))
        if pairs is not None:
            a = tk.Volume(ctx, shape, np.float32, h=H)
            b = tk.Volume(ctx, shape, np.float32, h=H)
            tk._check(tk.lib.tk_vol_copy(a.h, cr.h))
            tk._check(tk.lib.tk_vol_copy(b.h, ce.h))
            pairs.append((a, b))
        if net is not None:
            net.predict(cr, cg)
        tk._check(tk.lib.tk_vol_prolong_add(ctx.h, (cg if net is not None else ce).h, phi.h,
                                             1.0))
        tk._check(tk.lib.tk_poisson_dir_sweeps(ctx.h, phi.h, rhs.h, 1.0, 0.0, tk.TK_GS_REDBLACK,
                                                pre))
        total += 2 * pre
        res = tk.poisson_residual(ctx, phi, rhs)
        if res <= res_tol:
            break
    for v in (resid, cr, ce) + ((cg,) if cg is not None else ()):
        v.free()
    return total, res


def _c4_run(tk, ctx, seed, shape, sigma, n_ell=2, phantom=None, **kw):
    """the C1 generator (n_ell ellipsoids, scaled to the shape), image rescaled to [0, 255],
    FP64, red-black sweeps; phantom: the generator's (u0, phi0) if already made"""
    u0, phi0 = phantom if phantom is not None else \
        tk.gen_standard(shape, seed, max(1, shape[0] // 64), n_ell)
    vu = ctx.from_numpy(u0)
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0)
    r = tk.Rsfe(ctx, vu, vi, tk.default_params(gs_order=tk.TK_GS_REDBLACK, sigma=sigma, tol=0.0,
                                               **kw))
    vu.free()
    vi.free()
    return r


def collect_coarse_pairs(ctx, seeds, shape=(256, 256, 256), f=4, sigma=6.0, iters=(1,),
                         net=None, rounds=32, presweeps=32, judge=32, pres=(16, 32, 64),
                         max_cycles=10, max_sweeps=10000, on_policy=False, log=None):
    """(restricted residual, coarse solution) pairs from 256^3 Split-Bregman runs of the C1
    generator (FP64, red-black): at each outer iteration t in iters, the Poisson problem the
    solver built (rhs_t) is solved from the previous iterate phi_{t-1} by two-grid cycles, and
    every cycle's restricted residual and the coarse problem's converged solution form a pair
    (on_policy: the cycles apply net's correction, see coarse_cycles).
    Iteration 1's problem needs no converged solve before it (phi_0 is the initial contour);
    later ones need the gated loop, driven by `net` (a network that already serves as the
    coarse correction) so that it runs in seconds.  pres: the sweeps per cycle, one value
    per pass over the problem.  Pairs stay on the device (FP32 volumes of the coarse dims)."""
    import concurrent.futures as cf
    tk = _tk()
    pairs = []
    last = max(iters)
    t0 = time.perf_counter()
    scale = max(1, shape[0] // 64)
    # the host generator runs one seed ahead of the device work (ctypes releases the GIL)
    pool = cf.ThreadPoolExecutor(1)
    gen = lambda sd: tk.gen_standard(shape, sd, scale, 1 + sd % 4)
    nxt = pool.submit(gen, seeds[0]) if seeds else None
    for q, seed in enumerate(seeds):
        ph = nxt.result()
        nxt = pool.submit(gen, seeds[q + 1]) if q + 1 < len(seeds) else None
        if last == 1 or net is None:
            # rhs_1 depends only on phi_0: one cheap iteration yields it
            r = _c4_run(tk, ctx, seed, shape, sigma, phantom=ph, maxiter=1, ls_steps=1,
                        ls_tol=0.0)
            prev = [(1, None)]
        else:
            # the sweep cap bounds a solve the network fails to shorten (the data only need
            # the right-hand sides the gated loop builds)
            r = _c4_run(tk, ctx, seed, shape, sigma, phantom=ph, maxiter=last, ls_res_tol=1e-6,
                        res_check_every=64, max_sweeps=max_sweeps, warm_start=rounds,
                        ws_presweeps=presweeps, ws_judge_sweeps=judge, ws_continue=1)
            tk._check(tk.lib.tk_rsfe_set_warm_start(r.h, net.h))
            prev = [(t, None) for t in range(1, last + 1)]
        phi0 = _vol_copy(tk, ctx, tk.lib.tk_rsfe_phi(r.h), shape)
        for t, _ in prev:
            phi_prev = _vol_copy(tk, ctx, tk.lib.tk_rsfe_phi(r.h), shape) if t > 1 else phi0
            st = r.iterate(1)
            if log and net is not None and st:
                log(f"  seed {seed} iteration {t}: sweeps {st[0]['sweeps']}, rounds kept "
                    f"{st[0]['guess_used']}, first round residual plain {st[0]['res_plain']:.3e} "
                    f"learned {st[0]['res_guess']:.3e}")
            if t not in iters:
                if t > 1:
                    phi_prev.free()
                continue
            rhs = _vol_copy(tk, ctx, tk.lib.tk_rsfe_rhs(r.h), shape)
            for pre in pres:
                phi = _vol_copy(tk, ctx, phi_prev.h, shape)
                coarse_cycles(tk, ctx, phi, rhs, f, pre, max_cycles, pairs=pairs,
                              net=net if on_policy else None)
                phi.free()
            rhs.free()
            if t > 1:
                phi_prev.free()
        phi0.free()
        r.close()
        if log and (q + 1) % 10 == 0:
            log(f"coarse pairs: {q + 1}/{len(seeds)} seeds, {len(pairs)} pairs, "
                f"{time.perf_counter() - t0:.0f} s")
    pool.shutdown()
    return pairs


def train_pairs(ctx, net, pairs, steps, lr=1e-3, seed=7, log=None, set_scale=True):
    """Adam on random device pairs (input volume, target volume) with a cosine decay"""
    if set_scale:
        ratios = []
        for a, b in pairs[:: max(1, len(pairs) // 256)]:
            ra = np.sqrt(np.mean(a.numpy().astype(np.float64) ** 2))
            rb = np.sqrt(np.mean(b.numpy().astype(np.float64) ** 2))
            if ra > 0:
                ratios.append(rb / ra)
        net.out_scale = float(np.mean(ratios))
    rng = np.random.default_rng(seed)
    hist = []
    t0 = time.perf_counter()
    for k in range(steps):
        a, b = pairs[int(rng.integers(len(pairs)))]
        hist.append(net.train_step(a, b, lr * 0.5 * (1 + np.cos(np.pi * k / steps))))
        if log and (k + 1) % max(1, steps // 10) == 0:
            log(f"train: step {k + 1}/{steps}, loss (last 100) {np.mean(hist[-100:]):.4e}, "
                f"{time.perf_counter() - t0:.1f} s")
    return hist


def coarse_quality(ctx, net, pairs):
    """relative L2 error of the network's coarse correction against the coarse solution"""
    tk = _tk()
    out = []
    for a, b in pairs:
        g = tk.Volume(ctx, a.shape, np.float32)
        net.predict(a, g)
        x, y = g.numpy().astype(np.float64), b.numpy().astype(np.float64)
        out.append(np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-300))
        g.free()
    return np.array(out)


def evaluate_256(ctx, net, seeds, shape=(256, 256, 256), sigma=6.0, n_iter=3, res_tol=1e-6,
                 rounds=32, presweeps=32, judge=32, coarse=256, log=None):
    """per seed: the C1-generator Split-Bregman run with every Poisson step gated on a
    relative residual of res_tol, once from the previous iterate (plain) and once with the
    learned coarse correction; times (device-synchronised, per outer iteration), sweeps, and
    per outer iteration the Dice of the masks phi >= 0.5 and max |phi_learned - phi_plain|
    (phi after cut_phi)"""
    tk = _tk()
    rows = []
    for seed in seeds:
        res = {}
        for mode in ("plain", "learned"):
            r = _c4_run(tk, ctx, seed, shape, sigma, maxiter=n_iter, ls_res_tol=res_tol,
                        res_check_every=64, max_sweeps=1000000,
                        warm_start=rounds if mode == "learned" else 0, ws_presweeps=presweeps,
                        ws_judge_sweeps=judge, ws_continue=1, ws_coarse_sweeps=coarse)
            if mode == "learned":
                tk._check(tk.lib.tk_rsfe_set_warm_start(r.h, net.h))
            ctx.sync()
            secs, sweeps, used, coarse_sw, phis = 0.0, [], [], [], []
            for _ in range(n_iter):
                t0 = time.perf_counter()
                st = r.iterate(1)
                ctx.sync()
                secs += time.perf_counter() - t0
                if not st:
                    break
                sweeps.append(st[0]["sweeps"])
                used.append(st[0]["guess_used"])
                coarse_sw.append(st[0]["coarse_sweeps"])
                phis.append(r.phi())
            r.close()
            res[mode] = dict(seconds=secs, sweeps=sweeps, used=used, coarse=coarse_sw, phis=phis)
        dices, errs, sizes = [], [], []
        for a, b in zip(res["plain"]["phis"], res["learned"]["phis"]):
            ma, mb = a >= 0.5, b >= 0.5
            tot = int(ma.sum()) + int(mb.sum())
            dices.append(1.0 if tot == 0 else 2 * int((ma & mb).sum()) / tot)
            errs.append(float(np.max(np.abs(a - b))))
            sizes.append(int(ma.sum()))
        row = dict(seed=seed, s_plain=res["plain"]["seconds"], s_learned=res["learned"]["seconds"],
                   sweeps_plain=res["plain"]["sweeps"], sweeps_learned=res["learned"]["sweeps"],
                   rounds_used=res["learned"]["used"], coarse_sweeps=res["learned"]["coarse"],
                   dice=dices, max_abs=errs, mask_voxels_plain=sizes)
        rows.append(row)
        if log:
            log(f"eval256 seed {seed}: {row['s_plain']:.2f} s -> {row['s_learned']:.2f} s, sweeps "
                f"{row['sweeps_plain']} -> {row['sweeps_learned']}, rounds {row['rounds_used']}, "
                f"Dice {dices}, max|dphi| {errs}, mask voxels {sizes}")
    return rows


WEIGHTS_256 = __import__("os").path.join(
    __import__("os").path.dirname(__import__("os").path.abspath(__file__)), "weights",
    "warmstart_256c4.npz")


def main256(argv=None):
    """train the 256^3 coarse correction on the GPU: python -m ... warmstart.py train256 OUT.npz
    [--seeds N] [--seeds2 N] [--steps K] [--eval N]"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--seeds", type=int, default=100, help="seeds of the iteration-1 pairs")
    ap.add_argument("--seeds2", type=int, default=100,
                    help="seeds of the pairs of iterations 1-4 (driven by the first network)")
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--steps2", type=int, default=4000)
    ap.add_argument("--eval", type=int, default=2)
    a = ap.parse_args(argv)
    tk = _tk()

    def log(m):
        print(m, flush=True)

    ctx = tk.Context(0)
    f = 4
    cdims = (64, 64, 64)
    t0 = time.time()
    seeds1 = [SEED_TRAIN_256 + q for q in range(a.seeds)]
    pairs = collect_coarse_pairs(ctx, seeds1, f=f, iters=(1,), log=log)
    t_p1 = time.time() - t0
    log(f"stage 1: {len(pairs)} pairs from {a.seeds} seeds in {t_p1:.0f} s")
    net = tk.UNet(ctx, cdims, seed=1)
    t0 = time.time()
    h1 = train_pairs(ctx, net, pairs, a.steps, log=log)
    t_t1 = time.time() - t0
    hold = collect_coarse_pairs(ctx, [SEED_EVAL_256 - 1], f=f, iters=(1,), pres=(32,))
    log(f"stage 1 network, held-out relative error per cycle: "
        f"{np.round(coarse_quality(ctx, net, hold), 4).tolist()}")
    for a_, b_ in hold:
        a_.free()
        b_.free()
    np.savez_compressed(a.out + ".stage1.npz", params=net.params(), out_scale=net.out_scale,
                        dims=np.array(net.dims), coarsen=f)
    seeds2 = [SEED_TRAIN_256 + a.seeds + q for q in range(a.seeds2)]
    t0 = time.time()
    pairs2 = collect_coarse_pairs(ctx, seeds2, f=f, iters=(1, 2, 3, 4), net=net, pres=(32,),
                                  on_policy=True, log=log)
    t_p2 = time.time() - t0
    log(f"stage 2: {len(pairs2)} pairs from {a.seeds2} seeds in {t_p2:.0f} s")
    allp = pairs + pairs2
    t0 = time.time()
    h2 = train_pairs(ctx, net, allp, a.steps2, lr=5e-4, seed=8, log=log)
    t_t2 = time.time() - t0
    hold = collect_coarse_pairs(ctx, [SEED_EVAL_256 - 1], f=f, iters=(1,), pres=(32,))
    q = coarse_quality(ctx, net, hold)
    log(f"held-out relative error of the coarse correction per cycle: {np.round(q, 4).tolist()}")
    rows = evaluate_256(ctx, net, [SEED_EVAL_256 - 2 - k for k in range(a.eval)], log=log) \
        if a.eval else []
    np.savez_compressed(a.out, params=net.params(), out_scale=net.out_scale,
                        dims=np.array(net.dims), coarsen=f, meta=json.dumps(dict(
                            run_dims=[256, 256, 256], coarsen=f, first_seed=SEED_TRAIN_256,
                            seeds=a.seeds + a.seeds2, pairs=len(allp), pairs_stage1=len(pairs),
                            steps=a.steps + a.steps2, pairs_seconds=t_p1 + t_p2,
                            train_seconds=t_t1 + t_t2,
                            final_loss=float(np.mean(h2[-100:])),
                            holdout_rel_err=[float(x) for x in q],
                            evaluation=[{k: v for k, v in r.items()} for r in rows])))
    for a_, b_ in allp + hold:
        a_.free()
        b_.free()
    net.close()


def load256(ctx, path=None):
    """the 256^3 coarse-correction network shipped with the package (a 64^3 network) and the
    coarse sweeps it was evaluated with"""
    path = path or WEIGHTS_256
    d = np.load(path)
    return load(ctx, path), int(d["coarse_sweeps"]) if "coarse_sweeps" in d.files else 0


def make_coarse_weights(dst=None, src=None, f=4, coarse_sweeps=256):
    """the 256^3 weights from the 64^3 network: the same parameters with out_scale times f^2
    (the coarse problem has spacing f, and its correction is f^2 times that of the same
    right-hand side at unit spacing); the network then reads the residual that
    coarse_sweeps red-black sweeps from zero leave on the coarse problem"""
    import json
    d = np.load(src or WEIGHTS)
    meta = json.loads(str(d["meta"]))
    rec = dict(source="warmstart_64.npz (the 64^3 network, trained on the solver's own 64^3 "
                      "pairs; its record under 'trained')", coarsen=f,
               run_dims=[int(x) * f for x in d["dims"]], ws_coarse_sweeps=coarse_sweeps,
               out_scale_factor=f * f, trained=meta)
    np.savez_compressed(dst or WEIGHTS_256, params=d["params"],
                        out_scale=float(d["out_scale"]) * f * f, dims=d["dims"], coarsen=f,
                        coarse_sweeps=coarse_sweeps, meta=json.dumps(rec))


WEIGHTS = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)),
                                    "weights", "warmstart_64.npz")


def load(ctx, path=WEIGHTS):
    """the trained network shipped with the package (64^3)"""
    tk = _tk()
    d = np.load(path)
    net = tk.UNet(ctx, tuple(int(x) for x in d["dims"]), seed=1)
    net.set_params(d["params"])
    net.out_scale = float(d["out_scale"])
    return net


def main(argv=None):
    """train on the GPU and evaluate: python 3dimagetoolkit_amd/warmstart.py OUT.npz
    [--seeds N] [--steps K] [--eval N]"""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--seeds", type=int, default=100)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--eval", type=int, default=50)
    a = ap.parse_args(argv)
    tk = _tk()

    def log(m):
        print(m, flush=True)

    ctx = tk.Context(0)
    t0 = time.time()
    rhs, sol, sw = collect_pairs(ctx, a.seeds, log=log)
    t_pairs = time.time() - t0
    log(f"{len(rhs)} pairs from {a.seeds} seeds in {t_pairs:.1f} s")
    net = tk.UNet(ctx, (64, 64, 64), seed=1)
    t0 = time.time()
    hist = train(ctx, net, rhs, sol, a.steps, log=log)
    t_train = time.time() - t0
    del rhs, sol
    hr, hs, _ = collect_pairs(ctx, 3, seed0=40000)
    q = guess_quality(ctx, net, hr, hs)
    kinds = ["rhs", "r16", "r128", "r512", "r1024", "r2048"]
    qual = {k: float(q[i::len(kinds)].mean()) for i, k in enumerate(kinds)}
    log(f"held-out relative error of the prediction per input kind: {qual}")
    rows = evaluate(ctx, net, a.eval, log=log) if a.eval else []
    summ = {}
    if rows:
        sp = sum(r["sweeps_plain"] for r in rows)
        sl = sum(r["sweeps_learned"] for r in rows)
        tp = sum(r["s_plain"] for r in rows)
        tl = sum(r["s_learned"] for r in rows)
        summ = dict(seeds=len(rows), sweeps_plain=sp, sweeps_learned=sl, sweep_ratio=sl / sp,
                    seconds_plain=tp, seconds_learned=tl, time_ratio=tl / tp,
                    dice_min=min(r["dice"] for r in rows))
        log(f"evaluation: {json.dumps(summ)}")
    np.savez_compressed(a.out, params=net.params(), out_scale=net.out_scale,
                        dims=np.array(net.dims), meta=json.dumps(dict(
                            seeds=a.seeds, first_seed=SEED_TRAIN, steps=a.steps,
                            pairs_seconds=t_pairs, train_seconds=t_train,
                            final_loss=float(np.mean(hist[-100:])), holdout=qual,
                            evaluation=summ)))
    net.close()


if __name__ == "__main__":
    import os
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    if len(sys.argv) > 1 and sys.argv[1] == "train256":
        main256(sys.argv[2:])
    else:
        main()
