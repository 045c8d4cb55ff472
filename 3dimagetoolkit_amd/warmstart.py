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
             seed0=SEED_EVAL, rounds=8, presweeps=512, log=None):
    """per seed: sweeps and seconds of the residual-gated loop with the plain start and with
    the learned start, the Dice of their masks, and how often the guess was kept"""
    tk = _tk()
    rows = []
    for q in range(n_seeds):
        res = {}
        for mode in ("plain", "learned"):
            r = _run(tk, ctx, seed0 + q, shape, np.float64, tk.TK_GS_REDBLACK, sigma=sigma,
                     maxiter=n_iter, tol=0.0, ls_res_tol=res_tol, res_check_every=64,
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
        a, b = res["plain"]["mask"], res["learned"]["mask"]
        tot = int(a.sum()) + int(b.sum())
        dice = 1.0 if tot == 0 else 2 * int((a & b).sum()) / tot
        rows.append(dict(seed=seed0 + q, sweeps_plain=res["plain"]["sweeps"],
                         sweeps_learned=res["learned"]["sweeps"],
                         s_plain=res["plain"]["seconds"], s_learned=res["learned"]["seconds"],
                         guess_used=res["learned"]["used"], dice=float(dice)))
        if log:
            log(f"eval seed {seed0 + q}: sweeps {rows[-1]['sweeps_plain']} -> "
                f"{rows[-1]['sweeps_learned']}, {rows[-1]['s_plain']:.3f} s -> "
                f"{rows[-1]['s_learned']:.3f} s, guess kept {rows[-1]['guess_used']}/{n_iter}, "
                f"Dice {dice:.6f}")
    return rows


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
    main()
