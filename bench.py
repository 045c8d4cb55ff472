"""Benchmark: RSFE Split-Bregman voxel-iterations/s at 256^3 on one MI355X.

Workload (a "step" = one outer Split-Bregman iteration of rsfe_splitbregman::apply):
256^3 volume from the seeded two-ellipsoid generator scaled x4 (background N(0.2,0.05),
ellipsoids 0.8, x bias 0.7 -> 1.3, noise 0.05), rescaled to [0,255] as the reference's
driver does, Gaussian sigma 6 (37 taps) with circular boundaries, phi0 a sphere of radius
80, reference default lambda/lambda1/lambda2/beta/nu/epsilon, Dirichlet Gauss-Seidel
with the reference default Poisson setting (ls_steps 5, ls_tol 0: S = 5 sweeps per
iteration).  FP32 arithmetic with FP64 reductions and red-black sweeps (default), or
--dtype f64 for the FP64 plain mode with lexicographic sweeps.

value = N * K / t_loop, inputs resident in HBM; set-up (filter spectrum, K*1, K*u0,
edge detector) is outside the timed region.  The roofline object describes the kernel
with the largest total time, from HIP events recorded around every launch on the
context's stream in a second pass over K steps.  cpu_baseline times the CPU FP64
oracle (OpenMP) on the same workload for a bounded number of iterations.
"""
import os

# the CPU baseline's OpenMP threads stay on neighbouring cores; set before any library that
# carries an OpenMP runtime is loaded
os.environ.setdefault("OMP_PROC_BIND", "close")

import argparse  # noqa: E402
import importlib  # noqa: E402
import json  # noqa: E402
import sys  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED = 20240602

# algorithmic bytes per voxel per launch, in element sizes w
BYTES_PER_VOXEL_W = {
    "conv_x1": 5,   # read phi, u0; write complex (2); write phiold
    "conv_y": 4,    # read complex (2), write complex (2)
    "conv_z2": 6,   # read complex (2), kmyim, kones; write complex (2)
    "conv_x2": 6,   # read complex (2), u0, kones, divdb; write rhs
    "gs": 3,        # one red-black sweep: read phi, rhs; write phi
    "sb": 13,       # read phi, phiold, b(3), g; write phi, b(3), q(3)
    "div": 4,       # read q(3), write divdb
    "sbdiv": 11,    # update and div in one pass: read phi, phiold, b(3), g; write phi, b(3),
                    # divdb (q never reaches memory)
    "gs2": 6,       # two red-black sweeps in one launch: 2 x (read phi, rhs; write phi)
}

# HBM bytes per launch from rocprofv3 FETCH_SIZE and WRITE_SIZE passes over this bench's
# default workload, each counter scaled by what it reads for k_copy16 over 512 MiB (FETCH x2.0,
# WRITE x1.0; the same scale holds for a read-only reduction and a fill of 512 MiB); keyed by
# (size, dtype, sweep order, kernel tag); the sweep's tag carries its tile rows.
TRAFFIC_SOURCE = "profiles/rsfe256_f32_traffic.md"
_TRAFFIC_256_F32_RB = {
    "conv_x1": 336325877,   # 1.002 x algorithmic
    "conv_y": 268573577,    # 1.001
    "conv_z2": 402889322,   # 1.001
    "conv_x2": 407915765,   # 1.013
    "gs8": 221799404,       # 1.102 (the autotuner's trial sweeps)
    "gs16": 218014840,      # 1.083 (the autotuner's trial sweeps)
    "gs32": 221424851,      # 1.100 (the loop's sweeps)
    "sbdiv": 898419882,     # 1.217
}
MEASURED_TRAFFIC = {(256, "f32", "rb", k): v for k, v in _TRAFFIC_256_F32_RB.items()}


def copy_probe(tk, steps):
    """512^3 FP32 (512 MiB) volume copy, larger than the last-level cache: the plain
    copy of known size (512 MiB read + 512 MiB written per launch)."""
    ctx = tk.Context(0)
    shape = (512, 512, 512)
    a = ctx.volume(shape, np.float32)
    b = ctx.volume(shape, np.float32)
    tk._check(tk.lib.tk_vol_fill(a.h, 1.0))
    for _ in range(2):
        tk._check(tk.lib.tk_vol_copy(b.h, a.h))
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        tk._check(tk.lib.tk_vol_copy(b.h, a.h))
    ctx.sync()
    t = (time.perf_counter() - t0) / steps
    nbytes = 2 * 512 ** 3 * 4
    print(json.dumps({"probe": "k_copy16 512 MiB", "bytes_per_launch": nbytes,
                      "us_per_launch": round(t * 1e6, 1), "GBps": round(nbytes / t / 1e9, 1)}),
          flush=True)
    import ctypes as C
    # a second shape for the counters' scale: a read-only reduction (sum of squares, 512 MiB
    # read) and a write-only fill (512 MiB written)
    out = C.c_double()
    for name, fn, nb in (("k_sumsq_partial 512 MiB read", lambda: tk.lib.tk_norm_l2(ctx.h, a.h, C.byref(out)), 512 ** 3 * 4),
                         ("k_fill 512 MiB write", lambda: tk.lib.tk_vol_fill(b.h, 2.0), 512 ** 3 * 4)):
        tk._check(fn())
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            tk._check(fn())
        ctx.sync()
        t = (time.perf_counter() - t0) / steps
        print(json.dumps({"probe": name, "bytes_per_launch": nb, "us_per_launch": round(t * 1e6, 1),
                          "GBps": round(nb / t / 1e9, 1)}), flush=True)


BATCH_SEED = 20240604


def bench_batch(tk, a):
    """64 volumes, each edge from {48,64,80,96,128,160}, ellipsoid generator with 1-4
    ellipsoids (seed BATCH_SEED + v), rescaled to [0,255], sigma 3, ls_steps 5, ls_tol 0,
    FP32 red-black; a step is one outer iteration of every volume.  value = sum N_v * K / t."""
    dims, nell = tk.gen_batch_dims(BATCH_SEED, 64)
    ctx = tk.Context(0)
    p = tk.default_params(sigma=3.0, maxiter=1 << 30, tol=0.0, ls_steps=5, ls_tol=0.0,
                          gs_order=tk.TK_GS_REDBLACK)
    b = tk.Batch(ctx, dims, p, np.float32)
    vox = 0
    t0 = time.perf_counter()
    for v in range(64):
        u0, phi0 = tk.gen_standard(tuple(int(x) for x in dims[v]), BATCH_SEED + v, 1, int(nell[v]))
        mn, mx = u0.min(), u0.max()
        b.set_volume(v, (u0 - mn) / ((mx - mn) / 255.0), phi0)
        vox += int(np.prod(dims[v]))
    t_setup = time.perf_counter() - t0
    graph = not a.no_graph
    steps = a.steps + (a.steps & 1 if graph else 0)
    b.iterate(max(2, a.warmup + (a.warmup & 1)), use_graph=graph)
    t0 = time.perf_counter()
    b.iterate(steps, use_graph=graph)
    t = time.perf_counter() - t0
    info = b.info()
    b.close()
    value = vox * steps / t
    ref_model = vox * (58 * 4 + 3 * 4 * 5) * steps / t / 1e9
    out = {"metric": "Split-Bregman voxel-iters/s, 64 uneven volumes (aggregate)",
           "value": value, "unit": "voxel-iters/s", "n_gpus": 1, "steps": steps,
           "warmup": a.warmup, "ms_per_step": 1e3 * t / steps, "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64-reductions",
           "data": "synthetic (seeded ellipsoid phantoms, rescaled to [0,255])",
           "config": {"workload": "rsfe_splitbregman batch: 64 volumes, edges from "
                                  "{48,64,80,96,128,160}, sigma 3, red-black ls_steps 5 ls_tol 0",
                      "voxels": vox, "lanes": info["lanes"], "graph": graph,
                      "graph_kernel_nodes": info["graph_nodes"], "t_setup_s": round(t_setup, 3),
                      "loop_ref_model_GBps": round(ref_model, 1),
                      "loop_ref_model_frac": round(ref_model / HBM_PEAK_GBS, 4)},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def bench_warmstart(tk, a):
    """64^3 ellipsoid phantoms (seeds 50200 + i, unseen in training), FP64 Split-Bregman with
    red-black sweeps and every Poisson step gated on a relative residual of 1e-6, 15 outer
    iterations, once from the previous iterate (the reference's start) and once with the
    shipped network's correction rounds; a step is one seed.  value = plain time / learned
    time over all steps (higher is better); the sweep counts are in config."""
    ws = importlib.import_module("3dimagetoolkit_amd.warmstart")
    ctx = tk.Context(0)
    if a.size == 256:
        bench_warmstart_256(tk, ws, ctx, a)
        return
    net = ws.load(ctx)
    if a.warmup:
        ws.evaluate(ctx, net, 1, seed0=50190)
    rows = ws.evaluate(ctx, net, a.steps, seed0=50200)
    net.close()
    sp = sum(r["sweeps_plain"] for r in rows)
    sl = sum(r["sweeps_learned"] for r in rows)
    tp = sum(r["s_plain"] for r in rows)
    tl = sum(r["s_learned"] for r in rows)
    dices = [r["dice"] for r in rows]
    out = {"metric": "time to converged Poisson solves, plain start / learned start",
           "value": tp / tl, "unit": "x", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * tl / a.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 solver, bf16 network with fp32 accumulation",
           "data": "synthetic (seeded ellipsoid phantoms, rescaled to [0,255])",
           "config": {"workload": "rsfe_splitbregman 64^3 f64, sigma 3, red-black sweeps to "
                                  "relative residual 1e-6, 15 outer iterations per seed, "
                                  "learned rounds after 512 sweeps each (at most 8)",
                      "sweeps_plain": sp, "sweeps_learned": sl, "sweep_ratio": sl / sp,
                      "seconds_plain": tp, "seconds_learned": tl,
                      "dice_median": float(np.median(dices)), "dice_min": float(min(dices))},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def bench_unet(tk, a):
    """The warm start's network on the matrix cores: tk_unet_predict of the shipped 64^3
    network (29.9 GFLOP per forward pass), and single 3x3x3 convolution layers on random BF16
    data (2 * 27 * ci * co FLOP per voxel), timed over `steps` calls after `warmup`."""
    ws = importlib.import_module("3dimagetoolkit_amd.warmstart")
    ctx = tk.Context(0)
    rng = np.random.default_rng(SEED)
    steps, warm = max(a.steps, 1), max(a.warmup, 1)
    peak = 2.5e15

    def timed(f):
        for _ in range(warm):
            f()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            f()
        ctx.sync()
        return (time.perf_counter() - t0) / steps

    def bf16(x):
        u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
        return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)

    def dev(arr=None, nbytes=None):
        nb = arr.nbytes if arr is not None else nbytes
        v = ctx.volume((int(nb), 1, 1), np.uint8)
        if arr is not None:
            v.upload(np.frombuffer(np.ascontiguousarray(arr).tobytes(), np.uint8).reshape(-1, 1, 1))
        return v, tk.lib.tk_vol_device_ptr(v.h)

    layers = []
    for shape, ci, co in (((64, 64, 64), 1, 16), ((64, 64, 64), 16, 16), ((64, 64, 64), 32, 16),
                          ((32, 32, 32), 64, 32), ((128, 128, 128), 32, 16),
                          ((256, 256, 256), 32, 16)):
        n = int(np.prod(shape))
        kp = (27 * ci + 15) // 16 * 16
        x, xp = dev(bf16(rng.standard_normal((n, ci), dtype=np.float32)))
        w, wp = dev(bf16(rng.standard_normal((co, kp)) / np.sqrt(27 * ci)))
        b, bp = dev(rng.standard_normal(co).astype(np.float32))
        y, yp = dev(nbytes=2 * n * co)
        t = timed(lambda: tk._check(tk.lib.tk_nn_conv3d_fwd(
            ctx.h, xp, ci, 0, ci, wp, kp, bp, yp, co, 0, co,
            *shape, 1, 0)))
        fl = 2.0 * 27 * ci * co * n
        layers.append({"shape": list(shape), "ci": ci, "co": co, "us": round(t * 1e6, 2),
                       "TFLOPs": round(fl / t / 1e12, 1), "frac_bf16_peak": round(fl / t / peak, 4)})
        for v in (x, w, b, y):
            v.free()
    net = ws.load(ctx)
    rhs = ctx.from_numpy(rng.standard_normal(net.dims))
    guess = ctx.from_numpy(np.zeros(net.dims))
    t = timed(lambda: net.predict(rhs, guess))
    fl = 29.9e9
    out = {"metric": "U-Net forward pass @64^3 (BF16 on the matrix cores)", "value": fl / t / 1e12,
           "unit": "TFLOP/s", "n_gpus": 1, "steps": steps, "warmup": warm,
           "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "bf16 (fp32 accumulation)",
           "data": "synthetic (random rhs; the shipped 64^3 weights)",
           "config": {"workload": "tk_unet_predict, 3-level U-Net 16/32/64 channels, 64^3; "
                                  "single 3x3x3 layers on random BF16 data",
                      "forward_frac_bf16_peak": round(fl / t / peak, 4), "layers": layers},
           "roofline": {"bound": "mfma", "achieved": round(fl / t / 1e12, 1), "peak": 2500.0,
                        "unit": "TFLOP/s", "frac": round(fl / t / peak, 4), "traffic": None},
           "cpu_baseline": None}
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_c0(tk, a):
    """Config 0, time to converged mask: 64^3 two-ellipsoid phantom (seed of the committed
    oracle fixture), rescaled to [0,255], sigma 3, FP64, every Poisson step gated on the
    reference's relative-update rule at ls_tol 1e-8, 30 outer iterations; a step is one
    whole run (set-up included).  Modes: lexicographic sweeps (the reference's order, the
    parity mode) and red-black sweeps.  The masks are compared with the CPU FP64 oracle's
    fixture."""
    d = np.load(os.path.join(ROOT, "tests", "golden", "c0_64.npz"))
    meta = json.loads(str(d["meta"]))
    ref_mask = d["phi"] >= 0.5
    shape = tuple(meta["shape"])
    u0, phi0 = tk.gen_standard(shape, meta["seed"], 1, 2)
    ctx = tk.Context(0)

    def run(order, n_iter):
        ctx.sync()
        t0 = time.perf_counter()
        vu = ctx.from_numpy(u0)
        tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
        vi = ctx.from_numpy(phi0)
        p = tk.default_params(sigma=meta["sigma"], maxiter=meta["maxiter"], tol=0.0,
                              ls_steps=meta["ls_steps"], ls_tol=meta["ls_tol"], gs_order=order)
        r = tk.Rsfe(ctx, vu, vi, p)
        st = r.iterate(n_iter)
        phi = r.phi()
        ctx.sync()
        t = time.perf_counter() - t0
        r.close()
        vu.free()
        vi.free()
        return t, phi, [s["sweeps"] for s in st]

    modes = {}
    for name, order in (("f64_lex", tk.TK_GS_LEX), ("f64_redblack", tk.TK_GS_REDBLACK)):
        if a.warmup:
            run(order, 1)
        ts, sw = [], None
        for _ in range(max(1, a.steps)):
            t, phi, sw = run(order, meta["maxiter"])
            ts.append(t)
        mask = phi >= 0.5
        inter = int(np.count_nonzero(mask & ref_mask))
        dice = 2.0 * inter / max(1, int(mask.sum()) + int(ref_mask.sum()))
        modes[name] = {"seconds": float(np.median(ts)), "runs": [round(x, 3) for x in ts],
                       "sweeps_total": int(sum(sw)), "sweeps_first": sw[:3],
                       "sweeps_equal_oracle": sw == [int(x) for x in d["sweeps"]],
                       "max_abs_vs_oracle": float(np.max(np.abs(phi - d["phi"]))),
                       "dice_vs_oracle": round(dice, 7),
                       "mask_voxels_differing": int(np.count_nonzero(mask != ref_mask))}
    ctx.close()
    head = modes["f64_lex"]
    out = {"metric": "time to converged mask, 64^3 FP64, Poisson gate ls_tol 1e-8 (config 0)",
           "value": head["seconds"], "unit": "s", "n_gpus": 1, "steps": max(1, a.steps),
           "warmup": a.warmup, "ms_per_step": 1e3 * head["seconds"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded two-ellipsoid phantom, rescaled to [0,255])",
           "config": {"workload": "rsfe_splitbregman 64^3 f64, sigma 3 cyclic, 30 outer "
                                  "iterations, relative-update gate 1e-8 (ls_steps 5); set-up "
                                  "and host transfers included",
                      "modes": modes,
                      "oracle_seconds_recorded": meta.get("seconds"),
                      "oracle_threads_recorded": meta.get("threads")},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def c1_parity(tk, ctx, u0, phi0, dt, order, sigma, golden):
    """a fresh run of the mode to the 50 outer iterations of the golden fixture; phi against
    the CPU FP64 oracle's: max-abs over the fixture's 65536 samples, mask Dice and voxels
    differing over every voxel"""
    vu = ctx.from_numpy(u0.astype(dt))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0.astype(dt))
    iters = int(json.loads(str(golden["meta"]))["maxiter"])
    p = tk.default_params(sigma=sigma, maxiter=iters, tol=0.0, ls_steps=5, ls_tol=0.0,
                          gs_order=order)
    run = tk.Rsfe(ctx, vu, vi, p)
    vi.free()
    vu.free()
    run.iterate(iters)
    flat = run.phi(np.float64).reshape(-1)
    run.close()
    err = float(np.max(np.abs(flat[golden["sample_idx"]] - golden["phi_samples"])))
    mask = flat >= 0.5
    ref = np.unpackbits(golden["mask_bits"])[:flat.size].astype(bool)
    inter = int(np.count_nonzero(mask & ref))
    dice = 2.0 * inter / max(1, int(mask.sum()) + int(ref.sum()))
    return {"iterations": iters, "max_abs_samples": err, "dice": round(dice, 7),
            "mask_voxels_differing": int(np.count_nonzero(mask != ref))}


def c1_mode(tk, ctx, u0, phi0, n, sigma, dtype_name, order, a, timing_pass):
    """one timed mode of the 256^3 workload: W untimed iterations, then K timed; with
    timing_pass a second pass of K iterations records HIP events around every launch"""
    N = n ** 3
    dt = np.float32 if dtype_name == "f32" else np.float64
    w = 4 if dtype_name == "f32" else 8
    vu = ctx.from_numpy(u0.astype(dt))
    tk._check(tk.lib.tk_change_range_of_intensity(ctx.h, vu.h, 255.0, 0.0))
    vi = ctx.from_numpy(phi0.astype(dt))
    p = tk.default_params(sigma=sigma, maxiter=1 << 30, tol=0.0, ls_steps=a.ls_steps,
                          ls_tol=a.ls_tol, gs_order=order)
    t0 = time.perf_counter()
    run = tk.Rsfe(ctx, vu, vi, p)
    ctx.sync()
    t_setup = time.perf_counter() - t0
    vi.free()
    vu.free()
    rb_rows = None
    if order == tk.TK_GS_REDBLACK:
        # tune the sweep once more and keep that choice for the loop, so the tile rows the
        # timed loop runs are known (the result does not depend on them)
        vt = ctx.volume((n, n, n), dt)
        rb_rows, _ = tk.poisson_rb_autotune(ctx, vt)
        if a.rb_tile:
            # fixed tile rows and one sweep per launch (profiling runs: the autotuner's timing
            # under counter collection picks other variants)
            tk.poisson_rb_set_tile(ctx, vt, a.rb_tile)
            tk.poisson_rb_set_pair(ctx, False)
            rb_rows = a.rb_tile
        vt.free()
    if a.warmup:
        run.iterate(a.warmup)
    ctx.sync()
    t0 = time.perf_counter()
    stats = run.iterate(a.steps)
    ctx.sync()
    t_loop = time.perf_counter() - t0
    sweeps = [s["sweeps"] for s in stats]
    kernels = {}
    if timing_pass:
        ctx.set_timing(True)
        ctx.kernel_times(reset=True)
        run.iterate(a.steps)
        ctx.sync()
        kt = ctx.kernel_times(reset=True)
        ctx.set_timing(False)
        for tag, (ms, cnt) in kt.items():
            if cnt and tag in BYTES_PER_VOXEL_W:
                avg = ms / cnt
                per_launch = BYTES_PER_VOXEL_W[tag] * w * N
                if tag == "gs" and order == tk.TK_GS_LEX:
                    # a lexicographic launch is one pipeline step over part of the tiles:
                    # count the sweeps' bytes over all steps
                    per_launch = 3 * w * N * a.steps * a.ls_steps / cnt
                gbs = per_launch / (avg * 1e-3) / 1e9
                kernels[tag] = {"total_ms": round(ms, 3), "launches": cnt,
                                "avg_us": round(avg * 1e3, 2), "GBps": round(gbs, 1),
                                "bytes_per_launch": int(per_launch)}
    run.close()
    # whole-loop models.  Fused passes of this implementation: conv 5 passes (25w) + S sweeps
    # (3w each) + the update with div (11w) = N (36w + 3wS).  Reference data-flow model of
    # the metric's definition: N (58w + 3wS) (every temporary of rsfe_splitbregman_imp.hxx
    # :1547-1995 read and written once).
    S = float(np.mean(sweeps))
    loop_gbs = N * (36 * w + 3 * w * S) * a.steps / t_loop / 1e9
    ref_model_gbs = N * (58 * w + 3 * w * S) * a.steps / t_loop / 1e9
    return {"value": N * a.steps / t_loop, "ms_per_step": 1e3 * t_loop / a.steps,
            "t_setup_s": round(t_setup, 4), "sweeps": sweeps,
            "loop_model_GBps": round(loop_gbs, 1),
            "loop_model_frac": round(loop_gbs / HBM_PEAK_GBS, 4),
            "loop_ref_model_GBps": round(ref_model_gbs, 1),
            "loop_ref_model_frac": round(ref_model_gbs / HBM_PEAK_GBS, 4),
            "rb_tile_rows": rb_rows, "kernels": kernels}


def cpu_baseline(n, sigma, u0, phi0, a):
    """the CPU FP64 oracle (OpenMP, the project's own FFT in place of FFTW) on the same
    workload: a.cpu_runs runs of a.cpu_iters outer iterations each, median reported"""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as orc
    import ctypes as C
    N = n ** 3
    po = orc.default_params(sigma=sigma, ls_steps=a.ls_steps, ls_tol=0.0, tol=0.0)
    u255 = orc.change_range(u0, 255.0, 0.0)
    rates, fft_share = [], []
    for _ in range(a.cpu_runs):
        st = orc.OrStats()
        phi = phi0.copy()
        orc.lib().or_rsfe_time_iters(orc.dp(u255), C.c_int(n), C.c_int(n), C.c_int(n),
                                     C.byref(po), C.c_int(a.cpu_iters), orc.dp(phi), C.byref(st))
        rates.append(N * st.iters / st.t_loop)
        fft_share.append(st.t_fft / st.t_loop)
    med = float(np.median(rates))
    return {"value": med, "unit": "voxel-iters/s", "cores": orc.lib().or_num_threads(),
            "kind": "port",
            "sample": f"{n}^3 FP64 CPU oracle (OpenMP, OMP_PROC_BIND="
                      f"{os.environ.get('OMP_PROC_BIND', '')}), median of {a.cpu_runs} runs of "
                      f"{a.cpu_iters} outer iterations at S={a.ls_steps}, set-up excluded; the "
                      f"oracle's own mixed-radix FFT stands in for FFTW (absent), convolution "
                      f"share {float(np.median(fft_share)):.2f} of the loop",
            "runs": [round(r, 1) for r in rates]}


def bench_warmstart_256(tk, ws, ctx, a):
    """C4 at 256^3: C1-generator phantoms (seeds 90100 + i, unseen in training), sigma 6, FP64
    Split-Bregman with red-black sweeps and every Poisson step gated on a relative residual of
    1e-6, 3 outer iterations, once from the previous iterate and once with the learned
    coarse-grid correction (the 64^3 network on the residual restricted 4x, after 256 coarse
    sweeps); a step is one seed.  value = plain time / learned time over all steps."""
    net, coarse = ws.load256(ctx)
    if a.warmup:
        ws.evaluate_256(ctx, net, [90099], n_iter=1, coarse=coarse)
    rows = ws.evaluate_256(ctx, net, [90100 + q for q in range(a.steps)], n_iter=3,
                           coarse=coarse, log=lambda m: print(m, file=sys.stderr, flush=True))
    net.close()
    tp = sum(r["s_plain"] for r in rows)
    tl = sum(r["s_learned"] for r in rows)
    sp = sum(sum(r["sweeps_plain"]) for r in rows)
    sl = sum(sum(r["sweeps_learned"]) for r in rows)
    dices = [d for r in rows for d in r["dice"]]
    out = {"metric": "time to converged mask @256^3, plain start / learned start",
           "value": tp / tl, "unit": "x", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": 1e3 * tl / a.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64 solver, bf16 network with fp32 accumulation",
           "data": "synthetic (seeded two-ellipsoid phantoms, rescaled to [0,255])",
           "config": {"workload": "rsfe_splitbregman 256^3 f64, sigma 6, red-black sweeps to "
                                  "relative residual 1e-6, 3 outer iterations per seed; learned "
                                  "coarse-grid correction (64^3 network, 256 coarse sweeps, 32 "
                                  "sweeps before and 32 judge sweeps per round, up to 32 rounds)",
                      "seconds_plain": round(tp, 3), "seconds_learned": round(tl, 3),
                      "time_ratio": tl / tp, "sweeps_plain": sp, "sweeps_learned": sl,
                      "dice_min_per_iteration": min(dices),
                      "max_abs_phi_diff": max(m for r in rows for m in r["max_abs"]),
                      "per_seed": [{"seed": r["seed"], "s_plain": round(r["s_plain"], 3),
                                    "s_learned": round(r["s_learned"], 3),
                                    "dice": [round(d, 7) for d in r["dice"]],
                                    "mask_voxels": r["mask_voxels_plain"]} for r in rows]},
           "roofline": None, "cpu_baseline": None}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--dtype", choices=["f32", "f64"], default="f32")
    ap.add_argument("--ls-steps", type=int, default=5)
    ap.add_argument("--ls-tol", type=float, default=0.0)
    ap.add_argument("--cpu-iters", type=int, default=2)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--rb-tile", type=int, default=0,
                    help="red-black sweeps: these tile rows (8, 16 or 32) and one sweep per "
                         "launch instead of the autotuner's choice")
    ap.add_argument("--no-timing-pass", action="store_true")
    ap.add_argument("--no-parity", action="store_true",
                    help="skip the 50-iteration runs compared with the FP64 oracle's fixture")
    ap.add_argument("--only-headline", action="store_true",
                    help="time the headline mode only (not the two lexicographic modes)")
    ap.add_argument("--gs", choices=["default", "rb", "lex"], default="default",
                    help="Gauss-Seidel ordering (default: red-black for f32, lexicographic for f64)")
    ap.add_argument("--workload", choices=["c1", "batch", "warmstart", "c0", "unet"], default="c1",
                    help="c1: one 256^3 volume (default); batch: 64 uneven volumes; "
                         "warmstart: learned vs plain Poisson start (64^3, residual 1e-6); "
                         "c0: time to converged mask, 64^3 FP64 gated at 1e-8; "
                         "unet: the warm start's network forward pass and layers")
    ap.add_argument("--no-graph", action="store_true", help="batch: eager launches")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time a 512 MiB volume copy (k_copy16): the known-size kernel "
                         "that scales rocprofv3 FETCH_SIZE / WRITE_SIZE counts")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (every box here has "
              "one GPU, and this path has no multi-GPU sharding built)", file=sys.stderr)
        sys.exit(2)

    tk = importlib.import_module("3dimagetoolkit_amd")
    if a.copy_probe:
        copy_probe(tk, a.steps)
        return
    if a.workload == "batch":
        bench_batch(tk, a)
        return
    if a.workload == "warmstart":
        bench_warmstart(tk, a)
        return
    if a.workload == "c0":
        bench_c0(tk, a)
        return
    if a.workload == "unet":
        bench_unet(tk, a)
        return
    n = a.size
    shape = (n, n, n)
    scale = max(1, n // 64)
    sigma = 6.0 * n / 256.0 if n != 64 else 3.0
    order = tk.TK_GS_REDBLACK if a.dtype == "f32" else tk.TK_GS_LEX
    if a.gs != "default":
        order = tk.TK_GS_REDBLACK if a.gs == "rb" else tk.TK_GS_LEX
    oname = "redblack" if order == tk.TK_GS_REDBLACK else "lex"

    u0, phi0 = tk.gen_standard(shape, SEED, scale, 2)
    ctx = tk.Context(0)
    golden = None
    gpath = os.path.join(ROOT, "tests", "golden", "c1_256.npz")
    if (not a.no_parity and n == 256 and a.ls_steps == 5 and a.ls_tol == 0.0
            and os.path.exists(gpath)):
        d = np.load(gpath)
        golden = {k: d[k] for k in d.files}

    head = c1_mode(tk, ctx, u0, phi0, n, sigma, a.dtype, order, a, not a.no_timing_pass)
    modes = {f"{a.dtype}_{oname}": head}
    if not a.only_headline:
        for dn, o in (("f32", tk.TK_GS_LEX), ("f64", tk.TK_GS_LEX)):
            key = f"{dn}_lex"
            if key not in modes:
                modes[key] = c1_mode(tk, ctx, u0, phi0, n, sigma, dn, o, a, False)
    if golden is not None:
        for key, m in modes.items():
            dn, on = key.split("_")
            m["parity_vs_fp64_oracle"] = c1_parity(
                tk, ctx, u0, phi0, np.float32 if dn == "f32" else np.float64,
                tk.TK_GS_REDBLACK if on == "redblack" else tk.TK_GS_LEX, sigma, golden)
    ctx.close()

    roof = None
    kernels = head["kernels"]
    w = 4 if a.dtype == "f32" else 8
    if kernels:
        dom = max(kernels, key=lambda t: kernels[t]["total_ms"])
        k = kernels[dom]
        tag = f"gs{head['rb_tile_rows']}" if dom == "gs" and head["rb_tile_rows"] else dom
        key = (n, a.dtype, "rb" if order == tk.TK_GS_REDBLACK else "lex", tag)
        traffic = MEASURED_TRAFFIC.get(key)
        roof = {"bound": "hbm", "kernel": dom, "achieved": k["GBps"], "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(k["GBps"] / HBM_PEAK_GBS, 4),
                "traffic": traffic, "bytes_per_launch": k["bytes_per_launch"]}
        if traffic is not None:
            roof["traffic_source"] = TRAFFIC_SOURCE

    cpu = None
    if not a.no_cpu and a.cpu_iters > 0 and a.cpu_runs > 0:
        cpu = cpu_baseline(n, sigma, u0, phi0, a)

    sweeps = head["sweeps"]
    other = {}
    for key, m in modes.items():
        if m is head:
            continue
        other[key] = {"value": m["value"], "ms_per_step": round(m["ms_per_step"], 4),
                      "loop_ref_model_frac": m["loop_ref_model_frac"]}
        if "parity_vs_fp64_oracle" in m:
            other[key]["parity_vs_fp64_oracle"] = m["parity_vs_fp64_oracle"]
    config = {"workload": f"rsfe_splitbregman {n}^3 {a.dtype}, sigma {sigma:g} cyclic, "
                          f"Gauss-Seidel {'red-black' if order else 'lexicographic'} "
                          f"ls_steps {a.ls_steps} ls_tol {a.ls_tol:g}",
              "voxels": n ** 3,
              "sweeps_per_iter": sweeps[:4] + (["..."] if len(sweeps) > 4 else []),
              "t_setup_s": head["t_setup_s"],
              "loop_model_GBps": head["loop_model_GBps"],
              "loop_model_frac": head["loop_model_frac"],
              "loop_ref_model_GBps": head["loop_ref_model_GBps"],
              "loop_ref_model_frac": head["loop_ref_model_frac"],
              "rb_tile_rows": head["rb_tile_rows"],
              "kernels": kernels}
    if "parity_vs_fp64_oracle" in head:
        config["parity_vs_fp64_oracle"] = head["parity_vs_fp64_oracle"]
    if other:
        config["other_modes"] = other
    out = {
        "metric": "Split-Bregman voxel-iters/s @256^3",
        "value": head["value"],
        "unit": "voxel-iters/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64-reductions" if a.dtype == "f32" else "f64",
        "data": "synthetic (seeded ellipsoid phantom, rescaled to [0,255])",
        "config": config,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
