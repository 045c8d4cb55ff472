"""GPU parity of each kernel of libtk.so against the CPU FP64 oracle (and, for the FP32
red-black sweep, against a numpy float32 restatement).  Integer/index work and the
FP64 lexicographic sweep are compared bit for bit; FP64 FFT work to 1e-13 relative;
FP32 work within the tolerances written in each test."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


# ---- volumes ------------------------------------------------------------------------
def test_volume_roundtrip_and_convert(tk, ctx, rng):
    a = rng.standard_normal((7, 9, 11))
    v = ctx.from_numpy(a)
    assert np.array_equal(v.numpy(), a)
    assert np.array_equal(v.numpy(np.float32), a.astype(np.float32))
    w = ctx.volume(a.shape, np.float32)
    w.upload(a)  # F64 host -> F32 device
    assert np.array_equal(w.numpy(), a.astype(np.float32))
    v.fill(2.5)
    assert np.all(v.numpy() == 2.5)


# ---- R4 / R5 ----------------------------------------------------------------------
CONV_CASES = [((16, 16, 16), 1.5, (1.0, 1.0, 1.0)), ((48, 64, 80), 3.0, (1.0, 1.0, 1.0)),
              ((32, 40, 24), 2.0, (0.5, 1.0, 2.0)), ((20, 30, 25), 1.0, (1.0, 1.0, 1.0)),
              ((8, 9, 10), 2.0, (1.0, 1.0, 1.0)), ((64, 64, 64), 3.0, (1.0, 1.0, 1.0)),
              # lengths that are not 2/3/5-smooth take the direct path; Nh > L wraps
              ((28, 21, 26), 1.5, (1.0, 1.0, 1.0)), ((14, 11, 13), 3.0, (1.0, 1.0, 1.0))]


@pytest.mark.parametrize("shape,sigma,h", CONV_CASES)
def test_filter_spectrum_matches_oracle(tk, orc, ctx, shape, sigma, h):
    say("spectrum", shape, sigma, h)
    f = tk.Filter(ctx, shape, sigma, h=h)
    tf, nh, spec, nc = orc.gaussian(shape, sigma, 3.0, False, h)
    assert tuple(nh) == f.taps()
    H = [f.axis_spectrum(a) for a in range(3)]
    if tuple(nc) == shape:
        prod = np.einsum("i,j,k->ijk", *H) / (h[0] * h[1] * h[2])
        assert np.max(np.abs(prod - spec)) < 1e-13 / (h[0] * h[1] * h[2])


@pytest.mark.parametrize("shape,sigma,h", CONV_CASES)
def test_conv_cyclic_f64(tk, orc, ctx, shape, sigma, h, rng):
    say("conv f64", shape, sigma, h)
    x = rng.random(shape)
    tf, nh, spec, nc = orc.gaussian(shape, sigma, 3.0, False, h)
    ref = orc.cyclic_conv(spec, nc, x, h)
    f = tk.Filter(ctx, shape, sigma, h=h)
    vx = ctx.from_numpy(x, h=h)
    vy = ctx.volume(shape, h=h)
    f.conv(vy, vx)
    got = vy.numpy()
    assert np.max(np.abs(got - ref)) < 2e-15 * np.max(np.abs(x)) * np.log2(x.size)
    # pair: two fields through one complex transform, outputs aliasing inputs
    y2 = rng.standard_normal(shape)
    va, vb = ctx.from_numpy(x, h=h), ctx.from_numpy(y2, h=h)
    f.conv_pair(va, vb, va, vb)
    ref2 = orc.cyclic_conv(spec, nc, y2, h)
    assert np.max(np.abs(va.numpy() - ref)) < 4e-15 * np.log2(x.size) * np.max(np.abs(y2))
    assert np.max(np.abs(vb.numpy() - ref2)) < 4e-15 * np.log2(x.size) * np.max(np.abs(y2))


# FFT pass, direct circular convolution on the vector ALUs, the same on the matrix cores
@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("shape,sigma", [((64, 64, 64), 3.0), ((48, 80, 96), 3.0),
                                         ((128, 64, 160), 4.0), ((32, 48, 64), 6.0)])
def test_conv_cyclic_f32(tk, orc, ctx, shape, sigma, method, rng):
    say("conv f32", shape, sigma, method)
    x = rng.random(shape)
    tf, nh, spec, nc = orc.gaussian(shape, sigma)
    ref = orc.cyclic_conv(spec, nc, x)
    f = tk.Filter(ctx, shape, sigma, dtype=np.float32)
    for a in range(3):
        f.set_method(a, method)
    assert f.methods() == [method] * 3
    vx = ctx.from_numpy(x.astype(np.float32))
    vy = ctx.volume(shape, np.float32)
    f.conv(vy, vx)
    # FP32 FFT error grows like eps * log2(N): bound 1e-6 of the input scale
    assert np.max(np.abs(vy.numpy().astype(np.float64) - ref)) < 1e-6


def test_conv_f32_matrix_cores_accuracy(tk, orc, ctx, rng):
    """the banded matrix products on the matrix cores against the vector-ALU window: the
    same FP32 sums of 2s+1 products in another order, so the error against the FP64
    oracle has the same size and no bias (a truncating accumulation would show as a mean
    error of the order of the RMS error)"""
    shape = (256, 48, 64)
    x = rng.random(shape)
    tf, nh, spec, nc = orc.gaussian(shape, 6.0)
    ref = orc.cyclic_conv(spec, nc, x)
    stats = {}
    for method in (1, 2):
        f = tk.Filter(ctx, shape, 6.0, dtype=np.float32)
        for a in range(3):
            f.set_method(a, method)
        vx = ctx.from_numpy(x.astype(np.float32))
        vy = ctx.volume(shape, np.float32)
        f.conv(vy, vx)
        e = vy.numpy().astype(np.float64) - ref
        stats[method] = (np.abs(e).max(), np.sqrt(np.mean(e ** 2)), e.mean())
        say(f"method {method}: max {stats[method][0]:.3e} rms {stats[method][1]:.3e} "
            f"mean {stats[method][2]:.3e}")
    (m1, r1, b1), (m2, r2, b2) = stats[1], stats[2]
    assert m2 < 1e-6 and r2 < 1.5 * r1 and abs(b2) < 0.2 * r2


# acyclic_fftw_convolution (the reference's default cv): zero outside the volume.  The
# oracle pads to dims + nh - 1 and runs the FFT; the device runs direct passes with zero
# halos (the generic taps pass in FP64 unless a register window is compiled for the width).
@pytest.mark.parametrize("shape,sigma,h", CONV_CASES)
def test_conv_acyclic_f64(tk, orc, ctx, shape, sigma, h, rng):
    say("conv acyclic f64", shape, sigma, h)
    x = rng.random(shape)
    tf, nh, _, _ = orc.gaussian(shape, sigma, 3.0, False, h)
    spec, na = orc.acyclic_spectrum(tf, nh, shape)
    ref = orc.acyclic_conv(spec, na, x, h)
    f = tk.Filter(ctx, shape, sigma, h=h)
    vx = ctx.from_numpy(x, h=h)
    vy = ctx.volume(shape, h=h)
    f.conv(vy, vx, acyclic=True)
    n_pad = int(np.prod(na))
    assert np.max(np.abs(vy.numpy() - ref)) < 4e-15 * np.log2(n_pad) * np.max(np.abs(x))
    # pair, outputs aliasing inputs
    y2 = rng.standard_normal(shape)
    va, vb = ctx.from_numpy(x, h=h), ctx.from_numpy(y2, h=h)
    f.conv_pair(va, vb, va, vb, acyclic=True)
    ref2 = orc.acyclic_conv(spec, na, y2, h)
    assert np.max(np.abs(va.numpy() - ref)) < 4e-15 * np.log2(n_pad) * np.max(np.abs(y2))
    assert np.max(np.abs(vb.numpy() - ref2)) < 4e-15 * np.log2(n_pad) * np.max(np.abs(y2))
    # the boundary matters: the cyclic result differs by far more than the tolerance
    vc = ctx.volume(shape, h=h)
    f.conv(vc, vx)
    assert np.max(np.abs(vc.numpy() - ref)) > 1e-6


@pytest.mark.parametrize("method", [0, 1, 2])
@pytest.mark.parametrize("shape,sigma", [((64, 64, 64), 3.0), ((48, 80, 96), 3.0),
                                         ((32, 48, 64), 6.0), ((28, 21, 26), 1.5)])
def test_conv_acyclic_f32(tk, orc, ctx, shape, sigma, method, rng):
    say("conv acyclic f32", shape, sigma, method)
    x = rng.random(shape)
    tf, nh, _, _ = orc.gaussian(shape, sigma)
    spec, na = orc.acyclic_spectrum(tf, nh, shape)
    ref = orc.acyclic_conv(spec, na, x)
    f = tk.Filter(ctx, shape, sigma, dtype=np.float32)
    for a in range(3):
        try:
            f.set_method(a, method)
        except tk.TkError:
            pass  # no FFT plan (lengths with a factor 7) or no window for this width/shape
    vx = ctx.from_numpy(x.astype(np.float32))
    vy = ctx.volume(shape, np.float32)
    f.conv(vy, vx, acyclic=True)
    assert np.max(np.abs(vy.numpy().astype(np.float64) - ref)) < 1e-6


def test_filter_autotune_keeps_results(tk, orc, ctx, rng):
    shape = (64, 48, 80)
    f = tk.Filter(ctx, shape, 3.0, dtype=np.float32)
    ms = f.autotune()
    say("autotune ms (fft, direct, direct on the matrix cores) per axis:", ms, "->", f.methods())
    assert all(t[0] > 0 and t[1] > 0 and t[2] > 0 for t in ms)
    x = rng.random(shape)
    tf, nh, spec, nc = orc.gaussian(shape, 3.0)
    vx = ctx.from_numpy(x.astype(np.float32))
    vy = ctx.volume(shape, np.float32)
    f.conv(vy, vx)
    assert np.max(np.abs(vy.numpy() - orc.cyclic_conv(spec, nc, x))) < 1e-6


# ---- R9 / R8: Gauss-Seidel -------------------------------------------------------
GS_SHAPES = [(20, 18, 40), (11, 37, 70), (64, 64, 64), (5, 6, 7)]


@pytest.mark.parametrize("shape", GS_SHAPES)
@pytest.mark.parametrize("nsw", [1, 3])
def test_gs_lex_bitwise(tk, orc, ctx, shape, nsw, rng):
    say("gs lex", shape, nsw)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    ref = phi
    for _ in range(nsw):
        ref = orc.gs_sweep(ref, rhs, 1.0, 0.0)
    vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
    tk.poisson_sweeps(ctx, vp, vr, nsw, 1.0, 0.0, tk.TK_GS_LEX)
    assert np.array_equal(vp.numpy(), ref)


def test_gs_lex_bitwise_general_spacing(tk, orc, ctx, rng):
    shape = (19, 23, 45)
    h = (0.5, 1.0, 2.0)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    ref = orc.gs_sweep(orc.gs_sweep(phi, rhs, 0.7, 0.3, h), rhs, 0.7, 0.3, h)
    vp, vr = ctx.from_numpy(phi, h=h), ctx.from_numpy(rhs, h=h)
    tk.poisson_sweeps(ctx, vp, vr, 2, 0.7, 0.3, tk.TK_GS_LEX)
    assert np.array_equal(vp.numpy(), ref)


@pytest.mark.parametrize("ls_steps,ls_tol", [(5, 1e-6), (3, 1e-9), (40, 1e-3)])
def test_gs_solve_lex_matches_oracle_loop(tk, orc, ctx, ls_steps, ls_tol, rng):
    say("gs solve lex", ls_steps, ls_tol)
    shape = (24, 22, 26)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    ref, s_ref, e_ref = orc.poisson_loop(phi, rhs, ls_steps, ls_tol)
    vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
    s, e = tk.poisson_solve(ctx, vp, vr, ls_steps, ls_tol, order=tk.TK_GS_LEX)
    assert s == s_ref
    assert np.array_equal(vp.numpy(), ref)
    assert abs(e - e_ref) <= 1e-10 * e_ref


# NeuGaussSeidel (poisson_imp.hxx:115-334): every voxel updated, the derivative bc on the
# boundary; FP64 bit for bit against the oracle's sweep in the reference's loop order.
NEU_SHAPES = [(2, 2, 2), (3, 4, 2), (5, 6, 7), (9, 17, 33), (16, 8, 64), (20, 30, 70)]


@pytest.mark.parametrize("shape", NEU_SHAPES)
@pytest.mark.parametrize("dt,bc,h", [(1.0, 0.0, (1.0, 1.0, 1.0)), (0.7, 0.3, (0.5, 1.0, 2.0)),
                                     (1.0, -1.2, (1.0, 1.0, 1.0))])
def test_gs_neumann_bitwise(tk, orc, ctx, shape, dt, bc, h, rng):
    say("gs neumann", shape, dt, bc, h)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    ref = phi
    for _ in range(3):
        ref = orc.neu_sweep(ref, rhs, dt, bc, h)
    vp, vr = ctx.from_numpy(phi, h=h), ctx.from_numpy(rhs, h=h)
    tk.poisson_sweeps(ctx, vp, vr, 3, dt, bc, tk.TK_GS_LEX, neumann=True)
    assert np.array_equal(vp.numpy(), ref)


def test_gs_neumann_f32_and_refusals(tk, orc, ctx, rng):
    """FP32 storage against the FP64 oracle (error stated against the field's scale), and
    the orders and shapes the Neumann sweep refuses"""
    shape = (33, 20, 40)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    ref = phi
    for _ in range(4):
        ref = orc.neu_sweep(ref, rhs, 1.0, 0.25)
    vp, vr = ctx.from_numpy(phi.astype(np.float32)), ctx.from_numpy(rhs.astype(np.float32))
    tk.poisson_sweeps(ctx, vp, vr, 4, 1.0, 0.25, tk.TK_GS_LEX, neumann=True)
    err = np.max(np.abs(vp.numpy().astype(np.float64) - ref)) / np.max(np.abs(ref))
    say(f"neumann FP32 4 sweeps: max rel err {err:.2e}")
    assert err < 2e-6
    with pytest.raises(tk.TkError):
        tk.poisson_sweeps(ctx, vp, vr, 1, 1.0, 0.0, tk.TK_GS_REDBLACK, neumann=True)
    flat = ctx.from_numpy(rng.random((8, 8, 1)))
    with pytest.raises(tk.TkError):
        tk.poisson_sweeps(ctx, flat, flat, 1, 1.0, 0.0, tk.TK_GS_LEX, neumann=True)


@pytest.mark.parametrize("ls_steps,ls_tol", [(5, 1e-6), (40, 1e-3)])
def test_gs_neumann_solve_matches_oracle_loop(tk, orc, ctx, ls_steps, ls_tol, rng):
    """the pure Neumann problem is singular: a right-hand side with a non-zero mean has no
    steady state (phi drifts and the relative update falls like 1/sweeps, 1e6 sweeps to
    reach 1e-6), so the gated cases use a zero-mean one, which converges (~2000 sweeps)"""
    say("gs neumann solve", ls_steps, ls_tol)
    shape = (24, 22, 26)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    rhs -= rhs.mean()
    ref, s_ref, e_ref = orc.poisson_loop(phi, rhs, ls_steps, ls_tol, neumann=True)
    vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
    s, e = tk.poisson_solve(ctx, vp, vr, ls_steps, ls_tol, order=tk.TK_GS_LEX, neumann=True)
    assert s == s_ref
    assert np.array_equal(vp.numpy(), ref)
    assert abs(e - e_ref) <= 1e-10 * e_ref


def rb_sweep(phi, rhs, dt=np.float32):
    """red (i+j+k even) then black, in dt throughout, operand order of the kernel."""
    X, Y, Z = phi.shape
    p = phi.astype(dt).copy()
    b = rhs.astype(dt)
    ht = dt(1.0 / 7.0)
    i, j, k = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    par = (i + j + k) % 2
    p[0] = 0; p[-1] = 0; p[:, 0] = 0; p[:, -1] = 0; p[:, :, 0] = 0; p[:, :, -1] = 0
    for c in (0, 1):
        s = slice(1, -1)
        v = (((p[2:, s, s] + p[:-2, s, s]) + (p[s, 2:, s] + p[s, :-2, s])) +
             (p[s, s, 2:] + p[s, s, :-2]) + b[s, s, s] + p[s, s, s]) * ht
        m = par[s, s, s] == c
        inner = p[s, s, s]
        inner[m] = v[m]
        p[s, s, s] = inner
    return p


def rb_sweep_f32(phi, rhs):
    return rb_sweep(phi, rhs, np.float32)


@pytest.mark.parametrize("rows", [8, 16, 32])
@pytest.mark.parametrize("shape", [(16, 18, 20), (33, 17, 9), (40, 37, 136), (20, 75, 72)])
def test_gs_redblack_f32_bitwise_vs_numpy(tk, ctx, shape, rows, rng):
    say("gs rb", shape, "tile rows", rows)
    phi = rng.random(shape).astype(np.float32)
    rhs = rng.standard_normal(shape).astype(np.float32)
    ref = phi
    for _ in range(3):
        ref = rb_sweep_f32(ref, rhs)
    vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
    tk.poisson_rb_set_tile(ctx, vr, rows)
    tk.poisson_sweeps(ctx, vp, vr, 3, 1.0, 0.0, tk.TK_GS_REDBLACK)
    assert np.array_equal(vp.numpy(), ref)


@pytest.mark.parametrize("rows", [8, 16, 32])
@pytest.mark.parametrize("shape,sweeps,h", [((16, 18, 20), 2, 1.0), ((40, 37, 136), 5, 1.0),
                                            ((9, 70, 64), 4, 1.0), ((70, 45, 200), 3, 0.7),
                                            ((128, 96, 256), 5, 1.0)])
def test_gs_redblack_pair_launch_bitwise(tk, ctx, shape, sweeps, h, rows, rng):
    """two sweeps per launch (the default for fixed counts) against one launch per sweep:
    bitwise equal, at tile edges in j and k, chunk edges in i, odd counts (a pair and a
    single) and general spacing"""
    say("gs rb pair", shape, sweeps, h, "tile rows", rows)
    phi = rng.random(shape).astype(np.float32)
    rhs = rng.standard_normal(shape).astype(np.float32)
    out = []
    for pair in (False, True):
        vp, vr = ctx.from_numpy(phi, h=(h, 1.0, h)), ctx.from_numpy(rhs, h=(h, 1.0, h))
        tk.poisson_rb_set_tile(ctx, vr, rows)
        tk.poisson_rb_set_pair(ctx, pair)
        tk.poisson_sweeps(ctx, vp, vr, sweeps, 1.0, 0.0, tk.TK_GS_REDBLACK)
        out.append(vp.numpy())
    tk.poisson_rb_set_pair(ctx, True)
    assert np.array_equal(out[0], out[1])
    if h == 1.0 and np.prod(shape) < 1e6:
        ref = phi
        for _ in range(sweeps):
            ref = rb_sweep_f32(ref, rhs)
        assert np.array_equal(out[1], ref)


@pytest.mark.parametrize("shape,sweeps", [((16, 18, 20), 2), ((40, 37, 136), 5), ((9, 70, 64), 4),
                                          ((64, 64, 64), 3)])
def test_gs_redblack_pair_launch_bitwise_f64(tk, ctx, shape, sweeps, rng):
    """FP64 two-sweep launches (6-row tiles) against one launch per sweep and numpy"""
    say("gs rb pair f64", shape, sweeps)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    out = []
    for pair in (False, True):
        vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
        tk.poisson_rb_set_tile(ctx, vr, 8)
        tk.poisson_rb_set_pair(ctx, pair)
        tk.poisson_sweeps(ctx, vp, vr, sweeps, 1.0, 0.0, tk.TK_GS_REDBLACK)
        out.append(vp.numpy())
    tk.poisson_rb_set_pair(ctx, True)
    assert np.array_equal(out[0], out[1])
    ref = phi
    for _ in range(sweeps):
        ref = rb_sweep(ref, rhs, np.float64)
    assert np.array_equal(out[1], ref)


def rb_gate_replay(phi, rhs, dt, ls_steps, tol, cap=100000):
    """sweeps until the relative update (whole volume, faces included) is <= tol after at
    least ls_steps sweeps; returns the field, the sweep count and the last error"""
    q = phi
    s = 0
    while True:
        prev = q.astype(np.float64)
        q = rb_sweep(q, rhs, dt)
        s += 1
        d = q.astype(np.float64) - prev
        err = np.sqrt((d * d).sum()) / np.sqrt((prev * prev).sum())
        if (s >= ls_steps and not err > tol) or s >= cap:
            return q, s, err


@pytest.mark.parametrize("dtype,shape,tols", [
    (np.float32, (20, 22, 24), (3e-3, 1e-3, 3e-4, 1e-4)),
    (np.float64, (18, 20, 22), (1e-3, 3e-4, 1e-4, 3e-5)),
    (np.float64, (24, 26, 32), (2e-6,)),  # more than one batch of 256 sweeps
])
def test_gs_redblack_gate_pairs_match_replay(tk, ctx, dtype, shape, tols, rng):
    """gated solves with two sweeps per launch and with one: the same sweep count as the
    numpy replay, the field bitwise, the error to rounding; the tolerances put the stop on
    the first and on the second sweep of a launch"""
    phi = rng.random(shape).astype(dtype)
    rhs = rng.standard_normal(shape).astype(dtype)
    parities = set()
    for tol in tols:
        q, s, err = rb_gate_replay(phi, rhs, dtype, 5, tol)
        parities.add(s % 2)
        for pair in (True, False):
            say("gs rb gate", np.dtype(dtype).name, shape, "tol", tol, "pair", pair, "sweeps", s)
            vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
            tk.poisson_rb_set_tile(ctx, vr, 8)
            tk.poisson_rb_set_pair(ctx, pair)
            sg, eg = tk.poisson_solve(ctx, vp, vr, 5, tol, order=tk.TK_GS_REDBLACK)
            assert sg == s
            assert np.array_equal(vp.numpy(), q)
            assert abs(eg - err) < 1e-6 * err
    tk.poisson_rb_set_pair(ctx, True)
    if len(tols) > 1:
        assert parities == {0, 1}


def test_gs_redblack_autotune(tk, ctx, rng):
    shape = (64, 66, 128)
    rhs = ctx.from_numpy(rng.standard_normal(shape).astype(np.float32))
    rows, ms = tk.poisson_rb_autotune(ctx, rhs)
    say("rb autotune ms (8, 16, 32 rows; two-sweep launch 8, 16):", ms, "->", rows)
    assert rows in (8, 16, 32) and min(ms) > 0
    assert rows == (8, 16, 32)[int(np.argmin(ms[:3]))]
    r64 = ctx.from_numpy(rng.standard_normal(shape))
    rows64, ms64 = tk.poisson_rb_autotune(ctx, r64)
    assert rows64 in (8, 16) and ms64[2] == 0 and ms64[4] == 0 and min(ms64[:2]) > 0
    with pytest.raises(tk.TkError, match="32 in FP32"):
        tk.poisson_rb_set_tile(ctx, r64, 32)
    phi = rng.random(shape).astype(np.float32)
    vp = ctx.from_numpy(phi)
    tk.poisson_sweeps(ctx, vp, rhs, 2, 1.0, 0.0, tk.TK_GS_REDBLACK)
    assert np.array_equal(vp.numpy(), rb_sweep_f32(rb_sweep_f32(phi, rhs.numpy()), rhs.numpy()))


@pytest.mark.parametrize("rows", [8, 16, 32])
def test_gs_redblack_gate_matches_replay(tk, ctx, rows, rng):
    shape = (20, 22, 24)
    phi = rng.random(shape).astype(np.float32)
    rhs = rng.standard_normal(shape).astype(np.float32)
    tol = 1e-4
    q = phi
    s = 0
    while True:
        prev = q.astype(np.float64)  # the whole volume before the sweep, faces included
        q = rb_sweep_f32(q, rhs)
        s += 1
        d = q.astype(np.float64) - prev
        err = np.sqrt((d * d).sum()) / np.sqrt((prev * prev).sum())
        if s >= 5 and not err > tol:
            break
    vp, vr = ctx.from_numpy(phi), ctx.from_numpy(rhs)
    tk.poisson_rb_set_tile(ctx, vr, rows)
    sg, eg = tk.poisson_solve(ctx, vp, vr, 5, tol, order=tk.TK_GS_REDBLACK)
    assert sg == s
    assert np.array_equal(vp.numpy(), q)
    assert abs(eg - err) < 1e-6 * err


def test_poisson_residual(tk, ctx, rng):
    shape = (12, 13, 14)
    phi = rng.random(shape)
    rhs = rng.standard_normal(shape)
    s = slice(1, -1)
    lap = ((phi[2:, s, s] + phi[:-2, s, s] - 2 * phi[s, s, s]) +
           (phi[s, 2:, s] + phi[s, :-2, s] - 2 * phi[s, s, s]) +
           (phi[s, s, 2:] + phi[s, s, :-2] - 2 * phi[s, s, s]))
    ref = np.linalg.norm(rhs[s, s, s] + lap) / np.linalg.norm(rhs[s, s, s])
    got = tk.poisson_residual(ctx, ctx.from_numpy(phi), ctx.from_numpy(rhs))
    assert abs(got - ref) < 1e-12 * ref


# ---- R6, R10-R13 --------------------------------------------------------------------
@pytest.mark.parametrize("shape", [(13, 17, 19), (12, 13, 16)])  # per-voxel and 16-byte-row paths
def test_pointwise_and_stencils_f64(tk, orc, ctx, shape, rng):
    lib = tk.lib
    h = (0.5, 1.0, 2.0)
    phi = rng.random(shape) * 1.2 - 0.1
    v = ctx.from_numpy(phi, h=h)
    out = ctx.volume(shape, h=h)
    tk._check(lib.tk_heaviside(ctx.h, v.h, 0.0, 1.0, 1e-3, out.h))
    assert np.allclose(out.numpy(), orc.heaviside(phi), rtol=4e-16, atol=4e-16)
    g = [ctx.volume(shape, h=h) for _ in range(3)]
    tk._check(lib.tk_grad(ctx.h, v.h, g[0].h, g[1].h, g[2].h))
    rg = orc.grad(phi, h)
    for a in range(3):
        assert np.array_equal(g[a].numpy(), rg[a])
    tk._check(lib.tk_div(ctx.h, g[0].h, g[1].h, g[2].h, out.h))
    assert np.array_equal(out.numpy(), orc.div(*rg, h))
    # fields that expose the two div quirks
    i, j, k = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    f = [ctx.from_numpy(i * 1.0 + 0.1 * k, h=h), ctx.from_numpy(j * 2.0, h=h),
         ctx.from_numpy(k * 3.0 + i * 0.5, h=h)]
    tk._check(lib.tk_div(ctx.h, f[0].h, f[1].h, f[2].h, out.h))
    assert np.array_equal(out.numpy(), orc.div(i * 1.0 + 0.1 * k, j * 2.0, k * 3.0 + i * 0.5, h))
    tk._check(lib.tk_edge_detect(ctx.h, v.h, 100.0, 5e-2, 1.0, out.h))
    assert np.array_equal(out.numpy(), orc.edge_detect(phi, 100.0, 5e-2, 1.0, h))
    b = rng.standard_normal(shape)
    gg = rng.random(shape)
    vb, vg = ctx.from_numpy(b, h=h), ctx.from_numpy(gg, h=h)
    tk._check(lib.tk_shrink(ctx.h, vb.h, vg.h, out.h))
    assert np.array_equal(out.numpy(), orc.shrink(b, gg))


def test_div_f32_rows(tk, orc, ctx, rng):
    """FP32 divergence on the 16-byte-row path against the FP64 oracle, including the
    two quirk locations."""
    shape = (10, 11, 24)
    h = (0.5, 1.0, 2.0)
    f = [rng.standard_normal(shape) for _ in range(3)]
    vf = [ctx.from_numpy(a.astype(np.float32), h=h) for a in f]
    out = ctx.volume(shape, np.float32, h=h)
    tk._check(tk.lib.tk_div(ctx.h, vf[0].h, vf[1].h, vf[2].h, out.h))
    ref = orc.div(*[a.astype(np.float32).astype(np.float64) for a in f], h)
    assert np.max(np.abs(out.numpy() - ref)) < 2e-5 * np.max(np.abs(ref))


def test_norms_minmax_range_cut(tk, orc, ctx, rng):
    import ctypes as C
    lib = tk.lib
    a = rng.standard_normal((21, 22, 23))
    b = rng.standard_normal(a.shape)
    va, vb = ctx.from_numpy(a), ctx.from_numpy(b)
    r = C.c_double()
    tk._check(lib.tk_norm_l2(ctx.h, va.h, C.byref(r)))
    assert abs(r.value - orc.norm_l2(a)) < 1e-14 * r.value
    tk._check(lib.tk_norm_l2_diff(ctx.h, va.h, vb.h, C.byref(r)))
    assert abs(r.value - orc.norm_l2(a - b)) < 1e-14 * r.value
    mn, mx = C.c_double(), C.c_double()
    tk._check(lib.tk_minmax(ctx.h, va.h, C.byref(mn), C.byref(mx)))
    assert mn.value == a.min() and mx.value == a.max()
    tk._check(lib.tk_change_range_of_intensity(ctx.h, va.h, 255.0, 0.0))
    assert np.array_equal(va.numpy(), orc.change_range(a, 255.0, 0.0))
    tk._check(lib.tk_cut(ctx.h, vb.h, 0.0, 1.0))
    assert np.array_equal(vb.numpy(), np.clip(b, 0.0, 1.0))


# ---- R14 ---------------------------------------------------------------------------
def enumerate_img(X, Y, Z):
    i, j, k = np.meshgrid(np.arange(X), np.arange(Y), np.arange(Z), indexing="ij")
    return (i + j * X + k * X * Y).astype(np.float64)


@pytest.mark.parametrize("full", [False, True])
def test_cc_seeded_bit_exact(tk, orc, ctx, full, rng):
    cases = [(enumerate_img(4, 4, 4), (3, 3, 3)),
             ((enumerate_img(3, 3, 3) % 2 == 0).astype(float), (3, 3, 3))]
    for q in range(3):
        img = rng.random((17, 19, 23))
        img = (img > 0.42 + 0.04 * q).astype(np.float64) + rng.random(img.shape) * 0.05
        cases.append((img, (8, 9, 10)))
    tube = np.zeros((40, 30, 50))
    tube[5:35, 14:17, 20:23] = 1.0
    tube[33:36, 14:17, 5:45] = 1.0
    cases.append((tube, (6, 15, 21)))
    for img, seed in cases:
        say("cc", img.shape, seed, full)
        ref, _ = orc.connected_component(img, seed, 0.5, full)
        vi = ctx.from_numpy(img)
        out = ctx.volume(img.shape, np.uint8)
        tk._check(tk.lib.tk_cc_seeded(ctx.h, vi.h, *seed, 0.5, int(full), out.h, None))
        assert np.array_equal(out.numpy(), ref.astype(np.uint8))


def test_cc_label_matches_scipy(tk, ctx, rng):
    from scipy import ndimage
    import ctypes as C
    m = (rng.random((30, 31, 32)) > 0.6).astype(np.uint8)
    vm = ctx.from_numpy(m)
    lab_dev = ctx.volume(m.shape, np.float32)  # 4-byte storage for int32 labels
    tk._check(tk.lib.tk_cc_label(ctx.h, vm.h, tk.lib.tk_vol_device_ptr(lab_dev.h), None))
    raw = lab_dev.numpy().view(np.int32)
    ref, n = ndimage.label(m)
    flat = np.arange(m.size).reshape(m.shape)
    expect = np.zeros(m.shape, np.int32)
    for c in range(1, n + 1):
        sel = ref == c
        expect[sel] = flat[sel].min() + 1
    assert np.array_equal(raw, expect)


@pytest.mark.parametrize("shape,h", [((3, 3, 8), (1.0, 1.0, 1.0)), ((5, 7, 8), (0.7, 1.3, 1.0)),
                                     ((16, 18, 20), (1.0, 1.0, 1.0)),
                                     ((33, 17, 12), (1.0, 2.0, 0.5)),
                                     ((40, 37, 136), (1.0, 1.0, 1.0)), ((9, 10, 11), (1.0, 1.0, 1.0))])
def test_sb_update_vs_oracle(tk, orc, ctx, shape, h, rng):
    """the Split-Bregman update and div of the run loop (tk_sb_update) in FP64 against the
    oracle's cut_phi, grad, shrink and div composed in the reference's order
    (rsfe_splitbregman_imp.hxx:1868-1932), bit for bit; rows of 16 bytes and scalar voxels"""
    a0, b0 = 0.0, 1.0
    phi = rng.uniform(-0.3, 1.3, shape)
    phiold = rng.random(shape)
    g = rng.uniform(0.0, 0.5, shape)
    b = [0.3 * rng.standard_normal(shape) for _ in range(3)]
    vp = ctx.from_numpy(phi, h=h)
    vo = ctx.from_numpy(phiold, h=h)
    vg = ctx.from_numpy(g, h=h)
    vb = [ctx.from_numpy(x, h=h) for x in b]
    vq = [ctx.volume(shape, np.float64, h=h) for _ in range(3)]
    vd = ctx.volume(shape, np.float64, h=h)
    n1 = tk.sb_update(ctx, vp, vo, vg, vb, vq, vd, a0, b0)
    pc = np.clip(phi, a0, b0)
    dgr = orc.grad(pc, h)
    qs, bs = [], []
    for a in range(3):
        bb = b[a] + dgr[a]
        dn = orc.shrink(bb, g)
        bb = bb - dn
        qs.append(dn - bb)
        bs.append(bb)
    assert np.array_equal(vp.numpy(), pc)
    assert all(np.array_equal(v.numpy(), y) for v, y in zip(vb, bs))
    assert all(np.array_equal(v.numpy(), y) for v, y in zip(vq, qs))
    assert np.array_equal(vd.numpy(), orc.div(*qs, h))
    assert np.isclose(n1[0], np.sum((pc - phiold) ** 2), rtol=1e-12)
    assert np.isclose(n1[1], np.sum(phiold ** 2), rtol=1e-12)
