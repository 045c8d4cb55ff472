"""BF16 network layers of the learned warm start (tk_nn_*) against numpy restatements.

Inputs and weights are BF16-rounded first; numpy then computes in float64, so the
differences are the FP32 accumulation order and the final BF16 rounding of the output
(relative 2^-8), bounded in each assertion.  The convolution cases include partial tiles
(nz not a multiple of 16, co not a multiple of 16), channel slices of wider buffers and
the 1-channel input and output layers.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def to_bf16(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bf16(u):
    return (np.asarray(u, np.uint16).astype(np.uint32) << 16).view(np.float32)


class Buf:
    """raw device bytes (a U8 volume), for the BF16 / FP32 arrays of the layer tests"""

    def __init__(self, tk, ctx, arr=None, nbytes=None):
        self.tk = tk
        nbytes = arr.nbytes if arr is not None else nbytes
        self.v = ctx.volume((int(nbytes), 1, 1), np.uint8)
        self.n = int(nbytes)
        if arr is not None:
            self.v.upload(np.frombuffer(np.ascontiguousarray(arr).tobytes(), np.uint8).reshape(-1, 1, 1))

    @property
    def p(self):
        return self.tk.lib.tk_vol_device_ptr(self.v.h)

    def get(self, dtype, shape):
        return np.frombuffer(self.v.numpy().tobytes(), dtype).reshape(shape)


TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def shift(x, d):
    """x[i+dx, j+dy, k+dz] with zeros outside (x: X, Y, Z, C)"""
    out = np.zeros_like(x)
    X, Y, Z = x.shape[:3]
    src = [slice(max(0, e), n + min(0, e)) for e, n in zip(d, (X, Y, Z))]
    dst = [slice(max(0, -e), n - max(0, e)) for e, n in zip(d, (X, Y, Z))]
    out[dst[0], dst[1], dst[2]] = x[src[0], src[1], src[2]]
    return out


def conv_ref(x, w, b, ci, co):
    """x: (X,Y,Z,ci) float64, w: (co, kpad) float64 with k = tap*ci + c"""
    out = np.zeros(x.shape[:3] + (co,))
    for t, d in enumerate(TAPS):
        out += shift(x, d) @ w[:, t * ci:(t + 1) * ci].T
    return out + (b if b is not None else 0.0)


def kpad_of(ci):
    return (27 * ci + 15) // 16 * 16


@pytest.mark.parametrize("shape,ci,co,relu,out_f32,xs_extra,ys_extra",
                         [((4, 5, 20), 16, 32, True, False, 0, 0),
                          ((3, 4, 16), 1, 16, True, False, 3, 8),
                          ((5, 3, 9), 32, 16, False, False, 8, 16),
                          ((4, 4, 24), 16, 1, False, True, 0, 0),
                          ((2, 3, 40), 64, 64, True, False, 0, 0),
                          # co = 16, ci = 16 or 32, z a multiple of 16: the LDS-staged march
                          # (partial row tiles, channel slices, one-plane volumes, no relu)
                          ((6, 11, 32), 32, 16, True, False, 0, 0),
                          ((9, 8, 16), 16, 16, True, False, 16, 16),
                          ((5, 17, 48), 32, 16, False, False, 32, 0),
                          ((1, 3, 16), 16, 16, True, False, 0, 0),
                          ((3, 9, 32), 16, 16, False, False, 0, 32),
                          ((64, 64, 64), 32, 16, True, False, 0, 0),
                          # several 16-channel output slices, ci = 64
                          ((4, 9, 32), 64, 32, True, False, 0, 0),
                          ((6, 8, 16), 32, 32, True, False, 32, 32),
                          ((3, 5, 16), 16, 32, False, False, 0, 0),
                          ((5, 7, 16), 64, 64, True, False, 64, 0),
                          ((32, 32, 32), 64, 32, True, False, 0, 0),
                          # a partial last slice, one FP32 output channel (the output layer)
                          ((3, 4, 16), 16, 20, True, False, 0, 0),
                          ((4, 6, 32), 16, 1, False, True, 0, 0),
                          ((3, 5, 16), 32, 1, True, True, 0, 0),
                          # one input channel, 16 output channels in 16-byte stores
                          ((5, 6, 18), 1, 16, True, False, 0, 0),
                          ((4, 3, 7), 1, 12, False, False, 2, 0)])
def test_conv3d_fwd(tk, ctx, rng, shape, ci, co, relu, out_f32, xs_extra, ys_extra):
    say("conv3d fwd", shape, ci, co, relu, out_f32)
    X, Y, Z = shape
    n = X * Y * Z
    xs, xo = ci + xs_extra, xs_extra // 2
    ys, yo = co + ys_extra, ys_extra // 2
    kp = kpad_of(ci)
    x = to_bf16(rng.standard_normal((n, xs)))
    w = np.zeros((co, kp), np.float32)
    w[:, :27 * ci] = rng.standard_normal((co, 27 * ci)) / np.sqrt(27 * ci)
    w = to_bf16(w)
    b = rng.standard_normal(co).astype(np.float32)
    ysz = 4 if out_f32 else 2
    y0 = np.zeros((n, ys), np.float32 if out_f32 else np.uint16)
    bx, bw, bb, by = Buf(tk, ctx, x), Buf(tk, ctx, w), Buf(tk, ctx, b), Buf(tk, ctx, y0)
    tk._check(tk.lib.tk_nn_conv3d_fwd(ctx.h, bx.p, xs, xo, ci, bw.p, kp, bb.p, by.p, ys, yo, co,
                                      X, Y, Z, int(relu), int(out_f32)))
    got = by.get(np.float32 if out_f32 else np.uint16, (n, ys))
    got = got.astype(np.float64) if out_f32 else from_bf16(got).astype(np.float64)
    xf = from_bf16(x).astype(np.float64).reshape(X, Y, Z, xs)[..., xo:xo + ci]
    ref = conv_ref(xf, from_bf16(w).astype(np.float64), b.astype(np.float64), ci, co)
    if relu:
        ref = np.maximum(ref, 0)
    ref = ref.reshape(n, co)
    scale = np.max(np.abs(ref))
    err = np.max(np.abs(got[:, yo:yo + co] - ref))
    tol = (1e-5 if out_f32 else 8e-3) * scale
    say(f"  max err {err:.3e} (scale {scale:.3e}, tol {tol:.3e})")
    assert err <= tol
    # channels outside the slice untouched
    rest = np.delete(got, np.s_[yo:yo + co], axis=1)
    assert np.all(rest == 0)


@pytest.mark.parametrize("shape,ci,co", [((4, 5, 20), 16, 32), ((3, 4, 16), 1, 16),
                                          ((4, 3, 12), 32, 16)])
def test_conv3d_wgrad_and_dgrad(tk, ctx, rng, shape, ci, co):
    say("conv3d grads", shape, ci, co)
    X, Y, Z = shape
    n = X * Y * Z
    kp, kpt = kpad_of(ci), kpad_of(co)
    x = to_bf16(rng.standard_normal((n, ci)))
    dy = to_bf16(rng.standard_normal((n, co)))
    w = np.zeros((co, kp), np.float32)
    w[:, :27 * ci] = rng.standard_normal((co, 27 * ci))
    w = to_bf16(w)
    vchunk = 64
    nchunk = (n + vchunk - 1) // vchunk
    bx, bdy, bw = Buf(tk, ctx, x), Buf(tk, ctx, dy), Buf(tk, ctx, w)
    part = Buf(tk, ctx, nbytes=4 * nchunk * co * kp)
    bdw, bdb = Buf(tk, ctx, nbytes=4 * co * kp), Buf(tk, ctx, nbytes=4 * co)
    tk._check(tk.lib.tk_nn_conv3d_wgrad(ctx.h, bx.p, ci, 0, ci, bdy.p, co, 0, co, X, Y, Z, kp,
                                        part.p, vchunk, bdw.p, bdb.p))
    dw = bdw.get(np.float32, (co, kp)).astype(np.float64)
    db = bdb.get(np.float32, (co,)).astype(np.float64)
    xf = from_bf16(x).astype(np.float64).reshape(X, Y, Z, ci)
    dyf = from_bf16(dy).astype(np.float64).reshape(n, co)
    ref = np.zeros((co, kp))
    for t, d in enumerate(TAPS):
        ref[:, t * ci:(t + 1) * ci] = dyf.T @ shift(xf, d).reshape(n, ci)
    assert np.max(np.abs(dw - ref)) <= 1e-5 * np.max(np.abs(ref))
    if kp > 27 * ci:
        assert np.max(np.abs(dw[:, 27 * ci:])) == 0
    assert np.allclose(db, dyf.sum(0), rtol=1e-5, atol=1e-4)
    # input gradient through the flipped weights and the forward kernel
    wt = Buf(tk, ctx, nbytes=2 * ci * kpt)
    tk._check(tk.lib.tk_nn_flip_weights(ctx.h, bw.p, ci, co, kp, wt.p, kpt))
    bdx = Buf(tk, ctx, nbytes=4 * n * ci)
    tk._check(tk.lib.tk_nn_conv3d_fwd(ctx.h, bdy.p, co, 0, co, wt.p, kpt, None, bdx.p, ci, 0, ci,
                                      X, Y, Z, 0, 1))
    dx = bdx.get(np.float32, (n, ci)).astype(np.float64)
    wf = from_bf16(w).astype(np.float64)
    dyv = dyf.reshape(X, Y, Z, co)
    refx = np.zeros((X, Y, Z, ci))
    for t, d in enumerate(TAPS):
        # y[v] = sum_t w_t x[v + d_t]  =>  dx[u] = sum_t w_t^T dy[u - d_t]
        refx += shift(dyv, tuple(-e for e in d)) @ wf[:, t * ci:(t + 1) * ci]
    refx = refx.reshape(n, ci)
    assert np.max(np.abs(dx - refx)) <= 1e-5 * np.max(np.abs(refx))


@pytest.mark.parametrize("X,Y,Z,C,extra,xo", [(4, 6, 8, 16, 4, 2), (4, 6, 8, 16, 16, 8),
                                                (6, 4, 10, 32, 0, 0)])
def test_pool_upsample_relu_add(tk, ctx, rng, X, Y, Z, C, extra, xo):
    """pool and upsample with 2-byte channels (offset 2) and with 16-byte parts of 8 channels
    (strides and offsets multiples of 8): the two pools agree bit for bit"""
    n = X * Y * Z
    x = to_bf16(rng.standard_normal((n, C + extra)))
    bx = Buf(tk, ctx, x)
    by = Buf(tk, ctx, nbytes=2 * (n // 8) * C)
    tk._check(tk.lib.tk_nn_avgpool2(ctx.h, bx.p, C + extra, xo, C, X, Y, Z, by.p, C, 0))
    xf = from_bf16(x).astype(np.float64).reshape(X, Y, Z, C + extra)[..., xo:xo + C]
    ref = xf.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean(axis=(1, 3, 5)).reshape(-1, C)
    got = from_bf16(by.get(np.uint16, (n // 8, C)))
    assert np.max(np.abs(got - ref)) <= 8e-3 * np.max(np.abs(ref))
    # the same channels at a stride and offset that only 2-byte accesses serve
    x2 = np.zeros((n, C + 6), np.uint16)
    x2[:, 3:3 + C] = x.reshape(n, C + extra)[:, xo:xo + C]
    bx2, by2 = Buf(tk, ctx, x2), Buf(tk, ctx, nbytes=2 * (n // 8) * C)
    tk._check(tk.lib.tk_nn_avgpool2(ctx.h, bx2.p, C + 6, 3, C, X, Y, Z, by2.p, C, 0))
    assert np.array_equal(by2.get(np.uint16, (n // 8, C)), by.get(np.uint16, (n // 8, C)))
    # upsample of the pooled field, and its backward (sum of the 8 children)
    bu = Buf(tk, ctx, nbytes=2 * n * C)
    tk._check(tk.lib.tk_nn_upsample2(ctx.h, by.p, C, 0, C, X, Y, Z, bu.p, C, 0))
    up = from_bf16(bu.get(np.uint16, (n, C))).reshape(X, Y, Z, C)
    pooled = got.reshape(X // 2, Y // 2, Z // 2, C)
    assert np.array_equal(up, pooled.repeat(2, 0).repeat(2, 1).repeat(2, 2))
    bd = Buf(tk, ctx, nbytes=2 * (n // 8) * C)
    tk._check(tk.lib.tk_nn_upsample2_bwd(ctx.h, bu.p, C, 0, C, X, Y, Z, bd.p, C, 0))
    s = from_bf16(bd.get(np.uint16, (n // 8, C)))
    assert np.max(np.abs(s - 8 * got)) <= 8e-3 * np.max(np.abs(8 * got))
    # pool backward with accumulation into a slice
    g = to_bf16(rng.standard_normal((n // 8, C)))
    acc0 = to_bf16(rng.standard_normal((n, C)))
    bg, bacc = Buf(tk, ctx, g), Buf(tk, ctx, acc0)
    tk._check(tk.lib.tk_nn_avgpool2_bwd(ctx.h, bg.p, C, 0, C, X, Y, Z, bacc.p, C, 0, 1))
    gf = from_bf16(g).reshape(X // 2, Y // 2, Z // 2, C)
    refb = from_bf16(acc0).reshape(X, Y, Z, C) + gf.repeat(2, 0).repeat(2, 1).repeat(2, 2) / 8
    gotb = from_bf16(bacc.get(np.uint16, (n, C))).reshape(X, Y, Z, C)
    assert np.max(np.abs(gotb - refb)) <= 8e-3 * np.max(np.abs(refb))
    # relu backward masks where the output is not positive; add of slices
    yv = to_bf16(rng.standard_normal((n, C)))
    gv = to_bf16(rng.standard_normal((n, C)))
    byv, bgv = Buf(tk, ctx, yv), Buf(tk, ctx, gv)
    tk._check(tk.lib.tk_nn_relu_bwd(ctx.h, bgv.p, C, 0, byv.p, C, 0, C, n))
    gm = bgv.get(np.uint16, (n, C))
    assert np.array_equal(gm, np.where(from_bf16(yv) > 0, gv, 0).astype(np.uint16))
    tk._check(tk.lib.tk_nn_add_slice(ctx.h, bgv.p, C, 0, byv.p, C, 0, C, n))
    ga = from_bf16(bgv.get(np.uint16, (n, C)))
    assert np.max(np.abs(ga - (from_bf16(gm) + from_bf16(yv)))) <= 8e-3 * 4
