"""The learned warm start network on the device against the CPU FP64 network oracle
(oracle/unet_oracle.py, itself pinned by finite differences).

The device keeps weights, activations and activation gradients in BF16 (8-bit mantissa) and
accumulates in FP32; the oracle runs the same network in float64 on the same BF16-rounded
weights and input.  The differences are therefore the BF16 roundings of every stored
activation and gradient, and the bars below are stated for that: per layer, the relative L2
error of the gradient (< 0.15; measured 0.7-9 %) and its cosine with the oracle's (> 0.99,
mean over layers > 0.998); the output and the loss agree to a few 1e-3.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                "oracle"))
import unet_oracle as uo  # noqa: E402


def say(*a):
    print(*a, flush=True)


def bf16_round(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint32) << 16
    return r.astype(np.uint32).view(np.float32)


def init_params(rng):
    return np.concatenate([np.concatenate([rng.standard_normal(co * 27 * ci) * np.sqrt(2 / (27 * ci)),
                                           0.05 * rng.standard_normal(co)])
                           for ci, co in uo.LAYERS]).astype(np.float32)


def layer_slices():
    out, o = [], 0
    for ci, co in uo.LAYERS:
        out.append((slice(o, o + co * 27 * ci), slice(o + co * 27 * ci, o + co * 27 * ci + co)))
        o += co * (27 * ci + 1)
    return out


@pytest.mark.parametrize("shape", [(8, 8, 8), (8, 12, 16)])
def test_unet_forward_and_gradients_vs_oracle(tk, ctx, rng, shape):
    say("unet grad check", shape)
    net = tk.UNet(ctx, shape, seed=1)
    p = init_params(rng)
    assert p.size == net.n_params == uo.n_params()
    net.set_params(p)
    assert np.array_equal(net.params(), p)
    rhs = rng.standard_normal(shape)
    tgt = rng.standard_normal(shape)
    vr, vt = ctx.from_numpy(rhs.astype(np.float32)), ctx.from_numpy(tgt.astype(np.float32))
    loss = net.train_step(vr, vt, 0.0)
    g_dev = net.params(grads=True).astype(np.float64)
    out_dev = net.output().astype(np.float64)
    net.close()
    # the oracle on the same rounded weights and input (biases stay FP32 on the device)
    pr = p.astype(np.float64).copy()
    for ws, _ in layer_slices():
        pr[ws] = bf16_round(p[ws])
    r32 = rhs.astype(np.float32).astype(np.float64)
    rms = np.sqrt(np.mean(r32 ** 2))
    inp = bf16_round((r32 / rms).astype(np.float32)).astype(np.float64)
    t = tgt.astype(np.float32).astype(np.float64) / rms
    loss_ref, g_ref, out_ref = uo.loss_and_grad(pr, inp, t)
    e_out = np.max(np.abs(out_dev - out_ref)) / np.max(np.abs(out_ref))
    say(f"  output rel max err {e_out:.3e}; loss {loss:.6e} vs {loss_ref:.6e}")
    assert e_out < 3e-2
    assert abs(loss - loss_ref) < 2e-2 * loss_ref
    coss = []
    for l, (ws, bs) in enumerate(layer_slices()):
        for name, s in (("w", ws), ("b", bs)):
            a, b = g_dev[s], g_ref[s]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
            coss.append(cos)
            say(f"  layer {l} d{name}: rel L2 err {rel:.3e}, cosine {cos:.6f}")
            # the deepest layers' gradients pass through the most BF16 roundings
            assert rel < 0.15 and cos > 0.99
    assert np.mean(coss) > 0.998


def test_unet_training_reduces_loss(tk, ctx, rng):
    """Adam steps on a fixed pair drive the loss down (the whole training path runs)"""
    shape = (16, 16, 16)
    net = tk.UNet(ctx, shape, seed=2)
    x = np.linspace(0, np.pi, shape[0])
    smooth = np.sin(x)[:, None, None] * np.sin(x)[None, :, None] * np.sin(x)[None, None, :]
    rhs = smooth + 0.1 * rng.standard_normal(shape)
    tgt = 3.0 * smooth
    vr, vt = ctx.from_numpy(rhs.astype(np.float32)), ctx.from_numpy(tgt.astype(np.float32))
    losses = [net.train_step(vr, vt, 1e-3) for _ in range(60)]
    net.close()
    say(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.2 * losses[0]
