"""CPU FP64 restatement of the learned warm start network (3-level 3-D U-Net).

TEST INFRASTRUCTURE ONLY: the checker of the device network's forward pass and gradients
(tests/test_gpu_unet.py).  The product never imports it.

The network has no counterpart in the reference (it is the build's addition on top of the
Poisson step of rsfe_splitbregman_imp.hxx:1762-1797).  Its definition, restated here in plain
float64 numpy: 11 convolutions with 3x3x3 kernels and zero padding, ReLU after all but the
last, 2x2x2 average pooling, nearest 2x upsampling and channel concatenation
(decoder channels first, skip channels second):

    e1a = relu(conv0(in))          e1b = relu(conv1(e1a))      p1 = pool(e1b)
    e2a = relu(conv2(p1))          e2b = relu(conv3(e2a))      p2 = pool(e2b)
    e3a = relu(conv4(p2))          e3b = relu(conv5(e3a))
    d2a = relu(conv6(up(e3b)))     d2b = relu(conv7([d2a, e2b]))
    d1a = relu(conv8(up(d2b)))     d1b = relu(conv9([d1a, e1b]))   out = conv10(d1b)

Parameters are flattened layer by layer as w[co][27*ci] (k = tap*ci + c, tap = (dx+1)*9 +
(dy+1)*3 + (dz+1)) followed by b[co].  The loss is mean((out - t)^2).  The backward pass is
pinned by central finite differences in tests/test_oracle_pins.py.
"""
import numpy as np

LAYERS = [(1, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 32), (64, 32),
          (32, 16), (32, 16), (16, 1)]
TAPS = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]


def n_params():
    return sum(co * (27 * ci + 1) for ci, co in LAYERS)


def unflatten(p):
    out, o = [], 0
    for ci, co in LAYERS:
        w = p[o:o + co * 27 * ci].reshape(co, 27 * ci)
        o += co * 27 * ci
        b = p[o:o + co]
        o += co
        out.append((w, b))
    return out


def shift(x, d):
    """x[i+dx, j+dy, k+dz] with zeros outside; x: (X, Y, Z, C)"""
    out = np.zeros_like(x)
    X, Y, Z = x.shape[:3]
    src = [slice(max(0, e), n + min(0, e)) for e, n in zip(d, (X, Y, Z))]
    dst = [slice(max(0, -e), n - max(0, e)) for e, n in zip(d, (X, Y, Z))]
    out[dst[0], dst[1], dst[2]] = x[src[0], src[1], src[2]]
    return out


def conv(x, w, b):
    ci = x.shape[3]
    y = np.zeros(x.shape[:3] + (w.shape[0],))
    for t, d in enumerate(TAPS):
        y += shift(x, d) @ w[:, t * ci:(t + 1) * ci].T
    return y + b


def conv_bwd(x, w, dy):
    ci = x.shape[3]
    n = x.shape[0] * x.shape[1] * x.shape[2]
    dyf = dy.reshape(n, -1)
    dw = np.zeros_like(w)
    dx = np.zeros_like(x)
    for t, d in enumerate(TAPS):
        dw[:, t * ci:(t + 1) * ci] = dyf.T @ shift(x, d).reshape(n, ci)
        dx += shift(dy, tuple(-e for e in d)) @ w[:, t * ci:(t + 1) * ci]
    return dx, dw, dyf.sum(0)


def pool(x):
    X, Y, Z, C = x.shape
    return x.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).mean(axis=(1, 3, 5))


def pool_bwd(g):
    return g.repeat(2, 0).repeat(2, 1).repeat(2, 2) / 8.0


def up(x):
    return x.repeat(2, 0).repeat(2, 1).repeat(2, 2)


def up_bwd(g):
    X, Y, Z, C = g.shape
    return g.reshape(X // 2, 2, Y // 2, 2, Z // 2, 2, C).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


def forward(params, inp):
    """inp: (X, Y, Z) normalised input; returns out (X, Y, Z) and the cache"""
    P = unflatten(params)
    c = {"in": inp[..., None]}
    c["e1a"] = relu(conv(c["in"], *P[0]))
    c["e1b"] = relu(conv(c["e1a"], *P[1]))
    c["p1"] = pool(c["e1b"])
    c["e2a"] = relu(conv(c["p1"], *P[2]))
    c["e2b"] = relu(conv(c["e2a"], *P[3]))
    c["p2"] = pool(c["e2b"])
    c["e3a"] = relu(conv(c["p2"], *P[4]))
    c["e3b"] = relu(conv(c["e3a"], *P[5]))
    c["u2"] = up(c["e3b"])
    c["d2a"] = relu(conv(c["u2"], *P[6]))
    c["cat2"] = np.concatenate([c["d2a"], c["e2b"]], axis=3)
    c["d2b"] = relu(conv(c["cat2"], *P[7]))
    c["u1"] = up(c["d2b"])
    c["d1a"] = relu(conv(c["u1"], *P[8]))
    c["cat1"] = np.concatenate([c["d1a"], c["e1b"]], axis=3)
    c["d1b"] = relu(conv(c["cat1"], *P[9]))
    out = conv(c["d1b"], *P[10])[..., 0]
    return out, c


def loss_and_grad(params, inp, target):
    """loss = mean((out - target)^2) and its gradient, flattened like params"""
    P = unflatten(params)
    out, c = forward(params, inp)
    n = out.size
    loss = float(np.mean((out - target) ** 2))
    g = [None] * 11
    gout = (2.0 * (out - target) / n)[..., None]
    gx, *g[10] = conv_bwd(c["d1b"], P[10][0], gout)
    gx = gx * (c["d1b"] > 0)
    gcat1, *g[9] = conv_bwd(c["cat1"], P[9][0], gx)
    gd1a = gcat1[..., :16] * (c["d1a"] > 0)
    ge1b = gcat1[..., 16:]
    gu1, *g[8] = conv_bwd(c["u1"], P[8][0], gd1a)
    gd2b = up_bwd(gu1) * (c["d2b"] > 0)
    gcat2, *g[7] = conv_bwd(c["cat2"], P[7][0], gd2b)
    gd2a = gcat2[..., :32] * (c["d2a"] > 0)
    ge2b = gcat2[..., 32:]
    gu2, *g[6] = conv_bwd(c["u2"], P[6][0], gd2a)
    ge3b = up_bwd(gu2) * (c["e3b"] > 0)
    ge3a, *g[5] = conv_bwd(c["e3a"], P[5][0], ge3b)
    ge3a = ge3a * (c["e3a"] > 0)
    gp2, *g[4] = conv_bwd(c["p2"], P[4][0], ge3a)
    ge2b = (ge2b + pool_bwd(gp2)) * (c["e2b"] > 0)
    ge2a, *g[3] = conv_bwd(c["e2a"], P[3][0], ge2b)
    ge2a = ge2a * (c["e2a"] > 0)
    gp1, *g[2] = conv_bwd(c["p1"], P[2][0], ge2a)
    ge1b = (ge1b + pool_bwd(gp1)) * (c["e1b"] > 0)
    ge1a, *g[1] = conv_bwd(c["e1a"], P[1][0], ge1b)
    ge1a = ge1a * (c["e1a"] > 0)
    _, *g[0] = conv_bwd(c["in"], P[0][0], ge1a)
    flat = np.concatenate([np.concatenate([dw.reshape(-1), db]) for dw, db in g])
    return loss, flat, out
