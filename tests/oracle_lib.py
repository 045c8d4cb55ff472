"""ctypes access to the CPU FP64 oracle (oracle/liboracle.so) -- test side only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")


class OrParams(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "sigma", "lamda", "lamda1", "lamda2", "beta", "dt", "epsilon", "a0", "b0", "nu", "tol",
        "ls_tol", "radius")] + [(n, C.c_int) for n in (
            "maxiter", "ls_steps", "bc_dir", "gaussian_pixel_approach", "verbosity",
            "acyclic", "neumann", "auto_cc", "auto_cc_freq")] + [
                ("cc_pixel", C.c_double * 3), ("cc_binary_output", C.c_int),
                ("cc_init_variables", C.c_int), ("edge_detector_sigma", C.c_double)]


class OrStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("cc_extractions", C.c_int), ("final_tollerance", C.c_double),
                ("t_setup", C.c_double),
                ("t_loop", C.c_double), ("t_fft", C.c_double)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(PATH):
            raise ImportError(f"{PATH} missing: run make -C oracle")
        L = C.CDLL(PATH)
        dp, ip, i, d, vp = C.POINTER(C.c_double), C.POINTER(C.c_int), C.c_int, C.c_double, C.c_void_p
        L.or_norm_l2.restype = d
        for fn in ("or_scalarprod", "or_norm1", "or_norminf", "or_max", "or_min"):
            getattr(L, fn).restype = d
        L.or_rsfe_apply.restype = i
        L.or_connected_component.restype = i
        L.or_poisson_loop.restype = i
        L.or_poisson_loop_bc.restype = i
        L.or_neu_gauss_seidel.restype = i
        _lib = L
    return _lib


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ip(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_int))


def gaussian(shape, sigma, radius=3.0, pixel=False, h=(1.0, 1.0, 1.0)):
    L = lib()
    nh = np.zeros(3, np.int32)
    L.or_gaussian_dims(C.c_int(shape[2]), C.c_double(h[0]), C.c_double(h[1]), C.c_double(h[2]),
                       C.c_double(sigma), C.c_double(radius), C.c_int(int(pixel)), ip(nh))
    tf = np.zeros(tuple(nh), np.float64)
    L.or_gaussian_time_filter(C.c_int(shape[2]), C.c_double(h[0]), C.c_double(h[1]),
                              C.c_double(h[2]), C.c_double(sigma), C.c_double(radius),
                              C.c_int(int(pixel)), dp(tf))
    nc = np.zeros(3, np.int32)
    L.or_cyclic_dims(C.c_int(shape[0]), C.c_int(shape[1]), C.c_int(shape[2]), ip(nh), ip(nc))
    spec = np.zeros(tuple(nc), np.float64)
    L.or_cyclic_spectrum(dp(tf), ip(nh), ip(nc), dp(spec))
    return tf, nh, spec, nc


def cyclic_conv(spec, nc, x, h=(1.0, 1.0, 1.0)):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().or_cyclic_conv(dp(spec), ip(nc), dp(x), C.c_int(x.shape[0]), C.c_int(x.shape[1]),
                         C.c_int(x.shape[2]), C.c_double(h[0]), C.c_double(h[1]),
                         C.c_double(h[2]), dp(out))
    return out


def direct_cyclic_conv(tf, nh, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().or_direct_cyclic_conv(dp(tf), ip(nh), dp(x), *[C.c_int(s) for s in x.shape], dp(out))
    return out


def acyclic_spectrum(tf, nh, shape):
    """spectrum of conv::acyclic_fftw_convolution on the padded dims shape + nh - 1"""
    na = np.zeros(3, np.int32)
    lib().or_acyclic_dims(C.c_int(shape[0]), C.c_int(shape[1]), C.c_int(shape[2]), ip(nh),
                          ip(na))
    spec = np.zeros(tuple(na), np.float64)
    lib().or_cyclic_spectrum(dp(tf), ip(nh), ip(na), dp(spec))
    return spec, na


def acyclic_conv(spec, na, x, h=(1.0, 1.0, 1.0)):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().or_acyclic_conv(dp(spec), ip(na), dp(x), C.c_int(x.shape[0]), C.c_int(x.shape[1]),
                          C.c_int(x.shape[2]), C.c_double(h[0]), C.c_double(h[1]),
                          C.c_double(h[2]), dp(out))
    return out


def direct_acyclic_conv(tf, nh, x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros_like(x)
    lib().or_direct_acyclic_conv(dp(tf), ip(nh), dp(x), *[C.c_int(s) for s in x.shape], dp(out))
    return out


def fft(re, im, sign=-1):
    re = np.array(re, np.float64)
    im = np.array(im, np.float64)
    lib().or_fft_inplace(dp(re), dp(im), C.c_int(len(re)), C.c_int(sign))
    return re + 1j * im


def gs_sweep(phi, rhs, dt=1.0, bc=0.0, h=(1.0, 1.0, 1.0)):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    rhs = np.ascontiguousarray(rhs, np.float64)
    lib().or_dir_gauss_seidel(dp(phi), dp(rhs), *[C.c_int(s) for s in phi.shape],
                              *[C.c_double(x) for x in h], C.c_double(dt), C.c_double(bc))
    return phi


def neu_sweep(phi, rhs, dt=1.0, bc=0.0, h=(1.0, 1.0, 1.0)):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    rhs = np.ascontiguousarray(rhs, np.float64)
    rc = lib().or_neu_gauss_seidel(dp(phi), dp(rhs), *[C.c_int(s) for s in phi.shape],
                                   *[C.c_double(x) for x in h], C.c_double(dt), C.c_double(bc))
    if rc != 0:
        raise ValueError("or_neu_gauss_seidel: 3-D volumes with every edge >= 2 only")
    return phi


def poisson_loop(phi, rhs, ls_steps, ls_tol, dt=1.0, bc=0.0, h=(1.0, 1.0, 1.0), neumann=False):
    phi = np.ascontiguousarray(phi, np.float64).copy()
    rhs = np.ascontiguousarray(rhs, np.float64)
    err = C.c_double()
    s = lib().or_poisson_loop_bc(dp(phi), dp(rhs), *[C.c_int(q) for q in phi.shape],
                                 *[C.c_double(x) for x in h], C.c_double(dt), C.c_double(bc),
                                 C.c_int(ls_steps), C.c_double(ls_tol), C.byref(err),
                                 C.c_int(int(bool(neumann))))
    return phi, s, err.value


def heaviside(phi, a0=0.0, b0=1.0, eps=1e-3):
    phi = np.ascontiguousarray(phi, np.float64)
    out = np.zeros_like(phi)
    lib().or_heaviside(dp(phi), C.c_long(phi.size), C.c_double(a0), C.c_double(b0),
                       C.c_double(eps), dp(out))
    return out


def grad(f, h=(1.0, 1.0, 1.0)):
    f = np.ascontiguousarray(f, np.float64)
    g = [np.zeros_like(f) for _ in range(3)]
    lib().or_grad(dp(f), *[C.c_int(s) for s in f.shape], *[C.c_double(x) for x in h],
                  dp(g[0]), dp(g[1]), dp(g[2]))
    return g


def div(f0, f1, f2, h=(1.0, 1.0, 1.0)):
    f = [np.ascontiguousarray(a, np.float64) for a in (f0, f1, f2)]
    out = np.zeros_like(f[0])
    lib().or_div(dp(f[0]), dp(f[1]), dp(f[2]), *[C.c_int(s) for s in f[0].shape],
                 *[C.c_double(x) for x in h], dp(out))
    return out


def edge_detect(u0, beta=100.0, lamda=5e-2, nu=1.0, h=(1.0, 1.0, 1.0)):
    u0 = np.ascontiguousarray(u0, np.float64)
    g = np.zeros_like(u0)
    lib().or_edge_detect(dp(u0), *[C.c_int(s) for s in u0.shape], *[C.c_double(x) for x in h],
                         C.c_double(beta), C.c_double(lamda), C.c_double(nu), dp(g))
    return g


def shrink(b, g):
    b = np.ascontiguousarray(b, np.float64)
    g = np.ascontiguousarray(g, np.float64)
    out = np.zeros_like(b)
    lib().or_shrink(dp(b), dp(g), C.c_long(b.size), dp(out))
    return out


def scalarprod(a, b):
    return lib().or_scalarprod(dp(np.ascontiguousarray(a, np.float64)),
                               dp(np.ascontiguousarray(b, np.float64)), C.c_long(a.size))


def norm_l2(a, h=(1.0, 1.0, 1.0)):
    a = np.ascontiguousarray(a, np.float64)
    return lib().or_norm_l2(dp(a), C.c_long(a.size), *[C.c_double(x) for x in h])


def norm1(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib().or_norm1(dp(a), C.c_long(a.size))


def norminf(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib().or_norminf(dp(a), C.c_long(a.size))


def vmax(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib().or_max(dp(a), C.c_long(a.size))


def vmin(a):
    a = np.ascontiguousarray(a, np.float64)
    return lib().or_min(dp(a), C.c_long(a.size))


def change_range(a, mx, mn):
    a = np.ascontiguousarray(a, np.float64).copy()
    lib().or_change_range_of_intensity(dp(a), C.c_long(a.size), C.c_double(mx), C.c_double(mn))
    return a


def black_and_white(a, thr, negative=False):
    a = np.ascontiguousarray(a, np.float64)
    out = np.zeros_like(a)
    lib().or_im_to_black_and_white(dp(a), C.c_long(a.size), C.c_double(thr), C.c_int(int(negative)),
                                   dp(out))
    return out


def connected_component(img, seed, thr=0.5, full=False):
    img = np.ascontiguousarray(img, np.float64)
    res = np.zeros_like(img)
    passes = lib().or_connected_component(dp(img), *[C.c_int(s) for s in img.shape],
                                          *[C.c_int(int(s)) for s in seed], C.c_double(thr),
                                          C.c_int(int(full)), dp(res))
    return res, passes


def init_phi(init, shape, a0=0.0, b0=1.0):
    """initialize_phi_with_init / _with_cube; init None gives the cube.  Returns (phi, case)."""
    X, Y, Z = shape
    phi = np.empty(shape)
    if init is None:
        case = lib().or_init_phi(None, 0, 0, 0, X, Y, Z, C.c_double(a0), C.c_double(b0), dp(phi))
    else:
        init = np.ascontiguousarray(init, dtype=np.float64)
        case = lib().or_init_phi(dp(init), *init.shape, X, Y, Z, C.c_double(a0), C.c_double(b0),
                                 dp(phi))
    return phi, case


def default_params(**kw):
    p = OrParams()
    lib().or_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        if k == "cc_pixel":
            v = (C.c_double * 3)(*v)
        setattr(p, k, v)
    return p


def rsfe_apply(u0, init, params, h=(1.0, 1.0, 1.0)):
    u0 = np.ascontiguousarray(u0, np.float64)
    init = np.ascontiguousarray(init, np.float64)
    phi = np.zeros_like(u0)
    sw = np.zeros(max(params.maxiter, 1), np.int32)
    tl = np.zeros(max(params.maxiter, 1), np.float64)
    st = OrStats()
    n = lib().or_rsfe_apply(dp(u0), dp(init), *[C.c_int(s) for s in u0.shape],
                            *[C.c_double(x) for x in h], C.byref(params), dp(phi), ip(sw), dp(tl),
                            C.byref(st))
    return phi, sw[:n].copy(), tl[:n].copy(), st


def rsfe_apply_onthego(u0, init, p1, t_switch, p2, end_now=False, h=(1.0, 1.0, 1.0)):
    """rsfe_apply with the on-the-go re-read once: after iteration t_switch the parameters
    become p2 (end_now: the loop stops after that iteration)"""
    u0 = np.ascontiguousarray(u0, np.float64)
    init = np.ascontiguousarray(init, np.float64)
    phi = np.zeros_like(u0)
    m = max(p1.maxiter, p2.maxiter, 1)
    sw = np.zeros(m, np.int32)
    tl = np.zeros(m, np.float64)
    st = OrStats()
    lib().or_rsfe_apply_onthego.restype = C.c_int
    n = lib().or_rsfe_apply_onthego(dp(u0), dp(init), *[C.c_int(s) for s in u0.shape],
                                    *[C.c_double(x) for x in h], C.byref(p1), C.c_int(t_switch),
                                    C.byref(p2), C.c_int(int(end_now)), dp(phi), ip(sw), dp(tl),
                                    C.byref(st))
    return phi, sw[:n].copy(), tl[:n].copy(), st


def rsfe_forcing(u0, phi, divdb, params, h=(1.0, 1.0, 1.0)):
    u0 = np.ascontiguousarray(u0, np.float64)
    phi = np.ascontiguousarray(phi, np.float64)
    divdb = np.ascontiguousarray(divdb, np.float64)
    rhs = np.zeros_like(u0)
    lib().or_rsfe_forcing(dp(u0), dp(phi), dp(divdb), *[C.c_int(s) for s in u0.shape],
                          *[C.c_double(x) for x in h], C.byref(params), dp(rhs))
    return rhs


def rsfe_apply_ckpt(u0, init, params, ckpts, h=(1.0, 1.0, 1.0)):
    """rsfe_apply that also returns phi after each outer iteration listed in ckpts (1-based)"""
    u0 = np.ascontiguousarray(u0, np.float64)
    init = np.ascontiguousarray(init, np.float64)
    phi = np.zeros_like(u0)
    sw = np.zeros(max(params.maxiter, 1), np.int32)
    tl = np.zeros(max(params.maxiter, 1), np.float64)
    st = OrStats()
    ck = np.ascontiguousarray(ckpts, np.int32)
    out = np.zeros((len(ck),) + u0.shape, np.float64)
    lib().or_rsfe_apply_ckpt.restype = C.c_int
    n = lib().or_rsfe_apply_ckpt(dp(u0), dp(init), *[C.c_int(s) for s in u0.shape],
                                 *[C.c_double(x) for x in h], C.byref(params), dp(phi), ip(sw),
                                 dp(tl), C.byref(st), C.c_int(len(ck)), ip(ck), dp(out))
    return phi, sw[:n].copy(), tl[:n].copy(), st, out
