"""Python binding of libtk.so, the MI355X RSFE Split-Bregman segmentation library.

A thin ctypes layer over the C ABI declared in include/tk.h.  Volumes are numpy
arrays of shape (nx, ny, nz) in C order, i.e. index (i*ny + j)*nz + k with z
fastest -- the layout of the reference's im3d::image3d (image3d_imp.hxx:141-152).

The package directory name starts with a digit, so import it with
``importlib.import_module("3dimagetoolkit_amd")``.

There is no CPU fallback: if libtk.so is missing or cannot be loaded, importing
the binding raises.
"""
import ctypes as C
import weakref
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtk.so")

TK_OK = 0
TK_ERR_ARG, TK_ERR_HIP, TK_ERR_NOMEM, TK_ERR_SHAPE = -1, -2, -3, -4
TK_ERR_DTYPE, TK_ERR_UNSUPPORTED, TK_ERR_SWEEPS, TK_ERR_IO = -5, -6, -7, -8
TK_F32, TK_F64, TK_U8 = 1, 2, 3
TK_GS_LEX, TK_GS_REDBLACK = 0, 1
# tk_kernel_tag order
KERNEL_TAGS = ("conv_x1", "conv_y", "conv_z2", "conv_x2", "gs", "sb", "div", "other", "gs2",
               "sbdiv")

_NP2TK = {np.dtype(np.float32): TK_F32, np.dtype(np.float64): TK_F64, np.dtype(np.uint8): TK_U8}
_TK2NP = {v: k for k, v in _NP2TK.items()}


class TkError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tk error {code}: {msg}")
        self.code = code


class RsfeParams(C.Structure):
    """tk_rsfe_params (include/tk.h); defaults of rsfe_splitbregman_imp.hxx:25-41."""
    _fields_ = [(n, C.c_double) for n in (
        "sigma", "lamda", "lamda1", "lamda2", "beta", "dt", "epsilon", "a0", "b0", "nu",
        "tol", "ls_tol", "radius")] + [(n, C.c_int) for n in (
            "maxiter", "ls_steps", "bc_dir", "gaussian_pixel_approach", "gs_order",
            "max_sweeps", "warm_start")] + [("ls_res_tol", C.c_double),
                                            ("res_check_every", C.c_int),
                                            ("ws_presweeps", C.c_int),
                                            ("acyclic", C.c_int),
                                            ("neumann", C.c_int),
                                            ("auto_extract_conn_comp", C.c_int),
                                            ("auto_extract_conn_comp_freq", C.c_int),
                                            ("cc_init_pixel", C.c_double * 3),
                                            ("cc_threshold", C.c_double),
                                            ("cc_binary_output", C.c_int),
                                            ("cc_init_variables", C.c_int),
                                            ("dumpfrequency", C.c_int),
                                            ("outputname", C.c_char * 256),
                                            ("edge_detector_sigma", C.c_double),
                                            ("ws_judge_sweeps", C.c_int),
                                            ("ws_continue", C.c_int),
                                            ("ws_coarse_sweeps", C.c_int),
                                            ("verbosity", C.c_int)]


class IterStats(C.Structure):
    _fields_ = [("sweeps", C.c_int), ("ls_err", C.c_double), ("tollerance", C.c_double),
                ("guess_used", C.c_int), ("res_plain", C.c_double), ("res_guess", C.c_double),
                ("cc_extracted", C.c_int), ("coarse_sweeps", C.c_int), ("tol_inf", C.c_double)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run python 3dimagetoolkit_amd/build.py")
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp, i, d = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int, C.c_double
    u64 = C.c_uint64
    sig = {
        "tk_last_error": (C.c_char_p, []),
        "tk_version": (C.c_char_p, []),
        "tk_ctx_create": (i, [i, C.POINTER(vp)]),
        "tk_ctx_destroy": (i, [vp]),
        "tk_ctx_sync": (i, [vp]),
        "tk_ctx_set_stream": (i, [vp, vp]),
        "tk_ctx_stream": (vp, [vp]),
        "tk_ctx_set_timing": (i, [vp, i]),
        "tk_ctx_kernel_times": (i, [vp, dp, C.POINTER(C.c_long), i]),
        "tk_ctx_set_fused_update": (i, [vp, i]),
        "tk_vol_alloc": (i, [vp, i, i, i, d, d, d, i, C.POINTER(vp)]),
        "tk_vol_wrap": (i, [vp, vp, i, i, i, d, d, d, i, C.POINTER(vp)]),
        "tk_vol_free": (i, [vp]),
        "tk_vol_info": (i, [vp, ip, dp, ip]),
        "tk_vol_set_spacing": (i, [vp, d, d, d]),
        "tk_vol_device_ptr": (vp, [vp]),
        "tk_vol_upload": (i, [vp, vp, i]),
        "tk_vol_download": (i, [vp, vp, i]),
        "tk_vol_fill": (i, [vp, d]),
        "tk_vol_copy": (i, [vp, vp]),
        "tk_vol_write_mhd": (i, [vp, C.c_char_p]),
        "tk_vol_read_mhd": (i, [vp, C.c_char_p, i, C.POINTER(vp)]),
        "tk_filter_gaussian": (i, [vp, i, i, i, d, d, d, d, d, i, i, C.POINTER(vp)]),
        "tk_filter_free": (i, [vp]),
        "tk_filter_taps": (i, [vp, ip]),
        "tk_filter_axis_spectrum": (i, [vp, i, dp]),
        "tk_filter_get_method": (i, [vp, i, ip]),
        "tk_filter_set_method": (i, [vp, i, i]),
        "tk_filter_autotune": (i, [vp, vp, dp]),
        "tk_rsfe_filter": (vp, [vp]),
        "tk_conv_cyclic": (i, [vp, vp, vp, vp]),
        "tk_conv_cyclic_pair": (i, [vp, vp, vp, vp, vp, vp]),
        "tk_conv_acyclic": (i, [vp, vp, vp, vp]),
        "tk_conv_acyclic_pair": (i, [vp, vp, vp, vp, vp, vp]),
        "tk_poisson_dir_sweeps": (i, [vp, vp, vp, d, d, i, i]),
        "tk_poisson_dir_solve": (i, [vp, vp, vp, d, d, i, i, d, i, ip, dp]),
        "tk_poisson_neu_sweeps": (i, [vp, vp, vp, d, d, i, i]),
        "tk_poisson_neu_solve": (i, [vp, vp, vp, d, d, i, i, d, i, ip, dp]),
        "tk_poisson_residual": (i, [vp, vp, vp, dp]),
        "tk_poisson_rb_autotune": (i, [vp, vp, ip, dp]),
        "tk_poisson_rb_set_pair": (i, [vp, i]),
        "tk_poisson_solve_residual": (i, [vp, vp, vp, d, d, i, d, i, i, ip, dp]),
        "tk_poisson_residual_field": (i, [vp, vp, vp, vp]),
        "tk_rsfe_rhs": (vp, [vp]),
        "tk_rsfe_set_warm_start": (i, [vp, vp]),
        "tk_poisson_rb_set_tile": (i, [vp, vp, i]),
        "tk_sb_update": (i, [vp, vp, vp, vp, vp, vp, vp, d, d, dp]),
        "tk_heaviside": (i, [vp, vp, d, d, d, vp]),
        "tk_grad": (i, [vp, vp, vp, vp, vp]),
        "tk_div": (i, [vp, vp, vp, vp, vp]),
        "tk_edge_detect": (i, [vp, vp, d, d, d, vp]),
        "tk_shrink": (i, [vp, vp, vp, vp]),
        "tk_norm_l2": (i, [vp, vp, dp]),
        "tk_norm_l2_diff": (i, [vp, vp, vp, dp]),
        "tk_minmax": (i, [vp, vp, dp, dp]),
        "tk_change_range_of_intensity": (i, [vp, vp, d, d]),
        "tk_cut": (i, [vp, vp, d, d]),
        "tk_im_to_black_and_white": (i, [vp, vp, d, i, vp]),
        "tk_cc_seeded": (i, [vp, vp, i, i, i, d, i, vp, ip]),
        "tk_cc_label": (i, [vp, vp, vp, ip]),
        "tk_rsfe_default_params": (i, [C.POINTER(RsfeParams)]),
        "tk_rsfe_create": (i, [vp, vp, vp, C.POINTER(RsfeParams), C.POINTER(vp)]),
        "tk_rsfe_iterate": (i, [vp, i, C.POINTER(IterStats), ip]),
        "tk_rsfe_onthego": (i, [vp, C.POINTER(RsfeParams), i, i]),
        "tk_rsfe_phi": (vp, [vp]),
        "tk_rsfe_destroy": (i, [vp]),
        "tk_nn_conv3d_fwd": (i, [vp, vp, i, i, i, vp, i, vp, vp, i, i, i, i, i, i, i, i]),
        "tk_nn_conv3d_wgrad": (i, [vp, vp, i, i, i, vp, i, i, i, i, i, i, i, vp, i, vp, vp]),
        "tk_nn_flip_weights": (i, [vp, vp, i, i, i, vp, i]),
        "tk_nn_relu_bwd": (i, [vp, vp, i, i, vp, i, i, i, C.c_long]),
        "tk_nn_avgpool2": (i, [vp, vp, i, i, i, i, i, i, vp, i, i]),
        "tk_nn_avgpool2_bwd": (i, [vp, vp, i, i, i, i, i, i, vp, i, i, i]),
        "tk_nn_upsample2": (i, [vp, vp, i, i, i, i, i, i, vp, i, i]),
        "tk_nn_upsample2_bwd": (i, [vp, vp, i, i, i, i, i, i, vp, i, i]),
        "tk_nn_add_slice": (i, [vp, vp, i, i, vp, i, i, i, C.c_long]),
        "tk_vol_op_scalar": (i, [vp, vp, i, d]),
        "tk_vol_op_vol": (i, [vp, vp, i, vp]),
        "tk_vol_get": (i, [vp, i, i, i, dp]),
        "tk_vol_set": (i, [vp, i, i, i, d]),
        "tk_scalarprod": (i, [vp, vp, vp, dp]),
        "tk_norm1": (i, [vp, vp, dp]),
        "tk_norminf": (i, [vp, vp, dp]),
        "tk_norminf_diff": (i, [vp, vp, vp, dp, dp]),
        "tk_vector_abs": (i, [vp, vp, vp, vp, vp]),
        "tk_norm_l2_diff_noalloc": (i, [vp, vp, vp, C.c_long, i, d, d, d, vp]),
        "tk_vol_restrict": (i, [vp, vp, vp]),
        "tk_vol_prolong_add": (i, [vp, vp, vp, d]),
        "tk_nn_f32_to_bf16": (i, [vp, vp, C.c_long, vp]),
        "tk_nn_bf16_to_f32": (i, [vp, vp, C.c_long, vp]),
        "tk_unet_create": (i, [vp, i, i, i, u64, C.POINTER(vp)]),
        "tk_unet_destroy": (i, [vp]),
        "tk_unet_num_params": (i, [vp, C.POINTER(C.c_long)]),
        "tk_unet_set_out_scale": (i, [vp, d]),
        "tk_unet_get_out_scale": (i, [vp, dp]),
        "tk_unet_predict": (i, [vp, vp, vp]),
        "tk_unet_train_step": (i, [vp, vp, vp, d, dp]),
        "tk_unet_get_params": (i, [vp, i, C.POINTER(C.c_float), C.c_long]),
        "tk_unet_set_params": (i, [vp, C.POINTER(C.c_float), C.c_long]),
        "tk_unet_get_output": (i, [vp, C.POINTER(C.c_float)]),
        "tk_batch_create": (i, [vp, i, ip, i, C.POINTER(RsfeParams), C.POINTER(vp)]),
        "tk_batch_set_volume": (i, [vp, i, vp, vp, i]),
        "tk_batch_iterate": (i, [vp, i, i]),
        "tk_batch_get_phi": (i, [vp, i, vp, i]),
        "tk_batch_info": (i, [vp, ip, ip, ip]),
        "tk_batch_destroy": (i, [vp]),
        "tk_gen_ellipsoids": (i, [i, i, i, u64, i, dp, i, d, dp, dp]),
        "tk_gen_standard": (i, [i, i, i, u64, i, i, dp, dp]),
        "tk_gen_batch_dims": (i, [u64, i, ip, ip]),
        "tk_gen_ct": (i, [i, i, i, u64, i, dp, ip]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(rc):
    if rc != TK_OK:
        raise TkError(rc, lib.tk_last_error().decode())
    return rc


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---- synthetic volumes (host) ----------------------------------------------------
def gen_standard(dims, seed, scale=1, n_ell=2):
    """The standard ellipsoid phantom and its sphere initial contour (float64)."""
    nx, ny, nz = dims
    img = np.empty((nx, ny, nz), np.float64)
    phi0 = np.empty((nx, ny, nz), np.float64)
    _check(lib.tk_gen_standard(nx, ny, nz, seed, scale, n_ell, _dptr(img), _dptr(phi0)))
    return img, phi0


def gen_ct(dims, seed, n_vessels=12):
    nx, ny, nz = dims
    img = np.empty((nx, ny, nz), np.float64)
    sd = np.zeros(3, np.int32)
    _check(lib.tk_gen_ct(nx, ny, nz, seed, n_vessels, _dptr(img),
                         sd.ctypes.data_as(C.POINTER(C.c_int))))
    return img, tuple(int(x) for x in sd)


def gen_batch_dims(seed, n_vol):
    dims = np.zeros((n_vol, 3), np.int32)
    nell = np.zeros(n_vol, np.int32)
    _check(lib.tk_gen_batch_dims(seed, n_vol, dims.ctypes.data_as(C.POINTER(C.c_int)),
                                 nell.ctypes.data_as(C.POINTER(C.c_int))))
    return dims, nell


def default_params(**kw):
    p = RsfeParams()
    _check(lib.tk_rsfe_default_params(C.byref(p)))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise AttributeError(k)
        if k == "cc_init_pixel":
            v = (C.c_double * 3)(*[float(x) for x in v])
        elif k == "outputname":
            v = v.encode() if isinstance(v, str) else bytes(v)
            if len(v) >= 256:
                raise ValueError("outputname: at most 255 bytes")
        setattr(p, k, v)
    return p


# ---- device objects ------------------------------------------------------------------
class Context:
    """tk_ctx: one device and one HIP stream."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _check(lib.tk_ctx_create(device, C.byref(h)))
        self.h = h
        # objects living on this context, released before the context itself
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def sync(self):
        _check(lib.tk_ctx_sync(self.h))

    def set_timing(self, on):
        _check(lib.tk_ctx_set_timing(self.h, int(bool(on))))

    def set_fused_update(self, on):
        """update and div of single-volume runs in one pass (True, the default) or two"""
        _check(lib.tk_ctx_set_fused_update(self.h, int(bool(on))))

    def kernel_times(self, reset=True):
        """{tag: (total ms, launches)} from HIP events on this context's stream."""
        ms = (C.c_double * len(KERNEL_TAGS))()
        n = (C.c_long * len(KERNEL_TAGS))()
        _check(lib.tk_ctx_kernel_times(self.h, ms, n, int(bool(reset))))
        return {t: (ms[q], n[q]) for q, t in enumerate(KERNEL_TAGS)}

    def close(self):
        if self.h:
            for ch in list(self._children):
                ch.free() if hasattr(ch, "free") else ch.close()
            lib.tk_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # convenience constructors
    def volume(self, shape, dtype=np.float64, h=(1.0, 1.0, 1.0)):
        return Volume(self, shape, dtype, h)

    def from_numpy(self, a, dtype=None, h=(1.0, 1.0, 1.0)):
        a = np.ascontiguousarray(a)
        v = Volume(self, a.shape, dtype or a.dtype, h)
        v.upload(a)
        return v

    def read_mhd(self, path, dtype=np.float64):
        """a MetaImage file (the reference's dump format) as a new volume"""
        hv = C.c_void_p()
        _check(lib.tk_vol_read_mhd(self.h, os.fsencode(path), _NP2TK[np.dtype(dtype)],
                                   C.byref(hv)))
        dims = (C.c_int * 3)()
        h = (C.c_double * 3)()
        dt = C.c_int()
        _check(lib.tk_vol_info(hv, dims, h, C.byref(dt)))
        return Volume(self, tuple(dims), _TK2NP[dt.value], tuple(h), handle=hv)


class Volume:
    """tk_vol: a device-resident volume (im3d::image3d storage)."""

    def __init__(self, ctx, shape, dtype=np.float64, h=(1.0, 1.0, 1.0), handle=None):
        self.ctx = ctx
        ctx._adopt(self)
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.h_spacing = tuple(float(x) for x in h)
        if handle is not None:
            self.h = handle
            return
        hv = C.c_void_p()
        _check(lib.tk_vol_alloc(ctx.h, *self.shape, *[float(x) for x in h],
                                _NP2TK[self.dtype], C.byref(hv)))
        self.h = hv

    def upload(self, a):
        a = np.ascontiguousarray(a)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} != {self.shape}")
        _check(lib.tk_vol_upload(self.h, _ptr(a), _NP2TK[a.dtype]))
        return self

    def numpy(self, dtype=None):
        out = np.empty(self.shape, np.dtype(dtype) if dtype else self.dtype)
        _check(lib.tk_vol_download(self.h, _ptr(out), _NP2TK[out.dtype]))
        return out

    def fill(self, value):
        _check(lib.tk_vol_fill(self.h, float(value)))
        return self

    def write_mhd(self, basename):
        """<basename>.mhd + <basename>.raw (MetaImage, double, x fastest)"""
        _check(lib.tk_vol_write_mhd(self.h, os.fsencode(basename)))

    def free(self):
        if self.h:
            lib.tk_vol_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Filter:
    """tk_filter: Gaussian time filter and cyclic spectrum (conv::filtering statics)."""

    def __init__(self, ctx, shape, sigma, radius=3.0, pixel_approach=False, dtype=np.float64,
                 h=(1.0, 1.0, 1.0)):
        hv = C.c_void_p()
        _check(lib.tk_filter_gaussian(ctx.h, *shape, *h, float(sigma), float(radius),
                                      int(bool(pixel_approach)), _NP2TK[np.dtype(dtype)],
                                      C.byref(hv)))
        self.h = hv
        self.ctx = ctx
        ctx._adopt(self)
        self.shape = tuple(shape)

    def taps(self):
        nh = (C.c_int * 3)()
        _check(lib.tk_filter_taps(self.h, nh))
        return tuple(nh)

    def axis_spectrum(self, axis):
        out = np.empty(self.shape[axis], np.float64)
        _check(lib.tk_filter_axis_spectrum(self.h, axis, _dptr(out)))
        return out

    def methods(self):
        """per-axis method: 0 FFT pass, 1 direct circular convolution (vector ALUs),
        2 direct circular convolution on the matrix cores"""
        out = []
        for a in range(3):
            m = C.c_int()
            _check(lib.tk_filter_get_method(self.h, a, C.byref(m)))
            out.append(m.value)
        return out

    def set_method(self, axis, method):
        _check(lib.tk_filter_set_method(self.h, axis, method))

    def autotune(self):
        ms = (C.c_double * 9)()
        _check(lib.tk_filter_autotune(self.ctx.h, self.h, ms))
        return [(ms[3 * a], ms[3 * a + 1], ms[3 * a + 2]) for a in range(3)]

    def conv(self, out, inp, acyclic=False):
        """cyclic_fftw_convolution, or acyclic_fftw_convolution (zero outside the volume)"""
        fn = lib.tk_conv_acyclic if acyclic else lib.tk_conv_cyclic
        _check(fn(self.ctx.h, self.h, out.h, inp.h))
        return out

    def conv_pair(self, oa, ob, ia, ib, acyclic=False):
        fn = lib.tk_conv_acyclic_pair if acyclic else lib.tk_conv_cyclic_pair
        _check(fn(self.ctx.h, self.h, oa.h, ob.h, ia.h, ib.h))

    def free(self):
        if self.h:
            lib.tk_filter_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def poisson_sweeps(ctx, phi, rhs, n, dt=1.0, bc=0.0, order=TK_GS_LEX, neumann=False):
    """n DirGaussSeidel sweeps (bc: boundary value), or NeuGaussSeidel (bc: derivative)"""
    fn = lib.tk_poisson_neu_sweeps if neumann else lib.tk_poisson_dir_sweeps
    _check(fn(ctx.h, phi.h, rhs.h, dt, bc, order, n))


def poisson_solve(ctx, phi, rhs, ls_steps, ls_tol, dt=1.0, bc=0.0, order=TK_GS_LEX,
                  max_sweeps=0, allow_cap=False, neumann=False):
    s = C.c_int()
    e = C.c_double()
    fn = lib.tk_poisson_neu_solve if neumann else lib.tk_poisson_dir_solve
    rc = fn(ctx.h, phi.h, rhs.h, dt, bc, order, ls_steps, ls_tol, max_sweeps, C.byref(s),
            C.byref(e))
    if not (allow_cap and rc == TK_ERR_SWEEPS):
        _check(rc)
    return s.value, e.value


def poisson_residual(ctx, phi, rhs):
    r = C.c_double()
    _check(lib.tk_poisson_residual(ctx.h, phi.h, rhs.h, C.byref(r)))
    return r.value


def cc_seeded(ctx, img, seed, threshold, full, out):
    """image3d::connected_component(res, i, j, k, threshold, full) into the U8 volume out;
    returns the number of device passes."""
    passes = C.c_int()
    i, j, k = (int(x) for x in seed)
    _check(lib.tk_cc_seeded(ctx.h, img.h, i, j, k, float(threshold), int(bool(full)), out.h,
                            C.byref(passes)))
    return passes.value


def poisson_rb_autotune(ctx, rhs):
    """Time the red-black tile variants on rhs's shape; returns (rows, (ms8, ms16, ms32 of
    one sweep, ms8 and ms16 of the two-sweep launch; 0 where a variant does not apply))."""
    rows = C.c_int()
    ms = (C.c_double * 5)()
    _check(lib.tk_poisson_rb_autotune(ctx.h, rhs.h, C.byref(rows), ms))
    return rows.value, tuple(ms)


def poisson_rb_set_pair(ctx, enable):
    """red-black sweeps: two per launch where the autotuner found that faster (True, the
    default) or always one; gated solves stop at the same sweep with the same field"""
    _check(lib.tk_poisson_rb_set_pair(ctx.h, int(bool(enable))))


def poisson_rb_set_tile(ctx, rhs, rows):
    _check(lib.tk_poisson_rb_set_tile(ctx.h, rhs.h, int(rows)))


def sb_update(ctx, phi, phiold, g, b, q, divdb, a0=0.0, b0=1.0):
    """The Split-Bregman update and div (tk_sb_update): phi clamped, b updated in place,
    q = d - b, divdb = div(q); returns (sum (phi-phiold)^2, sum phiold^2)."""
    arr = lambda vs: (C.c_void_p * 3)(*[v.h for v in vs])
    norms = (C.c_double * 2)()
    _check(lib.tk_sb_update(ctx.h, phi.h, phiold.h, g.h, arr(b), arr(q), divdb.h, a0, b0,
                            norms))
    return norms[0], norms[1]


class Rsfe:
    """tk_rsfe: one segm::rsfe_splitbregman run resident on the device."""

    def __init__(self, ctx, u0, init, params):
        """init: a Volume (initialize_phi_with_init) or None (initialize_phi_with_cube)."""
        hv = C.c_void_p()
        _check(lib.tk_rsfe_create(ctx.h, u0.h, init.h if init is not None else None,
                                  C.byref(params), C.byref(hv)))
        self.h = hv
        self.ctx = ctx
        ctx._adopt(self)
        self.shape = u0.shape
        self.dtype = u0.dtype

    def iterate(self, n):
        st = (IterStats * max(n, 1))()
        done = C.c_int()
        _check(lib.tk_rsfe_iterate(self.h, n, st, C.byref(done)))
        return [dict(sweeps=st[q].sweeps, ls_err=st[q].ls_err, tollerance=st[q].tollerance,
                     guess_used=st[q].guess_used, res_plain=st[q].res_plain,
                     res_guess=st[q].res_guess, cc_extracted=st[q].cc_extracted,
                     coarse_sweeps=st[q].coarse_sweeps, tol_inf=st[q].tol_inf)
                for q in range(done.value)]

    def onthego(self, params, save_current=False, end_now=False):
        """update_param_onthego: params for the iterations that follow (tk_rsfe_onthego)"""
        _check(lib.tk_rsfe_onthego(self.h, C.byref(params), int(bool(save_current)),
                                   int(bool(end_now))))

    def phi(self, dtype=np.float64):
        out = np.empty(self.shape, dtype)
        _check(lib.tk_vol_download(lib.tk_rsfe_phi(self.h), _ptr(out), _NP2TK[np.dtype(dtype)]))
        return out

    def rhs(self, dtype=np.float64):
        """the forcing term of the last iteration (it holds that iteration's divdb)"""
        out = np.empty(self.shape, dtype)
        _check(lib.tk_vol_download(lib.tk_rsfe_rhs(self.h), _ptr(out), _NP2TK[np.dtype(dtype)]))
        return out

    def close(self):
        if self.h:
            lib.tk_rsfe_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """tk_batch: independent rsfe_splitbregman runs over volumes of uneven sizes."""

    def __init__(self, ctx, dims, params, dtype=np.float32):
        dims = np.ascontiguousarray(dims, np.int32).reshape(-1, 3)
        self.dims = [tuple(int(x) for x in d) for d in dims]
        self.dtype = np.dtype(dtype)
        hv = C.c_void_p()
        _check(lib.tk_batch_create(ctx.h, len(self.dims), dims.ctypes.data_as(C.POINTER(C.c_int)),
                                   _NP2TK[self.dtype], C.byref(params), C.byref(hv)))
        self.h = hv
        self.ctx = ctx
        ctx._adopt(self)

    def set_volume(self, v, u0, init=None):
        u0 = np.ascontiguousarray(u0, np.float64)
        assert u0.shape == self.dims[v]
        ip = None
        if init is not None:
            init = np.ascontiguousarray(init, np.float64)
            assert init.shape == self.dims[v]
            ip = init.ctypes.data_as(C.c_void_p)
        _check(lib.tk_batch_set_volume(self.h, v, u0.ctypes.data_as(C.c_void_p), ip,
                                       _NP2TK[np.dtype(np.float64)]))

    def iterate(self, n, use_graph=True):
        _check(lib.tk_batch_iterate(self.h, int(n), int(bool(use_graph))))

    def phi(self, v):
        out = np.empty(self.dims[v], np.float64)
        _check(lib.tk_batch_get_phi(self.h, v, out.ctypes.data_as(C.c_void_p),
                                    _NP2TK[np.dtype(np.float64)]))
        return out

    def info(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib.tk_batch_info(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return dict(lanes=a.value, graph_nodes=b.value, iterations=c.value)

    def close(self):
        if self.h:
            _check(lib.tk_batch_destroy(self.h))
            self.h = None


class UNet:
    """tk_unet: the learned Poisson warm start (3-level 3-D U-Net, BF16 on MFMA)."""

    def __init__(self, ctx, dims, seed=1):
        hv = C.c_void_p()
        _check(lib.tk_unet_create(ctx.h, *[int(x) for x in dims], int(seed), C.byref(hv)))
        self.h = hv
        self.ctx = ctx
        ctx._adopt(self)
        self.dims = tuple(int(x) for x in dims)
        c = C.c_long()
        _check(lib.tk_unet_num_params(self.h, C.byref(c)))
        self.n_params = c.value

    def _fp(self, a):
        return a.ctypes.data_as(C.POINTER(C.c_float))

    def params(self, grads=False):
        out = np.empty(self.n_params, np.float32)
        _check(lib.tk_unet_get_params(self.h, int(grads), self._fp(out), self.n_params))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(lib.tk_unet_set_params(self.h, self._fp(p), p.size))

    def output(self):
        out = np.empty(self.dims, np.float32)
        _check(lib.tk_unet_get_output(self.h, self._fp(out)))
        return out

    @property
    def out_scale(self):
        s = C.c_double()
        _check(lib.tk_unet_get_out_scale(self.h, C.byref(s)))
        return s.value

    @out_scale.setter
    def out_scale(self, s):
        _check(lib.tk_unet_set_out_scale(self.h, float(s)))

    def predict(self, rhs, guess):
        _check(lib.tk_unet_predict(self.h, rhs.h, guess.h))

    def train_step(self, rhs, target, lr):
        loss = C.c_double()
        _check(lib.tk_unet_train_step(self.h, rhs.h, target.h, float(lr), C.byref(loss)))
        return loss.value

    def close(self):
        if self.h:
            _check(lib.tk_unet_destroy(self.h))
            self.h = None
