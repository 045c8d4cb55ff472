/*
 * tk.h -- C ABI of the MI355X RSFE Split-Bregman segmentation library (libtk.so).
 *
 * Plain C: opaque handles, pointers and sizes; every function returns an int status
 * (TK_OK == 0, negative on error) and never aborts.  The C++ class layer in
 * include/tk/ (im3d::image3d, conv::filtering, lapl::unsteady_poisson_functor,
 * segm::rsfe_splitbregman) is built on these entry points, and a C, ctypes, cgo or
 * JNI caller can bind them directly.  Reference citations are to
 * marco-fedele/3dimagetoolkit (header/...).
 *
 * Volumes: nx*ny*nz voxels, index (i*ny + j)*nz + k (z fastest), exactly the
 * reference's image3d order (image3d_imp.hxx:141-152), so uploads and downloads
 * are plain copies.  Element type TK_F64 (the reference's double instantiation)
 * or TK_F32 (FP32 arithmetic with FP64 reductions); masks are TK_U8.
 */
#ifndef TK_H
#define TK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ---------------------------------------------------------------- */
enum tk_status {
    TK_OK = 0,
    TK_ERR_ARG = -1,         /* bad argument (null handle, bad size, bad enum) */
    TK_ERR_HIP = -2,         /* HIP runtime error; tk_last_error() has the text */
    TK_ERR_NOMEM = -3,       /* device or host allocation failed */
    TK_ERR_SHAPE = -4,       /* dimensions of the operands do not agree */
    TK_ERR_DTYPE = -5,       /* element types of the operands do not agree */
    TK_ERR_UNSUPPORTED = -6, /* valid request outside what this build implements */
    TK_ERR_SWEEPS = -7,      /* Poisson solve hit max_sweeps before ls_tol */
    TK_ERR_IO = -8           /* a file could not be opened, read, parsed or written */
};

typedef enum tk_dtype { TK_F32 = 1, TK_F64 = 2, TK_U8 = 3 } tk_dtype;

typedef struct tk_ctx_s* tk_ctx;       /* one device + one HIP stream + scratch */
typedef struct tk_vol_s* tk_vol;       /* a volume resident in device memory */
typedef struct tk_filter_s* tk_filter; /* a Gaussian filter's cyclic spectrum */
typedef struct tk_rsfe_s* tk_rsfe;     /* state of one Split-Bregman run */
typedef struct tk_unet_s* tk_unet;     /* the learned Poisson warm start network */

/* message of the last error on the calling thread ("" if none) */
const char* tk_last_error(void);
const char* tk_version(void);

/* ---- context ----------------------------------------------------------------
 * No reference counterpart (the reference is a host-only library).  One context
 * per host thread; all work of a context is ordered on its stream. */
int tk_ctx_create(int device, tk_ctx* out);
/* every volume, filter, run, network and batch created on ctx must be freed first */
int tk_ctx_destroy(tk_ctx ctx);
int tk_ctx_sync(tk_ctx ctx);
/* run on a caller-owned hipStream_t (e.g. PyTorch's current stream); NULL restores
   the context's own stream */
int tk_ctx_set_stream(tk_ctx ctx, void* hip_stream);
void* tk_ctx_stream(tk_ctx ctx);

/* per-kernel timing with HIP events recorded on the context's stream around every
   launch of the Split-Bregman path (off by default; no reference counterpart) */
enum tk_kernel_tag {
    TK_K_CONV_X1 = 0, /* x pass of pair 1 with the Heaviside prologue */
    TK_K_CONV_Y = 1,  /* y passes (complex -> complex) */
    TK_K_CONV_Z2 = 2, /* z pass: pair 1, region fit, pair 2 */
    TK_K_CONV_X2 = 3, /* x pass of pair 2 with the forcing-term epilogue */
    TK_K_GS = 4,      /* one Gauss-Seidel launch */
    TK_K_SB = 5,      /* cut + norms + grad/shrink/Bregman update */
    TK_K_DIV = 6,     /* divergence */
    TK_K_OTHER = 7,
    TK_K_GS2 = 8,     /* two red-black sweeps in one launch */
    TK_K_SBDIV = 9,   /* cut + norms + grad/shrink/Bregman update and divergence in one pass */
    TK_K_COUNT = 10
};
int tk_ctx_set_timing(tk_ctx ctx, int enable);
/* single-volume Split-Bregman runs: the update and div in one pass (enable 1, the default)
   or in two (0), on every later iteration of runs on ctx; the two give bitwise the same
   phi, b and divdb (the tolerance sums are added in another order) */
int tk_ctx_set_fused_update(tk_ctx ctx, int enable);
/* totals since the last call with reset != 0: milliseconds and launches per tag */
int tk_ctx_kernel_times(tk_ctx ctx, double ms[TK_K_COUNT], long launches[TK_K_COUNT], int reset);

/* ---- volumes: im3d::image3d<T> storage ----------------------------------------
 * image3d(x,y,z,hx,hy,hz) ctor (image3d_imp.hxx:35-41), setdim/seth (93-135),
 * operator=(scalar) (158-167), operator=(image3d) (170-184). */
int tk_vol_alloc(tk_ctx ctx, int nx, int ny, int nz, double hx, double hy, double hz,
                 tk_dtype dtype, tk_vol* out);
/* wrap caller-owned device memory (not freed by tk_vol_free) */
int tk_vol_wrap(tk_ctx ctx, void* device_ptr, int nx, int ny, int nz, double hx, double hy,
                double hz, tk_dtype dtype, tk_vol* out);
int tk_vol_free(tk_vol vol);
int tk_vol_info(tk_vol vol, int dims[3], double h[3], tk_dtype* dtype);
/* image3d::seth (image3d.hxx:238): spacing only, voxels unchanged */
int tk_vol_set_spacing(tk_vol vol, double hx, double hy, double hz);
void* tk_vol_device_ptr(tk_vol vol);
/* host <-> device; host_dtype may differ from the volume's (converted on device) */
int tk_vol_upload(tk_vol vol, const void* host, tk_dtype host_dtype);
int tk_vol_download(tk_vol vol, void* host, tk_dtype host_dtype);
int tk_vol_fill(tk_vol vol, double value);
int tk_vol_copy(tk_vol dst, tk_vol src); /* converts F32 <-> F64 */

/* ---- MetaImage files: the reference's dump format -------------------------------------
 * im3d::interface<S>::write(name, ".mhd") (interface_imp.hxx:672-690) of the image
 * convertfromimage3d made (104-122): <basename>.mhd (text header: DimSize nx ny nz,
 * ElementSpacing hx hy hz, ElementType MET_DOUBLE, little-endian, uncompressed) and
 * <basename>.raw (the voxels as double, x fastest: i + nx (j + ny k)).  The reader takes
 * any uncompressed MetaImage (NDims 2 or 3, MET_CHAR .. MET_DOUBLE, either byte order,
 * ElementDataFile LOCAL or a file, HeaderSize, channels: component 0), as
 * interface<S>(name.mhd) + convert2image3d (interface_imp.hxx:33-47, 125-150) do, into a
 * new volume of dtype. */
int tk_vol_write_mhd(tk_vol vol, const char* basename);
int tk_vol_read_mhd(tk_ctx ctx, const char* path, tk_dtype dtype, tk_vol* out);

/* ---- R4: Gaussian filter and its cyclic spectrum ---------------------------------
 * conv::filtering<T>::build_gaussian_filter (convolution_imp.hxx:22-91) and the
 * cyclic part of build_freq_filters (218-264, 404-531).  The spectrum of the
 * separable normalised kernel is kept per axis (the 3-D real spectrum is their
 * product); Nh > dim wraps the kernel as the reference's multiple-of-dims padding. */
int tk_filter_gaussian(tk_ctx ctx, int nx, int ny, int nz, double hx, double hy, double hz,
                       double sigma, double radius, int pixel_approach, tk_dtype dtype,
                       tk_filter* out);
int tk_filter_free(tk_filter flt);
int tk_filter_taps(tk_filter flt, int nh[3]);
/* real spectrum along one axis (length dims[axis]), including nothing else */
int tk_filter_axis_spectrum(tk_filter flt, int axis, double* out);
/* per-axis method: 0 = FFT pass (forward, spectrum multiply, inverse in LDS),
   1 = direct circular convolution with the symmetric taps on the vector ALUs,
   2 = the same direct convolution as banded matrix products on the matrix cores (FP32,
   axis length a multiple of 16, pencil groups of a multiple of 16); all three give the
   same result up to rounding.  No reference counterpart: the reference always uses FFTW. */
int tk_filter_get_method(tk_filter flt, int axis, int* method);
int tk_filter_set_method(tk_filter flt, int axis, int method);
/* autotuner: time every method of every axis on the device (on the filter's own
   sizes) and keep the fastest; ms (may be NULL) receives {fft, direct, direct on the
   matrix cores} per axis, -1 where a method is unavailable */
int tk_filter_autotune(tk_ctx ctx, tk_filter flt, double ms[9]);

/* ---- R5: cyclic (periodic) convolution ------------------------------------------
 * conv::cyclic_fftw_convolution (convolution_imp.hxx:657-757), the function the
 * functor conv::filtering<T>::operator() (convolution.hxx:104-107) dispatches to
 * after set_cv(filtering<T>(cyclic_fftw_convolution)).  out may alias in. */
int tk_conv_cyclic(tk_ctx ctx, tk_filter flt, tk_vol out, tk_vol in);
/* two real fields through one complex transform (outputs may alias inputs) */
int tk_conv_cyclic_pair(tk_ctx ctx, tk_filter flt, tk_vol out_a, tk_vol out_b, tk_vol in_a,
                        tk_vol in_b);

/* ---- R5, acyclic: zero outside the volume -------------------------------------------
 * conv::acyclic_fftw_convolution (convolution_imp.hxx:535-653), the reference's default cv
 * (rsfe_splitbregman_imp.hxx:25-41): the image zero-padded to dims + nh - 1, the padded
 * spectrum applied, the first dims voxels kept.  That equals the separable convolution with
 * the normalised taps and zero values outside the volume (conv::direct_convolution,
 * convolution_imp.hxx:761-812), which is what runs here, for every method of the filter
 * (direct passes whose halos hold zeros).  out may alias in. */
int tk_conv_acyclic(tk_ctx ctx, tk_filter flt, tk_vol out, tk_vol in);
int tk_conv_acyclic_pair(tk_ctx ctx, tk_filter flt, tk_vol out_a, tk_vol out_b, tk_vol in_a,
                         tk_vol in_b);

/* ---- R8 / R9: Dirichlet Gauss-Seidel Poisson sweeps -------------------------------
 * lapl::DirGaussSeidel (poisson_imp.hxx:340-429) behind
 * lapl::unsteady_poisson_functor<T>::operator() (poisson.hxx:90-96), and the
 * sweep loop of rsfe_splitbregman::apply (rsfe_splitbregman_imp.hxx:1762-1797).
 * TK_GS_LEX reproduces the reference's lexicographic order bit for bit (pipelined
 * tile wavefront); TK_GS_REDBLACK is the red-black ordering (FP32 fast mode). */
enum tk_gs_order { TK_GS_LEX = 0, TK_GS_REDBLACK = 1 };
int tk_poisson_dir_sweeps(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int n_sweeps);
/* runs at least ls_steps sweeps, then until the relative update
   ||phi_s - phi_{s-1}||_L2 / ||phi_{s-1}||_L2 <= ls_tol (reference rule); ls_tol == 0
   runs exactly ls_steps.  max_sweeps caps the loop (TK_ERR_SWEEPS if reached). */
int tk_poisson_dir_solve(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                         int ls_steps, double ls_tol, int max_sweeps, int* sweeps_out,
                         double* ls_err_out);

/* ---- lapl::NeuGaussSeidel (poisson_imp.hxx:115-334, 3-D branch) --------------------
 * One Gauss-Seidel sweep over every voxel with the derivative bc prescribed on the
 * boundary: per axis (p(+1) - n bc h)/h^2 on a low face, (p(-1) + n bc h)/h^2 on a high face,
 * n = 1, 1/sqrt 2, 1/sqrt 3 on faces, edges, corners, and htilde with 1/h^2 for each
 * boundary axis; FP64 is bitwise the reference's sweep.  Installed by set_bc(NEU)
 * (rsfe_splitbregman_imp.hxx:1294-1322).  order must be TK_GS_LEX (the reference's order;
 * TK_GS_REDBLACK returns TK_ERR_UNSUPPORTED); every edge >= 2 (the 2-D branch is not
 * built).  The solve variant runs the loop rule of tk_poisson_dir_solve. */
int tk_poisson_neu_sweeps(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int n_sweeps);
int tk_poisson_neu_solve(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                         int ls_steps, double ls_tol, int max_sweeps, int* sweeps, double* err);
/* build-defined residual (no reference counterpart):
   ||rhs - (phi/dt - lap_h phi) + phi/dt||_2 over the interior / ||rhs||_2, i.e.
   ||rhs + lap_h phi|| / ||rhs|| with FP64 accumulation */
int tk_poisson_residual(tk_ctx ctx, tk_vol phi, tk_vol rhs, double* rel_res);
/* out = rhs + lap_h phi on the interior, 0 on the faces (no reference counterpart) */
int tk_poisson_residual_field(tk_ctx ctx, tk_vol phi, tk_vol rhs, tk_vol out);
/* sweeps (in batches of check_every) until tk_poisson_residual <= res_tol, after setting
   phi's faces to bc (no reference counterpart: the reference gates on the relative update) */
int tk_poisson_solve_residual(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc,
                              int order, double res_tol, int check_every, int max_sweeps,
                              int* sweeps_out, double* res_out);
/* autotuner of the red-black sweep (no reference counterpart: the reference has one fixed
   lexicographic loop): times the 8-row and 16-row tile variants on rhs's shape and type,
   and 32 rows in FP32, keeps the fastest for every later red-black sweep of that shape on
   ctx, returns the rows in *tile_rows and the times in ms[0..2] (8, 16, 32 rows; ms[2] is
   0 in FP64; either pointer may be null); it also times the two-sweep launch with 8- and
   16-row tiles (ms[3..4], 0 where the volume has fewer than 3 interior planes or, for 16
   rows, in FP64; ms holds 5 doubles) and keeps the faster, or one launch per sweep where
   that is faster than the pairs.  All variants give identical results.  Runs
   synchronously; tk_rsfe_create calls it for red-black runs. */
int tk_poisson_rb_autotune(tk_ctx ctx, tk_vol rhs, int* tile_rows, double* ms);
/* red-black sweeps run two per launch where the volume allows (16-byte rows: z a multiple
   of 4 in FP32, of 2 in FP64; the result is bitwise that of one launch per sweep) and the
   shape's tile rows for the two-sweep launch were set (the autotuner, or
   tk_poisson_rb_set_tile); gated solves (ls_tol > 0) test the update norm of both sweeps
   and, where the first of a pair meets the gate, run that sweep again by itself, so they
   stop at the same sweep with the same field: enable 1 (the default) or 0 for one launch
   per sweep, on every later solve on ctx */
int tk_poisson_rb_set_pair(tk_ctx ctx, int enable);
/* set the red-black tile rows (8 or 16, or 32 in FP32) for rhs's shape and type on ctx
   (the one-sweep launch; the two-sweep launch takes 16 rows for 32) */
int tk_poisson_rb_set_tile(tk_ctx ctx, tk_vol rhs, int tile_rows);

/* ---- R6, R10-R13: pointwise and stencil operators ----------------------------- */
/* rsfe_splitbregman::set_heaviside (rsfe_splitbregman_imp.hxx:174-192) */
int tk_heaviside(tk_ctx ctx, tk_vol phi, double a0, double b0, double epsilon, tk_vol H);
/* image3d::grad (image3d_imp.hxx:743-888) */
int tk_grad(tk_ctx ctx, tk_vol f, tk_vol g0, tk_vol g1, tk_vol g2);
/* im3d::div (image3d_imp.hxx:947-1383), including its two boundary quirks */
int tk_div(tk_ctx ctx, tk_vol f0, tk_vol f1, tk_vol f2, tk_vol out);
/* rsfe_splitbregman::edge_detect, GRAD branch, edge_detector_sigma == 0 (imp:196-316) */
int tk_edge_detect(tk_ctx ctx, tk_vol u0, double beta, double lamda, double nu, tk_vol g);
/* rsfe_splitbregman::shrink (imp:343-365): d = sign(b) max(|b| - g, 0) */
int tk_shrink(tk_ctx ctx, tk_vol b, tk_vol g, tk_vol d);
/* The Split-Bregman update after the Poisson step (rsfe_splitbregman_imp.hxx:1800-1932):
   cut_phi (phi clamped to [a0, b0] in place, imp:320-339), d = grad phi, b += d,
   d = shrink(b, g), b -= d, d -= b (b[3] updated in place, d - b stored in q[3]) and
   divdb = div(q) (image3d_imp.hxx:947-1383, with its quirks); norms (may be NULL) receives
   sum (phi - phiold)^2 and sum phiold^2 in FP64 over every voxel (imp:1832 without the
   spacing factor).  The two passes of the run loop. */
int tk_sb_update(tk_ctx ctx, tk_vol phi, tk_vol phiold, tk_vol g, const tk_vol* b,
                 const tk_vol* q, tk_vol divdb, double a0, double b0, double* norms);
/* ---- R1: image3d elementwise operators ------------------------------------------
 * operator+= -= *= /= with a scalar (image3d_imp.hxx:187-198, 229-240, 266-277, 304-315): the
 * scalar is cast to the volume's type first (static_cast<T>), then every voxel is updated in
 * that type.  With an image (201-225, 243-267, 280-304, 318-342): b's voxel is cast to a's
 * type; dims must agree, else TK_ERR_SHAPE and a unchanged (the reference prints its WARNING
 * and does nothing; the C++ layer include/tk/image3d.hxx does the same).  Binary + - * /
 * (349-378) copy the left operand and apply these. */
enum tk_op { TK_OP_ADD = 0, TK_OP_SUB = 1, TK_OP_MUL = 2, TK_OP_DIV = 3 };
int tk_vol_op_scalar(tk_ctx ctx, tk_vol a, int op, double s);
int tk_vol_op_vol(tk_ctx ctx, tk_vol a, int op, tk_vol b);
/* element access, image3d::operator() (image3d.hxx:252, 262): one voxel read or written
   (synchronous); (i, j, k) outside the volume: TK_ERR_ARG */
int tk_vol_get(tk_vol v, int i, int j, int k, double* out);
int tk_vol_set(tk_vol v, int i, int j, int k, double value);
/* im3d::scalarprod (image3d_imp.hxx:385-407): sum a * static_cast<T_a>(b), FP64 accumulation
   in a fixed order (the reference's OpenMP reduction order varies with the thread count);
   scalarprod_L2 (411-426) multiplies by hx hy hz in the C++ layer */
int tk_scalarprod(tk_ctx ctx, tk_vol a, tk_vol b, double* out);
/* image3d::norm1 (1387-1398): sum |a|, FP64 accumulation; normL1 (1403-1406) times hx hy hz */
int tk_norm1(tk_ctx ctx, tk_vol a, double* out);
/* image3d::norminf (1427-1452) with its quirk: a running value starts at |a(0,0,0)| and takes
   |a_i| only when a_i exceeds it, so the result is max(|a_0|, max_i a_i): large negative
   entries are ignored */
int tk_norminf(tk_ctx ctx, tk_vol a, double* out);
/* norminf(a - b) without a temporary, and norminf(b) in the same pass (out_b may be NULL; b
   NULL: norminf(a)); the ratio (phi - phiold).norminf() / phiold.norminf() of the verbose
   iteration record (rsfe_splitbregman_imp.hxx:1835-1844) */
int tk_norminf_diff(tk_ctx ctx, tk_vol a, tk_vol b, double* out_diff, double* out_b);
/* im3d::vector_abs (image3d_imp.hxx:892-943): res = sqrt(f0^2 + f1^2 + f2^2) in the volumes'
   type, f2 NULL for two components; every volume of one dims and dtype (res pre-sized) */
int tk_vector_abs(tk_ctx ctx, tk_vol f0, tk_vol f1, tk_vol f2, tk_vol res);
/* image3d::normL2 (image3d_imp.hxx:1419-1423) with FP64 accumulation */
int tk_norm_l2(tk_ctx ctx, tk_vol a, double* out);
/* ||a - b||_L2 without a temporary (rsfe_splitbregman_imp.hxx:1783,1832) */
int tk_norm_l2_diff(tk_ctx ctx, tk_vol a, tk_vol b, double* out);
/* tk_norm_l2_diff for callers that cannot allocate (no reference counterpart: for code under
   HIP graph capture, serving loops and harnesses that forbid hipMalloc).  Enqueued on
   hip_stream (NULL: the default stream) of the current device; a and b are device arrays of
   n elements of dtype (b may be NULL); out_device[0] = ||a - b||_L2 and out_device[1] =
   ||b||_L2 (||a||_L2 and 0 without b), FP64, with the spacing factor hx hy hz.  Nothing is
   allocated: the partial sums live in a 256 KiB device array of the library, one per GPU,
   held until the process ends, so n is limited to 1 .. 134217728 voxels (TK_ERR_ARG beyond).
   Calls that overlap on one GPU share that array and race. */
int tk_norm_l2_diff_noalloc(void* hip_stream, const void* a, const void* b, long n,
                            tk_dtype dtype, double hx, double hy, double hz, double* out_device);
/* image3d::min / max (image3d_imp.hxx:1456-1509) */
int tk_minmax(tk_ctx ctx, tk_vol a, double* mn, double* mx);
/* image3d::change_range_of_intensity (image3d_imp.hxx:1627-1635) */
int tk_change_range_of_intensity(tk_ctx ctx, tk_vol a, double max, double min);
/* rsfe_splitbregman::cut_phi (imp:320-339) */
int tk_cut(tk_ctx ctx, tk_vol phi, double a0, double b0);

/* ---- R14: threshold and seeded connected component ------------------------------
 * image3d::im_to_black_and_white (image3d_imp.hxx:1591-1623) and
 * image3d::connected_component(res, i, j, k, threshold, full, binary=true)
 * (1694-1757 + sweep core 1821-2118).  out is TK_U8 (0/1).  full: 26- instead of
 * 6-connectivity.  passes_out (may be NULL) receives the number of device passes. */
int tk_im_to_black_and_white(tk_ctx ctx, tk_vol img, double threshold, int negative,
                             tk_vol out);
int tk_cc_seeded(tk_ctx ctx, tk_vol img, int i, int j, int k, double threshold, int full,
                 tk_vol out, int* passes_out);
/* full 6-connected component labelling of a U8 mask (labels are int32, 0 = background,
   each component labelled by 1 + its smallest linear voxel index); no reference
   counterpart beyond the seeded extraction above */
int tk_cc_label(tk_ctx ctx, tk_vol mask_u8, int32_t* labels_device, int* iters_out);

/* ---- R2 / R3 / R15: the Split-Bregman solver -------------------------------------
 * segm::rsfe_splitbregman<T> (rsfe_splitbregman.hxx:99-894): constructor defaults
 * (rsfe_splitbregman_imp.hxx:25-41) and apply(myim, init) (1547-1995) with
 * cv = cyclic convolution and onestep_poisson = DirGaussSeidel. */
typedef struct tk_rsfe_params {
    double sigma, lamda, lamda1, lamda2, beta, dt, epsilon, a0, b0, nu, tol, ls_tol;
    double radius;            /* Gaussian radius argument (3, imp:1641-1644) */
    int maxiter, ls_steps;
    int bc_dir;               /* 0: bc flag NEU -> boundary value 0; 1: DIR -> a0 (imp:1735-1743) */
    int gaussian_pixel_approach;
    int gs_order;             /* TK_GS_LEX / TK_GS_REDBLACK */
    int max_sweeps;           /* cap per Poisson solve when ls_tol > 0 */
    int warm_start;           /* 0: previous iterate (reference); R >= 1: up to R learned
                                 correction rounds (R15), network attached with
                                 tk_rsfe_set_warm_start */
    double ls_res_tol;        /* > 0: the Poisson step instead sweeps until the relative
                                 residual ||rhs + lap phi|| / ||rhs|| <= ls_res_tol (no
                                 reference counterpart; the learned-start evaluation) */
    int res_check_every;      /* sweeps between residual checks (default 64) */
    int ws_presweeps;         /* R15: sweeps before each learned round (default 512) */
    int acyclic;              /* 1: set_cv(acyclic_fftw_convolution), the reference's default
                                 cv (imp:25-41); 0 (default here): cyclic_fftw_convolution */
    int neumann;              /* 1: onestep_poisson = NeuGaussSeidel, as set_bc(NEU) installs
                                 (imp:1294-1322; bc value 0 unless bc_dir); lexicographic
                                 only, no warm_start / ls_res_tol.  0: DirGaussSeidel */
    /* automatic connected component (extract_connected_component, imp:447-491, called at
       1808-1812): when auto_extract_conn_comp and iteration t % auto_extract_conn_comp_freq
       == 0, the cut phi is replaced by its component through the voxel
       floor(cc_init_pixel / h) at relative threshold cc_threshold ((alpha - a0)/(b0 - a0),
       0.5 for the default alpha), binary or not (cc_binary_output), rescaled to [a0, b0];
       cc_init_variables then zeroes b and divdb in place of that iteration's Split-Bregman
       update (1849-1859).  A seed outside the volume prints the reference's WARNING. */
    int auto_extract_conn_comp, auto_extract_conn_comp_freq;
    double cc_init_pixel[3];
    double cc_threshold;
    int cc_binary_output, cc_init_variables;
    /* dumps (imp:1939-1951, 1968-1978): every dumpfrequency iterations (not at maxiter)
       phi is written to <outputname>_<t>.mhd/.raw, and when the run ends to
       <outputname>_<t>_final.mhd/.raw (tk_vol_write_mhd).  0: no dumps */
    int dumpfrequency;
    char outputname[256];
    double edge_detector_sigma; /* > 0: the edge detector (GRAD) reads the image convolved by
                                   the run's cv with a Gaussian of this sigma, radius 3
                                   (imp:214-227); 0: the image itself (default) */
    int ws_judge_sweeps;      /* R15: sweeps each candidate gets before the residuals are
                                 compared (default 16) */
    int ws_continue;          /* R15: 0 (default): a round whose candidate loses ends the
                                 rounds; 1: the rounds go on from the plain candidate (up to
                                 warm_start of them, or until ls_res_tol holds) */
    int ws_coarse_sweeps;     /* R15, network on a coarsened grid (tk_rsfe_set_warm_start):
                                 > 0: the coarse correction starts as that many sweeps from
                                 zero on the coarse problem, and the network predicts the rest
                                 from the coarse residual they leave (smooth, where the network
                                 is accurate); 0 (default): the network reads the restricted
                                 residual itself */
    int verbosity;            /* 1: the reference's verbose record (imp:1786-1793, 1835-1844)
                                 in tk_rsfe_iter_stats: ls_err of the last sweep also when
                                 ls_tol == 0 (one extra copy of phi per iteration) and
                                 tol_inf; 0 (default; timing runs) */
} tk_rsfe_params;

typedef struct tk_rsfe_iter_stats {
    int sweeps;               /* S_t: Gauss-Seidel sweeps in this outer iteration */
    double ls_err;            /* last relative update of the Poisson loop (-1 if not computed) */
    double tollerance;        /* ||phi - phiold||_L2 / ||phiold||_L2 (imp:1832) */
    int guess_used;           /* R15: learned rounds accepted */
    double res_plain;         /* R15: first round: residual of the previous iterate after 16 sweeps */
    double res_guess;         /* R15: first round: residual of the candidate after 16 sweeps */
    int cc_extracted;         /* 1: the automatic extraction replaced phi in this iteration */
    int coarse_sweeps;        /* R15: sweeps on the coarse grid (ws_coarse_sweeps per round) */
    double tol_inf;           /* verbosity: (phi - phiold).norminf() / phiold.norminf()
                                 (imp:1835-1844, the norminf quirk included); -1 otherwise */
} tk_rsfe_iter_stats;

int tk_rsfe_default_params(tk_rsfe_params* p);
/* set-up: filter spectrum, kones = K*1, kmyim = K*u0, g = edge_detect(u0) (imp:1641-1655);
   phi := initialize_phi_with_init(init) (imp:99-170): init of the image's dims is rescaled
   to [a0,b0]; init one voxel short along some axes has its last faces extended, then is
   rescaled; init NULL (apply(myim)) or of other dims (with the reference's WARNING on
   stderr) gives initialize_phi_with_cube (imp:48-95).  init may be F32 or F64. */
int tk_rsfe_create(tk_ctx ctx, tk_vol u0, tk_vol init, const tk_rsfe_params* p, tk_rsfe* out);
/* run up to n_iters outer iterations (stops early on tollerance <= tol);
   stats may be NULL, else has room for n_iters entries */
int tk_rsfe_iterate(tk_rsfe s, int n_iters, tk_rsfe_iter_stats* stats, int* n_done);
tk_vol tk_rsfe_phi(tk_rsfe s);
/* the right-hand side of the last outer iteration's Poisson step (valid until the next).
 * With tol <= 0 the loop test `tollerance > tol` fails only for a tollerance of exactly 0
 * (a stationary phi) or NaN (0/0: phi and phiold both zero everywhere).  A tk_rsfe_iterate
 * call that stops on either, with another iteration still due in the call, leaves in rhs
 * the forcing of the iteration it did not run, enqueued while the tollerance was read back;
 * phi, the statistics and the iteration count are those of the iterations that ran. */
tk_vol tk_rsfe_rhs(tk_rsfe s);
/* the on-the-go re-read (rsfe_splitbregman::update_param_onthego, imp:545-578, called at
   1816-1829 every showfrequency iterations): between two tk_rsfe_iterate calls, *p becomes
   the run's parameter set for the iterations that follow.  A changed sigma rebuilds the filter
   (radius 3) and kones, kmyim (imp:1683-1696); g is computed again when the previous
   edge_detector_sigma taken as a bool differs from the new value and the edge detector's
   inputs changed (imp:1877-1882); a0, b0 and bc_dir set the Poisson faces again.
   save_current writes phi as <outputname>_<t> (t: iterations so far), as does the
   dumpfrequency just set when due at t (imp:1939-1951); end_now ends the run (later
   tk_rsfe_iterate calls run nothing) and writes the final dump when dumpfrequency > 0.
   The reference re-reads between the cut and the Split-Bregman update of iteration t, so a
   changed edge detector reaches that update there; here, the next iteration's.  The
   sweep order, cv, the Neumann solver and the warm start cannot change
   (TK_ERR_UNSUPPORTED), nor can a batched run. */
int tk_rsfe_onthego(tk_rsfe s, const tk_rsfe_params* p, int save_current, int end_now);
/* R15: attach the learned warm start (created on the run's context for its dims).  The
   Poisson problem is linear, so the network that maps a right-hand side to its solution
   also maps the current residual r = rhs + lap phi to the correction.  The network is
   accurate on smooth residuals (those the sweeps leave) and poor on rough ones, so each
   round first runs ws_presweeps sweeps, then candidate = phi + network(r); candidate and phi
   each get ws_judge_sweeps sweeps and the one with the lower relative residual continues; a
   round whose candidate loses ends the rounds (or, with ws_continue, the next round starts
   from the plain candidate), then the Poisson step continues as configured.  All
   sweeps are counted in tk_rsfe_iter_stats.sweeps. */
int tk_rsfe_set_warm_start(tk_rsfe s, tk_unet net);
/* (continued) The network may also be created for the run's dims divided by one integer
   factor f on every axis (e.g. 64^3 for a 256^3 run, f = 4): it then serves as a learned
   coarse-grid correction.  Each round restricts the residual to the network's grid
   (tk_vol_restrict), predicts the correction there and adds its trilinear interpolation to
   phi (tk_vol_prolong_add, gain 1: the network's out_scale carries the factor f^2 by which
   the correction of the coarse problem exceeds that of a unit-spacing one).  The sweeps that
   follow remove the interpolation's rough part. */
/* the run's filter (FP32 runs autotune its per-axis methods at creation) */
tk_filter tk_rsfe_filter(tk_rsfe s);
int tk_rsfe_destroy(tk_rsfe s);

/* ---- R15 building blocks: BF16 3-D convolution network layers on the matrix cores ---
 * No reference counterpart (the learned Poisson warm start is an addition).  Device
 * pointers; activations are channels-last BF16 (uint16 bits): voxel v = (i*ny + j)*nz + k,
 * channel c at v*stride + offset + c.  3x3x3 weights are BF16 [co][kpad], k = tap*ci + c,
 * tap = (dx+1)*9 + (dy+1)*3 + (dz+1), zero-padded to kpad (multiple of 16, >= 27*ci).
 * Zero padding outside the volume.  FP32 accumulation (v_mfma_f32_16x16x16_bf16). */
int tk_nn_conv3d_fwd(tk_ctx ctx, const void* x, int xs, int xo, int ci, const void* w, int kpad,
                     const float* bias, void* y, int ys, int yo, int co, int nx, int ny, int nz,
                     int relu, int out_f32);
/* dw[co][kpad] = sum_v dy[v][o] x[v + tap(k)][c(k)] (FP32), db[co] = sum_v dy[v][o] (may be
   NULL); part: FP32 scratch of ceil(nvox / vchunk) * co * kpad (vchunk multiple of 16);
   deterministic (fixed-order chunk sums) */
int tk_nn_conv3d_wgrad(tk_ctx ctx, const void* x, int xs, int xo, int ci, const void* dy, int ds,
                       int dof, int co, int nx, int ny, int nz, int kpad, float* part,
                       int vchunk, float* dw, float* db);
/* weights of the input gradient: wt[c][kpad_t] = w flipped in space with ci and co exchanged,
   so dx = tk_nn_conv3d_fwd(dy, wt) */
int tk_nn_flip_weights(tk_ctx ctx, const void* w, int ci, int co, int kpad, void* wt, int kpad_t);
int tk_nn_relu_bwd(tk_ctx ctx, void* g, int gs, int go, const void* y, int ys, int yo, int C,
                   long n);
int tk_nn_avgpool2(tk_ctx ctx, const void* x, int xs, int xo, int C, int nx, int ny, int nz,
                   void* y, int ys, int yo);
int tk_nn_avgpool2_bwd(tk_ctx ctx, const void* dy, int ds, int dof, int C, int nx, int ny, int nz,
                       void* dx, int xs, int xo, int accum);
int tk_nn_upsample2(tk_ctx ctx, const void* x, int xs, int xo, int C, int nx, int ny, int nz,
                    void* y, int ys, int yo);
int tk_nn_upsample2_bwd(tk_ctx ctx, const void* dy, int ds, int dof, int C, int nx, int ny,
                        int nz, void* dx, int xs, int xo);
int tk_nn_add_slice(tk_ctx ctx, void* a, int as, int ao, const void* b, int bs, int bo, int C,
                    long n);
/* coarse = the mean of each f x f x f block of fine voxels (f = fine dims / coarse dims, one
   integer on every axis; coarse voxel c covers fine voxels f c .. f c + f - 1); F32 or F64,
   both volumes the same */
int tk_vol_restrict(tk_ctx ctx, tk_vol fine, tk_vol coarse);
/* fine += gain * the trilinear interpolation of coarse at the fine voxel centres (coarse
   values at their block centres, constant beyond the outermost centres) */
int tk_vol_prolong_add(tk_ctx ctx, tk_vol coarse, tk_vol fine, double gain);
int tk_nn_f32_to_bf16(tk_ctx ctx, const float* s, long n, void* d);
int tk_nn_bf16_to_f32(tk_ctx ctx, const void* s, long n, float* d);

/* ---- R15: the learned Poisson warm start (3-level 3-D U-Net, 16/32/64 channels) -------
 * No reference counterpart.  Forward, backward and Adam steps on the device (tk_nn_*
 * layers, BF16 weights and activations, FP32 accumulation and FP32 master weights).  The
 * network maps rhs / rms(rhs) to phi* / (rms(rhs) * out_scale) for volumes of the dims it
 * was created for (every edge a multiple of 4). */
int tk_unet_create(tk_ctx ctx, int nx, int ny, int nz, uint64_t seed, tk_unet* out);
int tk_unet_destroy(tk_unet net);
int tk_unet_num_params(tk_unet net, long* count);
int tk_unet_set_out_scale(tk_unet net, double s);
int tk_unet_get_out_scale(tk_unet net, double* s);
/* guess := network(rhs) in rhs's units (the Poisson initial guess) */
int tk_unet_predict(tk_unet net, tk_vol rhs, tk_vol guess);
/* one Adam step (lr > 0) on the pair (rhs, phi*); lr <= 0 computes the gradients only;
   loss (may be NULL) = mean squared error in normalised units */
int tk_unet_train_step(tk_unet net, tk_vol rhs, tk_vol target, double lr, double* loss);
/* which 0: FP32 master parameters, 1: last gradients; per layer w[co][27*ci] then b[co] */
int tk_unet_get_params(tk_unet net, int which, float* host, long count);
int tk_unet_set_params(tk_unet net, const float* host, long count);
/* the FP32 output of the last forward pass, normalised units */
int tk_unet_get_output(tk_unet net, float* host);

/* ---- batched uneven volumes (config: many independent runs of uneven sizes) --------
 * No reference counterpart (the reference's static filter state serves one size at a time,
 * convolution.hxx:307-314).  Each volume is one rsfe_splitbregman::apply run with its own
 * state.  With red-black sweeps, ls_tol == 0 and tol == 0 (and 16-byte rows), every stage
 * of an outer iteration is ONE launch over all volumes (each convolution pass per filter
 * variant, each sweep, the update with its tolerance sums, div), with results bitwise those
 * of each volume run alone; use_graph != 0 captures two outer iterations into one hipGraph
 * and replays it.  Lexicographic sweeps with ls_tol == 0 and tol == 0 enqueue each
 * volume's own launches over up to four lanes (a context and HIP stream each, volumes
 * assigned by longest-processing-time), captured the same way.  Otherwise each volume runs
 * the reference's host-checked loop in turn. */
typedef struct tk_batch_s* tk_batch;
int tk_batch_create(tk_ctx ctx, int n_vol, const int* dims3, tk_dtype dtype,
                    const tk_rsfe_params* p, tk_batch* out);
/* volume v's u0 and init are host arrays in the given dtype; init NULL: cube contour */
int tk_batch_set_volume(tk_batch b, int v, const void* u0, const void* init, tk_dtype host_dtype);
int tk_batch_iterate(tk_batch b, int n_iters, int use_graph);
int tk_batch_get_phi(tk_batch b, int v, void* host, tk_dtype host_dtype);
/* lanes in use, kernel nodes of the captured graph (0 if none), outer iterations run */
int tk_batch_info(tk_batch b, int* n_lanes, int* graph_nodes, int* iterations);
int tk_batch_destroy(tk_batch b);

/* ---- synthetic volumes (host functions; no reference counterpart: the reference
 * reads images from disk through VTK, interface_imp.hxx:33-88) -------------------- */
int tk_gen_ellipsoids(int nx, int ny, int nz, uint64_t seed, int n_ell, const double* semi_axes,
                      int margin, double sphere_radius, double* img, double* phi0);
/* the standard phantom: n_ell ellipsoids with semi-axes cycling (12,9,7),(8,8,10)
   times scale, margin 15*scale, phi0 sphere of radius 20*scale (<= 0.3125 min edge) */
int tk_gen_standard(int nx, int ny, int nz, uint64_t seed, int scale, int n_ell, double* img,
                    double* phi0);
int tk_gen_batch_dims(uint64_t seed, int n_vol, int* dims3, int* n_ell);
int tk_gen_ct(int nx, int ny, int nz, uint64_t seed, int n_vessels, double* img, int* seed_ijk);

#ifdef __cplusplus
}
#endif
#endif /* TK_H */
