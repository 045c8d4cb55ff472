/*
 * rsfe_oracle.h -- CPU FP64 restatement of the RSFE Split-Bregman path of
 * marco-fedele/3dimagetoolkit.  TEST INFRASTRUCTURE ONLY: the oracle is the
 * checker for the HIP library and the CPU baseline of the benchmark.  Product
 * code never links, loads or calls it.
 *
 * Layout: every volume is a dense double array, index (i,j,k) -> (i*ny + j)*nz + k
 * (z fastest), as image3d_imp.hxx:141-152.  Spacing (hx,hy,hz) as image3d.
 *
 * Each function cites the reference file:line it restates.  Parity pins for each
 * function are listed in the header of oracle/rsfe_oracle.cpp.
 */
#ifndef TK_RSFE_ORACLE_H
#define TK_RSFE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- R4: Gaussian time filter + cyclic real spectrum ---------------------- */
/* convolution_imp.hxx:22-91: filter dims Nh (odd), values exp(-r^2/(2 sigma^2)) / normL1 */
void or_gaussian_dims(int nz, double hx, double hy, double hz, double sigma,
                      double radius, int pixel_approach, int nh_out[3]);
void or_gaussian_time_filter(int nz, double hx, double hy, double hz, double sigma,
                             double radius, int pixel_approach, double* tf);
/* convolution_imp.hxx:240-264,407-522: cyclic spectrum dims (multiple of dims >= Nh) */
void or_cyclic_dims(int nx, int ny, int nz, const int nh[3], int nc_out[3]);
/* 8-block wrapped placement + forward 3-D FFT, real part kept */
void or_cyclic_spectrum(const double* tf, const int nh[3], const int nc[3], double* spec);

/* ---- R5: cyclic FFT convolution (convolution_imp.hxx:657-757) -------------- */
void or_cyclic_conv(const double* spec, const int nc[3], const double* in,
                    int nx, int ny, int nz, double hx, double hy, double hz, double* out);
/* closed form used to pin R5: direct circular convolution with the 3-D time filter */
void or_direct_cyclic_conv(const double* tf, const int nh[3], const double* in,
                           int nx, int ny, int nz, double* out);

/* plain 1-D complex DFT helpers (forward sign -1), used by the pins */
void or_fft_inplace(double* re, double* im, int n, int sign);
/* R5 acyclic (convolution_imp.hxx:244-246, 278-402, 535-653): padded dims dims + Nh - 1;
   the spectrum comes from or_cyclic_spectrum on those dims */
void or_acyclic_dims(int nx, int ny, int nz, const int nh[3], int na_out[3]);
void or_acyclic_conv(const double* spec, const int na[3], const double* in,
                     int dx, int dy, int dz, double hx, double hy, double hz, double* out);
/* closed form of the acyclic R5 (convolution_imp.hxx:761-812), zero outside the volume */
void or_direct_acyclic_conv(const double* tf, const int nh[3], const double* in,
                            int nx, int ny, int nz, double* out);

/* ---- R9: one Dirichlet Gauss-Seidel sweep (poisson_imp.hxx:340-429) -------- */
void or_dir_gauss_seidel(double* phi, const double* rhs, int nx, int ny, int nz,
                         double hx, double hy, double hz, double dt, double bc);

/* ---- lapl::NeuGaussSeidel, 3-D branch (poisson_imp.hxx:115-334): one lexicographic
   sweep over every voxel (k outer, i fastest, the reference's loop order) with the
   outward normal derivative bc prescribed on the boundary; every edge must be >= 2.  The
   2-D branch (Z == 1, poisson_imp.hxx:32-113) is not restated; the call returns -1. */
int or_neu_gauss_seidel(double* phi, const double* rhs, int nx, int ny, int nz,
                        double hx, double hy, double hz, double dt, double bc);

/* ---- R8: Poisson driver loop (rsfe_splitbregman_imp.hxx:1762-1797) ---------- */
/* returns the number of sweeps run; *ls_err_out = last relative-update error */
int or_poisson_loop(double* phi, const double* rhs, int nx, int ny, int nz,
                    double hx, double hy, double hz, double dt, double bc,
                    int ls_steps, double ls_tol, double* ls_err_out);
/* the same loop around onestep_poisson = NeuGaussSeidel (neumann != 0) or DirGaussSeidel */
int or_poisson_loop_bc(double* phi, const double* rhs, int nx, int ny, int nz,
                       double hx, double hy, double hz, double dt, double bc,
                       int ls_steps, double ls_tol, double* ls_err_out, int neumann);

/* ---- R6: Heaviside (rsfe_splitbregman_imp.hxx:174-192) ---------------------- */
void or_heaviside(const double* phi, long n, double a0, double b0, double eps, double* H);

/* ---- R12: image3d::grad (image3d_imp.hxx:743-888), edge_detect (196-316) ---- */
void or_grad(const double* f, int nx, int ny, int nz, double hx, double hy, double hz,
             double* g0, double* g1, double* g2);
void or_edge_detect(const double* u0, int nx, int ny, int nz, double hx, double hy,
                    double hz, double beta, double lamda, double nu, double* g);

/* ---- R13: im3d::div (image3d_imp.hxx:947-1383), shrink (rsfe 343-365) ------- */
void or_div(const double* f0, const double* f1, const double* f2, int nx, int ny, int nz,
            double hx, double hy, double hz, double* out);
void or_shrink(const double* b, const double* g, long n, double* d);

/* ---- R10: norms (image3d_imp.hxx:385-452,1387-1509) ------------------------- */
double or_scalarprod(const double* a, const double* b, long n);
double or_norm_l2(const double* a, long n, double hx, double hy, double hz);
double or_norm1(const double* a, long n);
double or_norminf(const double* a, long n);
double or_max(const double* a, long n);
double or_min(const double* a, long n);

/* ---- R3 / R14: range, b&w threshold, seeded connected component ------------- */
void or_change_range_of_intensity(double* a, long n, double max, double min);
void or_im_to_black_and_white(const double* a, long n, double threshold, int negative,
                              double* out);
/* image3d_imp.hxx:1694-1757 + sweep core 1821-2118; returns sweeps-until-stable */
int or_connected_component(const double* img, int nx, int ny, int nz, int i, int j, int k,
                           double threshold, int full_connected, double* res);

/* ---- R2: the whole apply() loop (rsfe_splitbregman_imp.hxx:1547-1995) ------- */
typedef struct or_params {
    double sigma, lamda, lamda1, lamda2, beta, dt, epsilon, a0, b0, nu, tol, ls_tol;
    double radius;          /* Gaussian radius argument (3 in apply, imp:1641-1644) */
    int maxiter, ls_steps;
    int bc_dir;             /* 0: flag NEU (boundary value 0), 1: DIR (value a0) */
    int gaussian_pixel_approach;
    int verbosity;          /* 1 adds the extra phi copy/norm of imp:1776-1793 (timing only) */
    int acyclic;            /* cv: 0 cyclic_fftw_convolution, 1 acyclic_fftw_convolution */
    int neumann;            /* onestep_poisson: 0 DirGaussSeidel (the ctor's), 1 NeuGaussSeidel
                               (set_bc(NEU), rsfe_splitbregman_imp.hxx:1294-1322) */
    /* automatic connected-component extraction (imp:447-491, called at 1808-1812):
       every auto_cc_freq iterations when auto_cc, phi := its component through the voxel at
       the physical point cc_pixel, rescaled to [a0,b0]; cc_init_variables then zeroes b
       and divdb in place of that iteration's Split-Bregman update (1849-1859) */
    int auto_cc, auto_cc_freq;
    double cc_pixel[3];
    int cc_binary_output, cc_init_variables;
    double edge_detector_sigma; /* != 0: g from the image blurred by cv with this sigma
                                   (rsfe_splitbregman_imp.hxx:214-227) */
} or_params;

typedef struct or_stats {
    int iters;              /* outer iterations run */
    int cc_extractions;     /* iterations whose phi the automatic extraction replaced */
    double final_tollerance;
    double t_setup, t_loop, t_fft;   /* seconds */
} or_stats;

void or_default_params(or_params* p);
/* R3: initialize_phi_with_init (rsfe_splitbregman_imp.hxx:99-170) for an init of dims
   (ix,iy,iz) (NULL: initialize_phi_with_cube, 48-95, as apply(myim)); phi has the image's
   dims (X,Y,Z).  Returns 1 for the init cases, 0 when the cube was used. */
int or_init_phi(const double* init, int ix, int iy, int iz, int X, int Y, int Z, double a0,
                double b0, double* phi);
/* R6 + R7: forcing term rhs of one outer iteration for given phi and divdb */
void or_rsfe_forcing(const double* myim, const double* phi, const double* divdb, int nx, int ny,
                     int nz, double hx, double hy, double hz, const or_params* p, double* rhs);
/* sweeps_per_iter (length >= maxiter, may be NULL) receives S_t;
   tol_per_iter (may be NULL) receives the outer relative change */
int or_rsfe_apply(const double* u0, const double* init, int nx, int ny, int nz,
                  double hx, double hy, double hz, const or_params* p, double* phi_out,
                  int* sweeps_per_iter, double* tol_per_iter, or_stats* st);

/* the same loop; phi after outer iteration ckpt_iters[c] (1-based) is copied to
   ckpt_phi + c * nx*ny*nz for each of the n_ckpt checkpoints */
int or_rsfe_apply_ckpt(const double* u0, const double* init, int nx, int ny, int nz,
                       double hx, double hy, double hz, const or_params* p, double* phi_out,
                       int* sweeps_per_iter, double* tol_per_iter, or_stats* st, int n_ckpt,
                       const int* ckpt_iters, double* ckpt_phi);

/* the loop with the reference's on-the-go re-read (update_param_onthego,
   rsfe_splitbregman_imp.hxx:545-578, called at 1816-1829) once: after iteration t_switch's
   cut and connected component the parameters become *p2, as if section
   splitbregman/onthego/ of the parameter file held them; a changed sigma rebuilds the filter
   (radius 3) and kones / kmyim at the next iteration's start (1683-1696); end_now stops the
   loop after iteration t_switch */
int or_rsfe_apply_onthego(const double* myim, const double* init, int nx, int ny, int nz,
                          double hx, double hy, double hz, const or_params* p1, int t_switch,
                          const or_params* p2, int end_now, double* phi_out,
                          int* sweeps_per_iter, double* tol_per_iter, or_stats* st);

/* time n outer iterations with a fixed sweep count per iteration (CPU baseline) */
int or_rsfe_time_iters(const double* u0, int nx, int ny, int nz, const or_params* p,
                       int n_iters, double* phi_inout, or_stats* st);

int or_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
