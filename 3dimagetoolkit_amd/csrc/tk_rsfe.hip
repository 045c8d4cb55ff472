// tk_rsfe.hip -- the RSFE Split-Bregman outer loop on the device
// (segm::rsfe_splitbregman<T>::apply, rsfe_splitbregman_imp.hxx:1547-1995).
//
// One outer iteration = 5 convolution passes + the Poisson sweeps + 2 stencil passes:
//   x1  load phi,u0 -> (H*u0, H)         [Heaviside prologue, phiold copy]    imp:1667,1700
//   y1  complex -> complex
//   z   conv z of pair 1, then region fit f1,f2 -> (l1 f1 - l2 f2, l1 f1^2 - l2 f2^2),
//       then conv z of pair 2, in LDS                                          imp:1703-1731
//   y2  complex -> complex
//   x2  -> forcing term rhs                [epilogue]                         imp:1750-1759
//   Poisson sweeps (lexicographic or red-black)                              imp:1762-1797
//   sb  cut_phi + ||phi-phiold|| + grad/shrink/Bregman update                imp:1800-1906
//   div divdb = div(d - b)                                                   imp:1932
// Convolution passes commute (the Gaussian is separable), which is what lets pair 2
// start on the axis where pair 1 ends.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <utility>
#include <vector>

#include "tk_device.hpp"
#include "tk_fft_kernel.hpp"
#include "tk_stencil.hpp"

namespace tk {

template <class T>
int launch_div(tk_ctx, tk_vol, tk_vol, tk_vol, tk_vol);
int stencil_batch_grid(tk_dtype dt, int X, int Y, int Z, int* gy, int* nparts);
int launch_sb_batch(tk_ctx ctx, tk_dtype dt, const SbJob* jobs, int n, unsigned nblocks);
int launch_div_batch(tk_ctx ctx, tk_dtype dt, const DivJob* jobs, int n, unsigned nblocks);
template <class T>
int launch_edge(tk_ctx, tk_vol, double, double, double, tk_vol);
template <class T>
int launch_sb_update(tk_ctx, tk_vol, tk_vol, tk_vol*, tk_vol, tk_vol*, double, double, int*,
                     double*);
template <class T>
int launch_sb_div(tk_ctx, tk_vol, tk_vol, tk_vol, tk_vol*, tk_vol*, tk_vol, double, double, int*,
                  double*);

template <class T>
__device__ __forceinline__ T heaviside(T p, T center, T eps)
{
    if constexpr (std::is_same<T, double>::value)
        return 0.5 * (1. + 2. / M_PI * atan((p - center) / eps));
    else
        return 0.5f * (1.f + (float)(2. / M_PI) * atanf(fdiv_nr(p - center, eps)));
}

template <class T>
struct LdHeaviside {
    const T* phi;
    const T* u0;
    T* phiold;
    T center, eps;
    __device__ cplx<T> operator()(long v) const
    {
        const T p = phi[v];
        phiold[v] = p;
        const T H = heaviside(p, center, eps);
        return {H * u0[v], H};
    }
    __device__ void ld4(long v, cplx<T>* c) const
    {
        const Quad<T> p = ldq(phi + v), u = ldq(u0 + v);
        stq(phiold + v, p);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const T H = heaviside(p.v[e], center, eps);
            c[e] = {H * u.v[e], H};
        }
    }
};

// imp:1709-1731 on (num, den) -> (kr2 factor, kr1 factor)
template <class T>
struct MdRegionFit {
    const T* kmyim;
    const T* kones;
    T l1, l2;
    __device__ cplx<T> operator()(long v, cplx<T> c) const { return fit(c, kmyim[v], kones[v]); }
    __device__ void md4(long v, cplx<T>* c) const
    {
        const Quad<T> km = ldq(kmyim + v), ko = ldq(kones + v);
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = fit(c[e], km.v[e], ko.v[e]);
    }
    // FP64: IEEE quotients (the reference's arithmetic); FP32: div_nr (this pass is bound by
    // its vector ALUs; the divisors den = K*H in (0, 1] and den - K*1 in [-1, 0] are normal
    // numbers or zero)
    static __device__ __forceinline__ T quot(T n, T d) { return div_nr(n, d); }
    __device__ cplx<T> fit(cplx<T> c, T km, T ko) const
    {
        const T num = c.x, den = c.y;
        T f1 = quot(num, den);
        T nm = num, dn = den;
        nm -= km;
        dn -= ko;
        nm = quot(nm, dn);
        T f2 = nm;
        const T a = f1 * l1 - f2 * l2;
        f1 *= f1;
        f1 *= l1;
        f2 *= f2;
        f2 *= l2;
        f1 -= f2;
        return {a, f1};
    }
};

// imp:1750-1759 on (kr2, kr1)
template <class T>
struct StRhs {
    const T* u0;
    const T* kones;
    const T* divdb;
    T* rhs;
    T l12, lam;
    __device__ void operator()(long v, cplx<T> c) const
    {
        rhs[v] = forcing(c, u0[v], kones[v], divdb[v]);
    }
    __device__ void st4(long v, const cplx<T>* c) const
    {
        const Quad<T> u = ldq(u0 + v), ko = ldq(kones + v), dd = ldq(divdb + v);
        Quad<T> r;
#pragma unroll
        for (int e = 0; e < 4; ++e) r.v[e] = forcing(c[e], u.v[e], ko.v[e], dd.v[e]);
        stq(rhs + v, r);
    }
    __device__ T forcing(cplx<T> c, T u, T ko, T dd) const
    {
        T r = c.x;
        r *= u;
        r *= (T)2;
        T aux = l12;
        aux *= u;
        aux *= u;
        aux *= ko;
        r -= aux;
        r -= c.y;
        r = div_nr(r, lam);
        r -= dd;
        return r;
    }
};

// initialize_phi_with_init, 2nd case (rsfe_splitbregman_imp.hxx:117-160): init is one voxel
// short along some axes; the reference copies init, then the last-but-one face onto the last
// face, axis by axis, which equals reading init at the clamped index
template <class T, class S>
__global__ void k_init_extend(T* __restrict__ phi, const S* __restrict__ init, int X, int Y,
                              int Z, int ix, int iy, int iz)
{
    const long n = (long)X * Y * Z;
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const int k = (int)(v % Z);
        const long r = v / Z;
        const int j = (int)(r % Y), i = (int)(r / Y);
        const int si = i < ix ? i : ix - 1, sj = j < iy ? j : iy - 1, sk = k < iz ? k : iz - 1;
        phi[v] = (T)init[((long)si * iy + sj) * iz + sk];
    }
}

// initialize_phi_with_cube (rsfe_splitbregman_imp.hxx:48-95): a0 everywhere, b0 on
// [ix,fx) x [iy,fy) x [iz,fz)
template <class T>
__global__ void k_init_cube(T* __restrict__ phi, int X, int Y, int Z, int ix, int fx, int iy,
                            int fy, int iz, int fz, T a0, T b0)
{
    const long n = (long)X * Y * Z;
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const int k = (int)(v % Z);
        const long r = v / Z;
        const int j = (int)(r % Y), i = (int)(r / Y);
        const bool in = i >= ix && i < fx && j >= iy && j < fy && k >= iz && k < fz;
        phi[v] = in ? b0 : a0;
    }
}

template <class T>
__global__ void k_add_inplace(T* __restrict__ a, const T* __restrict__ b, long n)
{
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) a[v] += b[v];
}

// sum (phi - phiold)^2 and sum phiold^2 per block (the tollerance of imp:1832 when the
// Split-Bregman pass, which forms them otherwise, does not run)
template <class T>
__global__ __launch_bounds__(256) void k_tol_partial(const T* __restrict__ phi,
                                                     const T* __restrict__ old, long n,
                                                     double* __restrict__ part)
{
    double s1 = 0, s2 = 0;
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const double o = (double)old[v];
        const double d = (double)phi[v] - o;
        s1 += d * d;
        s2 += o * o;
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

// the non-binary output of connected_component (image3d_imp.hxx:1806-1816):
// phi := ((cc + negative bw) * (phi - min)) + min, in T
template <class T>
__global__ void k_cc_values(T* __restrict__ phi, const uint8_t* __restrict__ cc,
                            const uint8_t* __restrict__ neg, T mn, long n)
{
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        T r = (T)cc[v];
        r += (T)neg[v];
        r *= phi[v] - mn;
        phi[v] = r + mn;
    }
}

__global__ void k_sum_pairs_rsfe(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    double s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        s1 += part[2 * i];
        s2 += part[2 * i + 1];
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        out[0] = s1;
        out[1] = s2;
    }
}

}  // namespace tk

using namespace tk;

struct tk_rsfe_s {
    tk_ctx ctx = nullptr;
    tk_rsfe_params p{};
    tk_dtype dtype = TK_F64;
    int X = 0, Y = 0, Z = 0;
    double h[3] = {1, 1, 1};
    tk_filter flt = nullptr;
    tk_vol u0 = nullptr, phi = nullptr, phiold = nullptr, rhs = nullptr, kones = nullptr,
           kmyim = nullptr, g = nullptr, divdb = nullptr;
    tk_vol b[3] = {nullptr, nullptr, nullptr}, q[3] = {nullptr, nullptr, nullptr};
    bool in_batch = false;     // a batch's tables hold b and q: no fused update (it swaps them)
    void* C = nullptr;
    void* alt = nullptr;       // red-black ping-pong partner of phi (this run's own)
    tk_unet net = nullptr;     // learned warm start (same context), or none
    tk_vol guess = nullptr;    // its guess
    tk_vol resid = nullptr;    // the residual field it reads
    int ws_f = 1;              // the network runs on the run's grid coarsened ws_f times
    tk_vol cresid = nullptr;   // ws_f > 1: the residual restricted to the network's grid
    tk_vol cguess = nullptr;   // ws_f > 1: the network's correction on its grid
    tk_vol cnet = nullptr;     // ws_f > 1: the coarse residual the network reads
    tk_vol vaux = nullptr;     // verbosity: phi before the last sweep (aux, imp:1774-1777)
    int it_coarse = 0;         // coarse sweeps of the current iteration
    double it_tol_inf = -1;    // verbosity: norminf ratio of the current iteration
    tk_vol ccmask = nullptr;   // U8: the automatic extraction's black-and-white / component
    tk_vol ccneg = nullptr;    // U8: its negative black-and-white image (non-binary output)
    double* tol_dev = nullptr; // device: sum (phi-phiold)^2, sum phiold^2 of the last iteration
    hipEvent_t tol_ev = nullptr; // after the read-back of tol_dev, when work follows it
    double* h_tol = nullptr;     // the device's address of the context's pinned h_acc
    int iter = 0;
    double tollerance = 1000;
    // phi's and alt's faces hold the Poisson boundary value: set by the first solve; the
    // sweeps never write the i and j faces, copy the k faces, and cut_phi keeps a value in
    // [a0, b0] (bc is 0 or a0), so they stay set for the run
    bool faces_set = false;
    bool final_dumped = false;  // the end-of-run dump has been written
    bool ended = false;         // end_now set on the go (tk_rsfe_onthego): no more iterations
    // edge_detector_sigma at the last g, held in a bool as the reference does (imp:1656,
    // 1877-1882)
    bool old_eds = false;
};

namespace {

// a += b (same dims and type)
template <class T>
int launch_add(tk_ctx ctx, tk_vol a, tk_vol b)
{
    const long n = a->n();
    hipLaunchKernelGGL(k_add_inplace<T>, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, ctx->stream,
                       (T*)a->d, (const T*)b->d, n);
    TK_LAUNCHED();
    return TK_OK;
}

// extract_connected_component, automatic branch (rsfe_splitbregman_imp.hxx:447-491), on the
// cut phi: seed floor(cc_pixel / h); outside the volume -> the reference's WARNING, no
// change; a seed that is black at threshold (alpha - a0)/(b0 - a0) -> no change; else
// phi := connected_component(seed, 0.5, 6-connected, cc_binary_output) rescaled to
// [a0, b0].  Reads one voxel of the seed test back to the host (waits on the stream).
template <class T>
int auto_extract_cc(tk_rsfe s, int* extracted)
{
    *extracted = 0;
    tk_ctx ctx = s->ctx;
    const tk_rsfe_params& p = s->p;
    const int i = (int)std::floor(p.cc_init_pixel[0] / s->h[0]),
              j = (int)std::floor(p.cc_init_pixel[1] / s->h[1]),
              k = (int)std::floor(p.cc_init_pixel[2] / s->h[2]);
    if (i < 0 || j < 0 || k < 0 || i > s->X - 1 || j > s->Y - 1 || k > s->Z - 1) {
        std::fprintf(stderr, "WARNING::apply::extract_connected_component: coordinates outside "
                             "allowed range\n");
        return TK_OK;
    }
    if (!s->ccmask)
        TK_TRY(tk_vol_alloc(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], TK_U8, &s->ccmask));
    TK_TRY(tk_im_to_black_and_white(ctx, s->phi, p.cc_threshold, 0, s->ccmask));
    uint8_t seed_bw = 0;
    const long v = ((long)i * s->Y + j) * s->Z + k;
    TK_HIP(hipMemcpyAsync(&seed_bw, (const uint8_t*)s->ccmask->d + v, 1, hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    if (seed_bw == 0) return TK_OK;
    double mn = 0, mx = 0;
    if (!p.cc_binary_output) {
        if (!s->ccneg)
            TK_TRY(tk_vol_alloc(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], TK_U8,
                                &s->ccneg));
        TK_TRY(tk_im_to_black_and_white(ctx, s->phi, 0.5, 1, s->ccneg));
        TK_TRY(tk_minmax(ctx, s->phi, &mn, &mx));
    }
    TK_TRY(tk_cc_seeded(ctx, s->phi, i, j, k, 0.5, 0, s->ccmask, nullptr));
    if (p.cc_binary_output) {
        TK_TRY(tk_vol_copy(s->phi, s->ccmask));
    } else {
        const long n = s->phi->n();
        hipLaunchKernelGGL(k_cc_values<T>, dim3(blocks_for(n, 256, 8192)), dim3(256), 0,
                           ctx->stream, (T*)s->phi->d, (const uint8_t*)s->ccmask->d,
                           (const uint8_t*)s->ccneg->d, (T)mn, n);
        TK_LAUNCHED();
    }
    TK_TRY(tk_change_range_of_intensity(ctx, s->phi, p.b0, p.a0));
    *extracted = 1;
    return TK_OK;
}

// One outer iteration.  sync: read the tollerance back and wait (the reference's loop
// condition needs it); without sync nothing waits on the host and nothing but kernels is
// enqueued, so the iteration can be captured into a graph (the caller reads tol_dev later).
template <class T>
int rsfe_forcing(tk_rsfe s)
{
    tk_ctx ctx = s->ctx;
    const tk_rsfe_params& p = s->p;
    cplx<T>* C = (cplx<T>*)s->C;
    const tk_filter_s* f = s->flt;
    const bool acy = p.acyclic != 0;
    // conv pair 1: x, y, then z with the region fit in the middle
    LdHeaviside<T> ld{(const T*)s->phi->d, (const T*)s->u0->d, (T*)s->phiold->d,
                      (T)((p.a0 + p.b0) / 2.), (T)p.epsilon};
    {
        TimedLaunch tl(ctx, TK_K_CONV_X1);
        TK_TRY((launch_conv_axis<T>(ctx, f, 0, 1, ld, MdNone{}, StComplex<T>{C}, acy)));
    }
    {
        TimedLaunch tl(ctx, TK_K_CONV_Y);
        TK_TRY((launch_conv_axis<T>(ctx, f, 1, 1, LdComplex<T>{C}, MdNone{}, StComplex<T>{C},
                                    acy)));
    }
    MdRegionFit<T> md{(const T*)s->kmyim->d, (const T*)s->kones->d, (T)p.lamda1, (T)p.lamda2};
    {
        TimedLaunch tl(ctx, TK_K_CONV_Z2);
        TK_TRY((launch_conv_axis<T>(ctx, f, 2, 2, LdComplex<T>{C}, md, StComplex<T>{C}, acy)));
    }
    {
        TimedLaunch tl(ctx, TK_K_CONV_Y);
        TK_TRY((launch_conv_axis<T>(ctx, f, 1, 1, LdComplex<T>{C}, MdNone{}, StComplex<T>{C},
                                    acy)));
    }
    StRhs<T> sr{(const T*)s->u0->d, (const T*)s->kones->d, (const T*)s->divdb->d, (T*)s->rhs->d,
                (T)(p.lamda1 - p.lamda2), (T)p.lamda};
    {
        TimedLaunch tl(ctx, TK_K_CONV_X2);
        TK_TRY((launch_conv_axis<T>(ctx, f, 0, 1, LdComplex<T>{C}, MdNone{}, sr, acy)));
    }
    return TK_OK;
}

// after the Poisson step: automatic connected component, cut + tolerance sums + the
// Split-Bregman update and div (or the reset of b and divdb), iteration count
template <class T>
int rsfe_update(tk_rsfe s, int* extracted_out);
int rsfe_finish(tk_rsfe s, tk_rsfe_iter_stats* st, int sweeps, double ls_err, int used,
                double res_plain, double res_guess, int extracted, int (*next)(tk_rsfe) = nullptr);

// ahead (may be null): in, the forcing passes of this iteration are already enqueued; out,
// those of the next one are (enqueued when `speculate`, between the read-back of this
// iteration's tollerance and the wait for it: they write only phiold, the convolution
// buffer and rhs, so a loop that then stops has lost nothing but their time)
template <class T>
int rsfe_iteration(tk_rsfe s, tk_rsfe_iter_stats* st, bool sync = true, bool* ahead = nullptr,
                   bool speculate = false)
{
    tk_ctx ctx = s->ctx;
    const tk_rsfe_params& p = s->p;
    const bool ready = ahead && *ahead;
    s->it_coarse = 0;
    if (ahead) *ahead = false;
    if (!ready) TK_TRY(rsfe_forcing<T>(s));
    // Poisson loop (rsfe_splitbregman_imp.hxx:1735-1797): bc value, then the sweeps from
    // the previous iterate (or from the learned guess when its residual is lower)
    const double chosen_bc = p.bc_dir ? p.a0 : 0.;
    int sweeps = 0, used = 0;
    double ls_err = -1, res_plain = -1, res_guess = -1;
    int judge = 0;
    if (s->net) {
        // The residual weights high frequencies (the Laplacian amplifies them), which a few
        // sweeps remove, while the sweep count is set by the smooth error: both candidates
        // get kJudge sweeps before their residuals are compared; all of them are counted.
        const int kJudge = p.ws_judge_sweeps;
        const int pre = p.ws_presweeps > 0 ? p.ws_presweeps : 0;
        for (int round = 0; round < p.warm_start; ++round) {
            if (pre > 0) {
                TK_TRY(poisson_dir_solve_alt(ctx, s->phi, s->rhs, p.dt, chosen_bc, p.gs_order,
                                             pre, 0, 0, s->alt ? &s->alt : nullptr, nullptr,
                                             nullptr));
                judge += pre;
            }
            TK_TRY(poisson_set_boundary(ctx, s->phi, chosen_bc));
            TK_TRY(tk_poisson_residual_field(ctx, s->phi, s->rhs, s->resid));
            if (s->ws_f > 1) {
                // learned coarse-grid correction: the residual averaged over ws_f^3 blocks,
                // the network's correction on that grid, interpolated back onto phi
                TK_TRY(tk_vol_restrict(ctx, s->resid, s->cresid));
                if (p.ws_coarse_sweeps > 0) {
                    // sweeps from zero on the coarse problem (a dt of 1e30 relaxes it with
                    // omega 1; the sweep's fixed point does not depend on dt), then the
                    // network's estimate of what they leave, read from their residual
                    TK_TRY(tk_vol_fill(s->cguess, 0.0));
                    TK_TRY(tk_poisson_dir_sweeps(ctx, s->cguess, s->cresid, 1e30, 0.0,
                                                 TK_GS_REDBLACK, p.ws_coarse_sweeps));
                    s->it_coarse += p.ws_coarse_sweeps;
                    TK_TRY(tk_poisson_residual_field(ctx, s->cguess, s->cresid, s->cnet));
                    TK_TRY(tk_unet_predict(s->net, s->cnet, s->cnet));
                    TK_TRY(launch_add<T>(ctx, s->cguess, s->cnet));
                } else {
                    TK_TRY(tk_unet_predict(s->net, s->cresid, s->cguess));
                }
                TK_TRY(tk_vol_copy(s->guess, s->phi));
                TK_TRY(tk_vol_prolong_add(ctx, s->cguess, s->guess, 1.0));
            } else {
                TK_TRY(tk_unet_predict(s->net, s->resid, s->guess));
                TK_TRY(launch_add<T>(ctx, s->guess, s->phi));
            }
            TK_TRY(poisson_dir_solve_alt(ctx, s->phi, s->rhs, p.dt, chosen_bc, p.gs_order, kJudge,
                                         0, 0, s->alt ? &s->alt : nullptr, nullptr, nullptr));
            TK_TRY(poisson_dir_solve_alt(ctx, s->guess, s->rhs, p.dt, chosen_bc, p.gs_order,
                                         kJudge, 0, 0, nullptr, nullptr, nullptr));
            double rp = 0, rg = 0;
            TK_TRY(tk_poisson_residual(ctx, s->phi, s->rhs, &rp));
            TK_TRY(tk_poisson_residual(ctx, s->guess, s->rhs, &rg));
            if (round == 0) {
                res_plain = rp;
                res_guess = rg;
            }
            judge += 2 * kJudge;
            if (!(rg < rp)) {
                // the plain candidate continues; with ws_continue the next round proposes
                // again from it
                if (!p.ws_continue) break;
                if (p.ls_res_tol > 0 && rp <= p.ls_res_tol) break;
                continue;
            }
            TK_TRY(tk_vol_copy(s->phi, s->guess));
            ++used;
            // the ls_res_tol gate may already hold
            if (p.ls_res_tol > 0 && rg <= p.ls_res_tol) break;
        }
    }
    // the fixed-count sweeps (ls_tol == 0) in two calls when the verbose record wants the
    // relative update of the last sweep: aux = phi before it (imp:1774-1777, 1786-1793)
    const bool last_err = p.verbosity && p.ls_tol == 0 && p.ls_res_tol <= 0 && p.ls_steps >= 1;
    auto plain = [&](int steps, int* sw, double* err) {
        if (p.neumann)
            return poisson_neu_solve_alt(ctx, s->phi, s->rhs, p.dt, chosen_bc, p.gs_order, steps,
                                         p.ls_tol, p.max_sweeps, sw, err);
        return poisson_dir_solve_alt(ctx, s->phi, s->rhs, p.dt, chosen_bc, p.gs_order, steps,
                                     p.ls_tol, p.max_sweeps, s->alt ? &s->alt : nullptr, sw, err,
                                     s->faces_set);
    };
    int r;
    if (p.ls_res_tol > 0 && !p.neumann)
        r = poisson_solve_residual_alt(ctx, s->phi, s->rhs, p.dt, chosen_bc, p.gs_order,
                                       p.ls_res_tol, p.res_check_every > 0 ? p.res_check_every : 64,
                                       p.max_sweeps, s->alt ? &s->alt : nullptr, &sweeps, &ls_err);
    else if (last_err) {
        int s1 = 0, s2 = 0;
        double e1 = -1, e2 = -1;
        r = p.ls_steps > 1 ? plain(p.ls_steps - 1, &s1, &e1) : TK_OK;
        if (r == TK_OK) {
            if (!s->vaux)
                TK_TRY(tk_vol_alloc(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], s->dtype,
                                    &s->vaux));
            TK_TRY(tk_vol_copy(s->vaux, s->phi));
            r = plain(1, &s2, &e2);
        }
        sweeps = s1 + s2;
        if (r == TK_OK) {
            double dn = 0, an = 0;
            TK_TRY(tk_norm_l2_diff(ctx, s->phi, s->vaux, &dn));
            TK_TRY(tk_norm_l2(ctx, s->vaux, &an));
            ls_err = dn / an;
        }
    } else
        r = plain(p.ls_steps, &sweeps, &ls_err);
    if (r != TK_OK && r != TK_ERR_SWEEPS) return r;
    if (!p.neumann && chosen_bc >= p.a0 && chosen_bc <= p.b0) s->faces_set = true;
    int extracted = 0;
    TK_TRY(rsfe_update<T>(s, &extracted));
    s->it_tol_inf = -1;
    if (p.verbosity) {
        // (phi - phiold).norminf() / phiold.norminf() after cut_phi (imp:1835-1844)
        double dinf = 0, oinf = 0;
        TK_TRY(tk_norminf_diff(ctx, s->phi, s->phiold, &dinf, &oinf));
        s->it_tol_inf = dinf / oinf;
    }
    if (!sync) return TK_OK;
    const bool next = ahead && speculate;
    TK_TRY(rsfe_finish(s, st, sweeps + judge, ls_err, used, res_plain, res_guess, extracted,
                       next ? &rsfe_forcing<T> : nullptr));
    if (next) *ahead = true;
    return TK_OK;
}

template <class T>
int rsfe_update(tk_rsfe s, int* extracted_out)
{
    tk_ctx ctx = s->ctx;
    const tk_rsfe_params& p = s->p;
    // automatic connected component (imp:1808-1812) on the cut phi; the extracted phi is
    // not cut again, and with cc_init_variables b and divdb restart from zero in place of
    // this iteration's Split-Bregman update (imp:1849-1859)
    int extracted = 0;
    double lo = p.a0, hi = p.b0;
    if (p.auto_extract_conn_comp && p.auto_extract_conn_comp_freq > 0 &&
        (s->iter + 1) % p.auto_extract_conn_comp_freq == 0) {
        TK_TRY(tk_cut(ctx, s->phi, p.a0, p.b0));
        TK_TRY(auto_extract_cc<T>(s, &extracted));
        if (extracted) {
            s->faces_set = false;  // the component may reach the faces
            lo = -std::numeric_limits<double>::infinity();
            hi = std::numeric_limits<double>::infinity();
        }
    }
    int nb = 0;
    double* part = ctx->acc + 64;
    if (extracted && p.cc_init_variables) {
        const long n = s->phi->n();
        nb = (int)blocks_for(n, 256, 1024);
        hipLaunchKernelGGL(k_tol_partial<T>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const T*)s->phi->d, (const T*)s->phiold->d, n, part);
        TK_LAUNCHED();
        hipLaunchKernelGGL(k_sum_pairs_rsfe, dim3(1), dim3(1024), 0, ctx->stream, part, nb,
                           s->tol_dev);
        TK_LAUNCHED();
        for (int a = 0; a < 3; ++a) TK_TRY(tk_vol_fill(s->b[a], 0.));
        TK_TRY(tk_vol_fill(s->divdb, 0.));
    } else if (ctx->sb_fused && !s->in_batch && s->Z % (16 / (int)sizeof(T)) == 0 &&
               s->X >= 3 && s->Y >= 3 && s->Z >= 3) {
        // cut + norms + grad/shrink/Bregman and div in one pass: b -> q's buffers (q itself
        // never reaches memory), then the two swap so that b holds the new values
        {
            TimedLaunch tl(ctx, TK_K_SBDIV);
            TK_TRY(launch_sb_div<T>(ctx, s->phi, s->phiold, s->g, s->b, s->q, s->divdb, lo, hi,
                                    &nb, part));
        }
        for (int a = 0; a < 3; ++a) std::swap(s->b[a]->d, s->q[a]->d);
        hipLaunchKernelGGL(k_sum_pairs_rsfe, dim3(1), dim3(1024), 0, ctx->stream, part, nb,
                           s->tol_dev);
        TK_LAUNCHED();
    } else {
        // cut + norms + grad/shrink/Bregman, then div
        {
            TimedLaunch tl(ctx, TK_K_SB);
            TK_TRY(launch_sb_update<T>(ctx, s->phi, s->phiold, s->b, s->g, s->q, lo, hi, &nb,
                                       part));
        }
        hipLaunchKernelGGL(k_sum_pairs_rsfe, dim3(1), dim3(1024), 0, ctx->stream, part, nb,
                           s->tol_dev);
        TK_LAUNCHED();
        {
            TimedLaunch tl(ctx, TK_K_DIV);
            TK_TRY(launch_div<T>(ctx, s->q[0], s->q[1], s->q[2], s->divdb));
        }
    }
    s->iter++;
    *extracted_out = extracted;
    return TK_OK;
}

// the tollerance sums into pinned host memory (a copy into it returns only once done, so
// work enqueued after it would reach an idle device); the fence makes them visible to the
// host by the time the kernel has finished
__global__ void k_tol_to_host(const double* __restrict__ src, double* dst)
{
    if (threadIdx.x < 2) dst[threadIdx.x] = src[threadIdx.x];
    __threadfence_system();
}

// the reference's loop condition needs the tollerance: read it back and wait (enqueueing
// next(s) after the read-back, when given, so the device has work while the host waits);
// the stats
int rsfe_finish(tk_rsfe s, tk_rsfe_iter_stats* st, int sweeps, double ls_err, int used,
                double res_plain, double res_guess, int extracted, int (*next)(tk_rsfe))
{
    tk_ctx ctx = s->ctx;
    if (next) {
        if (!s->tol_ev) TK_HIP(hipEventCreateWithFlags(&s->tol_ev, hipEventDisableTiming));
        if (!s->h_tol) TK_HIP(hipHostGetDevicePointer((void**)&s->h_tol, ctx->h_acc, 0));
        hipLaunchKernelGGL(k_tol_to_host, dim3(1), dim3(64), 0, ctx->stream,
                           (const double*)s->tol_dev, s->h_tol);
        TK_LAUNCHED();
        TK_HIP(hipEventRecord(s->tol_ev, ctx->stream));
        TK_TRY(next(s));
        TK_HIP(hipEventSynchronize(s->tol_ev));
    } else {
        TK_HIP(hipMemcpyAsync(ctx->h_acc, s->tol_dev, 2 * sizeof(double),
                              hipMemcpyDeviceToHost, ctx->stream));
        TK_HIP(hipStreamSynchronize(ctx->stream));
    }
    TK_TRY(timing_flush(ctx));
    const double h3 = s->h[0] * s->h[1] * s->h[2];
    s->tollerance = std::sqrt(h3 * ctx->h_acc[0]) / std::sqrt(h3 * ctx->h_acc[1]);
    if (st) {
        st->sweeps = sweeps;
        st->ls_err = ls_err;
        st->tollerance = s->tollerance;
        st->guess_used = used;
        st->res_plain = res_plain;
        st->res_guess = res_guess;
        st->cc_extracted = extracted;
        st->coarse_sweeps = s->it_coarse;
        st->tol_inf = s->it_tol_inf;
    }
    return TK_OK;
}

// kones = K*1 and kmyim = K*u0 as one complex transform
template <class T>
int conv_passes_pair_ones(tk_ctx ctx, const tk_filter_s* f, tk_vol kones, tk_vol kmyim, tk_vol u0,
                          void* work, bool acy)
{
    cplx<T>* C = (cplx<T>*)work;
    LdPair<T> ld{nullptr, (const T*)u0->d, (T)1, (T)0};
    TK_TRY((launch_conv_axis<T>(ctx, f, 0, 1, ld, MdNone{}, StComplex<T>{C}, acy)));
    TK_TRY((launch_conv_axis<T>(ctx, f, 1, 1, LdComplex<T>{C}, MdNone{}, StComplex<T>{C}, acy)));
    StPair<T> st{(T*)kones->d, (T*)kmyim->d};
    TK_TRY((launch_conv_axis<T>(ctx, f, 2, 1, LdComplex<T>{C}, MdNone{}, st, acy)));
    return TK_OK;
}

void free_state(tk_rsfe s)
{
    if (!s) return;
    tk_vol* vols[] = {&s->u0, &s->phi, &s->phiold, &s->rhs, &s->kones, &s->kmyim, &s->g, &s->divdb,
                      &s->b[0], &s->b[1], &s->b[2], &s->q[0], &s->q[1], &s->q[2]};
    for (tk_vol* v : vols) {
        tk_vol_free(*v);
        *v = nullptr;
    }
    if (s->C) (void)hipFree(s->C);
    if (s->alt) (void)hipFree(s->alt);
    if (s->guess) tk_vol_free(s->guess);
    if (s->resid) tk_vol_free(s->resid);
    if (s->cresid) tk_vol_free(s->cresid);
    if (s->cguess) tk_vol_free(s->cguess);
    if (s->cnet) tk_vol_free(s->cnet);
    if (s->vaux) tk_vol_free(s->vaux);
    if (s->ccmask) tk_vol_free(s->ccmask);
    if (s->ccneg) tk_vol_free(s->ccneg);
    if (s->tol_dev) (void)hipFree(s->tol_dev);
    if (s->tol_ev) (void)hipEventDestroy(s->tol_ev);
    if (s->flt) tk_filter_free(s->flt);
    delete s;
}

// phi := init as initialize_phi_with_init (rsfe_splitbregman_imp.hxx:99-170); no init, as
// apply(myim), or init of other dims: initialize_phi_with_cube (48-95)
int init_phi(tk_rsfe s, tk_vol init)
{
    tk_ctx ctx = s->ctx;
    const int X = s->X, Y = s->Y, Z = s->Z;
    const unsigned nb = blocks_for((long)X * Y * Z, 256, 8192);
    if (init && same_dims(init, s->u0)) {
        // 1st case: phi = init; change_range_of_intensity(b0, a0)
        TK_TRY(tk_vol_copy(s->phi, init));
        return tk_change_range_of_intensity(ctx, s->phi, s->p.b0, s->p.a0);
    }
    auto near = [](int a, int d) { return a == d || a == d - 1; };
    if (init && near(init->nx, X) && near(init->ny, Y) && near(init->nz, Z)) {
        // 2nd case: copy, extend the last faces, then change_range_of_intensity(b0, a0)
        const int ix = init->nx, iy = init->ny, iz = init->nz;
        if (s->dtype == TK_F64) {
            if (init->dtype == TK_F64)
                hipLaunchKernelGGL((k_init_extend<double, double>), dim3(nb), dim3(256), 0,
                                   ctx->stream, (double*)s->phi->d, (const double*)init->d, X, Y,
                                   Z, ix, iy, iz);
            else
                hipLaunchKernelGGL((k_init_extend<double, float>), dim3(nb), dim3(256), 0,
                                   ctx->stream, (double*)s->phi->d, (const float*)init->d, X, Y,
                                   Z, ix, iy, iz);
        } else {
            if (init->dtype == TK_F64)
                hipLaunchKernelGGL((k_init_extend<float, double>), dim3(nb), dim3(256), 0,
                                   ctx->stream, (float*)s->phi->d, (const double*)init->d, X, Y,
                                   Z, ix, iy, iz);
            else
                hipLaunchKernelGGL((k_init_extend<float, float>), dim3(nb), dim3(256), 0,
                                   ctx->stream, (float*)s->phi->d, (const float*)init->d, X, Y,
                                   Z, ix, iy, iz);
        }
        TK_LAUNCHED();
        return tk_change_range_of_intensity(ctx, s->phi, s->p.b0, s->p.a0);
    }
    if (init)
        fprintf(stderr, "WARNING::apply: wrong dimensions of init, cube initial contour is used "
                        "instead.\n");
    // 3rd case / no init: the cube [0.33 d, 0.67 d) per axis (imp:53-60)
    const double frac = 0.33;
    const int ix = (int)(X * frac), iy = (int)(Y * frac), iz = (int)(Z * frac);
    const int fx = (int)(X * (1 - frac)), fy = (int)(Y * (1 - frac)), fz = (int)(Z * (1 - frac));
    if (s->dtype == TK_F64)
        hipLaunchKernelGGL(k_init_cube<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (double*)s->phi->d, X, Y, Z, ix, fx, iy, fy, iz, fz, s->p.a0, s->p.b0);
    else
        hipLaunchKernelGGL(k_init_cube<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (float*)s->phi->d, X, Y, Z, ix, fx, iy, fy, iz, fz, (float)s->p.a0,
                           (float)s->p.b0);
    TK_LAUNCHED();
    return TK_OK;
}

}  // namespace

extern "C" {

int tk_rsfe_default_params(tk_rsfe_params* p)
{
    TK_CHECK(p, TK_ERR_ARG, "tk_rsfe_default_params: null");
    // rsfe_splitbregman_imp.hxx:25-41
    p->sigma = 3;
    p->lamda = 5.e-2;
    p->lamda1 = 1.e-5;
    p->lamda2 = 1.e-5;
    p->beta = 100.;
    p->dt = 1;
    p->epsilon = 1.e-3;
    p->a0 = 0.;
    p->b0 = 1.;
    p->nu = 1.;
    p->tol = 1.e-5;
    p->ls_tol = 0;
    p->radius = 3;
    p->maxiter = 100;
    p->ls_steps = 5;
    p->bc_dir = 0;
    p->gaussian_pixel_approach = 0;
    p->gs_order = TK_GS_LEX;
    p->max_sweeps = 0;
    p->warm_start = 0;
    p->ls_res_tol = 0;
    p->res_check_every = 64;
    p->ws_presweeps = 512;
    p->acyclic = 0;
    p->neumann = 0;
    // rsfe_splitbregman_imp.hxx:25-41: auto_extract_conn_comp false, cc_init_pixel -1,
    // cc_binary_output and cc_init_variables true, freq 0, dumpfrequency 0, "result"
    p->auto_extract_conn_comp = 0;
    p->auto_extract_conn_comp_freq = 0;
    p->cc_init_pixel[0] = p->cc_init_pixel[1] = p->cc_init_pixel[2] = -1;
    p->cc_threshold = 0.5;
    p->cc_binary_output = 1;
    p->cc_init_variables = 1;
    p->dumpfrequency = 0;
    std::memset(p->outputname, 0, sizeof p->outputname);
    std::strcpy(p->outputname, "result");
    p->edge_detector_sigma = 0;
    p->ws_judge_sweeps = 16;
    p->ws_continue = 0;
    p->ws_coarse_sweeps = 0;
    p->verbosity = 0;
    return TK_OK;
}

// gforshrink = edge_detect(myim) with the run's current parameters (imp:196-229): with
// edge_detector_sigma != 0 the edge detector reads cv(myim) with a Gaussian of that sigma and
// radius 3 (imp:214-227)
static int compute_edge(tk_rsfe s)
{
    const tk_rsfe_params& p = s->p;
    tk_ctx ctx = s->ctx;
    const tk_dtype dt = s->dtype;
    int r = TK_OK;
    tk_vol edge_src = s->u0, blurred = nullptr;
    tk_filter fe = nullptr;
    if (p.edge_detector_sigma != 0) {
        r = tk_filter_gaussian(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2],
                               p.edge_detector_sigma, 3., p.gaussian_pixel_approach, dt, &fe);
        if (r == TK_OK)
            r = tk_vol_alloc(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], dt, &blurred);
        if (r == TK_OK)
            r = p.acyclic ? tk_conv_acyclic(ctx, fe, blurred, s->u0)
                          : tk_conv_cyclic(ctx, fe, blurred, s->u0);
        edge_src = blurred;
    }
    if (r == TK_OK)
        r = dt == TK_F64 ? launch_edge<double>(ctx, edge_src, p.beta, p.lamda, p.nu, s->g)
                         : launch_edge<float>(ctx, edge_src, p.beta, p.lamda, p.nu, s->g);
    if (fe || blurred) {
        if (r == TK_OK) r = tk_ctx_sync(ctx);
        if (blurred) tk_vol_free(blurred);
        if (fe) tk_filter_free(fe);
    }
    return r;
}

// kones = cv(1), kmyim = cv(myim) in one complex transform with the run's filter
// (imp:1648-1651)
static int compute_ones_and_image(tk_rsfe s)
{
    return s->dtype == TK_F64
               ? conv_passes_pair_ones<double>(s->ctx, s->flt, s->kones, s->kmyim, s->u0, s->C,
                                               s->p.acyclic != 0)
               : conv_passes_pair_ones<float>(s->ctx, s->flt, s->kones, s->kmyim, s->u0, s->C,
                                              s->p.acyclic != 0);
}

int tk_rsfe_create(tk_ctx ctx, tk_vol u0, tk_vol init, const tk_rsfe_params* p, tk_rsfe* out)
{
    TK_CHECK(ctx && u0 && p && out, TK_ERR_ARG, "tk_rsfe_create: null argument");
    TK_CHECK(u0->dtype == TK_F32 || u0->dtype == TK_F64, TK_ERR_DTYPE, "tk_rsfe_create: u0 dtype");
    TK_CHECK(!init || init->dtype == TK_F32 || init->dtype == TK_F64, TK_ERR_DTYPE,
             "tk_rsfe_create: init dtype");
    TK_CHECK(u0->nx >= 3 && u0->ny >= 3 && u0->nz >= 3, TK_ERR_UNSUPPORTED,
             "tk_rsfe_create: 3-D volumes with every edge >= 3 only");
    TK_CHECK(p->warm_start >= 0 && p->warm_start <= 64, TK_ERR_ARG, "tk_rsfe_create: warm_start");
    TK_CHECK(p->ws_judge_sweeps >= 1 && p->ws_judge_sweeps <= 4096 &&
                 (p->ws_continue == 0 || p->ws_continue == 1),
             TK_ERR_ARG, "tk_rsfe_create: ws_judge_sweeps in [1, 4096], ws_continue 0 or 1");
    TK_CHECK(p->ws_coarse_sweeps >= 0 && (p->verbosity == 0 || p->verbosity == 1), TK_ERR_ARG,
             "tk_rsfe_create: ws_coarse_sweeps >= 0, verbosity 0 or 1");
    TK_CHECK(p->gs_order == TK_GS_LEX || p->gs_order == TK_GS_REDBLACK, TK_ERR_ARG,
             "tk_rsfe_create: gs_order");
    TK_CHECK(p->acyclic == 0 || p->acyclic == 1, TK_ERR_ARG, "tk_rsfe_create: acyclic");
    TK_CHECK(p->neumann == 0 || p->neumann == 1, TK_ERR_ARG, "tk_rsfe_create: neumann");
    TK_CHECK(p->auto_extract_conn_comp_freq >= 0 && p->dumpfrequency >= 0, TK_ERR_ARG,
             "tk_rsfe_create: auto_extract_conn_comp_freq and dumpfrequency must be >= 0");
    TK_CHECK(p->cc_threshold >= 0 && p->cc_threshold <= 1, TK_ERR_ARG,
             "tk_rsfe_create: cc_threshold must be in [0, 1]");
    TK_CHECK(p->edge_detector_sigma >= 0, TK_ERR_ARG,
             "tk_rsfe_create: edge_detector_sigma must be >= 0");
    TK_CHECK(std::memchr(p->outputname, 0, sizeof p->outputname) != nullptr, TK_ERR_ARG,
             "tk_rsfe_create: outputname is not terminated");
    // the Neumann sweep is the reference's lexicographic one; the learned start and the
    // residual gate are built on the Dirichlet problem
    TK_CHECK(!p->neumann || (p->gs_order == TK_GS_LEX && p->warm_start == 0 &&
                             p->ls_res_tol <= 0),
             TK_ERR_UNSUPPORTED,
             "tk_rsfe_create: neumann needs gs_order TK_GS_LEX, no warm_start and no ls_res_tol");
    tk_rsfe s = new tk_rsfe_s();
    s->ctx = ctx;
    s->p = *p;
    s->dtype = u0->dtype;
    s->X = u0->nx; s->Y = u0->ny; s->Z = u0->nz;
    s->h[0] = u0->hx; s->h[1] = u0->hy; s->h[2] = u0->hz;
    const tk_dtype dt = s->dtype;
    int r = TK_OK;
    auto alloc = [&](tk_vol* v) {
        if (r == TK_OK)
            r = tk_vol_alloc(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], dt, v);
    };
    tk_vol* vols[] = {&s->u0, &s->phi, &s->phiold, &s->rhs, &s->kones, &s->kmyim, &s->g, &s->divdb,
                      &s->b[0], &s->b[1], &s->b[2], &s->q[0], &s->q[1], &s->q[2]};
    for (tk_vol* v : vols) alloc(v);
    if (r == TK_OK && hipMalloc(&s->C, 2 * dtype_size(dt) * (size_t)u0->n()) != hipSuccess) {
        set_error("tk_rsfe_create: complex buffer allocation failed");
        r = TK_ERR_NOMEM;
    }
    if (r == TK_OK && p->gs_order == TK_GS_REDBLACK &&
        hipMalloc(&s->alt, dtype_size(dt) * (size_t)u0->n()) != hipSuccess) {
        set_error("tk_rsfe_create: red-black buffer allocation failed");
        r = TK_ERR_NOMEM;
    }
    if (r == TK_OK && hipMalloc((void**)&s->tol_dev, 2 * sizeof(double)) != hipSuccess) {
        set_error("tk_rsfe_create: allocation failed");
        r = TK_ERR_NOMEM;
    }
    if (r == TK_OK) r = tk_vol_copy(s->u0, u0);
    if (r == TK_OK) r = init_phi(s, init);
    // build_gaussian_filter(phi, sigma, 3, gpa) (imp:1641-1644)
    if (r == TK_OK)
        r = tk_filter_gaussian(ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], p->sigma,
                               p->radius, p->gaussian_pixel_approach, dt, &s->flt);
    // FP32 convolves directly on every axis (accuracy of the region fit; the filter picks
    // the register-window or the generic taps pass), FP64 through the FFT passes
    // red-black runs: pick the sweep tile by timing (the result does not depend on it)
    if (r == TK_OK && p->gs_order == TK_GS_REDBLACK)
        r = tk_poisson_rb_autotune(ctx, s->u0, nullptr, nullptr);
    // kones = cv(1), kmyim = cv(myim) in one complex transform (imp:1648-1651)
    if (r == TK_OK) r = compute_ones_and_image(s);
    // gforshrink = edge_detect(myim) (imp:1655)
    if (r == TK_OK) r = compute_edge(s);
    s->old_eds = p->edge_detector_sigma != 0;
    if (r == TK_OK) r = tk_ctx_sync(ctx);
    if (r != TK_OK) {
        free_state(s);
        return r;
    }
    *out = s;
    return TK_OK;
}

int tk_rsfe_iterate(tk_rsfe s, int n_iters, tk_rsfe_iter_stats* stats, int* n_done)
{
    TK_CHECK(s, TK_ERR_ARG, "tk_rsfe_iterate: null state");
    TK_CHECK(s->p.warm_start == 0 || s->net, TK_ERR_ARG,
             "tk_rsfe_iterate: warm_start set but no network attached (tk_rsfe_set_warm_start)");
    int done = 0;
    const tk_rsfe_params& p = s->p;
    // The next iteration's forcing passes go out before the wait for this one's tollerance
    // when this call runs another iteration unless the tollerance stops it, and tol <= 0
    // (then only a tollerance of exactly 0, a stationary phi, or NaN, 0/0 with phi and
    // phiold zero everywhere, stops the loop: the two cases in which rhs is left holding the
    // forcing of an iteration that did not run); not with per-kernel timing,
    // whose read-out would wait for them.  Nothing is carried over to the next call.
    bool ahead = false;
    // for (t = 1; tollerance > tol && t <= maxiter; ++t)  (imp:1659)
    while (!s->ended && done < n_iters && s->tollerance > p.tol && s->iter < p.maxiter) {
        const bool spec = done + 1 < n_iters && s->iter + 1 < p.maxiter && p.tol <= 0 &&
                          !s->ctx->timing;
        tk_rsfe_iter_stats* st = stats ? &stats[done] : nullptr;
        int r = s->dtype == TK_F64 ? rsfe_iteration<double>(s, st, true, &ahead, spec)
                                   : rsfe_iteration<float>(s, st, true, &ahead, spec);
        // dump every dumpfrequency iterations, not at maxiter (imp:1939-1951)
        const int t = s->iter;
        if (r == TK_OK && p.dumpfrequency != 0 && t % p.dumpfrequency == 0 && t != p.maxiter)
            r = tk_vol_write_mhd(s->phi, (std::string(p.outputname) + "_" + std::to_string(t)).c_str());
        if (r != TK_OK) {
            if (n_done) *n_done = done;
            return r;
        }
        ++done;
    }
    // the loop has ended: the final dump (imp:1968-1978), once
    if (p.dumpfrequency > 0 && !s->final_dumped &&
        (s->tollerance <= p.tol || s->iter >= p.maxiter || s->ended)) {
        s->final_dumped = true;
        const int r = tk_vol_write_mhd(
            s->phi, (std::string(p.outputname) + "_" + std::to_string(s->iter) + "_final").c_str());
        if (r != TK_OK) {
            if (n_done) *n_done = done;
            return r;
        }
    }
    if (n_done) *n_done = done;
    return TK_OK;
}

}  // extern "C"

namespace tk {
int rsfe_step_async(tk_rsfe s)
{
    if (s->iter >= s->p.maxiter) return TK_OK;
    return s->dtype == TK_F64 ? rsfe_iteration<double>(s, nullptr, false)
                              : rsfe_iteration<float>(s, nullptr, false);
}
int rsfe_read_tollerance(tk_rsfe s, double* tol)
{
    double h[2];
    TK_HIP(hipMemcpyAsync(h, s->tol_dev, sizeof h, hipMemcpyDeviceToHost, s->ctx->stream));
    TK_HIP(hipStreamSynchronize(s->ctx->stream));
    const double h3 = s->h[0] * s->h[1] * s->h[2];
    s->tollerance = std::sqrt(h3 * h[0]) / std::sqrt(h3 * h[1]);
    if (tol) *tol = s->tollerance;
    return TK_OK;
}
int rsfe_iterations(tk_rsfe s) { return s->iter; }

// ---- the convolution passes of many runs, one launch per pass and variant -----------------
// Runs are grouped by the variant their filter axis plans (half-width, quads, matrix cores),
// so each run computes exactly what its own launch would; a pass then costs one launch per
// group.  The first x pass reads phi, whose buffer alternates with the red-black partner:
// its tables exist for both roles.
struct BatchGroup {
    ConvPlan like;
    void* jobs = nullptr;
    int njobs = 0;
    unsigned nblocks = 0;
    size_t lds = 0;
};

struct RsfeBatchConv {
    tk_dtype dt = TK_F32;
    std::vector<BatchGroup> x1[2], y, z2, x2;
    ~RsfeBatchConv()
    {
        for (auto* v : {&x1[0], &x1[1], &y, &z2, &x2})
            for (BatchGroup& g : *v)
                if (g.jobs) (void)hipFree(g.jobs);
    }
};

namespace {

template <class T, class Ld, class Md, class St, class Ops>
int build_stage(std::vector<BatchGroup>& out, tk_rsfe* runs, int n, int axis, int rounds, Ops ops)
{
    typedef ConvJob<Ld, Md, St> Job;
    std::vector<std::vector<Job>> jobs;
    for (int v = 0; v < n; ++v) {
        tk_rsfe s = runs[v];
        ConvPlan pl;
        TK_TRY(conv_plan<T>(s->flt, axis, s->p.acyclic != 0, &pl));
        size_t q = 0;
        while (q < out.size() && !(out[q].like.S == pl.S && out[q].like.quad == pl.quad &&
                                   out[q].like.mf == pl.mf && out[q].like.small == pl.small))
            ++q;
        if (q == out.size()) {
            out.emplace_back();
            out[q].like = pl;
            jobs.emplace_back();
        }
        Job j;
        j.ax = pl.ax;
        j.g = pl.g;
        j.rounds = rounds;
        ops(s, &j.ld, &j.md, &j.st);
        j.groups = (int)pl.nb;
        j.nblk = (int)((pl.nb + kNumXcd - 1) / kNumXcd * kNumXcd);
        j.first = (int)out[q].nblocks;
        out[q].nblocks += (unsigned)j.nblk;
        out[q].lds = std::max(out[q].lds, pl.lds);
        jobs[q].push_back(j);
    }
    for (size_t q = 0; q < out.size(); ++q) {
        out[q].njobs = (int)jobs[q].size();
        TK_CHECK(hipMalloc(&out[q].jobs, sizeof(Job) * jobs[q].size()) == hipSuccess,
                 TK_ERR_NOMEM, "batch: convolution job table allocation failed");
        TK_HIP(hipMemcpy(out[q].jobs, jobs[q].data(), sizeof(Job) * jobs[q].size(),
                         hipMemcpyHostToDevice));
    }
    return TK_OK;
}

template <class T, class Ld, class Md, class St>
int run_stage(tk_ctx ctx, const std::vector<BatchGroup>& groups, int tag)
{
    TimedLaunch tl(ctx, tag);
    for (const BatchGroup& g : groups)
        TK_TRY((launch_conv_batch<T, Ld, Md, St>(ctx, g.like,
                                                 (const ConvJob<Ld, Md, St>*)g.jobs, g.njobs,
                                                 g.nblocks, g.lds)));
    return TK_OK;
}

template <class T>
int batch_conv_build(tk_rsfe* runs, int n, RsfeBatchConv* t)
{
    for (int role = 0; role < 2; ++role)
        TK_TRY((build_stage<T, LdHeaviside<T>, MdNone, StComplex<T>>(
            t->x1[role], runs, n, 0, 1,
            [role](tk_rsfe s, LdHeaviside<T>* ld, MdNone*, StComplex<T>* st) {
                const tk_rsfe_params& p = s->p;
                const T* phi = (const T*)(role == 0 ? s->phi->d : s->alt);
                *ld = LdHeaviside<T>{phi, (const T*)s->u0->d, (T*)s->phiold->d,
                                     (T)((p.a0 + p.b0) / 2.), (T)p.epsilon};
                *st = StComplex<T>{(cplx<T>*)s->C};
            })));
    TK_TRY((build_stage<T, LdComplex<T>, MdNone, StComplex<T>>(
        t->y, runs, n, 1, 1, [](tk_rsfe s, LdComplex<T>* ld, MdNone*, StComplex<T>* st) {
            *ld = LdComplex<T>{(cplx<T>*)s->C};
            *st = StComplex<T>{(cplx<T>*)s->C};
        })));
    TK_TRY((build_stage<T, LdComplex<T>, MdRegionFit<T>, StComplex<T>>(
        t->z2, runs, n, 2, 2,
        [](tk_rsfe s, LdComplex<T>* ld, MdRegionFit<T>* md, StComplex<T>* st) {
            const tk_rsfe_params& p = s->p;
            *ld = LdComplex<T>{(cplx<T>*)s->C};
            *md = MdRegionFit<T>{(const T*)s->kmyim->d, (const T*)s->kones->d, (T)p.lamda1,
                                 (T)p.lamda2};
            *st = StComplex<T>{(cplx<T>*)s->C};
        })));
    TK_TRY((build_stage<T, LdComplex<T>, MdNone, StRhs<T>>(
        t->x2, runs, n, 0, 1, [](tk_rsfe s, LdComplex<T>* ld, MdNone*, StRhs<T>* st) {
            const tk_rsfe_params& p = s->p;
            *ld = LdComplex<T>{(cplx<T>*)s->C};
            *st = StRhs<T>{(const T*)s->u0->d, (const T*)s->kones->d, (const T*)s->divdb->d,
                           (T*)s->rhs->d, (T)(p.lamda1 - p.lamda2), (T)p.lamda};
        })));
    return TK_OK;
}

template <class T>
int batch_conv_run(tk_ctx ctx, const RsfeBatchConv* t, int role)
{
    TK_TRY((run_stage<T, LdHeaviside<T>, MdNone, StComplex<T>>(ctx, t->x1[role], TK_K_CONV_X1)));
    TK_TRY((run_stage<T, LdComplex<T>, MdNone, StComplex<T>>(ctx, t->y, TK_K_CONV_Y)));
    TK_TRY((run_stage<T, LdComplex<T>, MdRegionFit<T>, StComplex<T>>(ctx, t->z2, TK_K_CONV_Z2)));
    TK_TRY((run_stage<T, LdComplex<T>, MdNone, StComplex<T>>(ctx, t->y, TK_K_CONV_Y)));
    TK_TRY((run_stage<T, LdComplex<T>, MdNone, StRhs<T>>(ctx, t->x2, TK_K_CONV_X2)));
    return TK_OK;
}

}  // namespace

int rsfe_batch_conv_build(tk_rsfe* runs, int n, RsfeBatchConv** out)
{
    TK_CHECK(runs && n > 0 && out, TK_ERR_ARG, "batch: convolution tables: argument");
    for (int v = 0; v < n; ++v)
        TK_CHECK(runs[v] && runs[v]->alt && runs[v]->dtype == runs[0]->dtype, TK_ERR_ARG,
                 "batch: convolution tables: run %d", v);
    RsfeBatchConv* t = new RsfeBatchConv();
    t->dt = runs[0]->dtype;
    const int r = t->dt == TK_F64 ? batch_conv_build<double>(runs, n, t)
                                  : batch_conv_build<float>(runs, n, t);
    if (r != TK_OK) {
        delete t;
        return r;
    }
    *out = t;
    return TK_OK;
}
int rsfe_batch_conv_run(tk_ctx ctx, const RsfeBatchConv* t, int role)
{
    return t->dt == TK_F64 ? batch_conv_run<double>(ctx, t, role)
                           : batch_conv_run<float>(ctx, t, role);
}
void rsfe_batch_conv_free(RsfeBatchConv* t) { delete t; }

struct RsfeBatchUpdate {
    tk_dtype dt = TK_F32;
    void* sb[2] = {nullptr, nullptr};  // SbJob tables, phi in each role
    void* div = nullptr;               // DivJob table
    double* parts = nullptr;           // every run's block partials
    unsigned sb_blocks = 0, div_blocks = 0;
    int n = 0;
    ~RsfeBatchUpdate()
    {
        for (void* p : {sb[0], sb[1], div, (void*)parts})
            if (p) (void)hipFree(p);
    }
};

int rsfe_batch_update_build(tk_rsfe* runs, int n, RsfeBatchUpdate** out)
{
    TK_CHECK(runs && n > 0 && out, TK_ERR_ARG, "batch: update tables: argument");
    RsfeBatchUpdate* t = new RsfeBatchUpdate();
    t->dt = runs[0]->dtype;
    t->n = n;
    std::vector<SbJob> sb((size_t)n);
    std::vector<DivJob> dv((size_t)n);
    long parts = 0;
    int first = 0;
    for (int v = 0; v < n; ++v) {
        tk_rsfe s = runs[v];
        s->in_batch = true;
        SbJob& j = sb[(size_t)v];
        j.phiold = s->phiold->d;
        j.g = s->g->d;
        for (int a = 0; a < 3; ++a) {
            j.b[a] = s->b[a]->d;
            j.q[a] = s->q[a]->d;
        }
        j.X = s->X;
        j.Y = s->Y;
        j.Z = s->Z;
        j.hx = s->h[0];
        j.hy = s->h[1];
        j.hz = s->h[2];
        j.lo = s->p.a0;
        j.hi = s->p.b0;
        j.tol = s->tol_dev;
        stencil_batch_grid(t->dt, s->X, s->Y, s->Z, &j.gy, &j.nparts);
        j.nblk = (j.nparts + (int)kNumXcd - 1) / (int)kNumXcd * (int)kNumXcd;
        j.first = first;
        first += j.nblk;
        j.part = nullptr;
        j.phi = nullptr;
        parts += 2L * j.nparts;
        DivJob& d = dv[(size_t)v];
        for (int a = 0; a < 3; ++a) d.q[a] = s->q[a]->d;
        d.out = s->divdb->d;
        d.X = s->X;
        d.Y = s->Y;
        d.Z = s->Z;
        d.hx = s->h[0];
        d.hy = s->h[1];
        d.hz = s->h[2];
        d.gy = j.gy;
        d.nblk = j.nblk;
        d.first = j.first;
    }
    t->sb_blocks = t->div_blocks = (unsigned)first;
    int r = hipMalloc(&t->parts, sizeof(double) * (size_t)parts) == hipSuccess ? TK_OK : TK_ERR_NOMEM;
    long off = 0;
    for (int v = 0; v < n; ++v) {
        sb[(size_t)v].part = t->parts + off;
        off += 2L * sb[(size_t)v].nparts;
    }
    for (int role = 0; r == TK_OK && role < 2; ++role) {
        for (int v = 0; v < n; ++v) sb[(size_t)v].phi = role == 0 ? runs[v]->phi->d : runs[v]->alt;
        if (hipMalloc(&t->sb[role], sizeof(SbJob) * (size_t)n) != hipSuccess) r = TK_ERR_NOMEM;
        else if (hipMemcpy(t->sb[role], sb.data(), sizeof(SbJob) * (size_t)n,
                           hipMemcpyHostToDevice) != hipSuccess)
            r = TK_ERR_HIP;
    }
    if (r == TK_OK && hipMalloc(&t->div, sizeof(DivJob) * (size_t)n) != hipSuccess) r = TK_ERR_NOMEM;
    if (r == TK_OK &&
        hipMemcpy(t->div, dv.data(), sizeof(DivJob) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess)
        r = TK_ERR_HIP;
    if (r != TK_OK) {
        delete t;
        set_error("batch: update tables: allocation or upload failed");
        return r;
    }
    *out = t;
    return TK_OK;
}

// cut + tolerance sums + grad/shrink/Bregman, the sums per run, div; the iteration counts.
// (The staged path has no automatic extraction and no dumps: the batch refuses them.)
int rsfe_batch_update_run(tk_ctx ctx, const RsfeBatchUpdate* t, tk_rsfe* runs, int n, int role)
{
    TK_TRY(launch_sb_batch(ctx, t->dt, (const SbJob*)t->sb[role], t->n, t->sb_blocks));
    TK_TRY(launch_div_batch(ctx, t->dt, (const DivJob*)t->div, t->n, t->div_blocks));
    for (int v = 0; v < n; ++v) runs[v]->iter++;
    return TK_OK;
}
void rsfe_batch_update_free(RsfeBatchUpdate* t) { delete t; }

// the staged batch path: the red-black buffers (faces set once), the pointer exchange
// after an odd sweep count
int rsfe_rb_prepare(tk_rsfe s, void** phi, void** alt, const void** rhs)
{
    TK_CHECK(s->alt, TK_ERR_ARG, "batch: the run has no red-black partner buffer");
    const tk_rsfe_params& p = s->p;
    const double bc = p.bc_dir ? p.a0 : 0.;
    if (!s->faces_set) {
        // as the first fixed-sweep solve does: both buffers carry the boundary value
        tk_vol_s tmp = *s->phi;
        tmp.d = s->alt;
        tmp.owned = false;
        TK_TRY(poisson_set_boundary(s->ctx, s->phi, bc));
        TK_TRY(poisson_set_boundary(s->ctx, &tmp, bc));
        if (bc >= p.a0 && bc <= p.b0) s->faces_set = true;
    }
    *phi = s->phi->d;
    *alt = s->alt;
    *rhs = s->rhs->d;
    return TK_OK;
}
void rsfe_rb_exchange(tk_rsfe s) { std::swap(s->phi->d, s->alt); }
const tk_rsfe_params& rsfe_params(tk_rsfe s) { return s->p; }
void rsfe_set_iterations(tk_rsfe s, int iter) { s->iter = iter; }
}  // namespace tk

extern "C" {

tk_vol tk_rsfe_phi(tk_rsfe s) { return s ? s->phi : nullptr; }
tk_vol tk_rsfe_rhs(tk_rsfe s) { return s ? s->rhs : nullptr; }

int tk_rsfe_onthego(tk_rsfe s, const tk_rsfe_params* p, int save_current, int end_now)
{
    TK_CHECK(s && p, TK_ERR_ARG, "tk_rsfe_onthego: null argument");
    const tk_rsfe_params old = s->p;
    TK_CHECK(p->gs_order == old.gs_order && p->acyclic == old.acyclic &&
                 p->neumann == old.neumann && p->warm_start == old.warm_start &&
                 p->ls_res_tol == old.ls_res_tol && p->res_check_every == old.res_check_every &&
                 p->ws_presweeps == old.ws_presweeps && p->ws_judge_sweeps == old.ws_judge_sweeps &&
                 p->ws_continue == old.ws_continue && p->ws_coarse_sweeps == old.ws_coarse_sweeps,
             TK_ERR_UNSUPPORTED,
             "tk_rsfe_onthego: the sweep order, the convolution, the Neumann solver and the "
             "warm start cannot change on the go");
    TK_CHECK(p->sigma > 0 && p->maxiter >= 0 && p->ls_steps >= 0 && p->dumpfrequency >= 0 &&
                 p->auto_extract_conn_comp_freq >= 0 && p->edge_detector_sigma >= 0 &&
                 p->cc_threshold >= 0 && p->cc_threshold <= 1 &&
                 (p->verbosity == 0 || p->verbosity == 1),
             TK_ERR_ARG, "tk_rsfe_onthego: parameter out of range");
    TK_CHECK(std::memchr(p->outputname, 0, sizeof p->outputname) != nullptr, TK_ERR_ARG,
             "tk_rsfe_onthego: outputname is not terminated");
    TK_CHECK(!s->in_batch, TK_ERR_UNSUPPORTED, "tk_rsfe_onthego: the run is part of a batch");
    s->p = *p;
    int r = TK_OK;
    // a changed sigma: the filter (radius 3, as imp:1686-1689 builds it) and kones, kmyim
    // for the next iteration (imp:1683-1696)
    if (p->sigma != old.sigma) {
        tk_filter f = nullptr;
        r = tk_filter_gaussian(s->ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], p->sigma,
                               3., p->gaussian_pixel_approach, s->dtype, &f);
        if (r == TK_OK) {
            tk_filter_free(s->flt);
            s->flt = f;
            r = compute_ones_and_image(s);
        }
    }
    // g again when the bool of the previous edge_detector_sigma differs from the current one
    // (imp:1877-1882); with unchanged inputs that is the same g
    if (r == TK_OK && (double)s->old_eds != p->edge_detector_sigma &&
        (p->edge_detector_sigma != old.edge_detector_sigma || p->beta != old.beta ||
         p->lamda != old.lamda || p->nu != old.nu ||
         p->gaussian_pixel_approach != old.gaussian_pixel_approach))
        r = compute_edge(s);
    s->old_eds = p->edge_detector_sigma != 0;
    // the Poisson boundary value (bc flag, a0) is set again by the next solve
    if (p->a0 != old.a0 || p->b0 != old.b0 || p->bc_dir != old.bc_dir) s->faces_set = false;
    if (r == TK_OK) r = tk_ctx_sync(s->ctx);
    if (r != TK_OK) return r;
    // this iteration's dump with the settings just read (imp:1939-1951): where the previous
    // settings wrote one under the same name it is already there
    const int t = s->iter;
    auto dump_due = [t](const tk_rsfe_params& q) {
        return q.dumpfrequency != 0 && t % q.dumpfrequency == 0 && t != q.maxiter;
    };
    const bool written = t > 0 && dump_due(old) && std::strcmp(old.outputname, p->outputname) == 0;
    if (t > 0 && (dump_due(*p) || save_current) && !written)
        TK_TRY(tk_vol_write_mhd(s->phi, (std::string(p->outputname) + "_" + std::to_string(t)).c_str()));
    if (end_now) {
        s->ended = true;
        if (p->dumpfrequency > 0 && !s->final_dumped) {
            s->final_dumped = true;
            TK_TRY(tk_vol_write_mhd(
                s->phi, (std::string(p->outputname) + "_" + std::to_string(t) + "_final").c_str()));
        }
    }
    return TK_OK;
}

int tk_rsfe_set_warm_start(tk_rsfe s, tk_unet net)
{
    TK_CHECK(s, TK_ERR_ARG, "tk_rsfe_set_warm_start: null state");
    if (!net) {
        s->net = nullptr;
        s->p.warm_start = 0;
        return TK_OK;
    }
    TK_CHECK(!s->p.neumann, TK_ERR_UNSUPPORTED,
             "tk_rsfe_set_warm_start: the learned start serves the Dirichlet sweep only");
    TK_CHECK(unet_ctx(net) == s->ctx, TK_ERR_ARG,
             "tk_rsfe_set_warm_start: the network must live on the run's context");
    int d[3];
    unet_dims(net, d);
    const int f = d[0] > 0 ? s->X / d[0] : 0;
    TK_CHECK(f >= 1 && s->X == f * d[0] && s->Y == f * d[1] && s->Z == f * d[2], TK_ERR_SHAPE,
             "tk_rsfe_set_warm_start: the run's dims must be the network's times one integer "
             "factor on every axis");
    if (f > 1 && (!s->cresid || s->cresid->nx != d[0] || s->cresid->ny != d[1] ||
                  s->cresid->nz != d[2])) {
        if (s->cresid) tk_vol_free(s->cresid);
        if (s->cguess) tk_vol_free(s->cguess);
        if (s->cnet) tk_vol_free(s->cnet);
        s->cresid = s->cguess = s->cnet = nullptr;
        const double H[3] = {s->h[0] * f, s->h[1] * f, s->h[2] * f};
        TK_TRY(tk_vol_alloc(s->ctx, d[0], d[1], d[2], H[0], H[1], H[2], s->dtype, &s->cresid));
        TK_TRY(tk_vol_alloc(s->ctx, d[0], d[1], d[2], H[0], H[1], H[2], s->dtype, &s->cguess));
        TK_TRY(tk_vol_alloc(s->ctx, d[0], d[1], d[2], H[0], H[1], H[2], s->dtype, &s->cnet));
    }
    s->ws_f = f;
    if (!s->guess)
        TK_TRY(tk_vol_alloc(s->ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], s->dtype,
                            &s->guess));
    if (!s->resid)
        TK_TRY(tk_vol_alloc(s->ctx, s->X, s->Y, s->Z, s->h[0], s->h[1], s->h[2], s->dtype,
                            &s->resid));
    s->net = net;
    if (s->p.warm_start < 1) s->p.warm_start = 1;
    return TK_OK;
}
tk_filter tk_rsfe_filter(tk_rsfe s) { return s ? s->flt : nullptr; }

int tk_sb_update(tk_ctx ctx, tk_vol phi, tk_vol phiold, tk_vol g, const tk_vol* b,
                 const tk_vol* q, tk_vol divdb, double a0, double b0, double* norms)
{
    TK_CHECK(ctx && phi && phiold && g && b && q && divdb, TK_ERR_ARG,
             "tk_sb_update: null argument");
    tk_vol all[] = {phiold, g, divdb, b[0], b[1], b[2], q[0], q[1], q[2]};
    for (tk_vol v : all) {
        TK_CHECK(v && v->d, TK_ERR_ARG, "tk_sb_update: null volume");
        TK_CHECK(same_dims(v, phi) && v->dtype == phi->dtype, TK_ERR_SHAPE,
                 "tk_sb_update: volumes differ in dims or type");
    }
    TK_CHECK(phi->dtype != TK_U8, TK_ERR_DTYPE, "tk_sb_update: U8 volume");
    TK_CHECK(phi->nx >= 3 && phi->ny >= 3 && phi->nz >= 3, TK_ERR_UNSUPPORTED,
             "tk_sb_update: 3-D volumes with every edge >= 3 only");
    const bool dbl = phi->dtype == TK_F64;
    int nb = 0;
    double* part = ctx->acc + 64;
    tk_vol bb[3] = {b[0], b[1], b[2]}, qq[3] = {q[0], q[1], q[2]};
    TK_TRY(dbl ? launch_sb_update<double>(ctx, phi, phiold, bb, g, qq, a0, b0, &nb, part)
               : launch_sb_update<float>(ctx, phi, phiold, bb, g, qq, a0, b0, &nb, part));
    TK_TRY(dbl ? launch_div<double>(ctx, qq[0], qq[1], qq[2], divdb)
               : launch_div<float>(ctx, qq[0], qq[1], qq[2], divdb));
    hipLaunchKernelGGL(k_sum_pairs_rsfe, dim3(1), dim3(1024), 0, ctx->stream, part, nb,
                       ctx->acc + 8);
    TK_LAUNCHED();
    if (norms) {
        TK_HIP(hipMemcpyAsync(norms, ctx->acc + 8, 2 * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
        TK_HIP(hipStreamSynchronize(ctx->stream));
    }
    return TK_OK;
}

int tk_rsfe_destroy(tk_rsfe s)
{
    if (!s) return TK_OK;
    (void)hipStreamSynchronize(s->ctx->stream);
    free_state(s);
    return TK_OK;
}

}  // extern "C"
