// tk_fft.hip -- Gaussian filter spectra (R4) and cyclic convolution (R5).
#include <hip/hip_runtime.h>

#include <cmath>

#include "tk_fft_kernel.hpp"

namespace tk {

// radix plan for the Stockham FFT: the largest power-of-two radices first (16 for FP32,
// 8 for FP64, where a thread holds 16 or 8 elements), then 2, 3, 5; 0 if L has another
// prime factor
int fft_plan(int L, int* radix, int* nstage, int max_radix)
{
    int n = 0, m = L;
    if (max_radix >= 16)
        while (m % 16 == 0) { radix[n++] = 16; m /= 16; }
    if (max_radix >= 8)
        while (m % 8 == 0) { radix[n++] = 8; m /= 8; }
    while (m % 4 == 0) { radix[n++] = 4; m /= 4; }
    while (m % 2 == 0) { radix[n++] = 2; m /= 2; }
    while (m % 3 == 0) { radix[n++] = 3; m /= 3; }
    while (m % 5 == 0) { radix[n++] = 5; m /= 5; }
    if (m != 1 || n > kMaxStages) return 0;
    *nstage = n;
    return 1;
}

// Nh of convolution_imp.hxx:34-60 along one axis
static int taps(int dim_is_z_of_2d, double radius, double sigma, double h)
{
    if (dim_is_z_of_2d) return 1;
    int N = (int)std::ceil(2 * radius * sigma / h);
    if (N % 2 == 0) N++;
    return N;
}

// real spectrum of the 1-D normalised Gaussian wrapped onto length L:
// H(k) = sum_{n=-s..s} w(n) cos(2 pi k n / L),  w = g / sum g,  g(n) = exp(-n^2 h^2 / (2 sigma^2))
static void axis_spectrum(int L, int Nh, double sigma, double h, std::vector<double>& H)
{
    const int s = Nh / 2;
    std::vector<long double> w(Nh);
    long double sum = 0;
    for (int n = -s; n <= s; ++n) {
        long double e = -(long double)n * n * h * h / (2.0L * sigma * sigma);
        w[n + s] = expl(e);
        sum += w[n + s];
    }
    H.assign(L, 0.0);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < L; ++k) {
        long double acc = 0;
        for (int n = -s; n <= s; ++n) {
            long m = ((long)k * n) % L;
            if (m < 0) m += L;
            acc += w[n + s] * cosl(two_pi * m / L);
        }
        H[k] = (double)(acc / sum);
    }
}

template <class T>
static int upload_axis(tk_filter_s* f, int a)
{
    const int L = f->dims[a];
    tk::AxisDev& ax = f->ax[a];
    ax.L = L;
    ax.nh = f->nh[a];
    if (L == 1) return TK_OK;
    {
        // symmetric taps w(0..s) of the normalised 1-D Gaussian for the direct path
        const int nh = f->nh[a], s = nh / 2;
        const double hh = f->pixel ? 1. : f->h[a];
        std::vector<long double> w(s + 1);
        long double sum = 0;
        for (int n = 0; n <= s; ++n) {
            w[n] = expl(-(long double)n * n * hh * hh / (2.0L * f->sigma * f->sigma));
            sum += n ? 2 * w[n] : w[n];
        }
        std::vector<T> half(s + 1);
        for (int n = 0; n <= s; ++n) half[n] = (T)(w[n] / sum);
        TK_HIP(hipMalloc(&f->half_dev[a], sizeof(T) * (s + 1)));
        TK_HIP(hipMemcpy(f->half_dev[a], half.data(), sizeof(T) * (s + 1), hipMemcpyHostToDevice));
        ax.half = f->half_dev[a];
        // all taps, for the generic direct pass (any method, and every acyclic pass that
        // has no register window)
        std::vector<T> taps(nh);
        for (int n = -s; n <= s; ++n) taps[n + s] = half[n < 0 ? -n : n];
        TK_HIP(hipMalloc(&f->taps_dev[a], sizeof(T) * nh));
        TK_HIP(hipMemcpy(f->taps_dev[a], taps.data(), sizeof(T) * nh, hipMemcpyHostToDevice));
        ax.taps = f->taps_dev[a];
        if (sizeof(T) == 4 && direct_width_supported(nh)) {
            // A operand of the matrix-core window, lane l = n + 16 g of K-step ks:
            // w(|4 ks + g - s - n|) inside the band, 0 outside
            const int KS = mfma_ksteps(s);
            std::vector<float> band((size_t)KS * 64);
            for (int ks = 0; ks < KS; ++ks)
                for (int l = 0; l < 64; ++l) {
                    const int d = 4 * ks + (l >> 4) - s - (l & 15);
                    band[(size_t)ks * 64 + l] = (d >= -s && d <= s) ? (float)half[d < 0 ? -d : d] : 0.f;
                }
            TK_HIP(hipMalloc(&f->band_dev[a], sizeof(float) * band.size()));
            TK_HIP(hipMemcpy(f->band_dev[a], band.data(), sizeof(float) * band.size(),
                             hipMemcpyHostToDevice));
            ax.band = f->band_dev[a];
        }
        // FP32 computes every convolution directly.  An FFT pass has rounding errors
        // relative to the largest value along the line, while the region fit divides
        // small local quantities (f2 = (num - kmyim) / (den - kones) where H is near 0 or
        // 1): in the Split-Bregman loop the FP32 FFT pass was 2.5e-2 away from FP64 after
        // one outer iteration, the direct pass 3.2e-5.  The direct pass is also the faster
        // one (~2(2s+1) fused operations per complex element against ~140).  The register
        // window needs a compiled half-width and 16-byte quads of pencils (dims % 4 == 0,
        // at least four pencils per group); otherwise the generic direct taps pass runs.
        // The window runs on the vector ALUs (method 1) by default.  The matrix-core form
        // (method 2, banded products) measured the same time at 256^3 (67-68 us per y or
        // z pass, 89-93 us per x pass): both passes are bound by the global and LDS
        // traffic around the arithmetic, not by the arithmetic itself.
        if (sizeof(T) == 4) {
            const AxisGeom g = make_geom(a, f->dims[0], f->dims[1], f->dims[2], L,
                                         conv_ept<T>() * 256);
            const bool quad = f->dims[0] % 4 == 0 && f->dims[1] % 4 == 0 &&
                              f->dims[2] % 4 == 0 && g.P >= 4;
            ax.direct = direct_width_supported(nh) && quad ? 1 : 0;
        } else {
            ax.direct = 0;
        }
    }
    const bool fp32_taps = sizeof(T) == 4 && !ax.direct;
    // FP64 (the parity mode) keeps the radix-4 family: the loop amplifies last-bit
    // differences of the convolution, and radix 8 moved the 64^3 result by ~1e-9
    if (fp32_taps || !fft_plan(L, ax.radix, &ax.nstage, sizeof(T) == 4 ? 16 : 4)) {
        // FP32 without a register window, or not 2/3/5-smooth: direct circular
        // convolution with the normalised taps
        ax.nstage = 0;
        const int nh = f->nh[a], s = nh / 2;
        const double hh = f->pixel ? 1. : f->h[a];
        std::vector<long double> w(nh);
        long double sum = 0;
        for (int n = -s; n <= s; ++n) {
            w[n + s] = expl(-(long double)n * n * hh * hh / (2.0L * f->sigma * f->sigma));
            sum += w[n + s];
        }
        std::vector<T> taps(nh);
        for (int n = 0; n < nh; ++n) taps[n] = (T)(w[n] / sum);
        TK_HIP(hipMalloc(&f->spec_dev[a], sizeof(T) * nh));
        TK_HIP(hipMemcpy(f->spec_dev[a], taps.data(), sizeof(T) * nh, hipMemcpyHostToDevice));
        ax.spec = f->spec_dev[a];
        ax.tw = nullptr;
        return TK_OK;
    }
    std::vector<cplx<T>> tw(L);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int t = 0; t < L; ++t) {
        long double ang = two_pi * t / L;
        tw[t].x = (T)cosl(ang);
        tw[t].y = (T)(-sinl(ang));
    }
    std::vector<T> sp(L);
    for (int k = 0; k < L; ++k) sp[k] = (T)((long double)f->spec_host[a][k] / L);
    TK_HIP(hipMalloc(&f->tw_dev[a], sizeof(cplx<T>) * L));
    TK_HIP(hipMalloc(&f->spec_dev[a], sizeof(T) * L));
    TK_HIP(hipMemcpy(f->tw_dev[a], tw.data(), sizeof(cplx<T>) * L, hipMemcpyHostToDevice));
    TK_HIP(hipMemcpy(f->spec_dev[a], sp.data(), sizeof(T) * L, hipMemcpyHostToDevice));
    ax.tw = f->tw_dev[a];
    ax.spec = f->spec_dev[a];
    return TK_OK;
}

// run the convolution passes over the axes of length > 1.  The first pass loads with
// `ld`, the last stores with `st`; intermediate passes go through the complex buffer.
template <class T, class Ld, class St>
static int conv_passes(tk_ctx ctx, const tk_filter_s* f, Ld ld, St st, bool acyclic)
{
    int axes[3], na = 0;
    for (int a = 0; a < 3; ++a)
        if (f->dims[a] > 1) axes[na++] = a;
    TK_CHECK(na > 0, TK_ERR_UNSUPPORTED, "conv: 1x1x1 volume");
    const long n = (long)f->dims[0] * f->dims[1] * f->dims[2];
    void* wk;
    TK_TRY(ctx_work(ctx, 0, sizeof(cplx<T>) * n, &wk));
    cplx<T>* C = (cplx<T>*)wk;
    for (int q = 0; q < na; ++q) {
        const bool first = q == 0, last = q == na - 1;
        if (first && last)
            TK_TRY((launch_conv_axis<T>(ctx, f, axes[q], 1, ld, MdNone{}, st, acyclic)));
        else if (first)
            TK_TRY((launch_conv_axis<T>(ctx, f, axes[q], 1, ld, MdNone{}, StComplex<T>{C},
                                        acyclic)));
        else if (last)
            TK_TRY((launch_conv_axis<T>(ctx, f, axes[q], 1, LdComplex<T>{C}, MdNone{}, st,
                                        acyclic)));
        else
            TK_TRY((launch_conv_axis<T>(ctx, f, axes[q], 1, LdComplex<T>{C}, MdNone{},
                                        StComplex<T>{C}, acyclic)));
    }
    return TK_OK;
}

template <class T>
static int conv_pair(tk_ctx ctx, const tk_filter_s* f, tk_vol oa, tk_vol ob, tk_vol ia, tk_vol ib,
                     bool acyclic)
{
    LdPair<T> ld{ia ? (const T*)ia->d : nullptr, ib ? (const T*)ib->d : nullptr, (T)0, (T)0};
    StPair<T> st{oa ? (T*)oa->d : nullptr, ob ? (T*)ob->d : nullptr};
    return conv_passes<T>(ctx, f, ld, st, acyclic);
}

}  // namespace tk

using namespace tk;

extern "C" {

int tk_filter_gaussian(tk_ctx ctx, int nx, int ny, int nz, double hx, double hy, double hz,
                       double sigma, double radius, int pixel_approach, tk_dtype dtype,
                       tk_filter* out)
{
    TK_CHECK(ctx && out, TK_ERR_ARG, "tk_filter_gaussian: null argument");
    TK_CHECK(nx > 0 && ny > 0 && nz > 0, TK_ERR_ARG, "tk_filter_gaussian: bad dims");
    TK_CHECK(sigma > 0 && radius > 0, TK_ERR_ARG, "tk_filter_gaussian: sigma/radius <= 0");
    TK_CHECK(dtype == TK_F32 || dtype == TK_F64, TK_ERR_DTYPE, "tk_filter_gaussian: dtype");
    tk_filter_s* f = new tk_filter_s();
    f->ctx = ctx;
    f->dims[0] = nx; f->dims[1] = ny; f->dims[2] = nz;
    f->h[0] = hx; f->h[1] = hy; f->h[2] = hz;
    f->sigma = sigma;
    f->radius = radius;
    f->pixel = pixel_approach;
    f->dtype = dtype;
    const double hh[3] = {pixel_approach ? 1. : hx, pixel_approach ? 1. : hy,
                          pixel_approach ? 1. : hz};
    for (int a = 0; a < 3; ++a) {
        f->nh[a] = taps(a == 2 && nz == 1, radius, sigma, hh[a]);
        axis_spectrum(f->dims[a], f->nh[a], sigma, hh[a], f->spec_host[a]);
        int r = dtype == TK_F64 ? upload_axis<double>(f, a) : upload_axis<float>(f, a);
        if (r != TK_OK) {
            tk_filter_free(f);
            return r;
        }
    }
    *out = f;
    return TK_OK;
}

int tk_filter_free(tk_filter f)
{
    if (!f) return TK_OK;
    for (int a = 0; a < 3; ++a) {
        if (f->tw_dev[a]) (void)hipFree(f->tw_dev[a]);
        if (f->spec_dev[a]) (void)hipFree(f->spec_dev[a]);
        if (f->half_dev[a]) (void)hipFree(f->half_dev[a]);
        if (f->band_dev[a]) (void)hipFree(f->band_dev[a]);
        if (f->taps_dev[a]) (void)hipFree(f->taps_dev[a]);
    }
    delete f;
    return TK_OK;
}

int tk_filter_taps(tk_filter f, int nh[3])
{
    TK_CHECK(f && nh, TK_ERR_ARG, "tk_filter_taps: null argument");
    for (int a = 0; a < 3; ++a) nh[a] = f->nh[a];
    return TK_OK;
}

int tk_filter_axis_spectrum(tk_filter f, int axis, double* out)
{
    TK_CHECK(f && out && axis >= 0 && axis < 3, TK_ERR_ARG, "tk_filter_axis_spectrum: argument");
    for (int k = 0; k < f->dims[axis]; ++k) out[k] = f->spec_host[axis][k];
    return TK_OK;
}

int tk_filter_get_method(tk_filter f, int axis, int* method)
{
    TK_CHECK(f && method && axis >= 0 && axis < 3, TK_ERR_ARG, "tk_filter_get_method: argument");
    *method = f->ax[axis].direct;
    return TK_OK;
}

int tk_filter_set_method(tk_filter f, int axis, int method)
{
    TK_CHECK(f && axis >= 0 && axis < 3 && method >= 0 && method <= 2, TK_ERR_ARG,
             "tk_filter_set_method: argument");
    const AxisDev& ax = f->ax[axis];
    if (method >= 1)
        TK_CHECK(direct_width_supported(ax.nh) && f->dims[0] % 4 == 0 && f->dims[1] % 4 == 0 &&
                     f->dims[2] % 4 == 0,
                 TK_ERR_UNSUPPORTED, "tk_filter_set_method: no direct path for %d taps", ax.nh);
    if (method == 2)
        TK_CHECK(f->dtype == TK_F32 && mfma_window_fits(axis, f->dims), TK_ERR_UNSUPPORTED,
                 "tk_filter_set_method: the matrix-core window needs FP32, a length that is a "
                 "multiple of 16 and groups of a multiple of 16 pencils (axis %d, length %d)",
                 axis, ax.L);
    if (method == 0)
        TK_CHECK(ax.nstage > 0 || ax.L == 1, TK_ERR_UNSUPPORTED,
                 "tk_filter_set_method: length %d has no FFT plan", ax.L);
    f->ax[axis].direct = method;
    return TK_OK;
}

}  // extern "C"

template <class T>
static int autotune(tk_ctx ctx, tk_filter_s* f, double* ms)
{
    const long n = (long)f->dims[0] * f->dims[1] * f->dims[2];
    void* wk;
    TK_TRY(ctx_work(ctx, 0, sizeof(cplx<T>) * n, &wk));
    cplx<T>* C = (cplx<T>*)wk;
    hipEvent_t e0, e1;
    TK_HIP(hipEventCreate(&e0));
    TK_HIP(hipEventCreate(&e1));
    int r = TK_OK;
    for (int a = 0; a < 3 && r == TK_OK; ++a) {
        double t[3] = {-1, -1, -1};
        const int keep = f->ax[a].direct;
        for (int m = 0; m < 3 && r == TK_OK; ++m) {
            if (f->ax[a].L == 1 || tk_filter_set_method(f, a, m) != TK_OK) continue;
            // one warm-up pass, then three timed passes, complex -> complex in place
            for (int rep = 0; rep < 4 && r == TK_OK; ++rep) {
                if (rep == 1) (void)hipEventRecord(e0, ctx->stream);
                r = launch_conv_axis<T>(ctx, f, a, 1, LdComplex<T>{C}, MdNone{}, StComplex<T>{C});
            }
            (void)hipEventRecord(e1, ctx->stream);
            (void)hipEventSynchronize(e1);
            float el = 0;
            (void)hipEventElapsedTime(&el, e0, e1);
            t[m] = el / 3.0;
        }
        f->ax[a].direct = keep;
        int best = -1;
        for (int m = 0; m < 3; ++m)
            if (t[m] >= 0 && (best < 0 || t[m] < t[best])) best = m;
        if (best >= 0) f->ax[a].direct = best;
        if (ms)
            for (int m = 0; m < 3; ++m) ms[3 * a + m] = t[m];
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return r;
}

extern "C" {

int tk_filter_autotune(tk_ctx ctx, tk_filter f, double* ms)
{
    TK_CHECK(ctx && f, TK_ERR_ARG, "tk_filter_autotune: null argument");
    return f->dtype == TK_F64 ? autotune<double>(ctx, f, ms) : autotune<float>(ctx, f, ms);
}

static int check_conv(tk_filter f, tk_vol v)
{
    TK_CHECK(v && v->d, TK_ERR_ARG, "conv: null volume");
    TK_CHECK(v->nx == f->dims[0] && v->ny == f->dims[1] && v->nz == f->dims[2], TK_ERR_SHAPE,
             "conv: volume %dx%dx%d vs filter %dx%dx%d", v->nx, v->ny, v->nz, f->dims[0],
             f->dims[1], f->dims[2]);
    TK_CHECK(v->dtype == f->dtype, TK_ERR_DTYPE, "conv: dtype differs from the filter's");
    return TK_OK;
}

int tk_conv_cyclic(tk_ctx ctx, tk_filter f, tk_vol out, tk_vol in)
{
    TK_CHECK(ctx && f, TK_ERR_ARG, "tk_conv_cyclic: null argument");
    TK_TRY(check_conv(f, out));
    TK_TRY(check_conv(f, in));
    if (f->dtype == TK_F64) return conv_pair<double>(ctx, f, out, nullptr, in, nullptr, false);
    return conv_pair<float>(ctx, f, out, nullptr, in, nullptr, false);
}

int tk_conv_cyclic_pair(tk_ctx ctx, tk_filter f, tk_vol oa, tk_vol ob, tk_vol ia, tk_vol ib)
{
    TK_CHECK(ctx && f, TK_ERR_ARG, "tk_conv_cyclic_pair: null argument");
    TK_TRY(check_conv(f, oa));
    TK_TRY(check_conv(f, ob));
    TK_TRY(check_conv(f, ia));
    TK_TRY(check_conv(f, ib));
    if (f->dtype == TK_F64) return conv_pair<double>(ctx, f, oa, ob, ia, ib, false);
    return conv_pair<float>(ctx, f, oa, ob, ia, ib, false);
}

int tk_conv_acyclic(tk_ctx ctx, tk_filter f, tk_vol out, tk_vol in)
{
    TK_CHECK(ctx && f, TK_ERR_ARG, "tk_conv_acyclic: null argument");
    TK_TRY(check_conv(f, out));
    TK_TRY(check_conv(f, in));
    if (f->dtype == TK_F64) return conv_pair<double>(ctx, f, out, nullptr, in, nullptr, true);
    return conv_pair<float>(ctx, f, out, nullptr, in, nullptr, true);
}

int tk_conv_acyclic_pair(tk_ctx ctx, tk_filter f, tk_vol oa, tk_vol ob, tk_vol ia, tk_vol ib)
{
    TK_CHECK(ctx && f, TK_ERR_ARG, "tk_conv_acyclic_pair: null argument");
    TK_TRY(check_conv(f, oa));
    TK_TRY(check_conv(f, ob));
    TK_TRY(check_conv(f, ia));
    TK_TRY(check_conv(f, ib));
    if (f->dtype == TK_F64) return conv_pair<double>(ctx, f, oa, ob, ia, ib, true);
    return conv_pair<float>(ctx, f, oa, ob, ia, ib, true);
}

}  // extern "C"
