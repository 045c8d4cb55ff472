// tk_core.hip -- context, volumes, conversions and the reductions/pointwise
// operators of im3d::image3d used on the Split-Bregman path.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}

int ctx_work(tk_ctx ctx, int slot, size_t bytes, void** out)
{
    if (ctx->work_bytes[slot] < bytes) {
        if (ctx->work[slot]) {
            TK_HIP(hipStreamSynchronize(ctx->stream));
            TK_HIP(hipFree(ctx->work[slot]));
            ctx->work[slot] = nullptr;
            ctx->work_bytes[slot] = 0;
        }
        TK_HIP(hipMalloc(&ctx->work[slot], bytes));
        ctx->work_bytes[slot] = bytes;
    }
    *out = ctx->work[slot];
    return TK_OK;
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
template <class D, class S>
__global__ void k_convert(D* __restrict__ dst, const S* __restrict__ src, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        dst[i] = (D)src[i];
}

template <class T>
__global__ void k_fill(T* __restrict__ a, T v, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        a[i] = v;
}

// 16-byte vector copy (the kernel of a same-type volume copy)
__global__ __launch_bounds__(256) void k_copy16(const float4* __restrict__ src,
                                                float4* __restrict__ dst, long n16)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n16; i += (long)gridDim.x * 256)
        dst[i] = src[i];
}

constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;

// partial sums of (a-b)^2 and b^2 (b may be null: then a^2 and 0)
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_sumsq_partial(const T* __restrict__ a,
                                                               const T* __restrict__ b, long n,
                                                               double* __restrict__ part)
{
    double s1 = 0, s2 = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double x = (double)a[i];
        if (b) {
            double y = (double)b[i];
            double d = x - y;
            s1 += d * d;
            s2 += y * y;
        } else {
            s1 += x * x;
        }
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

// fixed-order final sum of `nb` pairs (deterministic)
__global__ __launch_bounds__(1024) void k_sum_pairs(const double* __restrict__ part, int nb,
                                                    double* __restrict__ out)
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

template <class T>
__global__ __launch_bounds__(kRedThreads) void k_minmax_partial(const T* __restrict__ a, long n,
                                                                double* __restrict__ part)
{
    __shared__ double smin[kRedThreads], smax[kRedThreads];
    double mn = (double)a[0], mx = (double)a[0];
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double x = (double)a[i];
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smin[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

// min and max of the block partials: one block, strided then a tree (min and max are exact,
// so any order gives the same pair; a single lane took 91 us over 1024 partials)
__global__ __launch_bounds__(kRedThreads) void k_minmax_final(const double* __restrict__ part,
                                                              int nb, double* __restrict__ out)
{
    __shared__ double smin[kRedThreads], smax[kRedThreads];
    double mn = part[0], mx = part[1];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        mn = fmin(mn, part[2 * i]);
        mx = fmax(mx, part[2 * i + 1]);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
            smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = smin[0];
        out[1] = smax[0];
    }
}

// image3d_imp.hxx:1627-1635 as one pass: x = (x - mn) / (mx' / (max - min)) + min,
// where mx' = max of (x - mn) -- computed by the caller from the same min/max pass
template <class T>
__global__ void k_change_range(T* __restrict__ a, long n, T mn, T div, T add)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        T x = a[i];
        x -= mn;
        x /= div;
        x += add;
        a[i] = x;
    }
}

template <class T>
__global__ void k_max_shifted_partial(const T* __restrict__ a, long n, T mn, double* part)
{
    __shared__ double smax[kRedThreads];
    double mx = (double)(a[0] - mn);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        double x = (double)(T)(a[i] - mn);
        mx = x > mx ? x : mx;
    }
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = smax[0];
        part[2 * blockIdx.x + 1] = smax[0];
    }
}

// rsfe_splitbregman_imp.hxx:320-339
template <class T>
__global__ void k_cut(T* __restrict__ a, long n, T a0, T b0)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        T x = a[i];
        if (x < a0) x = a0;
        else if (x > b0) x = b0;
        a[i] = x;
    }
}

// rsfe_splitbregman_imp.hxx:174-192
template <class T>
__global__ void k_heaviside(const T* __restrict__ phi, T* __restrict__ H, long n, T center, T eps)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        H[i] = (T)(0.5 * (1. + 2. / M_PI * atan((double)((phi[i] - center) / eps))));
}

__global__ void k_heaviside_f32(const float* __restrict__ phi, float* __restrict__ H, long n,
                                float center, float eps)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        H[i] = 0.5f * (1.f + (float)(2. / M_PI) * atanf((phi[i] - center) / eps));
}

// rsfe_splitbregman_imp.hxx:343-365
template <class T>
__global__ void k_shrink(const T* __restrict__ b, const T* __restrict__ g, T* __restrict__ d, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        T f = b[i];
        T a = fabs(f) - g[i];
        T z = 0;
        d[i] = (T)((int)(f > 0) - (int)(f < 0)) * ((a < z) ? z : a);
    }
}

// image3d_imp.hxx:1591-1623
template <class T>
__global__ void k_bw(const T* __restrict__ a, uint8_t* __restrict__ out, long n, T pivot, int neg)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x)
        out[i] = (a[i] < pivot) ? (uint8_t)(0 + neg) : (uint8_t)(1 - neg);
}

// ---- image3d elementwise operators (image3d_imp.hxx:187-345) -------------------------
// a op= s with s already cast to T (static_cast<T>(addend)), in T arithmetic
template <class T>
__global__ void k_op_scalar(T* __restrict__ a, long n, T s, int op)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const T x = a[i];
        a[i] = op == TK_OP_ADD ? x + s : op == TK_OP_SUB ? x - s : op == TK_OP_MUL ? x * s : x / s;
    }
}

// a op= static_cast<T>(b) element by element
template <class T, class S>
__global__ void k_op_vol(T* __restrict__ a, const S* __restrict__ b, long n, int op)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        const T x = a[i], y = (T)b[i];
        a[i] = op == TK_OP_ADD ? x + y : op == TK_OP_SUB ? x - y : op == TK_OP_MUL ? x * y : x / y;
    }
}

// per-block FP64 partials: MODE 0: sum a * static_cast<T>(b) (scalarprod,
// image3d_imp.hxx:385-407); MODE 1: sum |a| (norm1, 1387-1398).  part[2 b] holds the sum.
template <int MODE, class T, class S>
__global__ __launch_bounds__(kRedThreads) void k_sum_partial(const T* __restrict__ a,
                                                             const S* __restrict__ b, long n,
                                                             double* __restrict__ part)
{
    double s1 = 0, s2 = 0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        if constexpr (MODE == 0) s1 += (double)(a[i] * (T)b[i]);
        else s1 += fabs((double)a[i]);
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = 0;
    }
}

// norminf as image3d_imp.hxx:1427-1452: a running value starts at |x[0]| and takes |x[i]|
// when x[i] exceeds it, so entries below it (negative ones among them) never count; the
// result is max(|x[0]|, max_i x[i]) whatever the partition.  x = a - b (b may be null: x = a);
// part[2 blk] = that maximum over the block's share, part[2 blk + 1] = the same for x = b.
template <class T>
__global__ __launch_bounds__(kRedThreads) void k_norminf_partial(const T* __restrict__ a,
                                                                 const T* __restrict__ b, long n,
                                                                 double* __restrict__ part)
{
    __shared__ double s1[kRedThreads], s2[kRedThreads];
    // the difference is formed in T, as the temporary (phi - phiold) of the reference
    const double b0 = b ? (double)b[0] : 0.0;
    double m1 = fabs(b ? (double)(a[0] - b[0]) : (double)a[0]), m2 = fabs(b0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n;
         i += (long)gridDim.x * blockDim.x) {
        const double y = b ? (double)b[i] : 0.0, x = b ? (double)(a[i] - b[i]) : (double)a[i];
        if (m1 < x) m1 = fabs(x);
        if (m2 < y) m2 = fabs(y);
    }
    s1[threadIdx.x] = m1;
    s2[threadIdx.x] = m2;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + o]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1[0];
        part[2 * blockIdx.x + 1] = s2[0];
    }
}

__global__ __launch_bounds__(kRedThreads) void k_max_pairs(const double* __restrict__ part,
                                                          int nb, double* __restrict__ out)
{
    __shared__ double s1[kRedThreads], s2[kRedThreads];
    double m1 = part[0], m2 = part[1];
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        m1 = fmax(m1, part[2 * i]);
        m2 = fmax(m2, part[2 * i + 1]);
    }
    s1[threadIdx.x] = m1;
    s2[threadIdx.x] = m2;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            s1[threadIdx.x] = fmax(s1[threadIdx.x], s1[threadIdx.x + o]);
            s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + o]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s1[0];
        out[1] = s2[0];
    }
}

// vector_abs (image3d_imp.hxx:892-943): sqrt of the sum of squares in T, 2 or 3 components
template <class T>
__global__ void k_vector_abs(const T* __restrict__ f0, const T* __restrict__ f1,
                             const T* __restrict__ f2, T* __restrict__ res, long n)
{
    for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
        T s = f0[i] * f0[i] + f1[i] * f1[i];
        if (f2) s = s + f2[i] * f2[i];
        res[i] = sqrt(s);
    }
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
int reduce_sumsq(tk_ctx ctx, tk_vol a, tk_vol b, double* s1, double* s2)
{
    long n = a->n();
    int nb = (int)blocks_for(n, kRedThreads, kRedBlocks);
    double* part = ctx->acc + 64;  // slots [64, 64 + 2*kRedBlocks)
    if (a->dtype == TK_F64)
        hipLaunchKernelGGL(k_sumsq_partial<double>, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                           (const double*)a->d, b ? (const double*)b->d : nullptr, n, part);
    else
        hipLaunchKernelGGL(k_sumsq_partial<float>, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                           (const float*)a->d, b ? (const float*)b->d : nullptr, n, part);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(1024), 0, ctx->stream, part, nb, ctx->acc);
    TK_LAUNCHED();
    TK_HIP(hipMemcpyAsync(ctx->h_acc, ctx->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    *s1 = ctx->h_acc[0];
    if (s2) *s2 = ctx->h_acc[1];
    return TK_OK;
}

static int minmax(tk_ctx ctx, tk_vol a, double* mn, double* mx)
{
    long n = a->n();
    int nb = (int)blocks_for(n, kRedThreads, kRedBlocks);
    double* part = ctx->acc + 64;
    if (a->dtype == TK_F64)
        hipLaunchKernelGGL(k_minmax_partial<double>, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                           (const double*)a->d, n, part);
    else if (a->dtype == TK_F32)
        hipLaunchKernelGGL(k_minmax_partial<float>, dim3(nb), dim3(kRedThreads), 0, ctx->stream,
                           (const float*)a->d, n, part);
    else
        TK_CHECK(false, TK_ERR_DTYPE, "minmax: U8 volume");
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kRedThreads), 0, ctx->stream, part, nb, ctx->acc);
    TK_LAUNCHED();
    TK_HIP(hipMemcpyAsync(ctx->h_acc, ctx->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    *mn = ctx->h_acc[0];
    *mx = ctx->h_acc[1];
    return TK_OK;
}

// ---- per-kernel event timing -------------------------------------------------------
constexpr size_t kMaxPending = 4096;

int timing_flush(tk_ctx ctx)
{
    if (ctx->pending.empty()) return TK_OK;
    TK_HIP(hipEventSynchronize(ctx->pending.back().b));
    for (auto& p : ctx->pending) {
        float ms = 0;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            ctx->tag_ms[p.tag] += ms;
            ctx->tag_n[p.tag] += 1;
        }
        ctx->ev_pool.push_back(p.a);
        ctx->ev_pool.push_back(p.b);
    }
    ctx->pending.clear();
    return TK_OK;
}

static hipEvent_t take_event(tk_ctx ctx)
{
    if (ctx->ev_pool.empty()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    hipEvent_t e = ctx->ev_pool.back();
    ctx->ev_pool.pop_back();
    return e;
}

int timing_begin(tk_ctx ctx, int tag)
{
    if (!ctx->timing) return -1;
    if (ctx->pending.size() >= kMaxPending) timing_flush(ctx);
    hipEvent_t a = take_event(ctx), b = take_event(ctx);
    if (!a || !b) return -1;
    if (hipEventRecord(a, ctx->stream) != hipSuccess) return -1;
    ctx->pending.push_back({a, b, tag});
    return (int)ctx->pending.size() - 1;
}

void timing_end(tk_ctx ctx, int slot)
{
    if (slot < 0 || slot >= (int)ctx->pending.size()) return;
    (void)hipEventRecord(ctx->pending[slot].b, ctx->stream);
}

int launch_cut(tk_ctx ctx, tk_vol phi, double a0, double b0)
{
    long n = phi->n();
    unsigned nb = blocks_for(n, 256, 8192);
    if (phi->dtype == TK_F64)
        hipLaunchKernelGGL(k_cut<double>, dim3(nb), dim3(256), 0, ctx->stream, (double*)phi->d, n,
                           a0, b0);
    else
        hipLaunchKernelGGL(k_cut<float>, dim3(nb), dim3(256), 0, ctx->stream, (float*)phi->d, n,
                           (float)a0, (float)b0);
    TK_LAUNCHED();
    return TK_OK;
}

}  // namespace tk

using namespace tk;

#define CHECK_VOL(v) TK_CHECK((v) && (v)->d, TK_ERR_ARG, "%s: null volume", __func__)
#define CHECK_FP(v) \
    TK_CHECK((v)->dtype == TK_F32 || (v)->dtype == TK_F64, TK_ERR_DTYPE, "%s: not FP", __func__)

extern "C" {

const char* tk_last_error(void) { return tk::g_err; }
const char* tk_version(void) { return "tk 0.1 gfx950"; }

int tk_ctx_create(int device, tk_ctx* out)
{
    TK_CHECK(out, TK_ERR_ARG, "tk_ctx_create: null out");
    int ndev = 0;
    TK_HIP(hipGetDeviceCount(&ndev));
    TK_CHECK(device >= 0 && device < ndev, TK_ERR_ARG, "tk_ctx_create: device %d of %d", device,
             ndev);
    TK_HIP(hipSetDevice(device));
    tk_ctx c = new tk_ctx_s();
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->acc, sizeof(double) * kAccSlots) != hipSuccess ||
        hipMalloc(&c->cnt, sizeof(unsigned) * kCntSlots) != hipSuccess ||
        hipHostMalloc(&c->h_acc, sizeof(double) * kAccSlots, hipHostMallocDefault) != hipSuccess) {
        set_error("tk_ctx_create: allocation failed");
        tk_ctx_destroy(c);
        return TK_ERR_NOMEM;
    }
    c->stream = c->own_stream;
    TK_HIP(hipMemsetAsync(c->acc, 0, sizeof(double) * kAccSlots, c->stream));
    TK_HIP(hipMemsetAsync(c->cnt, 0, sizeof(unsigned) * kCntSlots, c->stream));
    TK_HIP(hipStreamSynchronize(c->stream));
    // per-device set-up of the allocation-free entry point, with the device current
    const int rs = noalloc_setup(device, nullptr);
    if (rs != TK_OK) {
        tk_ctx_destroy(c);
        return rs;
    }
    *out = c;
    return TK_OK;
}

int tk_ctx_destroy(tk_ctx c)
{
    if (!c) return TK_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (int i = 0; i < 6; ++i)
        if (c->work[i]) (void)hipFree(c->work[i]);
    if (c->acc) (void)hipFree(c->acc);
    if (c->cnt) (void)hipFree(c->cnt);
    if (c->h_acc) (void)hipHostFree(c->h_acc);
    timing_flush(c);
    for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return TK_OK;
}

int tk_ctx_sync(tk_ctx c)
{
    TK_CHECK(c, TK_ERR_ARG, "tk_ctx_sync: null ctx");
    TK_HIP(hipStreamSynchronize(c->stream));
    return timing_flush(c);
}

int tk_ctx_set_timing(tk_ctx c, int enable)
{
    TK_CHECK(c, TK_ERR_ARG, "tk_ctx_set_timing: null ctx");
    TK_TRY(timing_flush(c));
    c->timing = enable != 0;
    return TK_OK;
}

int tk_ctx_set_fused_update(tk_ctx c, int enable)
{
    TK_CHECK(c, TK_ERR_ARG, "tk_ctx_set_fused_update: null ctx");
    c->sb_fused = enable != 0;
    return TK_OK;
}

int tk_ctx_kernel_times(tk_ctx c, double* ms, long* n, int reset)
{
    TK_CHECK(c, TK_ERR_ARG, "tk_ctx_kernel_times: null ctx");
    TK_TRY(timing_flush(c));
    for (int t = 0; t < TK_K_COUNT; ++t) {
        if (ms) ms[t] = c->tag_ms[t];
        if (n) n[t] = c->tag_n[t];
        if (reset) {
            c->tag_ms[t] = 0;
            c->tag_n[t] = 0;
        }
    }
    return TK_OK;
}

int tk_ctx_set_stream(tk_ctx c, void* s)
{
    TK_CHECK(c, TK_ERR_ARG, "tk_ctx_set_stream: null ctx");
    TK_HIP(hipStreamSynchronize(c->stream));
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return TK_OK;
}

void* tk_ctx_stream(tk_ctx c) { return c ? (void*)c->stream : nullptr; }

int tk_vol_alloc(tk_ctx ctx, int nx, int ny, int nz, double hx, double hy, double hz,
                 tk_dtype dtype, tk_vol* out)
{
    TK_CHECK(ctx && out, TK_ERR_ARG, "tk_vol_alloc: null argument");
    TK_CHECK(nx > 0 && ny > 0 && nz > 0, TK_ERR_ARG, "tk_vol_alloc: bad dims %d %d %d", nx, ny,
             nz);
    TK_CHECK(dtype == TK_F32 || dtype == TK_F64 || dtype == TK_U8, TK_ERR_ARG,
             "tk_vol_alloc: bad dtype");
    tk_vol v = new tk_vol_s();
    v->ctx = ctx;
    v->nx = nx; v->ny = ny; v->nz = nz;
    v->hx = hx; v->hy = hy; v->hz = hz;
    v->dtype = dtype;
    // image3d storage is value-initialised (std::vector::resize, image3d_imp.hxx:40)
    if (hipMalloc(&v->d, v->bytes()) != hipSuccess) {
        delete v;
        set_error("tk_vol_alloc: hipMalloc of %zu bytes failed", (size_t)nx * ny * nz * dtype_size(dtype));
        return TK_ERR_NOMEM;
    }
    if (hipMemsetAsync(v->d, 0, v->bytes(), ctx->stream) != hipSuccess) {
        (void)hipFree(v->d);
        delete v;
        set_error("tk_vol_alloc: memset failed");
        return TK_ERR_HIP;
    }
    *out = v;
    return TK_OK;
}

int tk_vol_wrap(tk_ctx ctx, void* p, int nx, int ny, int nz, double hx, double hy, double hz,
                tk_dtype dtype, tk_vol* out)
{
    TK_CHECK(ctx && out && p, TK_ERR_ARG, "tk_vol_wrap: null argument");
    TK_CHECK(nx > 0 && ny > 0 && nz > 0, TK_ERR_ARG, "tk_vol_wrap: bad dims");
    tk_vol v = new tk_vol_s();
    v->ctx = ctx;
    v->nx = nx; v->ny = ny; v->nz = nz;
    v->hx = hx; v->hy = hy; v->hz = hz;
    v->dtype = dtype;
    v->d = p;
    v->owned = false;
    *out = v;
    return TK_OK;
}

int tk_vol_free(tk_vol v)
{
    if (!v) return TK_OK;
    if (v->owned && v->d) {
        (void)hipStreamSynchronize(v->ctx->stream);
        (void)hipFree(v->d);
    }
    delete v;
    return TK_OK;
}

int tk_vol_set_spacing(tk_vol v, double hx, double hy, double hz)
{
    TK_CHECK(v, TK_ERR_ARG, "tk_vol_set_spacing: null volume");
    TK_CHECK(hx > 0 && hy > 0 && hz > 0, TK_ERR_ARG, "tk_vol_set_spacing: spacing must be > 0");
    v->hx = hx;
    v->hy = hy;
    v->hz = hz;
    return TK_OK;
}

int tk_vol_info(tk_vol v, int dims[3], double h[3], tk_dtype* dtype)
{
    CHECK_VOL(v);
    if (dims) { dims[0] = v->nx; dims[1] = v->ny; dims[2] = v->nz; }
    if (h) { h[0] = v->hx; h[1] = v->hy; h[2] = v->hz; }
    if (dtype) *dtype = v->dtype;
    return TK_OK;
}

void* tk_vol_device_ptr(tk_vol v) { return v ? v->d : nullptr; }

static int convert_on_device(tk_ctx ctx, void* dst, tk_dtype dt, const void* src, tk_dtype st,
                             long n)
{
    unsigned nb = blocks_for(n, 256, 8192);
    if (dt == TK_F64 && st == TK_F32)
        hipLaunchKernelGGL((k_convert<double, float>), dim3(nb), dim3(256), 0, ctx->stream,
                           (double*)dst, (const float*)src, n);
    else if (dt == TK_F32 && st == TK_F64)
        hipLaunchKernelGGL((k_convert<float, double>), dim3(nb), dim3(256), 0, ctx->stream,
                           (float*)dst, (const double*)src, n);
    else if (dt == TK_F64 && st == TK_U8)
        hipLaunchKernelGGL((k_convert<double, uint8_t>), dim3(nb), dim3(256), 0, ctx->stream,
                           (double*)dst, (const uint8_t*)src, n);
    else if (dt == TK_F32 && st == TK_U8)
        hipLaunchKernelGGL((k_convert<float, uint8_t>), dim3(nb), dim3(256), 0, ctx->stream,
                           (float*)dst, (const uint8_t*)src, n);
    else if (dt == TK_U8 && st == TK_F64)
        hipLaunchKernelGGL((k_convert<uint8_t, double>), dim3(nb), dim3(256), 0, ctx->stream,
                           (uint8_t*)dst, (const double*)src, n);
    else if (dt == TK_U8 && st == TK_F32)
        hipLaunchKernelGGL((k_convert<uint8_t, float>), dim3(nb), dim3(256), 0, ctx->stream,
                           (uint8_t*)dst, (const float*)src, n);
    else
        TK_CHECK(false, TK_ERR_DTYPE, "convert: unsupported pair %d <- %d", (int)dt, (int)st);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_vol_upload(tk_vol v, const void* host, tk_dtype hd)
{
    CHECK_VOL(v);
    TK_CHECK(host, TK_ERR_ARG, "tk_vol_upload: null host pointer");
    tk_ctx ctx = v->ctx;
    long n = v->n();
    if (hd == v->dtype) {
        TK_HIP(hipMemcpyAsync(v->d, host, v->bytes(), hipMemcpyHostToDevice, ctx->stream));
    } else {
        void* stage;
        TK_TRY(ctx_work(ctx, 3, (size_t)n * dtype_size(hd), &stage));
        TK_HIP(hipMemcpyAsync(stage, host, (size_t)n * dtype_size(hd), hipMemcpyHostToDevice,
                              ctx->stream));
        TK_TRY(convert_on_device(ctx, v->d, v->dtype, stage, hd, n));
    }
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TK_OK;
}

int tk_vol_download(tk_vol v, void* host, tk_dtype hd)
{
    CHECK_VOL(v);
    TK_CHECK(host, TK_ERR_ARG, "tk_vol_download: null host pointer");
    tk_ctx ctx = v->ctx;
    long n = v->n();
    if (hd == v->dtype) {
        TK_HIP(hipMemcpyAsync(host, v->d, v->bytes(), hipMemcpyDeviceToHost, ctx->stream));
    } else {
        void* stage;
        TK_TRY(ctx_work(ctx, 3, (size_t)n * dtype_size(hd), &stage));
        TK_TRY(convert_on_device(ctx, stage, hd, v->d, v->dtype, n));
        TK_HIP(hipMemcpyAsync(host, stage, (size_t)n * dtype_size(hd), hipMemcpyDeviceToHost,
                              ctx->stream));
    }
    TK_HIP(hipStreamSynchronize(ctx->stream));
    return TK_OK;
}

int tk_vol_fill(tk_vol v, double value)
{
    CHECK_VOL(v);
    long n = v->n();
    unsigned nb = blocks_for(n, 256, 8192);
    hipStream_t s = v->ctx->stream;
    if (v->dtype == TK_F64)
        hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(256), 0, s, (double*)v->d, value, n);
    else if (v->dtype == TK_F32)
        hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(256), 0, s, (float*)v->d, (float)value, n);
    else
        hipLaunchKernelGGL(k_fill<uint8_t>, dim3(nb), dim3(256), 0, s, (uint8_t*)v->d,
                           (uint8_t)value, n);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_vol_copy(tk_vol dst, tk_vol src)
{
    CHECK_VOL(dst);
    CHECK_VOL(src);
    TK_CHECK(same_dims(dst, src), TK_ERR_SHAPE, "tk_vol_copy: dims differ");
    if (dst->dtype == src->dtype) {
        const size_t bytes = dst->bytes();
        if (bytes % 16 == 0 && (uintptr_t)dst->d % 16 == 0 && (uintptr_t)src->d % 16 == 0) {
            const long n16 = (long)(bytes / 16);
            hipLaunchKernelGGL(k_copy16, dim3((unsigned)std::min<long>((n16 + 255) / 256, 65536)),
                               dim3(256), 0, dst->ctx->stream, (const float4*)src->d,
                               (float4*)dst->d, n16);
            TK_LAUNCHED();
            return TK_OK;
        }
        TK_HIP(hipMemcpyAsync(dst->d, src->d, bytes, hipMemcpyDeviceToDevice, dst->ctx->stream));
        return TK_OK;
    }
    return convert_on_device(dst->ctx, dst->d, dst->dtype, src->d, src->dtype, dst->n());
}

int tk_norm_l2(tk_ctx ctx, tk_vol a, double* out)
{
    CHECK_VOL(a);
    CHECK_FP(a);
    TK_CHECK(out, TK_ERR_ARG, "tk_norm_l2: null out");
    double s;
    TK_TRY(reduce_sumsq(ctx, a, nullptr, &s, nullptr));
    *out = std::sqrt(a->hx * a->hy * a->hz * s);
    return TK_OK;
}

int tk_norm_l2_diff(tk_ctx ctx, tk_vol a, tk_vol b, double* out)
{
    CHECK_VOL(a);
    CHECK_VOL(b);
    CHECK_FP(a);
    TK_CHECK(same_dims(a, b), TK_ERR_SHAPE, "tk_norm_l2_diff: dims differ");
    TK_CHECK(a->dtype == b->dtype, TK_ERR_DTYPE, "tk_norm_l2_diff: dtypes differ");
    double s1, s2;
    TK_TRY(reduce_sumsq(ctx, a, b, &s1, &s2));
    *out = std::sqrt(a->hx * a->hy * a->hz * s1);
    return TK_OK;
}

int tk_minmax(tk_ctx ctx, tk_vol a, double* mn, double* mx)
{
    CHECK_VOL(a);
    return minmax(ctx, a, mn, mx);
}

int tk_change_range_of_intensity(tk_ctx ctx, tk_vol a, double max, double min)
{
    CHECK_VOL(a);
    CHECK_FP(a);
    double mn, mx;
    TK_TRY(minmax(ctx, a, &mn, &mx));
    long n = a->n();
    unsigned nb = blocks_for(n, kRedThreads, kRedBlocks);
    // max() of the shifted image, evaluated in T like the reference
    double* part = ctx->acc + 64;
    if (a->dtype == TK_F64)
        hipLaunchKernelGGL(k_max_shifted_partial<double>, dim3(nb), dim3(kRedThreads), 0,
                           ctx->stream, (const double*)a->d, n, mn, part);
    else
        hipLaunchKernelGGL(k_max_shifted_partial<float>, dim3(nb), dim3(kRedThreads), 0,
                           ctx->stream, (const float*)a->d, n, (float)mn, part);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(kRedThreads), 0, ctx->stream, part, (int)nb, ctx->acc);
    TK_LAUNCHED();
    TK_HIP(hipMemcpyAsync(ctx->h_acc, ctx->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    double smax = ctx->h_acc[1];
    unsigned nb2 = blocks_for(n, 256, 8192);
    if (a->dtype == TK_F64) {
        double div = smax / (max - min);
        hipLaunchKernelGGL(k_change_range<double>, dim3(nb2), dim3(256), 0, ctx->stream,
                           (double*)a->d, n, mn, div, min);
    } else {
        float div = (float)smax / ((float)max - (float)min);
        hipLaunchKernelGGL(k_change_range<float>, dim3(nb2), dim3(256), 0, ctx->stream,
                           (float*)a->d, n, (float)mn, div, (float)min);
    }
    TK_LAUNCHED();
    return TK_OK;
}

int tk_cut(tk_ctx ctx, tk_vol phi, double a0, double b0)
{
    CHECK_VOL(phi);
    CHECK_FP(phi);
    return launch_cut(ctx, phi, a0, b0);
}

int tk_heaviside(tk_ctx ctx, tk_vol phi, double a0, double b0, double eps, tk_vol H)
{
    CHECK_VOL(phi);
    CHECK_VOL(H);
    CHECK_FP(phi);
    TK_CHECK(same_dims(phi, H) && phi->dtype == H->dtype, TK_ERR_SHAPE,
             "tk_heaviside: operands differ");
    long n = phi->n();
    unsigned nb = blocks_for(n, 256, 8192);
    if (phi->dtype == TK_F64)
        hipLaunchKernelGGL(k_heaviside<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double*)phi->d, (double*)H->d, n, (a0 + b0) / 2., eps);
    else
        hipLaunchKernelGGL(k_heaviside_f32, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float*)phi->d, (float*)H->d, n, ((float)a0 + (float)b0) / 2.f,
                           (float)eps);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_shrink(tk_ctx ctx, tk_vol b, tk_vol g, tk_vol d)
{
    CHECK_VOL(b);
    CHECK_VOL(g);
    CHECK_VOL(d);
    CHECK_FP(b);
    TK_CHECK(same_dims(b, g) && same_dims(b, d), TK_ERR_SHAPE, "tk_shrink: dims differ");
    TK_CHECK(b->dtype == g->dtype && b->dtype == d->dtype, TK_ERR_DTYPE, "tk_shrink: dtypes");
    long n = b->n();
    unsigned nb = blocks_for(n, 256, 8192);
    if (b->dtype == TK_F64)
        hipLaunchKernelGGL(k_shrink<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double*)b->d, (const double*)g->d, (double*)d->d, n);
    else
        hipLaunchKernelGGL(k_shrink<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float*)b->d, (const float*)g->d, (float*)d->d, n);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_im_to_black_and_white(tk_ctx ctx, tk_vol img, double threshold, int negative, tk_vol out)
{
    CHECK_VOL(img);
    CHECK_VOL(out);
    CHECK_FP(img);
    TK_CHECK(out->dtype == TK_U8 && same_dims(img, out), TK_ERR_ARG,
             "tk_im_to_black_and_white: out must be U8 of the same dims");
    if (threshold > 1 || threshold < 0) threshold = 0.5;
    double mn, mx;
    TK_TRY(minmax(ctx, img, &mn, &mx));
    long n = img->n();
    unsigned nb = blocks_for(n, 256, 8192);
    if (img->dtype == TK_F64) {
        double pivot = mn + (mx - mn) * threshold;
        hipLaunchKernelGGL(k_bw<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double*)img->d, (uint8_t*)out->d, n, pivot, negative ? 1 : 0);
    } else {
        // T pivot(min + (max - min) * threshold): the product is formed in double
        float pivot = (float)((float)mn + (double)((float)mx - (float)mn) * threshold);
        hipLaunchKernelGGL(k_bw<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float*)img->d, (uint8_t*)out->d, n, pivot, negative ? 1 : 0);
    }
    TK_LAUNCHED();
    return TK_OK;
}

// ---- R1: image3d elementwise operators -------------------------------------------------
int tk_vol_op_scalar(tk_ctx ctx, tk_vol a, int op, double s)
{
    CHECK_VOL(a);
    CHECK_FP(a);
    TK_CHECK(op >= TK_OP_ADD && op <= TK_OP_DIV, TK_ERR_ARG, "tk_vol_op_scalar: op");
    const long n = a->n();
    const unsigned nb = blocks_for(n, 256, 8192);
    if (a->dtype == TK_F64)
        hipLaunchKernelGGL(k_op_scalar<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (double*)a->d, n, s, op);
    else
        hipLaunchKernelGGL(k_op_scalar<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (float*)a->d, n, (float)s, op);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_vol_op_vol(tk_ctx ctx, tk_vol a, int op, tk_vol b)
{
    CHECK_VOL(a);
    CHECK_VOL(b);
    CHECK_FP(a);
    CHECK_FP(b);
    TK_CHECK(op >= TK_OP_ADD && op <= TK_OP_DIV, TK_ERR_ARG, "tk_vol_op_vol: op");
    TK_CHECK(same_dims(a, b), TK_ERR_SHAPE, "tk_vol_op_vol: dimensions must agree");
    const long n = a->n();
    const unsigned nb = blocks_for(n, 256, 8192);
#define TK_OPV(T, S)                                                                          \
    hipLaunchKernelGGL((k_op_vol<T, S>), dim3(nb), dim3(256), 0, ctx->stream, (T*)a->d,       \
                       (const S*)b->d, n, op)
    if (a->dtype == TK_F64) {
        if (b->dtype == TK_F64) TK_OPV(double, double);
        else TK_OPV(double, float);
    } else {
        if (b->dtype == TK_F64) TK_OPV(float, double);
        else TK_OPV(float, float);
    }
#undef TK_OPV
    TK_LAUNCHED();
    return TK_OK;
}

static int sum_final(tk_ctx ctx, int nb, double* out)
{
    hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(1024), 0, ctx->stream, ctx->acc + 64, nb,
                       ctx->acc);
    TK_LAUNCHED();
    TK_HIP(hipMemcpyAsync(ctx->h_acc, ctx->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    *out = ctx->h_acc[0];
    return TK_OK;
}

int tk_scalarprod(tk_ctx ctx, tk_vol a, tk_vol b, double* out)
{
    CHECK_VOL(a);
    CHECK_VOL(b);
    CHECK_FP(a);
    CHECK_FP(b);
    TK_CHECK(out, TK_ERR_ARG, "tk_scalarprod: null out");
    TK_CHECK(same_dims(a, b), TK_ERR_SHAPE, "tk_scalarprod: dimensions must agree");
    const long n = a->n();
    const int nb = (int)blocks_for(n, kRedThreads, kRedBlocks);
    double* part = ctx->acc + 64;
#define TK_DOT(T, S)                                                                          \
    hipLaunchKernelGGL((k_sum_partial<0, T, S>), dim3(nb), dim3(kRedThreads), 0, ctx->stream,  \
                       (const T*)a->d, (const S*)b->d, n, part)
    if (a->dtype == TK_F64) {
        if (b->dtype == TK_F64) TK_DOT(double, double);
        else TK_DOT(double, float);
    } else {
        if (b->dtype == TK_F64) TK_DOT(float, double);
        else TK_DOT(float, float);
    }
#undef TK_DOT
    TK_LAUNCHED();
    return sum_final(ctx, nb, out);
}

int tk_norm1(tk_ctx ctx, tk_vol a, double* out)
{
    CHECK_VOL(a);
    CHECK_FP(a);
    TK_CHECK(out, TK_ERR_ARG, "tk_norm1: null out");
    const long n = a->n();
    const int nb = (int)blocks_for(n, kRedThreads, kRedBlocks);
    double* part = ctx->acc + 64;
    if (a->dtype == TK_F64)
        hipLaunchKernelGGL((k_sum_partial<1, double, double>), dim3(nb), dim3(kRedThreads), 0,
                           ctx->stream, (const double*)a->d, (const double*)nullptr, n, part);
    else
        hipLaunchKernelGGL((k_sum_partial<1, float, float>), dim3(nb), dim3(kRedThreads), 0,
                           ctx->stream, (const float*)a->d, (const float*)nullptr, n, part);
    TK_LAUNCHED();
    return sum_final(ctx, nb, out);
}

int tk_norminf_diff(tk_ctx ctx, tk_vol a, tk_vol b, double* out_diff, double* out_b)
{
    CHECK_VOL(a);
    CHECK_FP(a);
    TK_CHECK(out_diff, TK_ERR_ARG, "tk_norminf_diff: null out");
    if (b) {
        CHECK_VOL(b);
        TK_CHECK(same_dims(a, b), TK_ERR_SHAPE, "tk_norminf_diff: dimensions must agree");
        TK_CHECK(a->dtype == b->dtype, TK_ERR_DTYPE, "tk_norminf_diff: dtypes differ");
    }
    const long n = a->n();
    const int nb = (int)blocks_for(n, kRedThreads, kRedBlocks);
    double* part = ctx->acc + 64;
    if (a->dtype == TK_F64)
        hipLaunchKernelGGL(k_norminf_partial<double>, dim3(nb), dim3(kRedThreads), 0,
                           ctx->stream, (const double*)a->d,
                           b ? (const double*)b->d : nullptr, n, part);
    else
        hipLaunchKernelGGL(k_norminf_partial<float>, dim3(nb), dim3(kRedThreads), 0,
                           ctx->stream, (const float*)a->d, b ? (const float*)b->d : nullptr, n,
                           part);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_max_pairs, dim3(1), dim3(kRedThreads), 0, ctx->stream, part, nb,
                       ctx->acc);
    TK_LAUNCHED();
    TK_HIP(hipMemcpyAsync(ctx->h_acc, ctx->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    *out_diff = ctx->h_acc[0];
    if (out_b) *out_b = b ? ctx->h_acc[1] : 0.0;
    return TK_OK;
}

int tk_norminf(tk_ctx ctx, tk_vol a, double* out)
{
    return tk_norminf_diff(ctx, a, nullptr, out, nullptr);
}

int tk_vector_abs(tk_ctx ctx, tk_vol f0, tk_vol f1, tk_vol f2, tk_vol res)
{
    CHECK_VOL(f0);
    CHECK_VOL(f1);
    CHECK_VOL(res);
    CHECK_FP(res);
    TK_CHECK(same_dims(f0, f1) && same_dims(f0, res) && (!f2 || same_dims(f0, f2)),
             TK_ERR_SHAPE, "tk_vector_abs: dimensions of the components must agree");
    TK_CHECK(f0->dtype == res->dtype && f1->dtype == res->dtype &&
                 (!f2 || f2->dtype == res->dtype),
             TK_ERR_DTYPE, "tk_vector_abs: components and result of one dtype");
    const long n = res->n();
    const unsigned nb = blocks_for(n, 256, 8192);
    if (res->dtype == TK_F64)
        hipLaunchKernelGGL(k_vector_abs<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double*)f0->d, (const double*)f1->d,
                           f2 ? (const double*)f2->d : nullptr, (double*)res->d, n);
    else
        hipLaunchKernelGGL(k_vector_abs<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float*)f0->d, (const float*)f1->d,
                           f2 ? (const float*)f2->d : nullptr, (float*)res->d, n);
    TK_LAUNCHED();
    return TK_OK;
}

static int element_ptr(tk_vol v, int i, int j, int k, void** p)
{
    CHECK_VOL(v);
    TK_CHECK(i >= 0 && i < v->nx && j >= 0 && j < v->ny && k >= 0 && k < v->nz, TK_ERR_ARG,
             "element (%d, %d, %d) outside the volume", i, j, k);
    const long off = ((long)i * v->ny + j) * v->nz + k;
    *p = (char*)v->d + off * dtype_size(v->dtype);
    return TK_OK;
}

int tk_vol_get(tk_vol v, int i, int j, int k, double* out)
{
    void* p = nullptr;
    TK_TRY(element_ptr(v, i, j, k, &p));
    TK_CHECK(out, TK_ERR_ARG, "tk_vol_get: null out");
    union { double d; float f; uint8_t u; } x;
    TK_HIP(hipMemcpyAsync(&x, p, dtype_size(v->dtype), hipMemcpyDeviceToHost, v->ctx->stream));
    TK_HIP(hipStreamSynchronize(v->ctx->stream));
    *out = v->dtype == TK_F64 ? x.d : v->dtype == TK_F32 ? (double)x.f : (double)x.u;
    return TK_OK;
}

int tk_vol_set(tk_vol v, int i, int j, int k, double value)
{
    void* p = nullptr;
    TK_TRY(element_ptr(v, i, j, k, &p));
    union { double d; float f; uint8_t u; } x;
    if (v->dtype == TK_F64) x.d = value;
    else if (v->dtype == TK_F32) x.f = (float)value;
    else x.u = (uint8_t)value;
    TK_HIP(hipMemcpyAsync(p, &x, dtype_size(v->dtype), hipMemcpyHostToDevice, v->ctx->stream));
    TK_HIP(hipStreamSynchronize(v->ctx->stream));
    return TK_OK;
}

}  // extern "C"
