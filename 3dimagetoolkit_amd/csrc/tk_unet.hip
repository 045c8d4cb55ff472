// tk_unet.hip -- the learned Poisson warm start: a 3-level 3-D U-Net (16/32/64 channels,
// 3x3x3 convolutions, BF16 weights and activations, FP32 accumulation on the matrix
// cores), its forward pass, backward pass and Adam training step, all on the device with
// the layers of tk_nn.hip.
//
// Network (levels L1 = nx*ny*nz, L2 = L1/8, L3 = L1/64; every edge a multiple of 4):
//   e1a = relu(conv(in[1] -> 16))          e1b = relu(conv(e1a -> 16))   -> cat1[16:32]
//   p1 = avgpool(e1b)                      e2a = relu(conv(p1 -> 32))
//   e2b = relu(conv(e2a -> 32)) -> cat2[32:64]
//   p2 = avgpool(e2b)                      e3a = relu(conv(p2 -> 64))    e3b = relu(conv(e3a -> 64))
//   u2 = upsample(e3b)                     d2a = relu(conv(u2 -> 32))    -> cat2[0:32]
//   d2b = relu(conv(cat2[64] -> 32))       u1 = upsample(d2b)
//   d1a = relu(conv(u1 -> 16)) -> cat1[0:16]
//   d1b = relu(conv(cat1[32] -> 16))       out = conv(d1b -> 1), FP32
// Input in = rhs / rms(rhs); target t = phi* / (rms(rhs) * out_scale); the guess is
// out * out_scale * rms(rhs).  The solution operator of the Poisson problem is linear,
// so the normalisation by rms(rhs) loses nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {

__device__ __forceinline__ float ubf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
__device__ __forceinline__ uint16_t uf2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

// in[v] = bf16(src[v] * s)
template <class T>
__global__ void k_unet_input(const T* __restrict__ src, long n, double s, uint16_t* __restrict__ in)
{
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256)
        in[v] = uf2bf((float)((double)src[v] * s));
}

// dst[v] = out[v] * s
template <class T>
__global__ void k_unet_output(const float* __restrict__ out, long n, double s, T* __restrict__ dst)
{
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256)
        dst[v] = (T)((double)out[v] * s);
}

// coarse[c] = mean of the f x f x f fine voxels of coarse voxel c (fine voxels f*c .. f*c+f-1
// per axis); grid (z blocks, cy, cx), one thread per coarse voxel, FP64 sum
template <class T>
__global__ __launch_bounds__(256) void k_restrict_avg(const T* __restrict__ fine,
                                                      T* __restrict__ coarse, int cx, int cy,
                                                      int cz, int f)
{
    const int k = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if (k >= cz) return;
    const int FY = cy * f, FZ = cz * f;
    double s = 0;
    for (int a = 0; a < f; ++a)
        for (int b = 0; b < f; ++b) {
            const T* row = fine + ((long)(i * f + a) * FY + (j * f + b)) * FZ + (long)k * f;
            for (int e = 0; e < f; ++e) s += (double)row[e];
        }
    coarse[((long)i * cy + j) * cz + k] = (T)(s / ((double)f * f * f));
}

// interpolation weight along one axis: fine index q of a cell-centred coarse grid of n
// cells, f fine voxels per cell; constant beyond the first and last cell centres
__device__ __forceinline__ void prolong_w(int q, int f, int n, int& c0, double& w)
{
    double x = (q + 0.5) / f - 0.5;
    x = fmin(fmax(x, 0.0), (double)(n - 1));
    c0 = min((int)x, max(n - 2, 0));
    w = n > 1 ? x - c0 : 0.0;
}

// fine[v] += gain * trilinear(coarse) at fine voxel v; grid (z blocks, FY, FX)
template <class T>
__global__ __launch_bounds__(256) void k_prolong_add(const T* __restrict__ coarse,
                                                     T* __restrict__ fine, int cx, int cy, int cz,
                                                     int f, double gain)
{
    const int FZ = cz * f;
    const int k = blockIdx.x * 256 + threadIdx.x, j = blockIdx.y, i = blockIdx.z;
    if (k >= FZ) return;
    const long sy = cz, sx = (long)cy * cz;
    int a, b, c;
    double wa, wb, wc;
    prolong_w(i, f, cx, a, wa);
    prolong_w(j, f, cy, b, wb);
    prolong_w(k, f, cz, c, wc);
    const int da = cx > 1 ? 1 : 0, db = cy > 1 ? 1 : 0, dc = cz > 1 ? 1 : 0;
    const T* p = coarse + a * sx + b * sy + c;
    auto at = [&](int x, int y, int z) { return (double)p[x * da * sx + y * db * sy + z * dc]; };
    const double v00 = at(0, 0, 0) * (1 - wc) + at(0, 0, 1) * wc;
    const double v01 = at(0, 1, 0) * (1 - wc) + at(0, 1, 1) * wc;
    const double v10 = at(1, 0, 0) * (1 - wc) + at(1, 0, 1) * wc;
    const double v11 = at(1, 1, 0) * (1 - wc) + at(1, 1, 1) * wc;
    const double v0 = v00 * (1 - wb) + v01 * wb, v1 = v10 * (1 - wb) + v11 * wb;
    const long v = ((long)i * cy * f + j) * FZ + k;
    fine[v] = (T)((double)fine[v] + gain * (v0 * (1 - wa) + v1 * wa));
}

// MSE: partial sums of (out - t)^2 per block; g[v] = bf16(2 (out - t) / n)
template <class T>
__global__ __launch_bounds__(256) void k_unet_loss(const float* __restrict__ out,
                                                   const T* __restrict__ tgt, double ts, long n,
                                                   uint16_t* __restrict__ g,
                                                   double* __restrict__ part)
{
    double s = 0;
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const double d = (double)out[v] - (double)tgt[v] * ts;
        s += d * d;
        g[v] = uf2bf((float)(2.0 * d / (double)n));
    }
    double z = 0;
    block_sum2(s, z);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_unet_sum(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0;
        for (int i = 0; i < nb; ++i) s += part[i];
        out[0] = s;
    }
}

// Adam (Kingma & Ba 2015) on FP32 master weights; the BF16 copy is refreshed
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, uint16_t* __restrict__ w16)
{
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const float gi = g[e];
        const float mi = b1 * m[e] + (1.f - b1) * gi;
        const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
        m[e] = mi;
        v[e] = vi;
        const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
        w[e] = wi;
        if (w16) w16[e] = uf2bf(wi);
    }
}

}  // namespace tk

using namespace tk;

namespace {

// activation buffers (BF16, channels-last)
enum Buf {
    B_IN, B_E1A, B_CAT1, B_D1B, B_U1,  // L1
    B_P1, B_E2A, B_CAT2, B_D2B, B_U2,  // L2
    B_P2, B_E3A, B_E3B,                // L3
    B_COUNT
};
const int kBufLevel[B_COUNT] = {1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 3, 3, 3};
const int kBufChannels[B_COUNT] = {1, 16, 32, 16, 32, 16, 32, 64, 32, 64, 32, 64, 64};

struct ConvSpec {
    int in_buf, in_off, ci;   // input slice
    int out_buf, out_off, co; // output slice (out_buf < 0: the FP32 output)
    bool relu;
    int level;
};
const ConvSpec kConv[11] = {
    {B_IN, 0, 1, B_E1A, 0, 16, true, 1},     // 0  e1a
    {B_E1A, 0, 16, B_CAT1, 16, 16, true, 1}, // 1  e1b
    {B_P1, 0, 16, B_E2A, 0, 32, true, 2},    // 2  e2a
    {B_E2A, 0, 32, B_CAT2, 32, 32, true, 2}, // 3  e2b
    {B_P2, 0, 32, B_E3A, 0, 64, true, 3},    // 4  e3a
    {B_E3A, 0, 64, B_E3B, 0, 64, true, 3},   // 5  e3b
    {B_U2, 0, 64, B_CAT2, 0, 32, true, 2},   // 6  d2a
    {B_CAT2, 0, 64, B_D2B, 0, 32, true, 2},  // 7  d2b
    {B_U1, 0, 32, B_CAT1, 0, 16, true, 1},   // 8  d1a
    {B_CAT1, 0, 32, B_D1B, 0, 16, true, 1},  // 9  d1b
    {B_D1B, 0, 16, -1, 0, 1, false, 1},      // 10 out
};

inline int kpad_of(int ci) { return (27 * ci + 15) / 16 * 16; }

struct ConvParams {
    int kpad, kpad_t;
    float *w32 = nullptr, *b32 = nullptr, *dw = nullptr, *db = nullptr;
    float *mw = nullptr, *vw = nullptr, *mb = nullptr, *vb = nullptr;
    uint16_t *w16 = nullptr, *wt16 = nullptr;
};

}  // namespace

struct tk_unet_s {
    tk_ctx ctx = nullptr;
    int dims[3] = {0, 0, 0};
    long nlev[4] = {0, 0, 0, 0};
    int ldims[4][3];
    uint16_t* act[B_COUNT] = {};
    uint16_t* grad[B_COUNT] = {};
    float* out = nullptr;     // FP32 network output (L1, 1 channel)
    uint16_t* gout = nullptr; // its gradient (BF16)
    float* part = nullptr;    // weight-gradient chunk partials
    size_t part_bytes = 0;
    double* lpart = nullptr;  // loss partials + result
    ConvParams cp[11];
    double out_scale = 1.0;
    long step = 0;
    std::vector<void*> allocs;
};

namespace {

const int kVchunk = 1024;

int dalloc(tk_unet n, void** p, size_t bytes)
{
    if (hipMalloc(p, bytes) != hipSuccess) {
        set_error("tk_unet: device allocation of %zu bytes failed", bytes);
        return TK_ERR_NOMEM;
    }
    n->allocs.push_back(*p);
    TK_HIP(hipMemsetAsync(*p, 0, bytes, n->ctx->stream));
    return TK_OK;
}

inline unsigned nb_for(long n) { return (unsigned)std::min<long>((n + 255) / 256, 8192); }

const int* ldim(tk_unet n, int lvl) { return n->ldims[lvl]; }

int conv_fwd(tk_unet n, int l)
{
    const ConvSpec& s = kConv[l];
    const int* d = ldim(n, s.level);
    const ConvParams& p = n->cp[l];
    void* y = s.out_buf < 0 ? (void*)n->out : (void*)n->act[s.out_buf];
    const int ys = s.out_buf < 0 ? 1 : kBufChannels[s.out_buf];
    return tk_nn_conv3d_fwd(n->ctx, n->act[s.in_buf], kBufChannels[s.in_buf], s.in_off, s.ci,
                            p.w16, p.kpad, p.b32, y, ys, s.out_off, s.co, d[0], d[1], d[2],
                            s.relu ? 1 : 0, s.out_buf < 0 ? 1 : 0);
}

int forward(tk_unet n)
{
    tk_ctx c = n->ctx;
    const int* d1 = ldim(n, 1);
    const int* d2 = ldim(n, 2);
    TK_TRY(conv_fwd(n, 0));
    TK_TRY(conv_fwd(n, 1));
    TK_TRY(tk_nn_avgpool2(c, n->act[B_CAT1], 32, 16, 16, d1[0], d1[1], d1[2], n->act[B_P1], 16, 0));
    TK_TRY(conv_fwd(n, 2));
    TK_TRY(conv_fwd(n, 3));
    TK_TRY(tk_nn_avgpool2(c, n->act[B_CAT2], 64, 32, 32, d2[0], d2[1], d2[2], n->act[B_P2], 32, 0));
    TK_TRY(conv_fwd(n, 4));
    TK_TRY(conv_fwd(n, 5));
    TK_TRY(tk_nn_upsample2(c, n->act[B_E3B], 64, 0, 64, d2[0], d2[1], d2[2], n->act[B_U2], 64, 0));
    TK_TRY(conv_fwd(n, 6));
    TK_TRY(conv_fwd(n, 7));
    TK_TRY(tk_nn_upsample2(c, n->act[B_D2B], 32, 0, 32, d1[0], d1[1], d1[2], n->act[B_U1], 32, 0));
    TK_TRY(conv_fwd(n, 8));
    TK_TRY(conv_fwd(n, 9));
    TK_TRY(conv_fwd(n, 10));
    return TK_OK;
}

// gradient of conv l: dy is the (already ReLU-masked) output gradient slice; writes dw, db
// and, if gx != nullptr, the input gradient into the gx slice
int conv_bwd(tk_unet n, int l, const uint16_t* dy, int ds, int dof, uint16_t* gx, int gs, int go)
{
    const ConvSpec& s = kConv[l];
    const int* d = ldim(n, s.level);
    const ConvParams& p = n->cp[l];
    TK_TRY(tk_nn_conv3d_wgrad(n->ctx, n->act[s.in_buf], kBufChannels[s.in_buf], s.in_off, s.ci,
                              dy, ds, dof, s.co, d[0], d[1], d[2], p.kpad, n->part, kVchunk,
                              p.dw, p.db));
    if (gx)
        TK_TRY(tk_nn_conv3d_fwd(n->ctx, dy, ds, dof, s.co, p.wt16, p.kpad_t, nullptr, gx, gs, go,
                                s.ci, d[0], d[1], d[2], 0, 0));
    return TK_OK;
}

int relu_mask(tk_unet n, int buf, int off, int C)
{
    const int ch = kBufChannels[buf];
    return tk_nn_relu_bwd(n->ctx, n->grad[buf], ch, off, n->act[buf], ch, off, C,
                          n->nlev[kBufLevel[buf]]);
}

int backward(tk_unet n)
{
    tk_ctx c = n->ctx;
    const int* d1 = ldim(n, 1);
    const int* d2 = ldim(n, 2);
    uint16_t** g = n->grad;
    // out = conv(d1b)
    TK_TRY(conv_bwd(n, 10, n->gout, 1, 0, g[B_D1B], 16, 0));
    TK_TRY(relu_mask(n, B_D1B, 0, 16));
    // d1b = conv(cat1)
    TK_TRY(conv_bwd(n, 9, g[B_D1B], 16, 0, g[B_CAT1], 32, 0));
    TK_TRY(relu_mask(n, B_CAT1, 0, 16));  // d1a; the e1b half waits for the pool gradient
    // d1a = conv(u1)
    TK_TRY(conv_bwd(n, 8, g[B_CAT1], 32, 0, g[B_U1], 32, 0));
    TK_TRY(tk_nn_upsample2_bwd(c, g[B_U1], 32, 0, 32, d1[0], d1[1], d1[2], g[B_D2B], 32, 0));
    TK_TRY(relu_mask(n, B_D2B, 0, 32));
    // d2b = conv(cat2)
    TK_TRY(conv_bwd(n, 7, g[B_D2B], 32, 0, g[B_CAT2], 64, 0));
    TK_TRY(relu_mask(n, B_CAT2, 0, 32));  // d2a
    // d2a = conv(u2)
    TK_TRY(conv_bwd(n, 6, g[B_CAT2], 64, 0, g[B_U2], 64, 0));
    TK_TRY(tk_nn_upsample2_bwd(c, g[B_U2], 64, 0, 64, d2[0], d2[1], d2[2], g[B_E3B], 64, 0));
    TK_TRY(relu_mask(n, B_E3B, 0, 64));
    TK_TRY(conv_bwd(n, 5, g[B_E3B], 64, 0, g[B_E3A], 64, 0));
    TK_TRY(relu_mask(n, B_E3A, 0, 64));
    TK_TRY(conv_bwd(n, 4, g[B_E3A], 64, 0, g[B_P2], 32, 0));
    // p2 = avgpool(e2b): add to the e2b half of cat2's gradient, then its ReLU
    TK_TRY(tk_nn_avgpool2_bwd(c, g[B_P2], 32, 0, 32, d2[0], d2[1], d2[2], g[B_CAT2], 64, 32, 1));
    TK_TRY(relu_mask(n, B_CAT2, 32, 32));
    TK_TRY(conv_bwd(n, 3, g[B_CAT2], 64, 32, g[B_E2A], 32, 0));
    TK_TRY(relu_mask(n, B_E2A, 0, 32));
    TK_TRY(conv_bwd(n, 2, g[B_E2A], 32, 0, g[B_P1], 16, 0));
    TK_TRY(tk_nn_avgpool2_bwd(c, g[B_P1], 16, 0, 16, d1[0], d1[1], d1[2], g[B_CAT1], 32, 16, 1));
    TK_TRY(relu_mask(n, B_CAT1, 16, 16));
    TK_TRY(conv_bwd(n, 1, g[B_CAT1], 32, 16, g[B_E1A], 16, 0));
    TK_TRY(relu_mask(n, B_E1A, 0, 16));
    TK_TRY(conv_bwd(n, 0, g[B_E1A], 16, 0, nullptr, 0, 0));
    return TK_OK;
}

int refresh_weights(tk_unet n)
{
    for (int l = 0; l < 11; ++l) {
        ConvParams& p = n->cp[l];
        const long nw = (long)kConv[l].co * p.kpad;
        TK_TRY(tk_nn_f32_to_bf16(n->ctx, p.w32, nw, p.w16));
        TK_TRY(tk_nn_flip_weights(n->ctx, p.w16, kConv[l].ci, kConv[l].co, p.kpad, p.wt16,
                                  p.kpad_t));
    }
    return TK_OK;
}

int set_input(tk_unet n, tk_vol rhs, double* rms_out)
{
    double nrm = 0;
    TK_TRY(tk_norm_l2(n->ctx, rhs, &nrm));
    const double h3 = rhs->hx * rhs->hy * rhs->hz;
    const double rms = nrm / std::sqrt(h3 * (double)rhs->n());
    *rms_out = rms;
    const double s = rms > 0 ? 1.0 / rms : 0.0;
    const long nv = rhs->n();
    if (rhs->dtype == TK_F64)
        hipLaunchKernelGGL(k_unet_input<double>, dim3(nb_for(nv)), dim3(256), 0, n->ctx->stream,
                           (const double*)rhs->d, nv, s, n->act[B_IN]);
    else
        hipLaunchKernelGGL(k_unet_input<float>, dim3(nb_for(nv)), dim3(256), 0, n->ctx->stream,
                           (const float*)rhs->d, nv, s, n->act[B_IN]);
    TK_LAUNCHED();
    return TK_OK;
}

int check_vol(tk_unet n, tk_vol v, const char* what)
{
    TK_CHECK(v && v->d, TK_ERR_ARG, "%s: null volume", what);
    TK_CHECK(v->nx == n->dims[0] && v->ny == n->dims[1] && v->nz == n->dims[2], TK_ERR_SHAPE,
             "%s: volume dims differ from the network's", what);
    TK_CHECK(v->dtype == TK_F32 || v->dtype == TK_F64, TK_ERR_DTYPE, "%s: dtype", what);
    return TK_OK;
}

}  // namespace

namespace tk {
tk_ctx unet_ctx(tk_unet n) { return n->ctx; }
void unet_dims(tk_unet n, int d[3])
{
    d[0] = n->dims[0];
    d[1] = n->dims[1];
    d[2] = n->dims[2];
}
}  // namespace tk

extern "C" {

int tk_unet_create(tk_ctx ctx, int nx, int ny, int nz, uint64_t seed, tk_unet* out)
{
    TK_CHECK(ctx && out, TK_ERR_ARG, "tk_unet_create: argument");
    TK_CHECK(nx % 4 == 0 && ny % 4 == 0 && nz % 4 == 0 && nx >= 4 && ny >= 4 && nz >= 4,
             TK_ERR_SHAPE, "tk_unet_create: every edge must be a positive multiple of 4");
    tk_unet n = new tk_unet_s();
    n->ctx = ctx;
    n->dims[0] = nx;
    n->dims[1] = ny;
    n->dims[2] = nz;
    for (int lvl = 1; lvl <= 3; ++lvl) {
        const int f = 1 << (lvl - 1);
        n->ldims[lvl][0] = nx / f;
        n->ldims[lvl][1] = ny / f;
        n->ldims[lvl][2] = nz / f;
        n->nlev[lvl] = (long)(nx / f) * (ny / f) * (nz / f);
    }
    int r = TK_OK;
    for (int b = 0; b < B_COUNT && r == TK_OK; ++b) {
        const size_t bytes = 2 * (size_t)n->nlev[kBufLevel[b]] * kBufChannels[b];
        r = dalloc(n, (void**)&n->act[b], bytes);
        if (r == TK_OK) r = dalloc(n, (void**)&n->grad[b], bytes);
    }
    if (r == TK_OK) r = dalloc(n, (void**)&n->out, 4 * (size_t)n->nlev[1]);
    if (r == TK_OK) r = dalloc(n, (void**)&n->gout, 2 * (size_t)n->nlev[1]);
    if (r == TK_OK) r = dalloc(n, (void**)&n->lpart, 8 * 8200);
    // He-normal initialisation of the FP32 master weights (fixed seed), zero bias
    std::mt19937_64 gen(seed);
    size_t part_max = 0;
    for (int l = 0; l < 11 && r == TK_OK; ++l) {
        const ConvSpec& s = kConv[l];
        ConvParams& p = n->cp[l];
        p.kpad = kpad_of(s.ci);
        p.kpad_t = kpad_of(s.co);
        const long nw = (long)s.co * p.kpad;
        const long nchunk = (n->nlev[s.level] + kVchunk - 1) / kVchunk;
        part_max = std::max(part_max, (size_t)(nchunk * nw));
        float** f32[] = {&p.w32, &p.dw, &p.mw, &p.vw};
        for (float** q : f32)
            if (r == TK_OK) r = dalloc(n, (void**)q, 4 * (size_t)nw);
        float** fb[] = {&p.b32, &p.db, &p.mb, &p.vb};
        for (float** q : fb)
            if (r == TK_OK) r = dalloc(n, (void**)q, 4 * (size_t)s.co);
        if (r == TK_OK) r = dalloc(n, (void**)&p.w16, 2 * (size_t)nw);
        if (r == TK_OK) r = dalloc(n, (void**)&p.wt16, 2 * (size_t)s.ci * p.kpad_t);
        if (r != TK_OK) break;
        std::normal_distribution<double> N(0.0, std::sqrt(2.0 / (27.0 * s.ci)));
        std::vector<float> w((size_t)nw, 0.f);
        for (int o = 0; o < s.co; ++o)
            for (int k = 0; k < 27 * s.ci; ++k) w[(size_t)o * p.kpad + k] = (float)N(gen);
        if (hipMemcpy(p.w32, w.data(), 4 * (size_t)nw, hipMemcpyHostToDevice) != hipSuccess)
            r = TK_ERR_HIP;
    }
    if (r == TK_OK) r = dalloc(n, (void**)&n->part, 4 * part_max);
    if (r == TK_OK) r = refresh_weights(n);
    if (r == TK_OK) r = tk_ctx_sync(ctx);
    if (r != TK_OK) {
        tk_unet_destroy(n);
        return r;
    }
    *out = n;
    return TK_OK;
}

int tk_unet_destroy(tk_unet n)
{
    if (!n) return TK_OK;
    (void)hipStreamSynchronize(n->ctx->stream);
    for (void* p : n->allocs) (void)hipFree(p);
    delete n;
    return TK_OK;
}

int tk_unet_num_params(tk_unet n, long* count)
{
    TK_CHECK(n && count, TK_ERR_ARG, "tk_unet_num_params: argument");
    long c = 0;
    for (int l = 0; l < 11; ++l) c += (long)kConv[l].co * (27 * kConv[l].ci + 1);
    *count = c;
    return TK_OK;
}

int tk_unet_set_out_scale(tk_unet n, double s)
{
    TK_CHECK(n && s > 0, TK_ERR_ARG, "tk_unet_set_out_scale: argument");
    n->out_scale = s;
    return TK_OK;
}

int tk_unet_get_out_scale(tk_unet n, double* s)
{
    TK_CHECK(n && s, TK_ERR_ARG, "tk_unet_get_out_scale: argument");
    *s = n->out_scale;
    return TK_OK;
}

// guess := out * out_scale * rms(rhs)
int tk_unet_predict(tk_unet n, tk_vol rhs, tk_vol guess)
{
    TK_CHECK(n, TK_ERR_ARG, "tk_unet_predict: null network");
    TK_TRY(check_vol(n, rhs, "tk_unet_predict"));
    TK_TRY(check_vol(n, guess, "tk_unet_predict"));
    double rms = 0;
    TK_TRY(set_input(n, rhs, &rms));
    TK_TRY(forward(n));
    const double s = n->out_scale * rms;
    const long nv = guess->n();
    if (guess->dtype == TK_F64)
        hipLaunchKernelGGL(k_unet_output<double>, dim3(nb_for(nv)), dim3(256), 0, n->ctx->stream,
                           n->out, nv, s, (double*)guess->d);
    else
        hipLaunchKernelGGL(k_unet_output<float>, dim3(nb_for(nv)), dim3(256), 0, n->ctx->stream,
                           n->out, nv, s, (float*)guess->d);
    TK_LAUNCHED();
    return TK_OK;
}

// one step on the pair (rhs, target); loss = mean (out - target / (rms out_scale))^2.
// lr <= 0: gradients only (no update), for checking them.
int tk_unet_train_step(tk_unet n, tk_vol rhs, tk_vol target, double lr, double* loss)
{
    TK_CHECK(n, TK_ERR_ARG, "tk_unet_train_step: null network");
    TK_TRY(check_vol(n, rhs, "tk_unet_train_step"));
    TK_TRY(check_vol(n, target, "tk_unet_train_step"));
    double rms = 0;
    TK_TRY(set_input(n, rhs, &rms));
    TK_TRY(forward(n));
    const double ts = rms > 0 ? 1.0 / (rms * n->out_scale) : 0.0;
    const long nv = n->nlev[1];
    const unsigned nb = nb_for(nv);
    if (target->dtype == TK_F64)
        hipLaunchKernelGGL(k_unet_loss<double>, dim3(nb), dim3(256), 0, n->ctx->stream, n->out,
                           (const double*)target->d, ts, nv, n->gout, n->lpart);
    else
        hipLaunchKernelGGL(k_unet_loss<float>, dim3(nb), dim3(256), 0, n->ctx->stream, n->out,
                           (const float*)target->d, ts, nv, n->gout, n->lpart);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_unet_sum, dim3(1), dim3(64), 0, n->ctx->stream, n->lpart, (int)nb,
                       n->lpart + 8192);
    TK_LAUNCHED();
    TK_TRY(backward(n));
    if (lr > 0) {
        ++n->step;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)n->step), c2 = 1.f - std::pow(b2, (float)n->step);
        for (int l = 0; l < 11; ++l) {
            ConvParams& p = n->cp[l];
            const long nw = (long)kConv[l].co * p.kpad;
            hipLaunchKernelGGL(k_adam, dim3(nb_for(nw)), dim3(256), 0, n->ctx->stream, p.w32,
                               p.dw, p.mw, p.vw, nw, (float)lr, b1, b2, 1e-8f, c1, c2, p.w16);
            TK_LAUNCHED();
            hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, n->ctx->stream, p.b32, p.db, p.mb,
                               p.vb, (long)kConv[l].co, (float)lr, b1, b2, 1e-8f, c1, c2,
                               (uint16_t*)nullptr);
            TK_LAUNCHED();
            TK_TRY(tk_nn_flip_weights(n->ctx, p.w16, kConv[l].ci, kConv[l].co, p.kpad, p.wt16,
                                      p.kpad_t));
        }
    }
    if (loss) {
        double h = 0;
        TK_HIP(hipMemcpyAsync(&h, n->lpart + 8192, 8, hipMemcpyDeviceToHost, n->ctx->stream));
        TK_HIP(hipStreamSynchronize(n->ctx->stream));
        *loss = h / (double)nv;
    }
    return TK_OK;
}

// parameters (FP32 masters) or their last gradients, layer by layer: w[co][27*ci] then b[co]
int tk_unet_get_params(tk_unet n, int which, float* host, long count)
{
    TK_CHECK(n && host && (which == 0 || which == 1), TK_ERR_ARG, "tk_unet_get_params: argument");
    long need = 0;
    TK_TRY(tk_unet_num_params(n, &need));
    TK_CHECK(count >= need, TK_ERR_ARG, "tk_unet_get_params: %ld floats needed", need);
    TK_TRY(tk_ctx_sync(n->ctx));
    long o = 0;
    for (int l = 0; l < 11; ++l) {
        const ConvParams& p = n->cp[l];
        const int ci = kConv[l].ci, co = kConv[l].co;
        std::vector<float> w((size_t)co * p.kpad);
        TK_HIP(hipMemcpy(w.data(), which ? p.dw : p.w32, 4 * w.size(), hipMemcpyDeviceToHost));
        for (int c = 0; c < co; ++c)
            for (int k = 0; k < 27 * ci; ++k) host[o++] = w[(size_t)c * p.kpad + k];
        TK_HIP(hipMemcpy(host + o, which ? p.db : p.b32, 4 * (size_t)co, hipMemcpyDeviceToHost));
        o += co;
    }
    return TK_OK;
}

int tk_unet_set_params(tk_unet n, const float* host, long count)
{
    TK_CHECK(n && host, TK_ERR_ARG, "tk_unet_set_params: argument");
    long need = 0;
    TK_TRY(tk_unet_num_params(n, &need));
    TK_CHECK(count >= need, TK_ERR_ARG, "tk_unet_set_params: %ld floats needed", need);
    TK_TRY(tk_ctx_sync(n->ctx));
    long o = 0;
    for (int l = 0; l < 11; ++l) {
        const ConvParams& p = n->cp[l];
        const int ci = kConv[l].ci, co = kConv[l].co;
        std::vector<float> w((size_t)co * p.kpad, 0.f);
        for (int c = 0; c < co; ++c)
            for (int k = 0; k < 27 * ci; ++k) w[(size_t)c * p.kpad + k] = host[o++];
        TK_HIP(hipMemcpy(p.w32, w.data(), 4 * w.size(), hipMemcpyHostToDevice));
        TK_HIP(hipMemcpy(p.b32, host + o, 4 * (size_t)co, hipMemcpyHostToDevice));
        o += co;
    }
    n->step = 0;
    for (int l = 0; l < 11; ++l) {
        const ConvParams& p = n->cp[l];
        const long nw = (long)kConv[l].co * p.kpad;
        TK_HIP(hipMemsetAsync(p.mw, 0, 4 * nw, n->ctx->stream));
        TK_HIP(hipMemsetAsync(p.vw, 0, 4 * nw, n->ctx->stream));
        TK_HIP(hipMemsetAsync(p.mb, 0, 4 * (size_t)kConv[l].co, n->ctx->stream));
        TK_HIP(hipMemsetAsync(p.vb, 0, 4 * (size_t)kConv[l].co, n->ctx->stream));
    }
    TK_TRY(refresh_weights(n));
    return tk_ctx_sync(n->ctx);
}

// the FP32 output of the last forward pass (normalised units), nx*ny*nz floats
int tk_unet_get_output(tk_unet n, float* host)
{
    TK_CHECK(n && host, TK_ERR_ARG, "tk_unet_get_output: argument");
    TK_TRY(tk_ctx_sync(n->ctx));
    TK_HIP(hipMemcpy(host, n->out, 4 * (size_t)n->nlev[1], hipMemcpyDeviceToHost));
    return TK_OK;
}

// ---- coarse-grid application of the network (no reference counterpart) -------------
static int coarse_factor(tk_vol fine, tk_vol coarse, const char* what, int* f)
{
    TK_CHECK(fine && coarse && fine->d && coarse->d, TK_ERR_ARG, "%s: null volume", what);
    TK_CHECK(fine->dtype == coarse->dtype && (fine->dtype == TK_F32 || fine->dtype == TK_F64),
             TK_ERR_DTYPE, "%s: both volumes F32 or both F64", what);
    TK_CHECK(coarse->nx > 0 && fine->nx % coarse->nx == 0, TK_ERR_SHAPE, "%s: dims", what);
    const int q = fine->nx / coarse->nx;
    TK_CHECK(q >= 1 && fine->ny == q * coarse->ny && fine->nz == q * coarse->nz, TK_ERR_SHAPE,
             "%s: fine dims must be the same integer multiple of the coarse dims on every axis",
             what);
    TK_CHECK(fine->nx <= 65535 && fine->ny <= 65535, TK_ERR_UNSUPPORTED,
             "%s: x and y edges above 65535 voxels", what);
    *f = q;
    return TK_OK;
}

int tk_vol_restrict(tk_ctx ctx, tk_vol fine, tk_vol coarse)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_vol_restrict: null context");
    int f = 0;
    TK_TRY(coarse_factor(fine, coarse, "tk_vol_restrict", &f));
    const dim3 grid((coarse->nz + 255) / 256, coarse->ny, coarse->nx);
    if (fine->dtype == TK_F64)
        hipLaunchKernelGGL(k_restrict_avg<double>, grid, dim3(256), 0, ctx->stream,
                           (const double*)fine->d, (double*)coarse->d, coarse->nx, coarse->ny,
                           coarse->nz, f);
    else
        hipLaunchKernelGGL(k_restrict_avg<float>, grid, dim3(256), 0, ctx->stream,
                           (const float*)fine->d, (float*)coarse->d, coarse->nx, coarse->ny,
                           coarse->nz, f);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_vol_prolong_add(tk_ctx ctx, tk_vol coarse, tk_vol fine, double gain)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_vol_prolong_add: null context");
    int f = 0;
    TK_TRY(coarse_factor(fine, coarse, "tk_vol_prolong_add", &f));
    const dim3 grid((fine->nz + 255) / 256, fine->ny, fine->nx);
    if (fine->dtype == TK_F64)
        hipLaunchKernelGGL(k_prolong_add<double>, grid, dim3(256), 0, ctx->stream,
                           (const double*)coarse->d, (double*)fine->d, coarse->nx, coarse->ny,
                           coarse->nz, f, gain);
    else
        hipLaunchKernelGGL(k_prolong_add<float>, grid, dim3(256), 0, ctx->stream,
                           (const float*)coarse->d, (float*)fine->d, coarse->nx, coarse->ny,
                           coarse->nz, f, gain);
    TK_LAUNCHED();
    return TK_OK;
}

}  // extern "C"
