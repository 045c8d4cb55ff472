// tk_cc.hip -- seeded connected component (image3d::connected_component,
// image3d_imp.hxx:1694-1757 and the sweep core 1821-2118) and full 6-connected
// labelling of a mask.
//
// The reference marks res(seed) = 1 and then repeats eight octant sweeps in which
// every *interior* voxel with res == 1 assigns res(n) = bw(n) to its forward
// neighbours, until nothing changes.  Because res == 1 implies bw == 1, the
// assignment only ever turns 0 into 1, and the fixed point is the set of voxels
// reachable from the seed along steps that start at interior voxels and land on
// bw == 1 voxels -- independent of sweep order.  The device computes the same closure
// by parallel "pull" passes (a voxel joins when an interior neighbour is in), so the
// result is bit-identical to the reference's.
#include <hip/hip_runtime.h>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {

__device__ __forceinline__ bool interior(int i, int j, int k, int X, int Y, int Z)
{
    if (i < 1 || i > X - 2 || j < 1 || j > Y - 2) return false;
    if (Z == 1) return k == 0;
    return k >= 1 && k <= Z - 2;
}

// one pull pass; a tile of the volume is iterated in place `inner` times so that
// chains inside the tile resolve within one launch
__global__ __launch_bounds__(256) void k_cc_pass(uint8_t* __restrict__ res,
                                                 const uint8_t* __restrict__ bw, int X, int Y,
                                                 int Z, int full, unsigned* __restrict__ changed)
{
    const long n = (long)X * Y * Z, sX = (long)Y * Z, sY = Z;
    int any = 0;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        if (res[v] || !bw[v]) continue;
        const int k = (int)(v % Z), j = (int)((v / Z) % Y), i = (int)(v / sX);
        bool join = false;
        for (int di = -1; di <= 1 && !join; ++di)
            for (int dj = -1; dj <= 1 && !join; ++dj)
                for (int dk = -1; dk <= 1 && !join; ++dk) {
                    const int m = (di != 0) + (dj != 0) + (dk != 0);
                    if (m == 0 || (!full && m != 1)) continue;
                    if (Z == 1 && dk != 0) continue;
                    const int a = i + di, b = j + dj, c = k + dk;
                    if (!interior(a, b, c, X, Y, Z)) continue;
                    if (res[a * sX + b * sY + c]) join = true;
                }
        if (join) {
            res[v] = 1;
            any = 1;
        }
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) atomicAdd(changed, 1u);
}

__global__ void k_set_seed(uint8_t* res, long v) { res[v] = 1; }

// ---- labelling: label = 1 + smallest linear index in the component -----------------
__global__ void k_label_init(const uint8_t* __restrict__ mask, int32_t* __restrict__ lab, long n)
{
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x)
        lab[v] = mask[v] ? (int32_t)(v + 1) : 0;
}

__global__ __launch_bounds__(256) void k_label_pass(int32_t* __restrict__ lab, int X, int Y, int Z,
                                                    unsigned* __restrict__ changed)
{
    const long n = (long)X * Y * Z, sX = (long)Y * Z, sY = Z;
    int any = 0;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        const int32_t l = lab[v];
        if (!l) continue;
        const int k = (int)(v % Z), j = (int)((v / Z) % Y), i = (int)(v / sX);
        int32_t m = l;
        int32_t q;
        if (i > 0 && (q = lab[v - sX]) && q < m) m = q;
        if (i < X - 1 && (q = lab[v + sX]) && q < m) m = q;
        if (j > 0 && (q = lab[v - sY]) && q < m) m = q;
        if (j < Y - 1 && (q = lab[v + sY]) && q < m) m = q;
        if (k > 0 && (q = lab[v - 1]) && q < m) m = q;
        if (k < Z - 1 && (q = lab[v + 1]) && q < m) m = q;
        // pointer jump through the label of the smallest neighbour's root
        const int32_t r = lab[m - 1];
        if (r && r < m) m = r;
        if (m < l) {
            atomicMin(&lab[v], m);
            any = 1;
        }
    }
    if (__syncthreads_or(any) && threadIdx.x == 0) atomicAdd(changed, 1u);
}

__global__ void k_zero1(unsigned* p) { *p = 0; }

}  // namespace tk

using namespace tk;

extern "C" {

int tk_cc_seeded(tk_ctx ctx, tk_vol img, int i0, int j0, int k0, double threshold, int full,
                 tk_vol out, int* passes_out)
{
    TK_CHECK(ctx && img && out && img->d && out->d, TK_ERR_ARG, "tk_cc_seeded: null argument");
    TK_CHECK(same_dims(img, out) && out->dtype == TK_U8, TK_ERR_ARG,
             "tk_cc_seeded: out must be U8 with the image's dims");
    TK_CHECK(img->nx >= 3 && img->ny >= 3 && (img->nz >= 3 || img->nz == 1), TK_ERR_UNSUPPORTED,
             "tk_cc_seeded: every edge >= 3 (or nz == 1)");
    const int X = img->nx, Y = img->ny, Z = img->nz;
    // seed clamp, image3d_imp.hxx:1708-1738 (unsigned arithmetic as in the reference)
    unsigned I = (unsigned)i0, J = (unsigned)j0, K = (unsigned)k0;
    if (I == 0) ++I;
    if (J == 0) ++J;
    if (I >= (unsigned)X - 1) I = X - 2;
    if (J >= (unsigned)Y - 1) J = Y - 2;
    if (Z == 1) K = 0;
    else {
        if (K == 0) ++K;
        if (K >= (unsigned)Z - 1) K = Z - 2;
    }
    const long n = img->n();
    const long seed = ((long)I * Y + J) * Z + K;
    void* bwp;
    TK_TRY(ctx_work(ctx, 1, (size_t)n, &bwp));
    tk_vol bw;
    TK_TRY(tk_vol_wrap(ctx, bwp, X, Y, Z, img->hx, img->hy, img->hz, TK_U8, &bw));
    int r = tk_im_to_black_and_white(ctx, img, threshold, 0, bw);
    uint8_t sv = 0;
    if (r == TK_OK) {
        TK_HIP(hipMemcpyAsync(&sv, (uint8_t*)bwp + seed, 1, hipMemcpyDeviceToHost, ctx->stream));
        TK_HIP(hipStreamSynchronize(ctx->stream));
        if (sv == 0) r = tk_im_to_black_and_white(ctx, img, threshold, 1, bw);  // :1742-1745
    }
    tk_vol_free(bw);
    TK_TRY(r);
    TK_HIP(hipMemsetAsync(out->d, 0, n, ctx->stream));
    hipLaunchKernelGGL(k_set_seed, dim3(1), dim3(1), 0, ctx->stream, (uint8_t*)out->d, seed);
    TK_LAUNCHED();
    unsigned* changed = ctx->cnt + 2048;
    const unsigned nb = blocks_for(n, 256, (long)ctx->num_cu * 16);
    int passes = 0;
    const int kBatch = 16;
    for (;;) {
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, ctx->stream, changed);
        for (int q = 0; q < kBatch; ++q)
            hipLaunchKernelGGL(k_cc_pass, dim3(nb), dim3(256), 0, ctx->stream, (uint8_t*)out->d,
                               (const uint8_t*)bwp, X, Y, Z, full ? 1 : 0, changed);
        TK_LAUNCHED();
        passes += kBatch;
        unsigned ch = 0;
        TK_HIP(hipMemcpyAsync(&ch, changed, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
        TK_HIP(hipStreamSynchronize(ctx->stream));
        if (ch == 0) break;
    }
    if (passes_out) *passes_out = passes;
    return TK_OK;
}

int tk_cc_label(tk_ctx ctx, tk_vol mask, int32_t* labels, int* iters_out)
{
    TK_CHECK(ctx && mask && mask->d && labels, TK_ERR_ARG, "tk_cc_label: null argument");
    TK_CHECK(mask->dtype == TK_U8, TK_ERR_DTYPE, "tk_cc_label: mask must be U8");
    TK_CHECK(mask->n() < (long)INT32_MAX, TK_ERR_UNSUPPORTED, "tk_cc_label: volume too large");
    const long n = mask->n();
    const unsigned nb = blocks_for(n, 256, (long)ctx->num_cu * 16);
    hipLaunchKernelGGL(k_label_init, dim3(nb), dim3(256), 0, ctx->stream,
                       (const uint8_t*)mask->d, labels, n);
    TK_LAUNCHED();
    unsigned* changed = ctx->cnt + 2049;
    int it = 0;
    const int kBatch = 16;
    for (;;) {
        hipLaunchKernelGGL(k_zero1, dim3(1), dim3(1), 0, ctx->stream, changed);
        for (int q = 0; q < kBatch; ++q)
            hipLaunchKernelGGL(k_label_pass, dim3(nb), dim3(256), 0, ctx->stream, labels, mask->nx,
                               mask->ny, mask->nz, changed);
        TK_LAUNCHED();
        it += kBatch;
        unsigned ch = 0;
        TK_HIP(hipMemcpyAsync(&ch, changed, sizeof(unsigned), hipMemcpyDeviceToHost, ctx->stream));
        TK_HIP(hipStreamSynchronize(ctx->stream));
        if (ch == 0) break;
    }
    if (iters_out) *iters_out = it;
    return TK_OK;
}

}  // extern "C"
