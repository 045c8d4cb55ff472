// tk_noalloc.hip -- tk_norm_l2_diff for callers that cannot allocate.
//
// tk_norm_l2_diff reduces through per-block partial sums in the context's device scratch;
// this variant keeps its partials in a file-scope device array instead, takes the caller's
// stream and writes its two results to caller-owned device memory, so a call allocates
// nothing (no hipMalloc, hipFree or allocating library routine) and can be captured in a
// HIP graph.  Every call on one GPU shares that GPU's array: calls that overlap on one GPU
// race.  Each block writes its own partial pair before the final kernel reads it, so nothing
// has to start at zero.
#include <hip/hip_runtime.h>

#include <cmath>
#include <mutex>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {
constexpr int kNaParts = 16384;   // partial pairs in the scratch array (256 KiB)
constexpr int kNaThreads = 256;
constexpr int kNaPerThread = 32;  // elements per thread: a block covers 8192 voxels
constexpr long kNaMaxVoxels = (long)kNaParts * kNaThreads * kNaPerThread;  // 134,217,728
}  // namespace tk

// the partial sums of one call: [2 b] = sum (a - b)^2 and [2 b + 1] = sum b^2 over block b
static __device__ __align__(256) double scratch[2 * tk::kNaParts];

namespace tk {
namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
double* g_na_scratch[kNaMaxDevices] = {};  // each device's address of `scratch`

// block bi sums voxels [bi * 8192, (bi + 1) * 8192), lanes 256 apart (coalesced)
template <class T>
__global__ __launch_bounds__(kNaThreads) void k_na_partial(const T* __restrict__ a,
                                                           const T* __restrict__ b, long n,
                                                           double* __restrict__ part)
{
    const long base = (long)blockIdx.x * kNaThreads * kNaPerThread + threadIdx.x;
    double s1 = 0, s2 = 0;
#pragma unroll 4
    for (int u = 0; u < kNaPerThread; ++u) {
        const long i = base + (long)u * kNaThreads;
        if (i < n) {
            const double x = (double)a[i];
            if (b) {
                const double y = (double)b[i], d = x - y;
                s1 += d * d;
                s2 += y * y;
            } else {
                s1 += x * x;
            }
        }
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

// fixed-order sum of the nb partial pairs: out[0] = sqrt(h3 s1), out[1] = sqrt(h3 s2)
__global__ __launch_bounds__(1024) void k_na_final(const double* __restrict__ part, int nb,
                                                   double h3, double* __restrict__ out)
{
    double s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        s1 += part[2 * i];
        s2 += part[2 * i + 1];
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        out[0] = sqrt(h3 * s1);
        out[1] = sqrt(h3 * s2);
    }
}

}  // namespace

// the current device's address of `scratch`, looked up once per device with that device
// current (in host code the array's name is only a stand-in for it)
int noalloc_setup(int device, double** out)
{
    TK_CHECK(device >= 0 && device < kNaMaxDevices, TK_ERR_ARG,
             "tk_norm_l2_diff_noalloc: device %d beyond the %d supported", device, kNaMaxDevices);
    std::lock_guard<std::mutex> lk(g_na_mu);
    if (!g_na_scratch[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        TK_CHECK(e == hipSuccess && p, TK_ERR_HIP,
                 "tk_norm_l2_diff_noalloc: hipGetSymbolAddress of the scratch array on device "
                 "%d: %s", device, hipGetErrorString(e));
        g_na_scratch[device] = (double*)p;
    }
    if (out) *out = g_na_scratch[device];
    return TK_OK;
}

}  // namespace tk

using namespace tk;

extern "C" {

int tk_norm_l2_diff_noalloc(void* hip_stream, const void* a, const void* b, long n,
                            tk_dtype dtype, double hx, double hy, double hz, double* out_device)
{
    TK_CHECK(a && out_device, TK_ERR_ARG, "tk_norm_l2_diff_noalloc: null a or out_device");
    TK_CHECK(dtype == TK_F32 || dtype == TK_F64, TK_ERR_DTYPE,
             "tk_norm_l2_diff_noalloc: dtype must be TK_F32 or TK_F64");
    TK_CHECK(n >= 1 && n <= kNaMaxVoxels, TK_ERR_ARG,
             "tk_norm_l2_diff_noalloc: n = %ld voxels; 1 to %ld voxels are supported (the "
             "scratch array holds %d partial sums of %d voxels)",
             n, kNaMaxVoxels, kNaParts, kNaThreads * kNaPerThread);
    int dev = 0;
    TK_HIP(hipGetDevice(&dev));
    double* part = nullptr;
    TK_TRY(noalloc_setup(dev, &part));
    const hipStream_t st = (hipStream_t)hip_stream;
    const int nb = (int)((n + (long)kNaThreads * kNaPerThread - 1) / ((long)kNaThreads * kNaPerThread));
    if (dtype == TK_F64)
        hipLaunchKernelGGL(k_na_partial<double>, dim3(nb), dim3(kNaThreads), 0, st,
                           (const double*)a, (const double*)b, n, part);
    else
        hipLaunchKernelGGL(k_na_partial<float>, dim3(nb), dim3(kNaThreads), 0, st,
                           (const float*)a, (const float*)b, n, part);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_na_final, dim3(1), dim3(1024), 0, st, (const double*)part, nb,
                       hx * hy * hz, out_device);
    TK_LAUNCHED();
    return TK_OK;
}

}  // extern "C"
