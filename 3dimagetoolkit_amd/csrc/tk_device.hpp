// tk_device.hpp -- device helpers shared by the kernels of libtk.so.
#pragma once

#include <hip/hip_runtime.h>

namespace tk {

constexpr int kWave = 64;

// wave64 sum of a double (all lanes get the result)
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ inline float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}

// block sum of two doubles; result valid in thread 0.  blockDim.x <= 1024.
__device__ inline void block_sum2(double& a, double& b)
{
    __shared__ double sa[16], sb[16];
    a = wave_sum(a);
    b = wave_sum(b);
    const int w = threadIdx.x / kWave, l = threadIdx.x % kWave;
    const int nw = (blockDim.x + kWave - 1) / kWave;
    __syncthreads();
    if (l == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double x = 0, y = 0;
        for (int i = 0; i < nw; ++i) { x += sa[i]; y += sb[i]; }
        a = x;
        b = y;
    }
}

template <class T>
__device__ inline T load_nt(const T* p)
{
    return __builtin_nontemporal_load(p);
}

}  // namespace tk
