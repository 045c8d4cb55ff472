// tk_device.hpp -- device helpers shared by the kernels of libtk.so.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace tk {

constexpr int kWave = 64;

// exact quotient of small non-negative integers (n < 2^13, d <= 2^12) through a float
// reciprocal (the hardware has no integer divide): the distance of (n + 1/2)/d to an
// integer is >= 1/(2d), and the rounding of 1/d and of the product moves it by less than
// (n + 1/2)/d * 2^-23 < 2^-10/d, so the truncation is exact
__device__ __forceinline__ int qdiv(int n, float inv_d) { return (int)(((float)n + 0.5f) * inv_d); }

// wave64 sum of a double (all lanes get the result)
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// FP32 quotient n / d from the hardware reciprocal and one Newton correction of the
// quotient (Markstein): within 1 ulp of the IEEE quotient, in 6 instructions instead of
// the IEEE sequence's 10 (for the vector-ALU-bound convolution passes).  For d normal or
// zero; a zero or infinite operand makes the correction's residual NaN, and the product
// n * (1/d) is then the IEEE quotient (+-inf, 0 or NaN).  FP64 keeps the IEEE division.
__device__ __forceinline__ float fdiv_nr(float n, float d)
{
    const float r = __builtin_amdgcn_rcpf(d);
    const float q = n * r;
    const float e = fmaf(-d, q, n);
    return e == e ? fmaf(e, r, q) : q;
}
template <class T>
__device__ __forceinline__ T div_nr(T n, T d)
{
    if constexpr (std::is_same<T, float>::value) return fdiv_nr(n, d);
    else return n / d;
}

// wave64 sum of a double through DPP row moves (a few cycles each, where the shuffles above
// go through the LDS crossbar): per row of 16 lanes, quad sums (quad_perm xor 1, xor 2),
// then row rotations by 4 and 8; lanes 0, 16, 32, 48 then hold the row sums, added as
// (r0 + r1) + (r2 + r3).  The order is fixed, so equal inputs give equal sums in every
// wave.  All 64 lanes must be active; the result is uniform.
template <int CTRL>
__device__ __forceinline__ double dpp_mov_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double readlane_f64(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ inline double wave_sum_dpp(double v)
{
    v += dpp_mov_f64<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov_f64<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov_f64<0x124>(v);  // row_ror:4
    v += dpp_mov_f64<0x128>(v);  // row_ror:8
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// block sum of two doubles with the DPP wave sum (every wave full); result valid in the
// thread with threadIdx.x == 0 of a 1-D block of at most 1024 threads
__device__ inline void block_sum2_dpp(double& a, double& b)
{
    __shared__ double sa[16], sb[16];
    a = wave_sum_dpp(a);
    b = wave_sum_dpp(b);
    const int t = threadIdx.x, w = t / kWave, nw = blockDim.x / kWave;
    __syncthreads();
    if ((t & (kWave - 1)) == 0) {
        sa[w] = a;
        sb[w] = b;
    }
    __syncthreads();
    if (t == 0) {
        double x = 0, y = 0;
        for (int i = 0; i < nw; ++i) {
            x += sa[i];
            y += sb[i];
        }
        a = x;
        b = y;
    }
}

__device__ inline float wave_min(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}

// Workgroups are dispatched round-robin over the XCDs, each with its own L2.  Renumber
// them so that XCD x runs the contiguous range of logical ids starting at x*q + min(x, r)
// (q, r = G div/mod nxcd): neighbouring tiles, which share cache lines and halos, then run
// on the same XCD.  A bijection on [0, G) for any G.
constexpr unsigned kNumXcd = 8;
__device__ __forceinline__ unsigned xcd_remap(unsigned w, unsigned G)
{
    const unsigned q = G / kNumXcd, r = G % kNumXcd, x = w % kNumXcd;
    return x * q + (x < r ? x : r) + w / kNumXcd;
}

// logical (x, y, z) block index of a 3-D grid after xcd_remap of the x-fastest flat id
__device__ __forceinline__ uint3 xcd_block3()
{
    unsigned w = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    w = xcd_remap(w, gridDim.x * gridDim.y * gridDim.z);
    uint3 b;
    b.x = w % gridDim.x;
    w /= gridDim.x;
    b.y = w % gridDim.y;
    b.z = w / gridDim.y;
    return b;
}

__device__ inline int flat_tid()
{
    return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
}

// block sum of two doubles; result valid in the thread with flat_tid() == 0.
// At most 1024 threads per block.
__device__ inline void block_sum2(double& a, double& b)
{
    __shared__ double sa[16], sb[16];
    a = wave_sum(a);
    b = wave_sum(b);
    const int t = flat_tid();
    const int w = t / kWave, l = t % kWave;
    const int nw = (blockDim.x * blockDim.y * blockDim.z + kWave - 1) / kWave;
    __syncthreads();
    if (l == 0) { sa[w] = a; sb[w] = b; }
    __syncthreads();
    if (t == 0) {
        double x = 0, y = 0;
        for (int i = 0; i < nw; ++i) { x += sa[i]; y += sb[i]; }
        a = x;
        b = y;
    }
}

// Visit the voxels of an X*Y*Z volume with a 2-D block: threads along k (contiguous),
// block rows over the X*Y lines.  One 32-bit division per line, none per voxel.
#define TK_FOR_VOXELS(X, Y, Z, i, j, k, v)                                                   \
    for (int r_ = blockIdx.y * blockDim.y + threadIdx.y; r_ < (X) * (Y);                  \
         r_ += gridDim.y * blockDim.y)                                                       \
        for (int k = blockIdx.x * blockDim.x + threadIdx.x, i = r_ / (Y), j = r_ - i * (Y);  \
             k < (Z); k += gridDim.x * blockDim.x)                                           \
            for (long v = (long)r_ * (Z) + k, once_ = 1; once_; once_ = 0)

inline dim3 vox_block(int Z) { return Z >= 64 ? dim3(64, 4) : dim3(32, 8); }
// grid with at most max_blocks workgroups (the row loop covers the rest)
inline dim3 vox_grid(int X, int Y, int Z, int max_blocks)
{
    const dim3 b = vox_block(Z);
    const int gx = (Z + b.x - 1) / b.x;
    long gy = ((long)X * Y + b.y - 1) / b.y;
    long cap = max_blocks / gx;
    if (cap < 1) cap = 1;
    if (gy > cap) gy = cap;
    return dim3(gx, (unsigned)gy);
}

// 16-byte rows: V = 4 floats or 2 doubles per lane (global_load_dwordx4)
template <class T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float4 type;
    static constexpr int n = 4;
};
template <>
struct Vec16<double> {
    typedef double2 type;
    static constexpr int n = 2;
};
template <class T>
struct Row {
    T v[Vec16<T>::n];
};
template <class T>
__device__ __forceinline__ Row<T> ldrow(const T* __restrict__ p)
{
    typedef typename Vec16<T>::type VT;
    const VT x = *(const VT*)p;
    Row<T> r;
    __builtin_memcpy(r.v, &x, 16);
    return r;
}
template <class T>
__device__ __forceinline__ void strow(T* __restrict__ p, const Row<T>& r)
{
    typedef typename Vec16<T>::type VT;
    VT x;
    __builtin_memcpy(&x, r.v, 16);
    *(VT*)p = x;
}

// f(integral_constant<int, 0>), .., f(integral_constant<int, N-1>): a loop whose index is a
// compile-time constant in the body, so that arrays indexed by it stay in registers
template <int U, int N>
struct StaticUnroll {
    template <class F>
    __device__ __forceinline__ static void run(F& f)
    {
        if constexpr (U < N) {
            f(std::integral_constant<int, U>{});
            StaticUnroll<U + 1, N>::run(f);
        }
    }
};
template <int N, class F>
__device__ __forceinline__ void static_unroll(F&& f)
{
    StaticUnroll<0, N>::run(f);
}

// Visit an X*Y*Z volume in 16-byte rows (Z % V == 0): threads along k in steps of V.
// Blocks are renumbered with xcd_remap so that rows j-1, j+1 of a block's edge rows are
// read by blocks of the same XCD.
#define TK_FOR_ROWS(X, Y, Z, V, i, j, k0, v0)                                                \
    for (unsigned w_ = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y), \
                  bx_ = w_ % gridDim.x, by_ = w_ / gridDim.x, o0_ = 1;                       \
         o0_; o0_ = 0)                                                                       \
    for (int r_ = (int)(by_ * blockDim.y + threadIdx.y); r_ < (X) * (Y);                   \
         r_ += gridDim.y * blockDim.y)                                                       \
        for (int k0 = (int)(bx_ * blockDim.x + threadIdx.x) * (V), i = r_ / (Y),             \
                 j = r_ - i * (Y);                                                           \
             k0 < (Z); k0 += gridDim.x * blockDim.x * (V))                                   \
            for (long v0 = (long)r_ * (Z) + k0, once_ = 1; once_; once_ = 0)

inline dim3 row_block(int Zv) { return Zv >= 64 ? dim3(64, 4) : (Zv >= 32 ? dim3(32, 8) : dim3(16, 16)); }
inline dim3 row_grid(int X, int Y, int Zv, int max_blocks)
{
    const dim3 b = row_block(Zv);
    const int gx = (Zv + b.x - 1) / b.x;
    long gy = ((long)X * Y + b.y - 1) / b.y;
    long cap = max_blocks / gx;
    if (cap < 1) cap = 1;
    if (gy > cap) gy = cap;
    return dim3(gx, (unsigned)gy);
}

template <class T>
__device__ inline T load_nt(const T* p)
{
    return __builtin_nontemporal_load(p);
}

}  // namespace tk
