// tk_nn.hip -- BF16 building blocks of the learned Poisson warm start (3-D U-Net).
//
// Activations are channels-last BF16 arrays: voxel v = (i*Y + j)*Z + k of an X*Y*Z level,
// channel c at v*stride + offset + c, so a convolution can read or write a channel slice
// of a wider (concatenated) buffer.  Weights of a 3x3x3 convolution are BF16 [co][kpad]
// with k = tap*ci + c, tap = (dx+1)*9 + (dy+1)*3 + (dz+1), zero-padded from 27*ci up to kpad
// (a multiple of 16).  Outside the volume the input is zero.
//
// The convolutions run on the matrix cores: v_mfma_f32_16x16x16_bf16, FP32 accumulation.
// One wave computes a 16 (voxels along z) x 16 (output channels) tile of the forward pass,
// or a 16 (output channels) x 16 (k) tile of the weight gradient summed over a chunk of
// voxels.  Operand layout of the instruction for wave64 (lane l, element q = 0..3):
// A[m = l % 16][k = 4*(l / 16) + q], B[k = 4*(l / 16) + q][n = l % 16],
// D[m = 4*(l / 16) + q][n = l % 16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }
// round to nearest even (finite inputs)
__device__ __forceinline__ uint16_t f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Tap {
    int dx, dy, dz;
};
__device__ __forceinline__ Tap tap_of(int t) { return {t / 9 - 1, (t / 3) % 3 - 1, t % 3 - 1}; }

// forward: y[v][yo + o] = act(b[o] + sum_k w[o][k] x[v + tap(k)][xo + c(k)])
// CI16: ci % 16 == 0 and xs, xo % 4 == 0, so a 16-wide k chunk lies inside one tap and a
// lane's four k values are four consecutive channels: one 8-byte load per operand.
template <bool RELU, bool OUT_F32, bool CI16>
__global__ __launch_bounds__(256) void k_conv3_fwd(const uint16_t* __restrict__ x, int xs, int xo,
                                                   int ci, const uint16_t* __restrict__ w,
                                                   int kpad, const float* __restrict__ bias,
                                                   void* __restrict__ y, int ys, int yo, int co,
                                                   int X, int Y, int Z)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nzt = (Z + 15) / 16, nct = (co + 15) / 16;
    const long ntiles = (long)X * Y * nzt * nct;
    const long t = (long)blockIdx.x * 4 + wave;
    if (t >= ntiles) return;
    const int ct = (int)(t % nct);
    const long vt = t / nct;
    const int zt = (int)(vt % nzt);
    const long row = vt / nzt;
    const int j = (int)(row % Y), i = (int)(row / Y);
    const int z0 = zt * 16;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    const int za = z0 + m;           // A row: this lane's voxel
    const int ob = ct * 16 + m;      // B column: this lane's output channel
    const int K = 27 * ci;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    if constexpr (CI16) {
        // K = 27 ci is a multiple of 16 here, so kpad == K
        const uint16_t* wrow = w + (long)(ob < co ? ob : 0) * kpad + kq;
        for (int tp = 0; tp < 27; ++tp) {
            const Tap d = tap_of(tp);
            const int ii = i + d.dx, jj = j + d.dy, zz = za + d.dz;
            const bool in = za < Z && ii >= 0 && ii < X && jj >= 0 && jj < Y && zz >= 0 && zz < Z;
            const uint16_t* xrow = in ? x + (((long)ii * Y + jj) * Z + zz) * xs + xo + kq : x;
            for (int c0 = 0; c0 < ci; c0 += 16) {
                v4s a = {0, 0, 0, 0}, b = {0, 0, 0, 0};
                if (in) a = *(const v4s*)(xrow + c0);
                if (ob < co) b = *(const v4s*)(wrow + tp * ci + c0);
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
            }
        }
    } else {
        for (int kc = 0; kc < kpad; kc += 16) {
            v4s a, b;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = kc + kq + q;
                uint16_t av = 0;
                if (k < K && za < Z) {
                    const int tp = k / ci, c = k - tp * ci;
                    const Tap d = tap_of(tp);
                    const int ii = i + d.dx, jj = j + d.dy, zz = za + d.dz;
                    if (ii >= 0 && ii < X && jj >= 0 && jj < Y && zz >= 0 && zz < Z)
                        av = x[(((long)ii * Y + jj) * Z + zz) * xs + xo + c];
                }
                a[q] = (short)av;
                b[q] = (short)(ob < co ? w[(long)ob * kpad + k] : (uint16_t)0);
            }
            acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
        }
    }
    const int od = ct * 16 + (lane & 15);
    if (od >= co) return;
    const float bo = bias ? bias[od] : 0.f;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int z = z0 + kq + q;
        if (z >= Z) continue;
        float v = acc[q] + bo;
        if (RELU) v = fmaxf(v, 0.f);
        const long e = (((long)i * Y + j) * Z + z) * ys + yo + od;
        if (OUT_F32) ((float*)y)[e] = v;
        else ((uint16_t*)y)[e] = f2bf(v);
    }
}

// Forward convolution with ci = CI (16, 32 or 64) and any co, the input staged in LDS.  A
// workgroup of NW waves owns BY rows (y) x 16 voxels (z) x 16 output channels (one of
// ceil(co / 16) slices) and marches in x over a chunk of CX planes.  The input planes x-1, x,
// x+1 with a one-voxel halo in y and z ((BY + 2) x 18 voxels, stored part-major: the
// 8-channel part c of every voxel of the slot contiguous, the parts a multiple of 256 bytes
// apart, so that the lanes reading part c of 16 consecutive voxels read 256 contiguous
// bytes with no bank conflict) sit in three slots
// of a ring of four; plane x+2 is loaded into registers before the step's products and
// written to the fourth slot after them, so one barrier per step suffices.  The weights are
// B fragments held in registers for the whole march.  v_mfma_f32_16x16x32_bf16 (gfx950): K =
// 32 is half a tap (CI = 64), one tap (CI = 32) or two taps (CI = 16); operand layout for lane
// l (element e = 0..7): A[m = l % 16][k = 8 (l / 16) + e], B[k = 8 (l / 16) + e][n = l % 16],
// D[m = 4 (l / 16) + q][n = l % 16] (q = 0..3).  A wave computes MT = BY / NW rows; the A
// fragments that two rows share (row t with dy = +1 is row t + 1 with dy = 0) are one LDS
// address, read once.
// Forward convolution of a one-channel input to co <= 16 channels (the network's first
// layer): one thread per output voxel, the 27 inputs read through the cache (zero outside
// the volume), FP32 sums over the taps in order; the 27 x 16 weights in LDS as FP32, read as
// four 16-byte broadcasts per tap.
template <bool RELU, bool V16>
__global__ __launch_bounds__(256) void k_conv3_c1(const uint16_t* __restrict__ x, int xs, int xo,
                                                  const uint16_t* __restrict__ w, int kpad,
                                                  const float* __restrict__ bias,
                                                  uint16_t* __restrict__ y, int ys, int yo,
                                                  int co, int X, int Y, int Z)
{
    __shared__ __align__(16) float ws[27][16];
    for (int e = threadIdx.x; e < 27 * 16; e += 256) {
        const int t = e >> 4, o = e & 15;
        ws[t][o] = o < co ? bf2f(w[(long)o * kpad + t]) : 0.f;
    }
    __syncthreads();
    const long n = (long)X * Y * Z;
    for (long v = blockIdx.x * 256L + threadIdx.x; v < n; v += (long)gridDim.x * 256) {
        const int k = (int)(v % Z);
        const long r = v / Z;
        const int j = (int)(r % Y), i = (int)(r / Y);
        float acc[16];
#pragma unroll
        for (int o = 0; o < 16; ++o) acc[o] = 0.f;
#pragma unroll 3
        for (int t = 0; t < 27; ++t) {
            const int ii = i + t / 9 - 1, jj = j + (t / 3) % 3 - 1, kk = k + t % 3 - 1;
            float xv = 0.f;
            if (ii >= 0 && ii < X && jj >= 0 && jj < Y && kk >= 0 && kk < Z)
                xv = bf2f(x[(((long)ii * Y + jj) * Z + kk) * xs + xo]);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float4 wq = *(const float4*)&ws[t][4 * q];
                acc[4 * q] += xv * wq.x;
                acc[4 * q + 1] += xv * wq.y;
                acc[4 * q + 2] += xv * wq.z;
                acc[4 * q + 3] += xv * wq.w;
            }
        }
        uint16_t o16[16];
#pragma unroll
        for (int o = 0; o < 16; ++o) {
            float vv = acc[o] + (bias && o < co ? bias[o] : 0.f);
            if (RELU) vv = fmaxf(vv, 0.f);
            o16[o] = f2bf(vv);
        }
        uint16_t* yv = y + v * ys + yo;
        if (V16) {
            // the voxel's 16 channels as two 16-byte stores (one 2-byte store per channel
            // would scatter 64 voxels 32 bytes apart per instruction)
            uint4 p0, p1;
            __builtin_memcpy(&p0, o16, 16);
            __builtin_memcpy(&p1, o16 + 8, 16);
            *(uint4*)yv = p0;
            *(uint4*)(yv + 8) = p1;
        } else {
#pragma unroll
            for (int o = 0; o < 16; ++o)
                if (o < co) yv[o] = o16[o];
        }
    }
}

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
// output rows per workgroup (16 rows for CI = 32 leave one workgroup per CU: 20.9 against
// 16.1 us at 64^3) and waves per workgroup
constexpr int march_rows(int) { return 8; }
constexpr int kMarchNW = 4;

template <int CI, int BY, int NW>
struct MarchCfg {
    static constexpr int RY = BY + 2, RZ = 18;      // a plane slot: rows x columns of voxels
    static constexpr int NV = RY * RZ;              // voxels per slot
    // part stride in voxels: a multiple of 16, so that the 16-byte parts of the 4 lane
    // groups of an A-fragment read (16 consecutive voxels each, one part per group) start
    // a multiple of 256 bytes apart; at 180 voxels (2880 bytes) every read took 3.8 bank
    // conflict cycles (an LDS probe of that access: none for strides of 2816, 3072 or 4096
    // bytes, whatever the offset of the runs)
    static constexpr int NVP = (NV + 15) / 16 * 16;
    static constexpr int SLOT = NVP * CI;           // bf16 per slot (part-major)
    static constexpr int CH = CI / 8;               // 16-byte parts per voxel
    static constexpr int NCH = RY * RZ * CH;        // 16-byte parts per plane
    static constexpr int NT = 64 * NW;              // threads
    static constexpr int PF = (NCH + NT - 1) / NT;  // parts per thread
    static constexpr int KS = (27 * CI + 31) / 32;  // K steps of 32
    static constexpr int MT = BY / NW;              // rows per wave
    static constexpr int H = CI >= 32 ? CI / 32 : 1;  // K steps per tap (CI >= 32)
};

template <int CI, int BY, int NW>
constexpr long march_lds_bytes()
{
    return 2L * 4 * MarchCfg<CI, BY, NW>::SLOT;
}

// CI = 64 holds 54 B fragments (216 registers): one wave per SIMD, all 512 registers.  (CI =
// 32 bounded to three waves per SIMD, 165 registers: 17.5 against 16.0 us at 64^3.)
// Output channels past co (the last slice of a co that is not a multiple of 16) have zero
// weights and are not written; OUT_F32 writes FP32 instead of BF16.
template <int CI, int BY, int NW, bool RELU, bool OUT_F32>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(CI == 64 ? 1 : 2)))
void k_conv3_march(const uint16_t* __restrict__ x, int xs, int xo, const uint16_t* __restrict__ w,
                   const float* __restrict__ bias, void* __restrict__ y, int ys, int yo, int co,
                   int X, int Y, int Z, int CX, int nsl)
{
    using C = MarchCfg<CI, BY, NW>;
    static_assert(BY % NW == 0 && C::CH * 8 == CI, "tile");
    __shared__ __align__(16) uint16_t sm[4 * C::SLOT];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int lg = lane >> 4, m = lane & 15;
    const int z0 = blockIdx.x * 16, y0 = blockIdx.y * BY;
    // output channels 16 sl .. 16 sl + 15: weight rows, bias and output slice; this lane's
    // channel n = 16 sl + m exists if n < co
    const int sl = blockIdx.z % nsl, chunk = blockIdx.z / nsl;
    w += (long)16 * sl * (27 * CI);
    if (bias) bias += 16 * sl;
    yo += 16 * sl;
    const bool nok = 16 * sl + m < co;
    const int xa = chunk * CX, xb = min(xa + CX, X);
    // weights w[n][k] (kpad = 27 CI) as B fragments: step s holds k = 32 s + 8 lg + e for
    // output channel n = m, zero past 27 CI
    v8bf bw[C::KS];
#pragma unroll
    for (int s = 0; s < C::KS; ++s) {
        const int k = 32 * s + 8 * lg;
        uint4 raw = make_uint4(0u, 0u, 0u, 0u);
        if (k < 27 * CI && nok) raw = *(const uint4*)(w + (long)m * (27 * CI) + k);
        bw[s] = __builtin_bit_cast(v8bf, raw);
    }
    const float bo = bias && nok ? bias[m] : 0.f;
    // 16-byte part u of this thread in a plane: voxel r = row * RZ + col of the slot, part c
    // (the parts of a voxel on consecutive lanes, which load whole voxels: with 16 lanes on
    // one part of 16 voxels instead, the LDS writes have no bank conflicts, but the layers
    // took 2-12% longer)
    int poff[C::PF], soff[C::PF], gy[C::PF], gz[C::PF];
#pragma unroll
    for (int u = 0; u < C::PF; ++u) {
        const int q = min(tid + C::NT * u, C::NCH - 1);
        const int r = q / C::CH, c = q - r * C::CH;
        const int ry = r / C::RZ, rz = r - ry * C::RZ;
        gy[u] = y0 - 1 + ry;
        gz[u] = z0 - 1 + rz;
        soff[u] = (c * C::NVP + r) * 8;
        poff[u] = (min(max(gy[u], 0), Y - 1) * Z + min(max(gz[u], 0), Z - 1)) * xs + xo + 8 * c;
    }
    // every part is loaded from a clamped address; those outside the volume become zeros
    // when written to LDS, after the step's products (so no wait is spent on them before)
    auto plane_ld = [&](int p, uint4 (&b)[C::PF]) {
        const long pb = (long)min(max(p, 0), X - 1) * Y * Z * xs;
#pragma unroll
        for (int u = 0; u < C::PF; ++u) b[u] = *(const uint4*)(x + pb + poff[u]);
    };
    auto plane_st = [&](int p, const uint4 (&b)[C::PF]) {
        uint16_t* sl = sm + (p & 3) * C::SLOT;
#pragma unroll
        for (int u = 0; u < C::PF; ++u) {
            if (tid + C::NT * u >= C::NCH) continue;
            const bool in = p >= 0 && p < X && gy[u] >= 0 && gy[u] < Y && gz[u] >= 0 && gz[u] < Z;
            *(uint4*)(sl + soff[u]) = in ? b[u] : make_uint4(0u, 0u, 0u, 0u);
        }
    };
    {
        uint4 b0[C::PF], b1[C::PF], b2[C::PF];
        plane_ld(xa - 1, b0);
        plane_ld(xa, b1);
        plane_ld(xa + 1, b2);
        plane_st(xa - 1, b0);
        plane_st(xa, b1);
        plane_st(xa + 1, b2);
    }
    __syncthreads();
    for (int xc = xa; xc < xb; ++xc) {
        uint4 pf[C::PF];
        plane_ld(xc + 2, pf);
        v4f acc[C::MT];
#pragma unroll
        for (int t = 0; t < C::MT; ++t) acc[t] = v4f{0.f, 0.f, 0.f, 0.f};
        if constexpr (CI >= 32) {
            // step s = H tap + h (channels 32 h .. 32 h + 31 of the tap): for each (dx, dz, h)
            // the MT + 2 rows the three dy taps read, once
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx)
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz)
#pragma unroll
                    for (int h = 0; h < C::H; ++h) {
                        const uint16_t* a0 =
                            sm + ((xc + dx) & 3) * C::SLOT +
                            ((4 * h + lg) * C::NVP + wave * C::MT * C::RZ + 1 + dz + m) * 8;
                        v8bf a[C::MT + 2];
#pragma unroll
                        for (int r = 0; r < C::MT + 2; ++r)
                            a[r] = __builtin_bit_cast(v8bf, *(const uint4*)(a0 + r * C::RZ * 8));
#pragma unroll
                        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                            for (int t = 0; t < C::MT; ++t)
                                acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                    a[t + dy + 1],
                                    bw[((dx + 1) * 9 + (dy + 1) * 3 + dz + 1) * C::H + h], acc[t],
                                    0, 0, 0);
                    }
        } else {
#pragma unroll
            for (int s = 0; s < C::KS; ++s) {
                // two taps per step, 2 s for lanes 0-31 and 2 s + 1 for lanes 32-63 (the one
                // past the last has zero weights and reads tap 26's values)
                const int tp = min(2 * s + (lg >> 1), 26);
                const int dx = tp / 9 - 1, dy = (tp / 3) % 3 - 1, dz = tp % 3 - 1;
                const uint16_t* a0 =
                    sm + ((xc + dx) & 3) * C::SLOT +
                    ((lg & 1) * C::NVP + (1 + dy + wave * C::MT) * C::RZ + 1 + dz + m) * 8;
#pragma unroll
                for (int t = 0; t < C::MT; ++t) {
                    const uint4 a = *(const uint4*)(a0 + t * C::RZ * 8);
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        __builtin_bit_cast(v8bf, a), bw[s], acc[t], 0, 0, 0);
                }
            }
        }
#pragma unroll
        for (int t = 0; t < C::MT; ++t) {
            const int yy = y0 + wave * C::MT + t;
            if (yy >= Y || !nok) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int z = z0 + 4 * lg + q;
                float v = acc[t][q] + bo;
                if (RELU) v = fmaxf(v, 0.f);
                const long e = (((long)xc * Y + yy) * Z + z) * ys + yo + m;
                if (OUT_F32) ((float*)y)[e] = v;
                else ((uint16_t*)y)[e] = f2bf(v);
            }
        }
        if (xc + 1 < xb) plane_st(xc + 2, pf);
        __syncthreads();
    }
}

// weight gradient of one chunk of voxels:
// part[chunk][o][k] = sum_{v in chunk} dy[v][o] x[v + tap(k)][c(k)]
__global__ __launch_bounds__(256) void k_conv3_wgrad(const uint16_t* __restrict__ x, int xs,
                                                     int xo, int ci,
                                                     const uint16_t* __restrict__ dy, int ds,
                                                     int dof, int co, int X, int Y, int Z,
                                                     int kpad, int vchunk, int nchunk,
                                                     float* __restrict__ part)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nkt = kpad / 16, nct = (co + 15) / 16;
    const long ntiles = (long)nchunk * nct * nkt;
    const long t = (long)blockIdx.x * 4 + wave;
    if (t >= ntiles) return;
    const int kt = (int)(t % nkt);
    const long r = t / nkt;
    const int ct = (int)(r % nct), chunk = (int)(r / nct);
    const long n = (long)X * Y * Z;
    const long v0 = (long)chunk * vchunk;
    const long v1 = v0 + vchunk < n ? v0 + vchunk : n;
    const int m = lane & 15, kq = (lane >> 4) * 4;
    const int oa = ct * 16 + m;      // A row: output channel
    const int kb = kt * 16 + m;      // B column: k index
    const int K = 27 * ci;
    int tp = 0, c = 0;
    Tap d = {0, 0, 0};
    if (kb < K) {
        tp = kb / ci;
        c = kb - tp * ci;
        d = tap_of(tp);
    }
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (long vb = v0; vb < v1; vb += 16) {
        v4s a, b;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const long v = vb + kq + q;
            uint16_t av = 0, bv = 0;
            if (v < v1) {
                if (oa < co) av = dy[v * ds + dof + oa];
                if (kb < K) {
                    const int z = (int)(v % Z);
                    const long rr = v / Z;
                    const int jj = (int)(rr % Y) + d.dy, ii = (int)(rr / Y) + d.dx, zz = z + d.dz;
                    if (ii >= 0 && ii < X && jj >= 0 && jj < Y && zz >= 0 && zz < Z)
                        bv = x[(((long)ii * Y + jj) * Z + zz) * xs + xo + c];
                }
            }
            a[q] = (short)av;
            b[q] = (short)bv;
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    const int kd = kt * 16 + (lane & 15);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int o = ct * 16 + kq + q;
        if (o < co) part[((long)chunk * co + o) * kpad + kd] = acc[q];
    }
}

// dw[e] = sum over chunks of part[chunk][e] in chunk order (deterministic)
__global__ void k_sum_chunks(const float* __restrict__ part, int nchunk, long len,
                             float* __restrict__ dw)
{
    for (long e = blockIdx.x * 256L + threadIdx.x; e < len; e += (long)gridDim.x * 256) {
        float s = 0.f;
        for (int q = 0; q < nchunk; ++q) s += part[(long)q * len + e];
        dw[e] = s;
    }
}

// bias gradient partials: bpart[chunk][o] = sum_{v in chunk} dy[v][o]
__global__ void k_bias_grad(const uint16_t* __restrict__ dy, int ds, int dof, int co, long n,
                            int vchunk, float* __restrict__ bpart)
{
    const int chunk = blockIdx.x;
    const long v0 = (long)chunk * vchunk, v1 = v0 + vchunk < n ? v0 + vchunk : n;
    for (int o = threadIdx.y; o < co; o += blockDim.y) {
        float s = 0.f;
        for (long v = v0 + threadIdx.x; v < v1; v += blockDim.x) s += bf2f(dy[v * ds + dof + o]);
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
        if (threadIdx.x == 0) bpart[(long)chunk * co + o] = s;
    }
}

// backward weights of the input gradient: wt[c][tap'][o] = w[o][26 - tap'][c] (the 3x3x3
// kernel flipped, input and output channels exchanged), so dx = conv3_fwd(dy, wt)
__global__ void k_flip_weights(const uint16_t* __restrict__ w, int ci, int co, int kpad,
                               uint16_t* __restrict__ wt, int kpad_t)
{
    const long n = (long)ci * kpad_t;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int c = (int)(e / kpad_t), kt = (int)(e % kpad_t);
        uint16_t v = 0;
        if (kt < 27 * co) {
            const int tp = kt / co, o = kt - tp * co;
            v = w[(long)o * kpad + (26 - tp) * ci + c];
        }
        wt[e] = v;
    }
}

// g[v][c] *= (y[v][c] > 0): ReLU backward from the post-activation output
__global__ void k_relu_bwd(uint16_t* __restrict__ g, int gs, int go, const uint16_t* __restrict__ y,
                           int ysd, int yo, int C, long n)
{
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n * C; e += (long)gridDim.x * 256) {
        const long v = e / C;
        const int c = (int)(e - v * C);
        if (!(bf2f(y[v * ysd + yo + c]) > 0.f)) g[v * gs + go + c] = 0;
    }
}

// 2x2x2 average pool (X, Y, Z even): out level X/2 x Y/2 x Z/2
__global__ void k_avgpool2(const uint16_t* __restrict__ x, int xs, int xo, int C, int X, int Y,
                           int Z, uint16_t* __restrict__ y, int ys, int yo)
{
    const int X2 = X / 2, Y2 = Y / 2, Z2 = Z / 2;
    const long n = (long)X2 * Y2 * Z2 * C;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int c = (int)(e % C);
        const long v = e / C;
        const int k = (int)(v % Z2);
        const long r = v / Z2;
        const int j = (int)(r % Y2), i = (int)(r / Y2);
        float s = 0.f;
        for (int a = 0; a < 8; ++a) {
            const int ii = 2 * i + (a >> 2), jj = 2 * j + ((a >> 1) & 1), kk = 2 * k + (a & 1);
            s += bf2f(x[(((long)ii * Y + jj) * Z + kk) * xs + xo + c]);
        }
        y[v * ys + yo + c] = f2bf(s * 0.125f);
    }
}

// the same with 16-byte parts of 8 channels (C / 8 = G a power of two; strides, offsets and
// pointers multiples of 8 channels): one workgroup per output row (i, j), threads over
// (k, part); the 8 children are added in the same order as above
__global__ __launch_bounds__(256) void k_avgpool2_v8(const uint16_t* __restrict__ x, int xs,
                                                     int xo, int lgG, int X, int Y, int Z,
                                                     uint16_t* __restrict__ y, int ys, int yo)
{
    const int Y2 = Y / 2, Z2 = Z / 2, G = 1 << lgG;
    const int row = blockIdx.x, i = row / Y2, j = row - i * Y2;
    for (int e = threadIdx.x; e < Z2 << lgG; e += 256) {
        const int k = e >> lgG, g = e & (G - 1);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int a = 0; a < 8; ++a) {
            const int ii = 2 * i + (a >> 2), jj = 2 * j + ((a >> 1) & 1), kk = 2 * k + (a & 1);
            const uint4 q = *(const uint4*)(x + (((long)ii * Y + jj) * Z + kk) * xs + xo + 8 * g);
            const uint32_t u[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                s[2 * c] += bf2f((uint16_t)(u[c] & 0xFFFFu));
                s[2 * c + 1] += bf2f((uint16_t)(u[c] >> 16));
            }
        }
        uint32_t o[4];
#pragma unroll
        for (int c = 0; c < 4; ++c)
            o[c] = (uint32_t)f2bf(s[2 * c] * 0.125f) | ((uint32_t)f2bf(s[2 * c + 1] * 0.125f) << 16);
        *(uint4*)(y + ((long)row * Z2 + k) * ys + yo + 8 * g) = make_uint4(o[0], o[1], o[2], o[3]);
    }
}

// its backward: dx[child][c] (+)= dy[parent][c] / 8
__global__ void k_avgpool2_bwd(const uint16_t* __restrict__ dy, int ds, int dof, int C, int X,
                               int Y, int Z, uint16_t* __restrict__ dx, int xs, int xo, int accum)
{
    const long n = (long)X * Y * Z * C;
    const int Y2 = Y / 2, Z2 = Z / 2;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int c = (int)(e % C);
        const long v = e / C;
        const int k = (int)(v % Z);
        const long r = v / Z;
        const int j = (int)(r % Y), i = (int)(r / Y);
        const long p = ((long)(i / 2) * Y2 + j / 2) * Z2 + k / 2;
        float g = bf2f(dy[p * ds + dof + c]) * 0.125f;
        if (accum) g += bf2f(dx[v * xs + xo + c]);
        dx[v * xs + xo + c] = f2bf(g);
    }
}

// nearest 2x upsample: y[child][c] = x[parent][c] (output level X x Y x Z, all even)
__global__ void k_upsample2(const uint16_t* __restrict__ x, int xs, int xo, int C, int X, int Y,
                            int Z, uint16_t* __restrict__ y, int ys, int yo)
{
    const long n = (long)X * Y * Z * C;
    const int Y2 = Y / 2, Z2 = Z / 2;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int c = (int)(e % C);
        const long v = e / C;
        const int k = (int)(v % Z);
        const long r = v / Z;
        const int j = (int)(r % Y), i = (int)(r / Y);
        const long p = ((long)(i / 2) * Y2 + j / 2) * Z2 + k / 2;
        y[v * ys + yo + c] = x[p * xs + xo + c];
    }
}

// the same with 16-byte parts of 8 channels (as k_avgpool2_v8): one workgroup per output row
__global__ __launch_bounds__(256) void k_upsample2_v8(const uint16_t* __restrict__ x, int xs,
                                                      int xo, int lgG, int X, int Y, int Z,
                                                      uint16_t* __restrict__ y, int ys, int yo)
{
    const int Y2 = Y / 2, Z2 = Z / 2, G = 1 << lgG;
    const int row = blockIdx.x, i = row / Y, j = row - i * Y;
    const long pr = ((long)(i >> 1) * Y2 + (j >> 1)) * Z2;
    for (int e = threadIdx.x; e < Z << lgG; e += 256) {
        const int k = e >> lgG, g = e & (G - 1);
        *(uint4*)(y + ((long)row * Z + k) * ys + yo + 8 * g) =
            *(const uint4*)(x + (pr + (k >> 1)) * xs + xo + 8 * g);
    }
}

// its backward: dx[parent][c] = sum of the 8 children's dy (FP32 sum, one rounding)
__global__ void k_upsample2_bwd(const uint16_t* __restrict__ dy, int ds, int dof, int C, int X,
                                int Y, int Z, uint16_t* __restrict__ dx, int xs, int xo)
{
    const int X2 = X / 2, Y2 = Y / 2, Z2 = Z / 2;
    const long n = (long)X2 * Y2 * Z2 * C;
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256) {
        const int c = (int)(e % C);
        const long v = e / C;
        const int k = (int)(v % Z2);
        const long r = v / Z2;
        const int j = (int)(r % Y2), i = (int)(r / Y2);
        float s = 0.f;
        for (int a = 0; a < 8; ++a) {
            const int ii = 2 * i + (a >> 2), jj = 2 * j + ((a >> 1) & 1), kk = 2 * k + (a & 1);
            s += bf2f(dy[(((long)ii * Y + jj) * Z + kk) * ds + dof + c]);
        }
        dx[v * xs + xo + c] = f2bf(s);
    }
}

// a[v][ao + c] = bf16(a + b[v][bo + c])
__global__ void k_add_slice(uint16_t* __restrict__ a, int as, int ao,
                            const uint16_t* __restrict__ b, int bs, int bo, int C, long n)
{
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n * C; e += (long)gridDim.x * 256) {
        const long v = e / C;
        const int c = (int)(e - v * C);
        a[v * as + ao + c] = f2bf(bf2f(a[v * as + ao + c]) + bf2f(b[v * bs + bo + c]));
    }
}

__global__ void k_f32_to_bf16(const float* __restrict__ s, long n, uint16_t* __restrict__ d)
{
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256)
        d[e] = f2bf(s[e]);
}

__global__ void k_bf16_to_f32(const uint16_t* __restrict__ s, long n, float* __restrict__ d)
{
    for (long e = blockIdx.x * 256L + threadIdx.x; e < n; e += (long)gridDim.x * 256)
        d[e] = bf2f(s[e]);
}

}  // namespace tk

using namespace tk;

namespace {
inline unsigned nblk(long n) { return (unsigned)std::min<long>((n + 255) / 256, 16384); }
// log2 of the 16-byte parts per voxel when C is 8 times a power of two and every stride,
// offset and pointer allows 16-byte accesses, else -1
inline int v8_parts_lg(int C, const void* x, int xs, int xo, const void* y, int ys, int yo)
{
    if (C % 8 != 0 || xs % 8 || xo % 8 || ys % 8 || yo % 8 || (uintptr_t)x % 16 ||
        (uintptr_t)y % 16)
        return -1;
    const int G = C / 8;
    if (G & (G - 1)) return -1;
    int lg = 0;
    while ((1 << lg) < G) ++lg;
    return lg;
}
}  // namespace

extern "C" {

int tk_nn_conv3d_fwd(tk_ctx ctx, const void* x, int xs, int xo, int ci, const void* w, int kpad,
                     const float* bias, void* y, int ys, int yo, int co, int nx, int ny, int nz,
                     int relu, int out_f32)
{
    TK_CHECK(ctx && x && w && y && ci > 0 && co > 0 && nx > 0 && ny > 0 && nz > 0, TK_ERR_ARG,
             "tk_nn_conv3d_fwd: argument");
    TK_CHECK(kpad % 16 == 0 && kpad >= 27 * ci, TK_ERR_ARG, "tk_nn_conv3d_fwd: kpad");
    TK_CHECK(xo + ci <= xs && yo + co <= ys, TK_ERR_ARG, "tk_nn_conv3d_fwd: channel slice");
    const long tiles = (long)nx * ny * ((nz + 15) / 16) * ((co + 15) / 16);
    const unsigned nb = (unsigned)((tiles + 3) / 4);
    const uint16_t* xp = (const uint16_t*)x;
    const uint16_t* wp = (const uint16_t*)w;
    // one input channel (the network's first layer), co <= 16, BF16 output
    if (ci == 1 && co <= 16 && !out_f32 && kpad >= 27) {
        const long n = (long)nx * ny * nz;
        const bool v16 = co == 16 && ys % 8 == 0 && yo % 8 == 0 && (uintptr_t)y % 16 == 0;
#define TK_C1(R, V)                                                                            \
    hipLaunchKernelGGL((k_conv3_c1<R, V>), dim3(nblk(n)), dim3(256), 0, ctx->stream, xp, xs, xo,  \
                       wp, kpad, bias, (uint16_t*)y, ys, yo, co, nx, ny, nz)
        if (relu && v16) TK_C1(true, true);
        else if (relu) TK_C1(true, false);
        else if (v16) TK_C1(false, true);
        else TK_C1(false, false);
#undef TK_C1
        TK_LAUNCHED();
        return TK_OK;
    }
    // ci = 16, 32 or 64, z a multiple of 16, 16-byte channel parts: the LDS-staged march,
    // one 16-channel output slice per workgroup (the last one partial where co is not a
    // multiple of 16), with chunks of x planes for about two workgroups per CU
    if ((ci == 16 || ci == 32 || ci == 64) && nz % 16 == 0 && xs % 8 == 0 && xo % 8 == 0 &&
        kpad == 27 * ci && (uintptr_t)x % 16 == 0 && (uintptr_t)w % 16 == 0 &&
        (uintptr_t)y % (out_f32 ? 4 : 2) == 0) {
        constexpr int NW = kMarchNW;
        const int BY = march_rows(ci);
        const int nsl = (co + 15) / 16;
        const long nyz = (long)(nz / 16) * ((ny + BY - 1) / BY) * nsl;
        int cx = (int)std::max(2L, ((long)nx * nyz + 2L * ctx->num_cu - 1) / (2L * ctx->num_cu));
        cx = std::min(cx, nx);
        const dim3 grid(nz / 16, (ny + BY - 1) / BY, (unsigned)(((nx + cx - 1) / cx) * nsl));
#define TK_MARCH(CI_, R, F)                                                                    \
    hipLaunchKernelGGL((k_conv3_march<CI_, march_rows(CI_), NW, R, F>), grid, dim3(64 * NW), 0,  \
                       ctx->stream,                                                              \
                       xp, xs, xo, wp, bias, y, ys, yo, co, nx, ny, nz, cx, nsl)
#define TK_MARCH_CI(CI_)                                                                       \
    do {                                                                                       \
        if (relu && out_f32) TK_MARCH(CI_, true, true);                                        \
        else if (relu) TK_MARCH(CI_, true, false);                                             \
        else if (out_f32) TK_MARCH(CI_, false, true);                                          \
        else TK_MARCH(CI_, false, false);                                                      \
    } while (0)
        if (ci == 64) TK_MARCH_CI(64);
        else if (ci == 32) TK_MARCH_CI(32);
        else TK_MARCH_CI(16);
#undef TK_MARCH_CI
#undef TK_MARCH
        TK_LAUNCHED();
        return TK_OK;
    }
    const bool ci16 = ci % 16 == 0 && xs % 4 == 0 && xo % 4 == 0 && kpad == 27 * ci &&
                      (uintptr_t)x % 8 == 0 && (uintptr_t)w % 8 == 0;
#define TK_FWD(R, F, V)                                                                        \
    hipLaunchKernelGGL((k_conv3_fwd<R, F, V>), dim3(nb), dim3(256), 0, ctx->stream, xp, xs, xo,  \
                       ci, wp, kpad, bias, y, ys, yo, co, nx, ny, nz)
    if (ci16) {
        if (relu && out_f32) TK_FWD(true, true, true);
        else if (relu) TK_FWD(true, false, true);
        else if (out_f32) TK_FWD(false, true, true);
        else TK_FWD(false, false, true);
    } else {
        if (relu && out_f32) TK_FWD(true, true, false);
        else if (relu) TK_FWD(true, false, false);
        else if (out_f32) TK_FWD(false, true, false);
        else TK_FWD(false, false, false);
    }
#undef TK_FWD
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_conv3d_wgrad(tk_ctx ctx, const void* x, int xs, int xo, int ci, const void* dy, int ds,
                       int dof, int co, int nx, int ny, int nz, int kpad, float* part,
                       int vchunk, float* dw, float* db)
{
    TK_CHECK(ctx && x && dy && part && dw && vchunk > 0 && vchunk % 16 == 0, TK_ERR_ARG,
             "tk_nn_conv3d_wgrad: argument");
    TK_CHECK(kpad % 16 == 0 && kpad >= 27 * ci, TK_ERR_ARG, "tk_nn_conv3d_wgrad: kpad");
    const long n = (long)nx * ny * nz;
    const int nchunk = (int)((n + vchunk - 1) / vchunk);
    const long tiles = (long)nchunk * ((co + 15) / 16) * (kpad / 16);
    hipLaunchKernelGGL(k_conv3_wgrad, dim3((unsigned)((tiles + 3) / 4)), dim3(256), 0, ctx->stream,
                       (const uint16_t*)x, xs, xo, ci, (const uint16_t*)dy, ds, dof, co, nx, ny,
                       nz, kpad, vchunk, nchunk, part);
    TK_LAUNCHED();
    const long len = (long)co * kpad;
    hipLaunchKernelGGL(k_sum_chunks, dim3(nblk(len)), dim3(256), 0, ctx->stream, part, nchunk, len,
                       dw);
    TK_LAUNCHED();
    if (db) {
        // bias partials reuse the head of part (the weight partials are consumed above)
        hipLaunchKernelGGL(k_bias_grad, dim3(nchunk), dim3(64, 4), 0, ctx->stream,
                           (const uint16_t*)dy, ds, dof, co, n, vchunk, part);
        TK_LAUNCHED();
        hipLaunchKernelGGL(k_sum_chunks, dim3(1), dim3(256), 0, ctx->stream, part, nchunk,
                           (long)co, db);
        TK_LAUNCHED();
    }
    return TK_OK;
}

int tk_nn_flip_weights(tk_ctx ctx, const void* w, int ci, int co, int kpad, void* wt, int kpad_t)
{
    TK_CHECK(ctx && w && wt && kpad_t % 16 == 0 && kpad_t >= 27 * co, TK_ERR_ARG,
             "tk_nn_flip_weights: argument");
    hipLaunchKernelGGL(k_flip_weights, dim3(nblk((long)ci * kpad_t)), dim3(256), 0, ctx->stream,
                       (const uint16_t*)w, ci, co, kpad, (uint16_t*)wt, kpad_t);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_relu_bwd(tk_ctx ctx, void* g, int gs, int go, const void* y, int ys, int yo, int C,
                   long n)
{
    TK_CHECK(ctx && g && y, TK_ERR_ARG, "tk_nn_relu_bwd: argument");
    hipLaunchKernelGGL(k_relu_bwd, dim3(nblk(n * C)), dim3(256), 0, ctx->stream, (uint16_t*)g, gs,
                       go, (const uint16_t*)y, ys, yo, C, n);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_avgpool2(tk_ctx ctx, const void* x, int xs, int xo, int C, int nx, int ny, int nz,
                   void* y, int ys, int yo)
{
    TK_CHECK(ctx && x && y && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, TK_ERR_ARG,
             "tk_nn_avgpool2: argument");
    const int lg = v8_parts_lg(C, x, xs, xo, y, ys, yo);
    if (lg >= 0)
        hipLaunchKernelGGL(k_avgpool2_v8, dim3((unsigned)((long)nx / 2 * (ny / 2))), dim3(256), 0,
                           ctx->stream, (const uint16_t*)x, xs, xo, lg, nx, ny, nz, (uint16_t*)y,
                           ys, yo);
    else
        hipLaunchKernelGGL(k_avgpool2, dim3(nblk((long)nx * ny * nz / 8 * C)), dim3(256), 0,
                           ctx->stream, (const uint16_t*)x, xs, xo, C, nx, ny, nz, (uint16_t*)y,
                           ys, yo);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_avgpool2_bwd(tk_ctx ctx, const void* dy, int ds, int dof, int C, int nx, int ny, int nz,
                       void* dx, int xs, int xo, int accum)
{
    TK_CHECK(ctx && dy && dx && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, TK_ERR_ARG,
             "tk_nn_avgpool2_bwd: argument");
    hipLaunchKernelGGL(k_avgpool2_bwd, dim3(nblk((long)nx * ny * nz * C)), dim3(256), 0,
                       ctx->stream, (const uint16_t*)dy, ds, dof, C, nx, ny, nz, (uint16_t*)dx, xs,
                       xo, accum);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_upsample2(tk_ctx ctx, const void* x, int xs, int xo, int C, int nx, int ny, int nz,
                    void* y, int ys, int yo)
{
    TK_CHECK(ctx && x && y && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, TK_ERR_ARG,
             "tk_nn_upsample2: argument");
    const int lg = v8_parts_lg(C, x, xs, xo, y, ys, yo);
    if (lg >= 0)
        hipLaunchKernelGGL(k_upsample2_v8, dim3((unsigned)((long)nx * ny)), dim3(256), 0,
                           ctx->stream, (const uint16_t*)x, xs, xo, lg, nx, ny, nz, (uint16_t*)y,
                           ys, yo);
    else
        hipLaunchKernelGGL(k_upsample2, dim3(nblk((long)nx * ny * nz * C)), dim3(256), 0,
                           ctx->stream, (const uint16_t*)x, xs, xo, C, nx, ny, nz, (uint16_t*)y,
                           ys, yo);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_upsample2_bwd(tk_ctx ctx, const void* dy, int ds, int dof, int C, int nx, int ny,
                        int nz, void* dx, int xs, int xo)
{
    TK_CHECK(ctx && dy && dx && nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, TK_ERR_ARG,
             "tk_nn_upsample2_bwd: argument");
    hipLaunchKernelGGL(k_upsample2_bwd, dim3(nblk((long)nx * ny * nz / 8 * C)), dim3(256), 0,
                       ctx->stream, (const uint16_t*)dy, ds, dof, C, nx, ny, nz, (uint16_t*)dx, xs,
                       xo);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_add_slice(tk_ctx ctx, void* a, int as, int ao, const void* b, int bs, int bo, int C,
                    long n)
{
    TK_CHECK(ctx && a && b, TK_ERR_ARG, "tk_nn_add_slice: argument");
    hipLaunchKernelGGL(k_add_slice, dim3(nblk(n * C)), dim3(256), 0, ctx->stream, (uint16_t*)a, as,
                       ao, (const uint16_t*)b, bs, bo, C, n);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_f32_to_bf16(tk_ctx ctx, const float* s, long n, void* d)
{
    TK_CHECK(ctx && s && d, TK_ERR_ARG, "tk_nn_f32_to_bf16: argument");
    hipLaunchKernelGGL(k_f32_to_bf16, dim3(nblk(n)), dim3(256), 0, ctx->stream, s, n, (uint16_t*)d);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_nn_bf16_to_f32(tk_ctx ctx, const void* s, long n, float* d)
{
    TK_CHECK(ctx && s && d, TK_ERR_ARG, "tk_nn_bf16_to_f32: argument");
    hipLaunchKernelGGL(k_bf16_to_f32, dim3(nblk(n)), dim3(256), 0, ctx->stream,
                       (const uint16_t*)s, n, d);
    TK_LAUNCHED();
    return TK_OK;
}

}  // extern "C"
