// tk_fft_kernel.hpp -- the per-axis convolution kernel and its launcher, templated
// on the element type and on load / mid / store functors.  Included by every
// translation unit that fuses a stage into a convolution pass.
#pragma once

#include <type_traits>

#include "tk_device.hpp"
#include "tk_fft.hpp"

namespace tk {

// ---- size-R DFTs, sign s (-1 forward, +1 inverse) -------------------------------
template <class T>
__device__ __forceinline__ void dft2(cplx<T>* a, int)
{
    cplx<T> t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}

template <class T>
__device__ __forceinline__ void dft3(cplx<T>* a, int s)
{
    const T h = (T)0.86602540378443864676372317075293618;  // sqrt(3)/2
    cplx<T> t = cadd(a[1], a[2]);
    cplx<T> u = csub(a[1], a[2]);
    cplx<T> m = {a[0].x - (T)0.5 * t.x, a[0].y - (T)0.5 * t.y};
    cplx<T> v = cmul_si(cscale(u, h), s);
    a[0] = cadd(a[0], t);
    a[1] = cadd(m, v);
    a[2] = csub(m, v);
}

template <class T>
__device__ __forceinline__ void dft4(cplx<T>* a, int s)
{
    cplx<T> p0 = cadd(a[0], a[2]), m0 = csub(a[0], a[2]);
    cplx<T> p1 = cadd(a[1], a[3]), m1 = cmul_si(csub(a[1], a[3]), s);
    a[0] = cadd(p0, p1);
    a[2] = csub(p0, p1);
    a[1] = cadd(m0, m1);
    a[3] = csub(m0, m1);
}

template <class T>
__device__ __forceinline__ void dft5(cplx<T>* a, int s)
{
    const T c1 = (T)0.30901699437494742410229341718281906;   // cos(2pi/5)
    const T c2 = (T)-0.80901699437494742410229341718281906;  // cos(4pi/5)
    const T s1 = (T)0.95105651629515357211643933337938214;   // sin(2pi/5)
    const T s2 = (T)0.58778525229247312916870595463907277;   // sin(4pi/5)
    cplx<T> b1 = cadd(a[1], a[4]), d1 = csub(a[1], a[4]);
    cplx<T> b2 = cadd(a[2], a[3]), d2 = csub(a[2], a[3]);
    cplx<T> x0 = a[0];
    cplx<T> r1 = {x0.x + c1 * b1.x + c2 * b2.x, x0.y + c1 * b1.y + c2 * b2.y};
    cplx<T> r2 = {x0.x + c2 * b1.x + c1 * b2.x, x0.y + c2 * b1.y + c1 * b2.y};
    cplx<T> q1 = cmul_si(cplx<T>{s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y}, s);
    cplx<T> q2 = cmul_si(cplx<T>{s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y}, s);
    a[0] = cadd(x0, cadd(b1, b2));
    a[1] = cadd(r1, q1);
    a[4] = csub(r1, q1);
    a[2] = cadd(r2, q2);
    a[3] = csub(r2, q2);
}

// e^{s 2 pi i m / 16}
template <class T>
__device__ __forceinline__ cplx<T> w16(int m, int s)
{
    const T c1 = (T)0.92387953251128675612818318939678829;  // cos(pi/8)
    const T s1 = (T)0.38268343236508977172845998403039887;  // sin(pi/8)
    const T c2 = (T)0.70710678118654752440084436210484904;  // cos(pi/4)
    T re, im;
    switch (m & 15) {
        case 0: re = 1; im = 0; break;
        case 1: re = c1; im = s1; break;
        case 2: re = c2; im = c2; break;
        case 3: re = s1; im = c1; break;
        case 4: re = 0; im = 1; break;
        case 5: re = -s1; im = c1; break;
        case 6: re = -c2; im = c2; break;
        case 7: re = -c1; im = s1; break;
        case 8: re = -1; im = 0; break;
        case 9: re = -c1; im = -s1; break;
        case 10: re = -c2; im = -c2; break;
        case 11: re = -s1; im = -c1; break;
        case 12: re = 0; im = -1; break;
        case 13: re = s1; im = -c1; break;
        case 14: re = c2; im = -c2; break;
        default: re = c1; im = -s1; break;
    }
    return {re, s > 0 ? im : -im};
}

// 8 = 2 x 4 and 16 = 4 x 4 Cooley-Tukey in registers: column DFTs, internal twiddles
// W_R^{n2 k1}, row DFTs; output in natural order
template <class T>
__device__ __forceinline__ void dft8(cplx<T>* a, int s)
{
    cplx<T> e[4] = {a[0], a[2], a[4], a[6]}, o[4] = {a[1], a[3], a[5], a[7]};
    dft4(e, s);
    dft4(o, s);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const cplx<T> t = k == 0 ? o[0] : cmul(o[k], w16<T>(2 * k, s));
        a[k] = cadd(e[k], t);
        a[k + 4] = csub(e[k], t);
    }
}

template <class T>
__device__ __forceinline__ void dft16(cplx<T>* a, int s)
{
    cplx<T> b[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
#pragma unroll
        for (int n1 = 0; n1 < 4; ++n1) b[n2][n1] = a[4 * n1 + n2];
        dft4(b[n2], s);
#pragma unroll
        for (int k1 = 1; k1 < 4; ++k1)
            if (n2 > 0) b[n2][k1] = cmul(b[n2][k1], w16<T>(n2 * k1, s));
    }
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cplx<T> c[4] = {b[0][k1], b[1][k1], b[2][k1], b[3][k1]};
        dft4(c, s);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) a[k1 + 4 * k2] = c[k2];
    }
}

template <class T, int R>
__device__ __forceinline__ void dftR(cplx<T>* a, int s)
{
    if constexpr (R == 2) dft2(a, s);
    else if constexpr (R == 3) dft3(a, s);
    else if constexpr (R == 4) dft4(a, s);
    else if constexpr (R == 5) dft5(a, s);
    else if constexpr (R == 8) dft8(a, s);
    else dft16(a, s);
}

// LDS layout of a group of P pencils of length L: element (pos, p) at pos*SP + p*SQ.
// x/y passes use SP = P+1, SQ = 1 (pencils interleaved); z passes use SP = 1,
// SQ = L+1 (pencils as padded rows), so the contiguous global rows map to LDS rows.
struct Lay {
    int L, P, lp, SP, SQ;
    int off;   // LDS index of position 0 (the direct path keeps a halo before it)
    int zero;  // acyclic pass: the halo holds zeros instead of the wrapped positions
};

// one Stockham stage; butterfly b handles pencil p = b mod P at index j = b / P
template <class T, int R, int EPT, int NT>
__device__ __forceinline__ void fft_stage(cplx<T>* __restrict__ buf, const cplx<T>* __restrict__ tw,
                                          const Lay& y, int Ns, float inv_ns, int sign)
{
    constexpr int MAXB = (EPT + R - 1) / R;
    const int LR = y.L / R;
    const int nbf = y.P * LR;
    const int stw = y.L / (Ns * R);
    cplx<T> v[MAXB][R];
    int dst[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * NT;
        dst[q] = -1;
        if (b < nbf) {
            const int p = b & (y.P - 1), j = b >> y.lp;
            const int jq = qdiv(j, inv_ns);
            const int k = j - jq * Ns;
            const int src = j * y.SP + p * y.SQ;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[src + r * LR * y.SP];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cplx<T> w = tw[r * k * stw];
                if (sign > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            dftR<T, R>(v[q], sign);
            dst[q] = (jq * Ns * R + k) * y.SP + p * y.SQ;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns * y.SP] = v[q][r];
        }
    }
    __syncthreads();
}

template <class T, int EPT, int NT>
__device__ __forceinline__ void fft_lds(cplx<T>* buf, const cplx<T>* tw, const AxisDev& ax,
                                        const Lay& y, int sign)
{
    int Ns = 1;
    for (int s = 0; s < ax.nstage; ++s) {
        const int R = ax.radix[s];
        const float inv = 1.0f / (float)Ns;
        if (R == 16) {
            if constexpr (EPT >= 16) fft_stage<T, 16, EPT, NT>(buf, tw, y, Ns, inv, sign);
        } else if (R == 8) fft_stage<T, 8, EPT, NT>(buf, tw, y, Ns, inv, sign);
        else if (R == 4) fft_stage<T, 4, EPT, NT>(buf, tw, y, Ns, inv, sign);
        else if (R == 2) fft_stage<T, 2, EPT, NT>(buf, tw, y, Ns, inv, sign);
        else if (R == 3) fft_stage<T, 3, EPT, NT>(buf, tw, y, Ns, inv, sign);
        else fft_stage<T, 5, EPT, NT>(buf, tw, y, Ns, inv, sign);
        Ns *= R;
    }
}

// direct circular convolution along the pencil with nh taps (lengths that are not
// 2/3/5-smooth): y(i) = sum_n w(n) x((i - n) mod L), n in [-s, s]
template <class T, int EPT, int NT>
__device__ __forceinline__ void direct_lds(cplx<T>* buf, const T* w, const AxisDev& ax,
                                           const Lay& y)
{
    const int L = y.L, s = ax.nh / 2;
    const int E = L * y.P;
    cplx<T> out[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * NT;
        out[q] = {0, 0};
        if (e < E) {
            const int p = e & (y.P - 1), pos = e >> y.lp;
            T ax_ = 0, ay = 0;
            for (int n = -s; n <= s; ++n) {
                if (y.zero && (pos - n < 0 || pos - n >= L)) continue;
                int m = (pos - n) % L;
                if (m < 0) m += L;
                const cplx<T> v = buf[m * y.SP + p * y.SQ];
                ax_ += w[n + s] * v.x;
                ay += w[n + s] * v.y;
            }
            out[q] = {ax_, ay};
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * NT;
        if (e < E) buf[(e >> y.lp) * y.SP + (e & (y.P - 1)) * y.SQ] = out[q];
    }
    __syncthreads();
}

template <class T>
__device__ __forceinline__ T fmat(T a, T b, T c)
{
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// Direct circular convolution with the symmetric taps w(0..S), S compile-time:
//   y(i) = w(0) x(i) + sum_{n=1..S} w(n) (x(i-n) + x(i+n))
// The pencils sit in LDS with S wrapped halo positions on each side (filled here), a
// thread computes B consecutive outputs of one pencil from a register window of B+2S
// inputs; results stay in registers until every window has been read.
template <class T, int EPT, int NT, int S>
__device__ __forceinline__ void direct_win(cplx<T>* buf, const T* wh, const Lay& y)
{
    constexpr int B = 4, MAXI = (EPT + B - 1) / B;
    const int L = y.L;
    // halo: positions -S..-1 and L..L+S-1 hold x((pos mod L)), or zeros (acyclic)
    for (int e = threadIdx.x; e < 2 * S * y.P; e += NT) {
        const int p = e & (y.P - 1), t = e >> y.lp;
        const int pos = t < S ? t - S : L + (t - S);
        // one wrap when S <= L (no integer division); shorter pencils wrap repeatedly
        int src = S <= L ? (pos < 0 ? pos + L : pos - L) : pos % L;
        if (src < 0) src += L;
        buf[y.off + pos * y.SP + p * y.SQ] =
            y.zero ? cplx<T>{0, 0} : buf[y.off + src * y.SP + p * y.SQ];
    }
    __syncthreads();
    T w[S + 1];
#pragma unroll
    for (int n = 0; n <= S; ++n) w[n] = wh[n];
    const int items = y.P * (L / B);
    // z pass (pencils SQ apart, positions adjacent): a wave's 16- or 8-pencil lane groups
    // take blocks 4 or 2 apart inside each run of 8 blocks, so their reads fall 32 or 16
    // dwords apart in the LDS banks (a bijection on the blocks when L/B is a multiple of 8)
    const int G = 32 >> y.lp, lg = 5 - y.lp;
    const bool swz = y.SP == 1 && (G == 2 || G == 4) && (L / B) % 8 == 0;
    cplx<T> out[MAXI][B];
    int dst[MAXI];
#pragma unroll
    for (int q = 0; q < MAXI; ++q) {
        const int it = threadIdx.x + q * NT;
        dst[q] = -1;
        if (it < items) {
            const int p = it & (y.P - 1), br = it >> y.lp;
            // G = 2^lg: (br % G) * (8 / G) + (br / G) % (8 / G) + (br & ~7) with shifts
            const int b = swz ? ((br & (G - 1)) << (3 - lg)) + ((br >> lg) & ((8 >> lg) - 1)) +
                                    (br & ~7)
                              : br;
            const int base = y.off + (b * B - S) * y.SP + p * y.SQ;
            // Output u = w(0) x(u+S) + sum_{n=1..S} w(n) (x(u+S-n) + x(u+S+n)), n ascending.
            // For each n the B outputs take the pairs from a left window L(u) = x(u+S-n) and
            // a right window R(u) = x(u+S+n); going to n+1 both slide outwards by one
            // position, so a step reads two values from LDS and only 2B inputs are live
            // (not the B + 2S of the whole window).
            cplx<T> L[B], R[B], a[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                L[u] = buf[base + (S - 1 + u) * y.SP];
                R[u] = buf[base + (S + 1 + u) * y.SP];
            }
#pragma unroll
            for (int u = 0; u < B; ++u) {
                // x(u+S): inside L or R for every u but the two ends of B = 4
                const cplx<T> xc = u + 1 < B ? L[u + 1] : R[u - 1];
                a[u] = {w[0] * xc.x, w[0] * xc.y};
            }
#pragma unroll
            for (int n = 1; n <= S; ++n) {
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    a[u].x = fmat(w[n], L[u].x + R[u].x, a[u].x);
                    a[u].y = fmat(w[n], L[u].y + R[u].y, a[u].y);
                }
                if (n < S) {
#pragma unroll
                    for (int u = B - 1; u > 0; --u) L[u] = L[u - 1];
                    L[0] = buf[base + (S - n - 1) * y.SP];
#pragma unroll
                    for (int u = 0; u < B - 1; ++u) R[u] = R[u + 1];
                    R[B - 1] = buf[base + (S + n + B) * y.SP];
                }
            }
#pragma unroll
            for (int u = 0; u < B; ++u) out[q][u] = a[u];
            dst[q] = y.off + b * B * y.SP + p * y.SQ;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXI; ++q)
        if (dst[q] >= 0) {
#pragma unroll
            for (int u = 0; u < B; ++u) buf[dst[q] + u * y.SP] = out[q][u];
        }
    __syncthreads();
}

// The same circular convolution on the matrix cores (FP32, v_mfma_f32_16x16x4_f32), for
// pencil groups of a multiple of 8 and lengths L % 16 == 0.  A 16 x 16 output tile is 16
// consecutive positions of 8 pencils x 2 fields (the real and imaginary parts of the
// complex pencils), the product of the banded Toeplitz matrix
// Wb[m][k] = w(k - S - m) (16 x 4*KS, zero outside |k - S - m| <= S) with the 4*KS input
// positions from 16b - S of the same pencils and fields:
//   out(16b + m, c) = sum_k Wb[m][k] x(16b - S + k, c),   c = 2 * pencil + field
// Operand layouts of the instruction (lane l, n = l % 16, g = l / 16): A[m = n][k = g],
// B[k = g][c = n], D[m = 4g + q][c = n] in register q.  K-step ks covers the input
// positions 4ks + g - S, so the band Wb is the same for every tile (KS registers, read
// from the filter's table) and consecutive tiles of a wave share KS - 4 K-steps of inputs
// (a sliding register window).  Where the last K-step reaches past the S halo positions
// (S = 9: two positions), its reads are clamped to the last halo position, so that no
// unset LDS word enters a product (their weights are zero).
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ constexpr int mfma_ksteps(int S) { return (16 + 2 * S + 3) / 4; }

template <int EPT, int NT, int S>
__device__ __forceinline__ void mfma_win(cplx<float>* buf, const float* band, const Lay& y)
{
    constexpr int KS = mfma_ksteps(S);
    constexpr bool CLAMP = 4 * KS - 16 - S > S;
    constexpr int NW = NT / 64;
    constexpr int MAXT = (EPT * NT / 128 + NW - 1) / NW;  // tiles per wave
    const int L = y.L;
    for (int e = threadIdx.x; e < 2 * S * y.P; e += NT) {
        const int p = e & (y.P - 1), t = e >> y.lp;
        const int pos = t < S ? t - S : L + (t - S);
        int src = pos;
        while (src < 0) src += L;
        while (src >= L) src -= L;
        buf[y.off + pos * y.SP + p * y.SQ] =
            y.zero ? cplx<float>{0.f, 0.f} : buf[y.off + src * y.SP + p * y.SQ];
    }
    const int lane = threadIdx.x & 63;
    float A[KS];
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) A[ks] = band[ks * 64 + lane];
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n = lane & 15, g = lane >> 4;
    const int NB = L >> 4, T = NB * (y.P >> 3);
    const int t0 = wave * T / NW, t1 = (wave + 1) * T / NW;
    const int last = L + S - 1;
    const int sp2 = 2 * y.SP, st4 = 4 * sp2;
    float* fb = (float*)buf;
    // lane n reads field n & 1 of pencil n >> 1 of the tile's 8-pencil group
    const int cofs = 2 * (y.off + (n >> 1) * y.SQ) + (n & 1);
    f32x4 o[MAXT];
    float R[KS];
    int prev = -2;
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        const int t = t0 + j;
        if (t < t1) {
            const int pg = t / NB, b = t - pg * NB;
            const float* col = fb + cofs + 16 * pg * y.SQ;
            const int p0 = 16 * b + g - S;
            if (t == prev + 1 && b > 0) {
#pragma unroll
                for (int i = 0; i < KS - 4; ++i) R[i] = R[i + 4];
                int a = (p0 + 4 * (KS - 4)) * sp2;
#pragma unroll
                for (int i = KS - 4; i < KS; ++i, a += st4) {
                    if constexpr (CLAMP) R[i] = col[min(p0 + 4 * i, last) * sp2];
                    else R[i] = col[a];
                }
            } else {
                int a = p0 * sp2;
#pragma unroll
                for (int i = 0; i < KS; ++i, a += st4) {
                    if constexpr (CLAMP) R[i] = col[min(p0 + 4 * i, last) * sp2];
                    else R[i] = col[a];
                }
            }
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < KS; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(A[ks], R[ks], acc, 0, 0, 0);
            o[j] = acc;
            prev = t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < MAXT; ++j) {
        const int t = t0 + j;
        if (t < t1) {
            const int pg = t / NB, b = t - pg * NB;
            float* col = fb + cofs + 16 * pg * y.SQ + (16 * b + 4 * g) * sp2;
#pragma unroll
            for (int q = 0; q < 4; ++q) col[q * sp2] = o[j][q];
        }
    }
    __syncthreads();
}

// MF: the matrix-core window (FP32; the launcher checks the group shape), else the
// vector-ALU window.  Separate kernels, so each gets the register budget it needs.
template <class T, int EPT, int NT, int S, bool MF>
__device__ __forceinline__ void direct_conv(cplx<T>* buf, const T* wh, const AxisDev& ax,
                                            const Lay& y)
{
    if constexpr (MF) {
        static_assert(sizeof(T) == 4, "matrix-core window: FP32 only");
        mfma_win<EPT, NT, S>((cplx<float>*)buf, (const float*)ax.band, y);
    } else {
        direct_win<T, EPT, NT, S>(buf, wh, y);
    }
}

// forward FFT, multiply by the (real, 1/L-scaled) spectrum, inverse FFT
template <class T, int EPT, int NT>
__device__ __forceinline__ void conv_lds(cplx<T>* buf, const cplx<T>* tw, const T* spec,
                                         const AxisDev& ax, const Lay& y)
{
    if (ax.nstage == 0 || y.zero) {
        direct_lds<T, EPT, NT>(buf, spec, ax, y);
        return;
    }
    fft_lds<T, EPT, NT>(buf, tw, ax, y, -1);
    const int E = y.L * y.P;
    for (int e = threadIdx.x; e < E; e += NT) {
        const int p = e & (y.P - 1), pos = e >> y.lp;
        cplx<T>& v = buf[pos * y.SP + p * y.SQ];
        const T s = spec[pos];
        v = {v.x * s, v.y * s};
    }
    __syncthreads();
    fft_lds<T, EPT, NT>(buf, tw, ax, y, +1);
}

// ---- quads: 4 consecutive voxels moved with 16-byte accesses ------------------------
template <class T>
struct Quad {
    T v[4];
};
template <class T>
__device__ __forceinline__ Quad<T> ldq(const T* __restrict__ p)
{
    Quad<T> q;
    if constexpr (sizeof(T) == 4) {
        const float4 x = *(const float4*)p;
        __builtin_memcpy(q.v, &x, 16);
    } else {
        const double2 x = ((const double2*)p)[0], y = ((const double2*)p)[1];
        __builtin_memcpy(q.v, &x, 16);
        __builtin_memcpy(q.v + 2, &y, 16);
    }
    return q;
}
template <class T>
__device__ __forceinline__ void stq(T* __restrict__ p, const Quad<T>& q)
{
    if constexpr (sizeof(T) == 4) {
        float4 x;
        __builtin_memcpy(&x, q.v, 16);
        *(float4*)p = x;
    } else {
        double2 x, y;
        __builtin_memcpy(&x, q.v, 16);
        __builtin_memcpy(&y, q.v + 2, 16);
        ((double2*)p)[0] = x;
        ((double2*)p)[1] = y;
    }
}
// four consecutive complex values (32 or 64 bytes)
template <class T>
__device__ __forceinline__ void ldc4(const cplx<T>* __restrict__ p, cplx<T>* c)
{
    const T* s = (const T*)p;
    Quad<T> a = ldq(s), b = ldq(s + 4);
    c[0] = {a.v[0], a.v[1]};
    c[1] = {a.v[2], a.v[3]};
    c[2] = {b.v[0], b.v[1]};
    c[3] = {b.v[2], b.v[3]};
}
template <class T>
__device__ __forceinline__ void stc4(cplx<T>* __restrict__ p, const cplx<T>* c)
{
    T* s = (T*)p;
    stq(s, Quad<T>{{c[0].x, c[0].y, c[1].x, c[1].y}});
    stq(s + 4, Quad<T>{{c[2].x, c[2].y, c[3].x, c[3].y}});
}

// ---- functors ---------------------------------------------------------------------
// Each functor works on one voxel (operator()) or on a quad of four consecutive
// voxels starting at v (ld4 / md4 / st4), with identical arithmetic.
template <class T>
struct LdComplex {
    const cplx<T>* src;
    __device__ cplx<T> operator()(long v) const { return src[v]; }
    __device__ void ld4(long v, cplx<T>* c) const { ldc4(src + v, c); }
};
// (a, b) packed as one complex field; a null pointer reads as `fill`
template <class T>
struct LdPair {
    const T* a;
    const T* b;
    T fill_a, fill_b;
    __device__ cplx<T> operator()(long v) const
    {
        return {a ? a[v] : fill_a, b ? b[v] : fill_b};
    }
    __device__ void ld4(long v, cplx<T>* c) const
    {
        const Quad<T> qa = a ? ldq(a + v) : Quad<T>{{fill_a, fill_a, fill_a, fill_a}};
        const Quad<T> qb = b ? ldq(b + v) : Quad<T>{{fill_b, fill_b, fill_b, fill_b}};
#pragma unroll
        for (int e = 0; e < 4; ++e) c[e] = {qa.v[e], qb.v[e]};
    }
};
struct MdNone {
};
template <class T>
struct StComplex {
    cplx<T>* dst;
    __device__ void operator()(long v, cplx<T> c) const { dst[v] = c; }
    __device__ void st4(long v, const cplx<T>* c) const { stc4(dst + v, c); }
};
template <class T>
struct StPair {
    T* a;
    T* b;
    __device__ void operator()(long v, cplx<T> c) const
    {
        if (a) a[v] = c.x;
        if (b) b[v] = c.y;
    }
    __device__ void st4(long v, const cplx<T>* c) const
    {
        if (a) stq(a + v, Quad<T>{{c[0].x, c[1].x, c[2].x, c[3].x}});
        if (b) stq(b + v, Quad<T>{{c[0].y, c[1].y, c[2].y, c[3].y}});
    }
};

// Global <-> LDS element maps, chosen for coalescing.  x/y passes: thread t owns
// pencil p = t mod P for all its elements (one base address per thread), positions
// advance by NT/P.  z passes: the P pencils are P contiguous rows, element e of the
// group is voxel c0*L + e.
template <class T, class F>
__device__ __forceinline__ void for_group(const AxisGeom& g, const Lay& y, long c0, int NT, F f)
{
    const int E = y.L * y.P;
    if (g.axis == 2) {
        const float invL = 1.0f / (float)y.L;
        const long v0 = c0 * y.L;
        const long vend = g.nlines * (long)y.L;
        for (int e = threadIdx.x; e < E; e += NT) {
            const int p = qdiv(e, invL), pos = e - p * y.L;
            if (v0 + e < vend) f(v0 + e, y.off + pos * y.SP + p * y.SQ);
        }
    } else {
        const int p = threadIdx.x & (y.P - 1);
        const long c = c0 + p;
        if (c >= g.nlines) return;
        long base;
        if (g.axis == 1) {
            const int cc = (int)c;
            const int i = cc / g.nz;
            base = (long)i * g.ny * g.nz + (cc - i * g.nz);
        } else {
            base = c;
        }
        const int step = NT >> y.lp;
        for (int pos = threadIdx.x >> y.lp; pos < y.L; pos += step)
            f(base + (long)pos * g.stride, y.off + pos * y.SP + p);
    }
}

// Quad traversal (every axis length and nz a multiple of 4, P >= 4): x/y passes give a
// thread 4 adjacent pencils at one position (4 adjacent voxels in memory); z passes give
// it 4 adjacent positions of one pencil.  f(v, l, lstep): quad at voxel v, LDS elements
// l + q*lstep.
template <class T, class F>
__device__ __forceinline__ void for_quads(const AxisGeom& g, const Lay& y, long c0, int NT, F f)
{
    if (g.axis == 2) {
        const int L4 = y.L >> 2;
        const float inv = 1.0f / (float)L4;
        const int nq = y.P * L4;
        for (int e4 = threadIdx.x; e4 < nq; e4 += NT) {
            const int p = qdiv(e4, inv), pos0 = (e4 - p * L4) << 2;
            if (c0 + p < g.nlines) f((c0 + p) * y.L + pos0, y.off + pos0 + p * y.SQ, 1);
        }
    } else {
        const int PQ = y.P >> 2;
        const int p0 = (threadIdx.x & (PQ - 1)) << 2;
        const long c = c0 + p0;
        if (c >= g.nlines) return;
        long base;
        if (g.axis == 1) {
            const int cc = (int)c;
            const int i = cc / g.nz;
            base = (long)i * g.ny * g.nz + (cc - i * g.nz);
        } else {
            base = c;
        }
        const int lpq = y.lp - 2;
        const int step = NT >> lpq;
        for (int pos = threadIdx.x >> lpq; pos < y.L; pos += step)
            f(base + (long)pos * g.stride, y.off + pos * y.SP + p0, 1);
    }
}

// LDS stride between positions of the x/y passes (pencils adjacent).  The direct
// window's reads are conflict-free in the 64 LDS banks when consecutive 4-position blocks
// of a 16-pencil group start 32 dwords apart (P + 4) and of an 8-pencil group 16 dwords
// apart (P + 2); the FFT path keeps P + 1.
inline __host__ __device__ int conv_sp(int axis, int P, int S)
{
    if (axis == 2) return 1;
    if (S > 0 && P == 16) return P + 4;
    if (S > 0 && P == 8) return P + 2;
    return P + 1;
}

// LDS elements of a pencil group; the direct path (S > 0) adds S halo positions per side
inline __host__ __device__ int conv_nbuf(int axis, int L, int P, int S)
{
    return axis == 2 ? P * (L + 2 * S + 1) : (L + 2 * S) * conv_sp(axis, P, S);
}

// per-axis convolution of the pencil group starting at line c0: S == 0 FFT (or the generic
// direct fallback), S > 0 the register-window direct path with half-width S, on the matrix
// cores when MF
template <class T, int EPT, int NT, bool QUAD, int S, bool MF, class Ld, class Md, class St>
__device__ __forceinline__ void conv_group(const AxisDev& ax, const AxisGeom& g, int rounds,
                                           const Ld& ld, const Md& md, const St& st, long c0,
                                           unsigned char* smem)
{
    Lay y;
    y.L = ax.L;
    y.P = g.P;
    y.lp = g.logP;
    y.SP = conv_sp(g.axis, g.P, S);
    y.SQ = g.axis == 2 ? ax.L + 2 * S + 1 : 1;
    y.off = S * y.SP;
    y.zero = ax.zero_halo;
    const int nbuf = conv_nbuf(g.axis, ax.L, g.P, S);
    cplx<T>* buf = (cplx<T>*)smem;
    cplx<T>* tw = buf + nbuf;
    T* spec = (T*)(tw + ax.L);
    if constexpr (S > 0) {
        for (int t = threadIdx.x; t <= S; t += NT) spec[t] = ((const T*)ax.half)[t];
    } else if (ax.nstage > 0 && !ax.zero_halo) {
        for (int t = threadIdx.x; t < ax.L; t += NT) {
            tw[t] = ((const cplx<T>*)ax.tw)[t];
            spec[t] = ((const T*)ax.spec)[t];
        }
    } else {
        for (int t = threadIdx.x; t < ax.nh; t += NT) spec[t] = ((const T*)ax.taps)[t];
    }
    // pencils past the end of the volume stay zero
    if ((c0 + g.P) > g.nlines) {
        for (int e = threadIdx.x; e < nbuf; e += NT) buf[e] = {0, 0};
        __syncthreads();
    }
    if constexpr (QUAD) {
        for_quads<T>(g, y, c0, NT, [&](long v, int l, int ls) {
            cplx<T> c[4];
            ld.ld4(v, c);
#pragma unroll
            for (int q = 0; q < 4; ++q) buf[l + q * ls] = c[q];
        });
    } else {
        for_group<T>(g, y, c0, NT, [&](long v, int l) { buf[l] = ld(v); });
    }
    __syncthreads();
    if constexpr (S > 0) direct_conv<T, EPT, NT, S, MF>(buf, spec, ax, y);
    else conv_lds<T, EPT, NT>(buf, tw, spec, ax, y);
    if constexpr (!std::is_same<Md, MdNone>::value) {
        if (rounds > 1) {
            if constexpr (QUAD) {
                for_quads<T>(g, y, c0, NT, [&](long v, int l, int ls) {
                    cplx<T> c[4];
#pragma unroll
                    for (int q = 0; q < 4; ++q) c[q] = buf[l + q * ls];
                    md.md4(v, c);
#pragma unroll
                    for (int q = 0; q < 4; ++q) buf[l + q * ls] = c[q];
                });
            } else {
                for_group<T>(g, y, c0, NT, [&](long v, int l) { buf[l] = md(v, buf[l]); });
            }
            __syncthreads();
            if constexpr (S > 0) direct_conv<T, EPT, NT, S, MF>(buf, spec, ax, y);
            else conv_lds<T, EPT, NT>(buf, tw, spec, ax, y);
        }
    }
    if constexpr (QUAD) {
        for_quads<T>(g, y, c0, NT, [&](long v, int l, int ls) {
            cplx<T> c[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) c[q] = buf[l + q * ls];
            st.st4(v, c);
        });
    } else {
        for_group<T>(g, y, c0, NT, [&](long v, int l) { st(v, buf[l]); });
    }
}

template <class T, int EPT, int NT, bool QUAD, int S, bool MF, class Ld, class Md, class St>
__global__ __launch_bounds__(NT) void k_conv_axis(AxisDev ax, AxisGeom g, int rounds, Ld ld,
                                                  Md md, St st)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // consecutive groups share 128-byte lines along x and y: keep them on one XCD
    conv_group<T, EPT, NT, QUAD, S, MF>(ax, g, rounds, ld, md, st,
                                        (long)xcd_remap(blockIdx.x, gridDim.x) * g.P, smem);
}

// one volume of a batched convolution pass (k_conv_axis_batch): its axis, geometry, the
// stage's load / middle / store operators on its fields, and its block range (first a
// multiple of kNumXcd, nblk padded to one; groups: the real pencil groups)
template <class Ld, class Md, class St>
struct ConvJob {
    AxisDev ax;
    AxisGeom g;
    int rounds;
    Ld ld;
    Md md;
    St st;
    int first, nblk, groups;
};

// one convolution pass of many volumes in one launch; each block finds its volume by a
// binary search of the first blocks and runs that volume's group in the XCD-aware order
// of the volume's own launch (padding blocks exit at once)
template <class T, int EPT, int NT, bool QUAD, int S, bool MF, class Ld, class Md, class St>
__global__ __launch_bounds__(NT) void k_conv_axis_batch(const ConvJob<Ld, Md, St>* __restrict__ jobs,
                                                        int njobs)
{
    extern __shared__ __align__(16) unsigned char smem[];
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const ConvJob<Ld, Md, St> jb = jobs[lo];
    const unsigned w = xcd_remap(blockIdx.x - (unsigned)jb.first, (unsigned)jb.nblk);
    if (w >= (unsigned)jb.groups) return;
    conv_group<T, EPT, NT, QUAD, S, MF>(jb.ax, jb.g, jb.rounds, jb.ld, jb.md, jb.st,
                                        (long)w * jb.g.P, smem);
}

// elements per workgroup: 4096 complex FP32 / 2048 complex FP64 (<= ~60 KB of LDS)
template <class T>
constexpr int conv_ept()
{
    return sizeof(T) == 4 ? 16 : 8;
}

inline AxisGeom make_geom(int axis, int nx, int ny, int nz, int L, int elems)
{
    AxisGeom g;
    g.axis = axis;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.nlines = (long)nx * ny * nz / L;
    g.stride = axis == 2 ? 1 : (axis == 1 ? nz : (long)ny * nz);
    int P = 1, lp = 0;
    while (P * 2 * L <= elems && P < 256) {
        P *= 2;
        ++lp;
    }
    g.P = P;
    g.logP = lp;
    return g;
}

// half-widths with a compiled register-window direct path
inline bool direct_width_supported(int nh) { return nh == 19 || nh == 25 || nh == 37; }

// The matrix-core window runs groups of 2048 complex elements (half the vector-ALU
// window's), which halves the LDS of a workgroup and doubles the workgroups resident per
// CU.
constexpr int kMfmaEpt = 8;

// the matrix-core direct window fits this axis: FP32, length a multiple of 16, groups of
// a multiple of 8 pencils
inline bool mfma_window_fits(int axis, const int dims[3])
{
    const int L = dims[axis];
    const AxisGeom g = make_geom(axis, dims[0], dims[1], dims[2], L, kMfmaEpt * 256);
    return L % 16 == 0 && g.P % 8 == 0 && L * g.P <= kMfmaEpt * 256;
}

// what a convolution pass of one filter axis runs: the axis with its boundary mode, the
// pencil-group geometry, the variant (register window of half-width S on the vector ALUs
// or the matrix cores, or the FFT / generic direct pass), the grid and the LDS bytes
struct ConvPlan {
    AxisDev ax;
    AxisGeom g;
    bool mf = false, quad = false;
    bool small = false;  // FP32 register window in groups of 2048 elements on 256 threads
    int S = 0;
    unsigned nb = 0;
    size_t lds = 0;
};

template <class T>
int conv_plan(const tk_filter_s* f, int axis, bool acyclic, ConvPlan* out)
{
    constexpr int NT = 256;
    constexpr int EPT = conv_ept<T>();
    ConvPlan p;
    // acyclic (zero outside the volume): direct passes with zero halos; an axis on the FFT
    // method runs the register window where one is compiled, else the generic direct pass
    p.ax = f->ax[axis];
    p.ax.zero_halo = acyclic ? 1 : 0;
    const AxisDev& ax = p.ax;
    const bool dims4 = f->dims[0] % 4 == 0 && f->dims[1] % 4 == 0 && f->dims[2] % 4 == 0;
    // the matrix-core window (checked when the method was chosen; checked again here since
    // the group shape decides the tiling)
    p.mf = std::is_same<T, float>::value && ax.direct == 2 && dims4 &&
           direct_width_supported(ax.nh) && mfma_window_fits(axis, f->dims);
    // the FP32 register window of the z axis runs groups of 2048 elements on 256 threads,
    // like the matrix-core window (six workgroups per CU at different phases of their load,
    // arithmetic and store: the z pass with the region fit 107 -> 100 us at 256^3); x and y
    // keep groups of 4096 on 512 threads (x 76 -> 87 us with the smaller groups)
    p.small = std::is_same<T, float>::value && axis == 2 &&
              (ax.direct == 1 || (acyclic && direct_width_supported(ax.nh)));
    const int elems = (p.mf || p.small ? kMfmaEpt : EPT) * NT;
    p.g = make_geom(axis, f->dims[0], f->dims[1], f->dims[2], ax.L, elems);
    TK_CHECK(ax.L * p.g.P <= EPT * NT, TK_ERR_UNSUPPORTED, "conv: axis length %d exceeds %d",
             ax.L, EPT * NT);
    p.nb = (unsigned)((p.g.nlines + p.g.P - 1) / p.g.P);
    p.quad = dims4 && p.g.P >= 4;
    // register-window direct path for the instantiated half-widths (sigma 3, 4, 6 at h=1)
    p.S = (ax.direct || (acyclic && direct_width_supported(ax.nh))) && p.quad ? ax.nh / 2 : 0;
    p.lds = (size_t)conv_nbuf(axis, ax.L, p.g.P, p.S) * sizeof(cplx<T>) + ax.L * sizeof(cplx<T>) +
            (size_t)(ax.L > ax.nh ? ax.L : ax.nh) * sizeof(T);
    *out = p;
    return TK_OK;
}

// the kernel instance of a plan's variant, launched as KERNEL_LAUNCH(kernel<...>, threads)
#define TK_CONV_DISPATCH(plan, KERNEL_LAUNCH)                                                  \
    do {                                                                                       \
        constexpr int NT_ = 256, EPT_ = conv_ept<T>();                                         \
        const int S_ = (plan).S;                                                               \
        if constexpr (std::is_same<T, float>::value) {                                         \
            if ((plan).mf && S_ == 9) KERNEL_LAUNCH((T, kMfmaEpt, NT_, true, 9, true), NT_);    \
            else if ((plan).mf && S_ == 12) KERNEL_LAUNCH((T, kMfmaEpt, NT_, true, 12, true), NT_); \
            else if ((plan).mf && S_ == 18) KERNEL_LAUNCH((T, kMfmaEpt, NT_, true, 18, true), NT_); \
            else if ((plan).small && S_ == 9) KERNEL_LAUNCH((T, kMfmaEpt, NT_, true, 9, false), NT_); \
            else if ((plan).small && S_ == 12)                                                 \
                KERNEL_LAUNCH((T, kMfmaEpt, NT_, true, 12, false), NT_);                       \
            else if ((plan).small && S_ == 18)                                                 \
                KERNEL_LAUNCH((T, kMfmaEpt, NT_, true, 18, false), NT_);                       \
            else if (S_ == 9) KERNEL_LAUNCH((T, EPT_ / 2, 2 * NT_, true, 9, false), 2 * NT_);   \
            else if (S_ == 12) KERNEL_LAUNCH((T, EPT_ / 2, 2 * NT_, true, 12, false), 2 * NT_); \
            else if (S_ == 18) KERNEL_LAUNCH((T, EPT_ / 2, 2 * NT_, true, 18, false), 2 * NT_); \
            else if ((plan).quad) KERNEL_LAUNCH((T, EPT_, NT_, true, 0, false), NT_);          \
            else KERNEL_LAUNCH((T, EPT_, NT_, false, 0, false), NT_);                          \
        } else {                                                                               \
            if (S_ == 9) KERNEL_LAUNCH((T, EPT_, NT_, true, 9, false), NT_);                    \
            else if (S_ == 12) KERNEL_LAUNCH((T, EPT_, NT_, true, 12, false), NT_);             \
            else if (S_ == 18) KERNEL_LAUNCH((T, EPT_, NT_, true, 18, false), NT_);             \
            else if ((plan).quad) KERNEL_LAUNCH((T, EPT_, NT_, true, 0, false), NT_);          \
            else KERNEL_LAUNCH((T, EPT_, NT_, false, 0, false), NT_);                          \
        }                                                                                      \
    } while (0)

#define TK_UNPACK_(...) __VA_ARGS__

template <class T, class Ld, class Md, class St>
int launch_conv_axis(tk_ctx ctx, const tk_filter_s* f, int axis, int rounds, Ld ld, Md md, St st,
                     bool acyclic = false)
{
    ConvPlan p;
    TK_TRY(conv_plan<T>(f, axis, acyclic, &p));
#define TK_CONV_ONE(ARGS, THREADS)                                                             \
    hipLaunchKernelGGL((k_conv_axis<TK_UNPACK_ ARGS, Ld, Md, St>), dim3(p.nb), dim3(THREADS), \
                       p.lds, ctx->stream, p.ax, p.g, rounds, ld, md, st)
    TK_CONV_DISPATCH(p, TK_CONV_ONE);
#undef TK_CONV_ONE
    TK_LAUNCHED();
    return TK_OK;
}

// one pass of many volumes whose plans share the variant of `like` (jobs in device memory,
// nblocks blocks in all, lds the largest of their LDS sizes)
template <class T, class Ld, class Md, class St>
int launch_conv_batch(tk_ctx ctx, const ConvPlan& like, const ConvJob<Ld, Md, St>* jobs, int njobs,
                      unsigned nblocks, size_t lds)
{
#define TK_CONV_MANY(ARGS, THREADS)                                                            \
    hipLaunchKernelGGL((k_conv_axis_batch<TK_UNPACK_ ARGS, Ld, Md, St>), dim3(nblocks),       \
                       dim3(THREADS), lds, ctx->stream, jobs, njobs)
    TK_CONV_DISPATCH(like, TK_CONV_MANY);
#undef TK_CONV_MANY
    TK_LAUNCHED();
    return TK_OK;
}

}  // namespace tk
