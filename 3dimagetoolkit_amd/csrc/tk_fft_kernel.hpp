// tk_fft_kernel.hpp -- the per-axis convolution kernel and its launcher, templated
// on the element type and on load / mid / store functors.  Included by every
// translation unit that fuses a stage into a convolution pass.
#pragma once

#include <type_traits>

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

template <class T, int R>
__device__ __forceinline__ void dftR(cplx<T>* a, int s)
{
    if constexpr (R == 2) dft2(a, s);
    else if constexpr (R == 3) dft3(a, s);
    else if constexpr (R == 4) dft4(a, s);
    else dft5(a, s);
}

// one Stockham stage over P pencils of length L held in LDS as buf[pos*PP + p]
template <class T, int R, int EPT, int NT>
__device__ __forceinline__ void fft_stage(cplx<T>* __restrict__ buf, const cplx<T>* __restrict__ tw,
                                          int L, int P, int PP, int Ns, int sign)
{
    constexpr int MAXB = (EPT + R - 1) / R;
    const int LR = L / R;
    const int nbf = P * LR;
    const int stw = L / (Ns * R);
    cplx<T> v[MAXB][R];
    int dst[MAXB];
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        const int b = threadIdx.x + q * NT;
        dst[q] = -1;
        if (b < nbf) {
            const int p = b % P, j = b / P;
            const int k = j % Ns;
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = buf[(j + r * LR) * PP + p];
#pragma unroll
            for (int r = 1; r < R; ++r) {
                cplx<T> w = tw[r * k * stw];
                if (sign > 0) w.y = -w.y;
                v[q][r] = cmul(v[q][r], w);
            }
            dftR<T, R>(v[q], sign);
            dst[q] = ((j / Ns) * Ns * R + k) * PP + p;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXB; ++q) {
        if (dst[q] >= 0) {
#pragma unroll
            for (int r = 0; r < R; ++r) buf[dst[q] + r * Ns * PP] = v[q][r];
        }
    }
    __syncthreads();
}

template <class T, int EPT, int NT>
__device__ __forceinline__ void fft_lds(cplx<T>* buf, const cplx<T>* tw, const AxisDev& ax, int P,
                                        int PP, int sign)
{
    int Ns = 1;
    for (int s = 0; s < ax.nstage; ++s) {
        const int R = ax.radix[s];
        if (R == 4) fft_stage<T, 4, EPT, NT>(buf, tw, ax.L, P, PP, Ns, sign);
        else if (R == 2) fft_stage<T, 2, EPT, NT>(buf, tw, ax.L, P, PP, Ns, sign);
        else if (R == 3) fft_stage<T, 3, EPT, NT>(buf, tw, ax.L, P, PP, Ns, sign);
        else fft_stage<T, 5, EPT, NT>(buf, tw, ax.L, P, PP, Ns, sign);
        Ns *= R;
    }
}

// direct circular convolution along the pencil with nh taps (lengths that are not
// 2/3/5-smooth): y(i) = sum_n w(n) x((i - n) mod L), n in [-s, s]
template <class T, int EPT, int NT>
__device__ __forceinline__ void direct_lds(cplx<T>* buf, const T* w, const AxisDev& ax, int P,
                                           int PP)
{
    const int L = ax.L, s = ax.nh / 2;
    const int E = L * P;
    cplx<T> out[EPT];
#pragma unroll
    for (int q = 0; q < EPT; ++q) {
        const int e = threadIdx.x + q * NT;
        out[q] = {0, 0};
        if (e < E) {
            const int pos = e / P, p = e % P;
            T ax_ = 0, ay = 0;
            for (int n = -s; n <= s; ++n) {
                int m = (pos - n) % L;
                if (m < 0) m += L;
                const cplx<T> v = buf[m * PP + p];
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
        if (e < E) buf[(e / P) * PP + e % P] = out[q];
    }
    __syncthreads();
}

// forward FFT, multiply by the (real, 1/L-scaled) spectrum, inverse FFT
template <class T, int EPT, int NT>
__device__ __forceinline__ void conv_lds(cplx<T>* buf, const cplx<T>* tw, const T* spec,
                                         const AxisDev& ax, int P, int PP)
{
    if (ax.nstage == 0) {
        direct_lds<T, EPT, NT>(buf, spec, ax, P, PP);
        return;
    }
    fft_lds<T, EPT, NT>(buf, tw, ax, P, PP, -1);
    const int E = ax.L * P;
    for (int e = threadIdx.x; e < E; e += NT) {
        const int pos = e / P, p = e % P;
        cplx<T> v = buf[pos * PP + p];
        const T s = spec[pos];
        buf[pos * PP + p] = {v.x * s, v.y * s};
    }
    __syncthreads();
    fft_lds<T, EPT, NT>(buf, tw, ax, P, PP, +1);
}

// ---- functors ---------------------------------------------------------------------
template <class T>
struct LdComplex {
    const cplx<T>* src;
    __device__ cplx<T> operator()(long v) const { return src[v]; }
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
};
struct MdNone {
};
template <class T>
struct StComplex {
    cplx<T>* dst;
    __device__ void operator()(long v, cplx<T> c) const { dst[v] = c; }
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
};

template <class T, int EPT, int NT, class Ld, class Md, class St>
__global__ __launch_bounds__(NT) void k_conv_axis(AxisDev ax, AxisGeom g, int rounds, Ld ld,
                                                  Md md, St st)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const int L = ax.L, P = g.P, PP = g.P + 1;
    cplx<T>* buf = (cplx<T>*)smem;
    cplx<T>* tw = buf + L * PP;
    T* spec = (T*)(tw + L);
    if (ax.nstage > 0) {
        for (int t = threadIdx.x; t < L; t += NT) {
            tw[t] = ((const cplx<T>*)ax.tw)[t];
            spec[t] = ((const T*)ax.spec)[t];
        }
    } else {
        for (int t = threadIdx.x; t < ax.nh; t += NT) spec[t] = ((const T*)ax.spec)[t];
    }
    const long c0 = (long)blockIdx.x * P;
    const int E = L * P;
    // load: coalesced along memory; z pencils are contiguous rows, x/y pencils are
    // P adjacent columns of a plane
    for (int e = threadIdx.x; e < E; e += NT) {
        int pos, p;
        if (g.axis == 2) { p = e / L; pos = e % L; }
        else { pos = e / P; p = e % P; }
        cplx<T> v = {0, 0};
        if (c0 + p < g.nlines) v = ld(pencil_base(g, c0 + p) + pos * g.stride);
        buf[pos * PP + p] = v;
    }
    __syncthreads();
    conv_lds<T, EPT, NT>(buf, tw, spec, ax, P, PP);
    if constexpr (!std::is_same<Md, MdNone>::value) {
        if (rounds > 1) {
            for (int e = threadIdx.x; e < E; e += NT) {
                const int pos = e / P, p = e % P;
                if (c0 + p < g.nlines) {
                    const long v = pencil_base(g, c0 + p) + pos * g.stride;
                    buf[pos * PP + p] = md(v, buf[pos * PP + p]);
                }
            }
            __syncthreads();
            conv_lds<T, EPT, NT>(buf, tw, spec, ax, P, PP);
        }
    }
    for (int e = threadIdx.x; e < E; e += NT) {
        int pos, p;
        if (g.axis == 2) { p = e / L; pos = e % L; }
        else { pos = e / P; p = e % P; }
        if (c0 + p < g.nlines) st(pencil_base(g, c0 + p) + pos * g.stride, buf[pos * PP + p]);
    }
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
    int P = 1;
    while (P * 2 * L <= elems) P *= 2;
    g.P = P;
    return g;
}

template <class T, class Ld, class Md, class St>
int launch_conv_axis(tk_ctx ctx, const tk_filter_s* f, int axis, int rounds, Ld ld, Md md, St st)
{
    constexpr int NT = 256;
    constexpr int EPT = conv_ept<T>();
    const AxisDev& ax = f->ax[axis];
    AxisGeom g = make_geom(axis, f->dims[0], f->dims[1], f->dims[2], ax.L, EPT * NT);
    TK_CHECK(g.P >= 1 && ax.L * g.P <= EPT * NT, TK_ERR_UNSUPPORTED,
             "conv: axis length %d exceeds %d", ax.L, EPT * NT);
    size_t lds = (size_t)ax.L * (g.P + 1) * sizeof(cplx<T>) + ax.L * sizeof(cplx<T>) +
                 (size_t)(ax.L > ax.nh ? ax.L : ax.nh) * sizeof(T);
    unsigned nb = (unsigned)((g.nlines + g.P - 1) / g.P);
    hipLaunchKernelGGL((k_conv_axis<T, EPT, NT, Ld, Md, St>), dim3(nb), dim3(NT), lds,
                       ctx->stream, ax, g, rounds, ld, md, st);
    TK_LAUNCHED();
    return TK_OK;
}

}  // namespace tk
