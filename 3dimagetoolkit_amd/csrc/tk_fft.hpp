// tk_fft.hpp -- cyclic Gaussian convolution as per-axis FFT passes in LDS.
//
// The reference convolves through a 3-D double-complex FFT, a multiply by the real
// spectrum of the wrapped Gaussian and an inverse FFT (convolution_imp.hxx:657-757).
// The normalised Gaussian is separable and even on every axis, so its 3-D real
// spectrum is the product of three 1-D real spectra, and the 3-D convolution is the
// composition of three 1-D circular convolutions.  Each axis pass below therefore
// loads a group of pencils into LDS, runs forward FFT -> multiply by the axis
// spectrum -> inverse FFT in LDS and writes back once: three passes per convolution
// instead of an out-of-core 3-D FFT pair, with prologue/epilogue functors that fuse
// the pointwise stages of the Split-Bregman iteration into the first/last pass.
//
// FFT: Stockham autosort, radix 2/3/4/5 stages (Govindaraju et al., "High
// Performance Discrete Fourier Transforms on Graphics Processors", SC 2008),
// register-staged in place in LDS: every thread reads its butterflies, barrier,
// writes them to the autosort positions, barrier.  Twiddles come from a table
// computed on the host in long double and rounded once.
#pragma once

#include <hip/hip_runtime.h>

#include "tk_internal.hpp"

namespace tk {

template <class T>
struct cplx {
    T x, y;
};

template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * (s*i), s = +-1
template <class T>
__device__ __forceinline__ cplx<T> cmul_si(cplx<T> a, int s)
{
    return s > 0 ? cplx<T>{-a.y, a.x} : cplx<T>{a.y, -a.x};
}
template <class T>
__device__ __forceinline__ cplx<T> cscale(cplx<T> a, T s) { return {a.x * s, a.y * s}; }

constexpr int kMaxStages = 12;

// one axis of a filter: length, radix plan, and device tables
struct AxisDev {
    int L = 1;
    int nstage = 0;              // 0 with L > 1: length not 2/3/5-smooth, direct convolution
    int radix[kMaxStages] = {0};
    int nh = 1;                  // taps of the direct path (odd)
    int direct = 0;              // symmetric direct circular convolution in halo-padded LDS:
                                 // 1 register window on the vector ALUs, 2 on the matrix cores
    const void* tw = nullptr;    // cplx<T>[L], exp(-2 pi i t / L)
    const void* spec = nullptr;  // T[L], real spectrum / L (FFT path) or T[nh] taps (direct)
    const void* half = nullptr;  // T[nh/2 + 1]: w(0), w(1), .. w(s) of the direct path
    const void* band = nullptr;  // FP32 [KS][64]: the matrix-core window's band operand
    const void* taps = nullptr;  // T[nh]: the normalised taps w(-s) .. w(s)
    int zero_halo = 0;           // set per launch: 1 = acyclic (zero outside the volume)
};

// how the pencils of one axis sit in memory
struct AxisGeom {
    int axis;      // 0 = x (stride ny*nz), 1 = y (stride nz), 2 = z (contiguous)
    int nx, ny, nz;
    long nlines;   // number of pencils
    long stride;   // element stride along the axis
    int P;         // pencils per workgroup (a power of two)
    int logP;
};

// base voxel index of pencil c
__device__ __forceinline__ long pencil_base(const AxisGeom& g, long c)
{
    if (g.axis == 2) return c * g.nz;
    if (g.axis == 1) return (c / g.nz) * (long)g.ny * g.nz + (c % g.nz);
    return c;
}

int fft_plan(int L, int* radix, int* nstage, int max_radix);

}  // namespace tk

struct tk_filter_s {
    tk_ctx ctx = nullptr;
    int dims[3] = {1, 1, 1};
    double h[3] = {1, 1, 1};
    int nh[3] = {1, 1, 1};
    double sigma = 0, radius = 0;
    int pixel = 0;
    tk_dtype dtype = TK_F64;
    std::vector<double> spec_host[3];  // real spectrum along each axis (not divided by L)
    void* spec_dev[3] = {nullptr, nullptr, nullptr};
    void* tw_dev[3] = {nullptr, nullptr, nullptr};
    void* half_dev[3] = {nullptr, nullptr, nullptr};
    void* band_dev[3] = {nullptr, nullptr, nullptr};
    void* taps_dev[3] = {nullptr, nullptr, nullptr};
    tk::AxisDev ax[3];
};
