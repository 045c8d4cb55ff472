// tk_stencil.hpp -- per-voxel difference operators of im3d::image3d (device).
#pragma once

#include <hip/hip_runtime.h>

namespace tk {

// one volume of a batched Split-Bregman update launch (k_sb_update_batch): its fields, the
// cut bounds, its block partials and tolerance sums, its row grid (gy block rows, nparts
// blocks) and block range (first a multiple of the XCD count, nblk padded to one)
struct SbJob {
    void* phi;
    const void* phiold;
    void* b[3];
    const void* g;
    void* q[3];
    int X, Y, Z;
    double hx, hy, hz;
    double lo, hi;
    double* part;
    double* tol;
    int gy, nparts;
    int first, nblk;
};

// one volume of a batched div launch (k_div_batch)
struct DivJob {
    const void* q[3];
    void* out;
    int X, Y, Z;
    double hx, hy, hz;
    int gy;
    int first, nblk;
};

// central inside, one-sided second order on the first / last plane
// (image3d_imp.hxx:767-825); c = 0 first plane, 1 inside, 2 last plane
template <class T>
__device__ __forceinline__ T dterm(const T* __restrict__ f, long v, long st, int c, int n, int idx,
                                   double h)
{
    if (c == 1) return (f[v + st] - f[v - st]) / (T)(2. * h);
    const long base = v - (long)idx * st;  // element 0 of this line
    if (c == 0) return (4 * f[base + st] - 3 * f[base] - f[base + 2 * st]) / (T)(2. * h);
    return (3 * f[base + (long)(n - 1) * st] - 4 * f[base + (long)(n - 2) * st] +
            f[base + (long)(n - 3) * st]) /
           (T)(2. * h);
}

__device__ __forceinline__ int plane_class(int i, int n) { return i == 0 ? 0 : (i == n - 1 ? 2 : 1); }

// im3d::div at one voxel (image3d_imp.hxx:947-1383), 3-D, with its two quirks:
//  - face i = X-1 (j, k inside): x term "/ 2. * (hx)" (1092-1095)
//  - edge (i inside, 0, Z-1): z term uses Z-2 twice (1192-1194)
template <class T>
__device__ __forceinline__ T div_at(const T* __restrict__ f0, const T* __restrict__ f1,
                                    const T* __restrict__ f2, int i, int j, int k, int X, int Y,
                                    int Z, double hx, double hy, double hz)
{
    const long sX = (long)Y * Z, sY = Z;
    const long v = i * sX + j * sY + k;
    const int ci = plane_class(i, X), cj = plane_class(j, Y), ck = plane_class(k, Z);
    T tx, ty, tz;
    if (ci == 2 && cj == 1 && ck == 1) {
        const long b = v - (long)i * sX;
        tx = (3 * f0[b + (long)(X - 1) * sX] - 4 * f0[b + (long)(X - 2) * sX] +
              f0[b + (long)(X - 3) * sX]) /
             (T)2. * (T)hx;
    } else {
        tx = dterm(f0, v, sX, ci, X, i, hx);
    }
    ty = dterm(f1, v, sY, cj, Y, j, hy);
    if (ci == 1 && cj == 0 && ck == 2) {
        const long b = v - k;
        tz = (3 * f2[b + (Z - 1)] - 4 * f2[b + (Z - 2)] + f2[b + (Z - 2)]) / (T)(2. * hz);
    } else {
        tz = dterm(f2, v, 1, ck, Z, k, hz);
    }
    return tx + ty + tz;
}

}  // namespace tk
