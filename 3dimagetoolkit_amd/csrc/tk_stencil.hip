// tk_stencil.hip -- image3d::grad, im3d::div, the GRAD edge detector and the fused
// Split-Bregman update (rsfe_splitbregman_imp.hxx:1868-1936).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "tk_device.hpp"
#include "tk_internal.hpp"
#include "tk_stencil.hpp"

namespace tk {

template <class T>
__global__ void k_grad(const T* __restrict__ f, T* __restrict__ g0, T* __restrict__ g1,
                       T* __restrict__ g2, int X, int Y, int Z, double hx, double hy, double hz)
{
    const long sX = (long)Y * Z, sY = Z;
    TK_FOR_VOXELS(X, Y, Z, i, j, k, v)
    {
        g0[v] = dterm(f, v, sX, plane_class(i, X), X, i, hx);
        g1[v] = dterm(f, v, sY, plane_class(j, Y), Y, j, hy);
        g2[v] = dterm(f, v, 1, plane_class(k, Z), Z, k, hz);
    }
}

template <class T>
__global__ void k_div(const T* __restrict__ f0, const T* __restrict__ f1, const T* __restrict__ f2,
                      T* __restrict__ out, int X, int Y, int Z, double hx, double hy, double hz)
{
    TK_FOR_VOXELS(X, Y, Z, i, j, k, v)
    {
        out[v] = div_at(f0, f1, f2, i, j, k, X, Y, Z, hx, hy, hz);
    }
}

// rsfe_splitbregman_imp.hxx:233-242 and 308-313 (GRAD, edge_detector_sigma == 0)
template <class T>
__global__ void k_edge(const T* __restrict__ f, T* __restrict__ g, int X, int Y, int Z, double hx,
                       double hy, double hz, T beta, T lamda, T nu)
{
    const long sX = (long)Y * Z, sY = Z;
    TK_FOR_VOXELS(X, Y, Z, i, j, k, v)
    {
        const T a0 = dterm(f, v, sX, plane_class(i, X), X, i, hx);
        const T a1 = dterm(f, v, sY, plane_class(j, Y), Y, j, hy);
        const T a2 = dterm(f, v, 1, plane_class(k, Z), Z, k, hz);
        T aux = a0 * a0;
        aux += a1 * a1;
        aux += a2 * a2;
        aux *= beta;
        aux += (T)1.;
        T r = (T)1.;
        r /= aux;
        r /= lamda;
        r *= nu;
        g[v] = r;
    }
}

template <class T>
__device__ __forceinline__ T clampv(T x, T a0, T b0)
{
    if (x < a0) return a0;
    if (x > b0) return b0;
    return x;
}

// One pass over the volume after the Poisson loop:
//   cut_phi (imp:320-339) ; sums for ||phi - phiold||, ||phiold|| (imp:1832) ;
//   d = grad(phi) ; b += d ; d = shrink(b, g) ; b -= d ; q = d - b   (imp:1868-1906)
// phi is clamped on the fly when read as a neighbour (the clamp is idempotent), so the
// separate cut pass disappears.  Writes phi (clamped), b and q; q feeds im3d::div.
template <class T>
__global__ __launch_bounds__(256) void k_sb_update(T* __restrict__ phi, const T* __restrict__ phiold,
                                                   T* __restrict__ b0, T* __restrict__ b1,
                                                   T* __restrict__ b2, const T* __restrict__ g,
                                                   T* __restrict__ q0, T* __restrict__ q1,
                                                   T* __restrict__ q2, int X, int Y, int Z,
                                                   double hx, double hy, double hz, T lo, T hi,
                                                   double* __restrict__ part)
{
    const long sX = (long)Y * Z, sY = Z;
    double s1 = 0, s2 = 0;
    TK_FOR_VOXELS(X, Y, Z, i, j, k, v)
    {
        // clamped phi along the three lines through v
        auto cp = [&](long w) { return clampv(phi[w], lo, hi); };
        T d[3];
        {
            const int c = plane_class(i, X);
            const long base = v - (long)i * sX;
            if (c == 1) d[0] = (cp(v + sX) - cp(v - sX)) / (T)(2. * hx);
            else if (c == 0) d[0] = (4 * cp(base + sX) - 3 * cp(base) - cp(base + 2 * sX)) / (T)(2. * hx);
            else
                d[0] = (3 * cp(base + (long)(X - 1) * sX) - 4 * cp(base + (long)(X - 2) * sX) +
                        cp(base + (long)(X - 3) * sX)) /
                       (T)(2. * hx);
        }
        {
            const int c = plane_class(j, Y);
            const long base = v - (long)j * sY;
            if (c == 1) d[1] = (cp(v + sY) - cp(v - sY)) / (T)(2. * hy);
            else if (c == 0) d[1] = (4 * cp(base + sY) - 3 * cp(base) - cp(base + 2 * sY)) / (T)(2. * hy);
            else
                d[1] = (3 * cp(base + (long)(Y - 1) * sY) - 4 * cp(base + (long)(Y - 2) * sY) +
                        cp(base + (long)(Y - 3) * sY)) /
                       (T)(2. * hy);
        }
        {
            const int c = plane_class(k, Z);
            const long base = v - k;
            if (c == 1) d[2] = (cp(v + 1) - cp(v - 1)) / (T)(2. * hz);
            else if (c == 0) d[2] = (4 * cp(base + 1) - 3 * cp(base) - cp(base + 2)) / (T)(2. * hz);
            else
                d[2] = (3 * cp(base + (Z - 1)) - 4 * cp(base + (Z - 2)) + cp(base + (Z - 3))) /
                       (T)(2. * hz);
        }
        const T pc = cp(v);
        const T po = phiold[v];
        {
            const double dd = (double)pc - (double)po;
            s1 += dd * dd;
            s2 += (double)po * (double)po;
        }
        const T gv = g[v];
        T* bp[3] = {b0, b1, b2};
        T* qp[3] = {q0, q1, q2};
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            T bb = bp[a][v];
            bb += d[a];
            const T m = fabs(bb) - gv;
            const T z = 0;
            T dn = (T)((int)(bb > 0) - (int)(bb < 0)) * ((m < z) ? z : m);
            bb -= dn;
            dn -= bb;
            bp[a][v] = bb;
            qp[a][v] = dn;
        }
        // the clamped value is written after every read of this voxel in this thread;
        // other threads read it through cp(), which clamps again
        phi[v] = pc;
    }
    block_sum2(s1, s2);
    if (flat_tid() == 0) {
        const int b = blockIdx.x + gridDim.x * blockIdx.y;
        part[2 * b] = s1;
        part[2 * b + 1] = s2;
    }
}

// ---- 16-byte row versions (nz a multiple of V): each lane owns V consecutive voxels
// of a line; the x and y neighbours are whole rows, the z neighbours a (V+2)-window.
// Arithmetic and operand order are those of the per-voxel kernels above.

// x / (2h), the reference's difference quotient.  With P2 (2h a power of two, unit spacing
// included) the product with the reciprocal, itself exact, is the same number (both round
// the exact quotient once) for one instruction instead of a division sequence.
template <class T, bool P2>
__device__ __forceinline__ T q2h(T x, double h)
{
    if constexpr (P2) return x * (T)(0.5 / h);
    else return x / (T)(2. * h);
}

// 2h is a power of two well inside the range of either type
__host__ __device__ inline bool pow2_2h(double h)
{
    int e = 0;
    const double m = frexp(2. * h, &e);
    return m == 0.5 && e > -100 && e < 100;
}

__host__ __device__ inline bool pow2_2h3(double hx, double hy, double hz)
{
    return pow2_2h(hx) && pow2_2h(hy) && pow2_2h(hz);
}

// derivative along z of element e of a window c[0..V+1] (c[1+e] = f(k0+e))
template <class T, int V, bool P2>
__device__ __forceinline__ T dz_win(const T* c, int e, int k, int Z, double h)
{
    if (k == 0) return q2h<T, P2>(4 * c[2] - 3 * c[1] - c[3], h);
    if (k == Z - 1) return q2h<T, P2>(3 * c[e + 1] - 4 * c[e] + c[e - 1], h);
    return q2h<T, P2>(c[e + 2] - c[e], h);
}

// The rows of f that d/dx (or d/dy) of the row at index `idx` reads, along an axis of
// length n with line stride st: central -> (idx+1, idx-1); first -> (0, 1, 2); last ->
// (n-1, n-2, n-3).  Loading is split from the arithmetic so that a thread issues every
// load of its row (all axes, all fields) before it uses any of them: a load followed by
// its use inside a branch makes the thread wait for it before issuing the next axis.
template <class T>
struct DRows {
    Row<T> r[3];
    int c;  // plane_class(idx, n)
};

template <class T>
__device__ __forceinline__ DRows<T> drow_load(const T* __restrict__ f, long v0, long st, int idx,
                                              int n)
{
    DRows<T> d;
    d.c = plane_class(idx, n);
    const long base = v0 - (long)idx * st;
    const long o0 = d.c == 1 ? v0 + st : (d.c == 0 ? base : base + (long)(n - 1) * st);
    const long o1 = d.c == 1 ? v0 - st : (d.c == 0 ? base + st : base + (long)(n - 2) * st);
    d.r[0] = ldrow(f + o0);
    d.r[1] = ldrow(f + o1);
    if (d.c != 1) d.r[2] = ldrow(f + (d.c == 0 ? base + 2 * st : base + (long)(n - 3) * st));
    return d;
}

// central -> (f(+1) - f(-1)) / 2h; first -> (4 f(1) - 3 f(0) - f(2)) / 2h; last ->
// (3 f(n-1) - 4 f(n-2) + f(n-3)) / 2h, each value through clampf
template <class T, bool P2, class Fn>
__device__ __forceinline__ void drow_eval(const DRows<T>& d, double h, Fn clampf, T* out)
{
    constexpr int V = Vec16<T>::n;
    const Row<T>&a = d.r[0], &b = d.r[1], &c = d.r[2];
    if (d.c == 1) {
#pragma unroll
        for (int e = 0; e < V; ++e) out[e] = q2h<T, P2>(clampf(a.v[e]) - clampf(b.v[e]), h);
    } else if (d.c == 0) {
#pragma unroll
        for (int e = 0; e < V; ++e)
            out[e] = q2h<T, P2>(4 * clampf(b.v[e]) - 3 * clampf(a.v[e]) - clampf(c.v[e]), h);
    } else {
#pragma unroll
        for (int e = 0; e < V; ++e)
            out[e] = q2h<T, P2>(3 * clampf(a.v[e]) - 4 * clampf(b.v[e]) + clampf(c.v[e]), h);
    }
}

// The rows a thread visits in a gx x gy grid of bdx x bdy blocks, as thread (tx, ty) of
// block (bx, by): the grid of one launch, or one volume's share of a batched launch.
#define TK_FOR_ROWS_IN(X, Y, Z, V, i, j, k0, v0, TX, TY, BDX, BDY, BX, BY, GX, GY)            \
    for (int r_ = (int)((BY) * (BDY) + (TY)); r_ < (X) * (Y); r_ += (int)((GY) * (BDY)))      \
        for (int k0 = (int)((BX) * (BDX) + (TX)) * (V), i = r_ / (Y), j = r_ - i * (Y);      \
             k0 < (Z); k0 += (int)((GX) * (BDX)) * (V))                                       \
            for (long v0 = (long)r_ * (Z) + k0, once_ = 1; once_; once_ = 0)

// the Split-Bregman update on the rows of one thread (sums of the tolerance into s1, s2)
template <class T, bool P2>
__device__ __forceinline__ void sb_rows(T* __restrict__ phi, const T* __restrict__ phiold,
                                        T* __restrict__ b0, T* __restrict__ b1, T* __restrict__ b2,
                                        const T* __restrict__ g, T* __restrict__ q0,
                                        T* __restrict__ q1, T* __restrict__ q2, int X, int Y, int Z,
                                        double hx, double hy, double hz, T lo, T hi, double& s1,
                                        double& s2, int tx, int ty, int bdx, int bdy, unsigned bx,
                                        unsigned by, unsigned gx, unsigned gy)
{
    constexpr int V = Vec16<T>::n;
    const long sX = (long)Y * Z, sY = Z;
    auto cl = [&](T x) { return clampv(x, lo, hi); };
    TK_FOR_ROWS_IN(X, Y, Z, V, i, j, k0, v0, tx, ty, bdx, bdy, bx, by, gx, gy)
    {
        T* bp[3] = {b0, b1, b2};
        T* qp[3] = {q0, q1, q2};
        const DRows<T> rx = drow_load(phi, v0, sX, i, X), ry = drow_load(phi, v0, sY, j, Y);
        const Row<T> cc = ldrow(phi + v0);
        const T zl = k0 > 0 ? phi[v0 - 1] : (T)0, zr = k0 + V < Z ? phi[v0 + V] : (T)0;
        const Row<T> po = ldrow(phiold + v0);
        const Row<T> gv = ldrow(g + v0);
        Row<T> bb[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) bb[a] = ldrow(bp[a] + v0);
        T d[3][V];
        drow_eval<T, P2>(rx, hx, cl, d[0]);
        drow_eval<T, P2>(ry, hy, cl, d[1]);
        T c[V + 2];
        c[0] = k0 > 0 ? cl(zl) : (T)0;
        c[V + 1] = k0 + V < Z ? cl(zr) : (T)0;
#pragma unroll
        for (int e = 0; e < V; ++e) c[e + 1] = cl(cc.v[e]);
#pragma unroll
        for (int e = 0; e < V; ++e) d[2][e] = dz_win<T, V, P2>(c, e, k0 + e, Z, hz);
        Row<T> pc;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            pc.v[e] = c[e + 1];
            const double dd = (double)pc.v[e] - (double)po.v[e];
            s1 += dd * dd;
            s2 += (double)po.v[e] * (double)po.v[e];
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            Row<T> qq;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                T b = bb[a].v[e];
                b += d[a][e];
                const T m = fabs(b) - gv.v[e];
                const T z = 0;
                T dn = (T)((int)(b > 0) - (int)(b < 0)) * ((m < z) ? z : m);
                b -= dn;
                dn -= b;
                bb[a].v[e] = b;
                qq.v[e] = dn;
            }
            strow(bp[a] + v0, bb[a]);
            strow(qp[a] + v0, qq);
        }
        strow(phi + v0, pc);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_sb_update_v(T* __restrict__ phi, const T* __restrict__ phiold,
                                                     T* __restrict__ b0, T* __restrict__ b1,
                                                     T* __restrict__ b2, const T* __restrict__ g,
                                                     T* __restrict__ q0, T* __restrict__ q1,
                                                     T* __restrict__ q2, int X, int Y, int Z,
                                                     double hx, double hy, double hz, T lo, T hi,
                                                     double* __restrict__ part)
{
    // blocks renumbered so that rows j-1, j+1 of a block's edge rows are read on one XCD
    const unsigned w = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    double s1 = 0, s2 = 0;
    if (pow2_2h3(hx, hy, hz))
        sb_rows<T, true>(phi, phiold, b0, b1, b2, g, q0, q1, q2, X, Y, Z, hx, hy, hz, lo, hi, s1,
                         s2, threadIdx.x, threadIdx.y, blockDim.x, blockDim.y, w % gridDim.x,
                         w / gridDim.x, gridDim.x, gridDim.y);
    else
        sb_rows<T, false>(phi, phiold, b0, b1, b2, g, q0, q1, q2, X, Y, Z, hx, hy, hz, lo, hi, s1,
                          s2, threadIdx.x, threadIdx.y, blockDim.x, blockDim.y, w % gridDim.x,
                          w / gridDim.x, gridDim.x, gridDim.y);
    block_sum2(s1, s2);
    if (flat_tid() == 0) {
        const int b = blockIdx.x + gridDim.x * blockIdx.y;
        part[2 * b] = s1;
        part[2 * b + 1] = s2;
    }
}

template <class T, bool P2>
__device__ __forceinline__ void div_rows(const T* __restrict__ f0, const T* __restrict__ f1,
                                         const T* __restrict__ f2, T* __restrict__ out, int X,
                                         int Y, int Z, double hx, double hy, double hz, int tx,
                                         int ty, int bdx, int bdy, unsigned bx, unsigned by,
                                         unsigned gx, unsigned gy)
{
    constexpr int V = Vec16<T>::n;
    const long sX = (long)Y * Z, sY = Z;
    auto id = [](T x) { return x; };
    TK_FOR_ROWS_IN(X, Y, Z, V, i, j, k0, v0, tx, ty, bdx, bdy, bx, by, gx, gy)
    {
        const DRows<T> rx = drow_load(f0, v0, sX, i, X), ry = drow_load(f1, v0, sY, j, Y);
        const Row<T> cc = ldrow(f2 + v0);
        const T zl = k0 > 0 ? f2[v0 - 1] : (T)0, zr = k0 + V < Z ? f2[v0 + V] : (T)0;
        const int ci = rx.c, cj = ry.c;
        T tx[V], ty[V];
        drow_eval<T, P2>(rx, hx, id, tx);
        if (ci == 2 && cj == 1) {
            // face i = X-1 (j, k inside): "/ 2. * (hx)" (image3d_imp.hxx:1092-1095); the
            // rows X-1, X-2, X-3 are those of the last-plane derivative
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const int k = k0 + e;
                if (k >= 1 && k <= Z - 2)
                    // "/ 2." as a product with 0.5: the same number for every operand
                    tx[e] = (3 * rx.r[0].v[e] - 4 * rx.r[1].v[e] + rx.r[2].v[e]) * (T)0.5 * (T)hx;
            }
        }
        drow_eval<T, P2>(ry, hy, id, ty);
        T c[V + 2];
        c[0] = zl;
        c[V + 1] = zr;
#pragma unroll
        for (int e = 0; e < V; ++e) c[e + 1] = cc.v[e];
        Row<T> o;
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int k = k0 + e;
            T tz;
            if (ci == 1 && cj == 0 && k == Z - 1)
                // edge (i inside, 0, Z-1): Z-2 used twice (image3d_imp.hxx:1192-1194)
                tz = q2h<T, P2>(3 * c[e + 1] - 4 * c[e] + c[e], hz);
            else
                tz = dz_win<T, V, P2>(c, e, k, Z, hz);
            o.v[e] = tx[e] + ty[e] + tz;
        }
        strow(out + v0, o);
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_div_v(const T* __restrict__ f0, const T* __restrict__ f1,
                                               const T* __restrict__ f2, T* __restrict__ out, int X,
                                               int Y, int Z, double hx, double hy, double hz)
{
    const unsigned w = xcd_remap(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
    if (pow2_2h3(hx, hy, hz))
        div_rows<T, true>(f0, f1, f2, out, X, Y, Z, hx, hy, hz, threadIdx.x, threadIdx.y,
                          blockDim.x, blockDim.y, w % gridDim.x, w / gridDim.x, gridDim.x,
                          gridDim.y);
    else
        div_rows<T, false>(f0, f1, f2, out, X, Y, Z, hx, hy, hz, threadIdx.x, threadIdx.y,
                           blockDim.x, blockDim.y, w % gridDim.x, w / gridDim.x, gridDim.x,
                           gridDim.y);
}

// ---- many volumes per launch (the batch's staged path) --------------------------------
// Blocks of 256 threads laid out as bdx = Z/V quads (one row of 16-byte words) by
// bdy = 256 / bdx rows, so short rows leave few lanes idle (a 16 x 16 shape left 25-37 %
// idle for z = 48 or 80); each volume's block range starts at a multiple of the XCD count
// and is padded to one.  A block finds its volume by a binary search of the first blocks.
template <class J>
__device__ __forceinline__ const J& batch_job(const J* jobs, int njobs)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    return jobs[lo];
}

__host__ __device__ __forceinline__ void batch_block_shape(int Zv, int* bdx, int* bdy)
{
    *bdx = Zv >= 64 ? 64 : Zv;
    *bdy = 256 / *bdx;
}

template <class T>
__global__ __launch_bounds__(256) void k_sb_update_batch(const SbJob* __restrict__ jobs, int njobs)
{
    const SbJob& jb = batch_job(jobs, njobs);
    constexpr int V = Vec16<T>::n;
    int bdx, bdy;
    batch_block_shape(jb.Z / V, &bdx, &bdy);
    const unsigned gx = (unsigned)((jb.Z / V + bdx - 1) / bdx), gy = (unsigned)jb.gy;
    const unsigned w = xcd_remap(blockIdx.x - (unsigned)jb.first, (unsigned)jb.nblk);
    if (w >= gx * gy) return;
    double s1 = 0, s2 = 0;
    const int ty = threadIdx.x / bdx;
    if (ty < bdy) {
        if (pow2_2h3(jb.hx, jb.hy, jb.hz))
            sb_rows<T, true>((T*)jb.phi, (const T*)jb.phiold, (T*)jb.b[0], (T*)jb.b[1],
                             (T*)jb.b[2], (const T*)jb.g, (T*)jb.q[0], (T*)jb.q[1], (T*)jb.q[2],
                             jb.X, jb.Y, jb.Z, jb.hx, jb.hy, jb.hz, (T)jb.lo, (T)jb.hi, s1, s2,
                             threadIdx.x % bdx, ty, bdx, bdy, w % gx, w / gx, gx, gy);
        else
            sb_rows<T, false>((T*)jb.phi, (const T*)jb.phiold, (T*)jb.b[0], (T*)jb.b[1],
                              (T*)jb.b[2], (const T*)jb.g, (T*)jb.q[0], (T*)jb.q[1], (T*)jb.q[2],
                              jb.X, jb.Y, jb.Z, jb.hx, jb.hy, jb.hz, (T)jb.lo, (T)jb.hi, s1, s2,
                              threadIdx.x % bdx, ty, bdx, bdy, w % gx, w / gx, gx, gy);
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        jb.part[2 * w] = s1;
        jb.part[2 * w + 1] = s2;
    }
}

template <class T>
__global__ __launch_bounds__(256) void k_div_batch(const DivJob* __restrict__ jobs, int njobs)
{
    const DivJob& jb = batch_job(jobs, njobs);
    constexpr int V = Vec16<T>::n;
    int bdx, bdy;
    batch_block_shape(jb.Z / V, &bdx, &bdy);
    const unsigned gx = (unsigned)((jb.Z / V + bdx - 1) / bdx), gy = (unsigned)jb.gy;
    const unsigned w = xcd_remap(blockIdx.x - (unsigned)jb.first, (unsigned)jb.nblk);
    if (w >= gx * gy) return;
    const int ty = threadIdx.x / bdx;
    if (ty < bdy) {
        if (pow2_2h3(jb.hx, jb.hy, jb.hz))
            div_rows<T, true>((const T*)jb.q[0], (const T*)jb.q[1], (const T*)jb.q[2],
                              (T*)jb.out, jb.X, jb.Y, jb.Z, jb.hx, jb.hy, jb.hz,
                              threadIdx.x % bdx, ty, bdx, bdy, w % gx, w / gx, gx, gy);
        else
            div_rows<T, false>((const T*)jb.q[0], (const T*)jb.q[1], (const T*)jb.q[2],
                               (T*)jb.out, jb.X, jb.Y, jb.Z, jb.hx, jb.hy, jb.hz,
                               threadIdx.x % bdx, ty, bdx, bdy, w % gx, w / gx, gx, gy);
    }
}

// each volume's tolerance sums from its block partials, one block per volume, in the order
// of k_sum_pairs_rsfe
__global__ __launch_bounds__(1024) void k_sum_pairs_batch(const SbJob* __restrict__ jobs)
{
    const SbJob& jb = jobs[blockIdx.x];
    double s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < jb.nparts; i += blockDim.x) {
        s1 += jb.part[2 * i];
        s2 += jb.part[2 * i + 1];
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        jb.tol[0] = s1;
        jb.tol[1] = s2;
    }
}

template <class T>
int launch_grad(tk_ctx ctx, tk_vol f, tk_vol g0, tk_vol g1, tk_vol g2)
{
    hipLaunchKernelGGL(k_grad<T>, vox_grid(f->nx, f->ny, f->nz, 8192), vox_block(f->nz), 0,
                       ctx->stream,
                       (const T*)f->d, (T*)g0->d, (T*)g1->d, (T*)g2->d, f->nx, f->ny, f->nz, f->hx,
                       f->hy, f->hz);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
int launch_div(tk_ctx ctx, tk_vol f0, tk_vol f1, tk_vol f2, tk_vol out)
{
    constexpr int V = Vec16<T>::n;
    if (out->nz % V == 0 && out->nz > V) {
        const int zv = out->nz / V;
        hipLaunchKernelGGL(k_div_v<T>, row_grid(out->nx, out->ny, zv, 8192), row_block(zv), 0,
                           ctx->stream, (const T*)f0->d, (const T*)f1->d, (const T*)f2->d,
                           (T*)out->d, out->nx, out->ny, out->nz, out->hx, out->hy, out->hz);
        TK_LAUNCHED();
        return TK_OK;
    }
    hipLaunchKernelGGL(k_div<T>, vox_grid(out->nx, out->ny, out->nz, 8192), vox_block(out->nz),
                       0, ctx->stream,
                       (const T*)f0->d, (const T*)f1->d, (const T*)f2->d, (T*)out->d, out->nx,
                       out->ny, out->nz, out->hx, out->hy, out->hz);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
int launch_edge(tk_ctx ctx, tk_vol u0, double beta, double lamda, double nu, tk_vol g)
{
    hipLaunchKernelGGL(k_edge<T>, vox_grid(u0->nx, u0->ny, u0->nz, 8192), vox_block(u0->nz), 0,
                       ctx->stream,
                       (const T*)u0->d, (T*)g->d, u0->nx, u0->ny, u0->nz, u0->hx, u0->hy, u0->hz,
                       (T)beta, (T)lamda, (T)nu);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
int launch_sb_update(tk_ctx ctx, tk_vol phi, tk_vol phiold, tk_vol* b, tk_vol g, tk_vol* q,
                     double a0, double b0, int* nblocks, double* part)
{
    constexpr int V = Vec16<T>::n;
    if (phi->nz % V == 0 && phi->nz > V) {
        const int zv = phi->nz / V;
        const dim3 grid = row_grid(phi->nx, phi->ny, zv, kMaxPart);
        hipLaunchKernelGGL(k_sb_update_v<T>, grid, row_block(zv), 0, ctx->stream, (T*)phi->d,
                           (const T*)phiold->d, (T*)b[0]->d, (T*)b[1]->d, (T*)b[2]->d,
                           (const T*)g->d, (T*)q[0]->d, (T*)q[1]->d, (T*)q[2]->d, phi->nx, phi->ny,
                           phi->nz, phi->hx, phi->hy, phi->hz, (T)a0, (T)b0, part);
        TK_LAUNCHED();
        *nblocks = (int)(grid.x * grid.y);
        return TK_OK;
    }
    const dim3 grid = vox_grid(phi->nx, phi->ny, phi->nz, 1024);
    const unsigned nb = grid.x * grid.y;
    hipLaunchKernelGGL(k_sb_update<T>, grid, vox_block(phi->nz), 0, ctx->stream, (T*)phi->d,
                       (const T*)phiold->d, (T*)b[0]->d, (T*)b[1]->d, (T*)b[2]->d, (const T*)g->d,
                       (T*)q[0]->d, (T*)q[1]->d, (T*)q[2]->d, phi->nx, phi->ny, phi->nz, phi->hx,
                       phi->hy, phi->hz, (T)a0, (T)b0, part);
    TK_LAUNCHED();
    *nblocks = (int)nb;
    return TK_OK;
}

int stencil_batch_grid(tk_dtype dt, int X, int Y, int Z, int* gy, int* nparts)
{
    const int V = dt == TK_F64 ? 2 : 4;
    const int zv = Z / V;
    int bdx, bdy;
    batch_block_shape(zv, &bdx, &bdy);
    const long gx = (zv + bdx - 1) / bdx;
    long rows = ((long)X * Y + bdy - 1) / bdy;
    rows = std::min(rows, std::max(1L, (long)kMaxPart / gx));
    *gy = (int)rows;
    *nparts = (int)(gx * rows);
    return TK_OK;
}

// ---- Split-Bregman update and div in one pass -------------------------------------------
// The update's q = d - b is read only by div (imp:1906-1932: d is resized away after it), so
// one pass can form q and its divergence without q going through memory.  div at a voxel
// needs q_x of the planes before and after, q_y of the rows before and after and q_z of the
// columns before and after, and q at a voxel needs phi around it; a workgroup owns a tile of
// TJ rows x 64 columns over a chunk of planes and marches in x.  Thread u owns one 16-byte
// row segment (row r, quad c4) of the tile's frame: rows j0-2 .. j0+TJ+1 and columns
// kb-4 .. kb+67.  At step q it forms q on plane q for the frame rows j0-1 .. j0+TJ (every
// column), with phi of its own column along x from registers and phi's y and z neighbours
// from the plane staged in LDS; q_x stays in the thread's registers, q_y and q_z go to LDS;
// then div on plane q-1 for the tile's own voxels (plane 0 waits one more step for q_x of
// plane 2, its one-sided derivative).  The ring rows, the halo columns and the planes next
// to a chunk are recomputed from the same inputs by the neighbouring workgroups, which
// agree bit for bit; since those read b where its owner writes the new b, the pass reads
// b_in and writes b_out (the caller swaps them).  Every value is the one the separate
// update and div passes compute, with the same expressions in the same order; only the
// tolerance sums are added in another order.
constexpr int kSdTK = 64, kSdW = kSdTK + 8;

template <class T, int TJ>
constexpr int sd_threads()
{
    return ((TJ + 4) * (kSdW / Vec16<T>::n) + 63) / 64 * 64;
}

template <class T>
struct SdArgs {
    T* phi;
    const T* phiold;
    const T* g;
    const T* bin[3];
    T* bout[3];
    T* divdb;
    int X, Y, Z, CI;
    double hx, hy, hz;
    T lo, hi;
    double* part;  // per-workgroup pairs (sum (phi-phiold)^2, sum phiold^2)
};

// three waves per SIMD: the march is latency-bound, and the registers of three waves hold
// its rings
template <class T, bool P2, int TJ>
__global__ __launch_bounds__((sd_threads<T, TJ>())) __attribute__((amdgpu_waves_per_eu(3)))
void k_sb_div(SdArgs<T> a)
{
    constexpr int V = Vec16<T>::n, W = kSdW, WV = W / V;
    constexpr int RJ = TJ + 4;
    constexpr int K = 4;  // ring period of the per-plane registers
    __shared__ __align__(16) T Ph[2][RJ * W];  // phi of the plane being updated (raw)
    __shared__ __align__(16) T Qy[2][RJ * W];  // q_y, q_z of the plane before
    __shared__ __align__(16) T Qz[2][RJ * W];
    const int X = a.X, Y = a.Y, Z = a.Z;
    const uint3 bl = xcd_block3();  // tiles sharing rings run on one XCD
    const int kb = bl.x * kSdTK, kbase = kb - 4;
    // owned rows [jo0, jo1); the frame starts at j0, the last tile's frame ends at Y so that
    // its own rows keep two rows of frame on each side inside the volume
    const int jo0 = bl.y * TJ, jo1 = min(jo0 + TJ, Y);
    const int j0 = max(0, min(jo0, Y - TJ));
    const int ia = bl.z * a.CI, ib = min(ia + a.CI, X);  // owned planes (the host keeps the
                                                          // first and last chunks >= 2 planes)
    const int tid = threadIdx.x;
    const bool act = tid < RJ * WV;
    const int r = act ? tid / WV : 0, c4 = act ? tid - r * WV : 0;
    const int j = j0 - 2 + r, k0 = kbase + c4 * V;
    const bool inside = act && j >= 0 && j < Y && k0 >= 0 && k0 < Z;
    const bool qunit = inside && r >= 1 && r <= RJ - 2;
    const bool own = inside && j >= jo0 && j < jo1 && c4 >= 4 / V && c4 < (4 + kSdTK) / V;
    const long sX = (long)Y * Z;
    // every lane loads on every step (clamped rows outside the volume, never used); b and g
    // are used on the rows where q is formed and phiold on the tile's own voxels, so lanes
    // outside those read a row that another lane of the workgroup reads (no extra bytes
    // from memory, and the load count stays fixed)
    const long off = (long)min(max(j, 0), Y - 1) * Z + min(max(k0, 0), Z - V);
    const int jq = min(max(j, max(j0 - 1, 0)), min(j0 + TJ, Y - 1));
    const long offq = (long)jq * Z + min(max(k0, 0), Z - V);
    const int jw = min(max(j, jo0), jo1 - 1);
    const int kw = kb + min(max(c4 * V - 4, 0), kSdTK - V);
    const long offw = (long)jw * Z + min(kw, Z - V);
    const int li = r * W + c4 * V;
    const int cj = plane_class(j, Y);
    auto ldo = [&](const T* f, int p, long o) {
        return ldrow(f + (long)min(max(p, 0), X - 1) * sX + o);
    };
    auto ld = [&](const T* f, int p) { return ldo(f, p, off); };
    auto cl = [&](T x) { return clampv(x, a.lo, a.hi); };
    auto id = [](T x) { return x; };
    // y rows of a staged plane for row class cj (central: r+1, r-1; first: r, r+1, r+2;
    // last: r, r-1, r-2), as the derivative helpers take them
    auto yrows = [&](const T* pl) {
        DRows<T> d;
        d.c = cj;
        if (cj == 1) {
            d.r[0] = ldrow(pl + li + W);
            d.r[1] = ldrow(pl + li - W);
        } else if (cj == 0) {
            d.r[0] = ldrow(pl + li);
            d.r[1] = ldrow(pl + li + W);
            d.r[2] = ldrow(pl + li + 2 * W);
        } else {
            d.r[0] = ldrow(pl + li);
            d.r[1] = ldrow(pl + li - W);
            d.r[2] = ldrow(pl + li - 2 * W);
        }
        return d;
    };
    const int qs = ia - 1, qe = ib;  // steps qs .. qe: update of plane q, div of plane q-1
    // rings (plane p in slot (p - qs) mod K; the march is unrolled by K, so slots are
    // constants): phi planes q-1 .. q+1 in use and q+2 in flight; b_in, g, phiold of plane
    // q in use and q+1 in flight; q_x of planes q-3 .. q
    Row<T> ph[K], rb0[K], rb1[K], rb2[K], rg[K], rpo[K], qx[K];
#pragma unroll
    for (int d = -1; d <= 2; ++d) ph[(d + K) % K] = ld(a.phi, qs + d);
#pragma unroll
    for (int d = 0; d <= 1; ++d) {
        rb0[d] = ldo(a.bin[0], qs + d, offq);
        rb1[d] = ldo(a.bin[1], qs + d, offq);
        rb2[d] = ldo(a.bin[2], qs + d, offq);
        rg[d] = ldo(a.g, qs + d, offq);
        rpo[d] = ldo(a.phiold, qs + d, offw);
    }
#pragma unroll
    for (int m = 0; m < K; ++m) qx[m] = Row<T>{};
    Row<T> ty0{}, tz0{};  // plane 0's y and z terms, kept one step for its x term
    double s1 = 0, s2 = 0;
    // (inlined into every unrolled step: the rings must stay registers)
    auto step = [&](auto u_c, int q) __attribute__((always_inline)) {
        constexpr int u = decltype(u_c)::value;
        auto sl = [](int d) { return ((u + d) % K + K) % K; };  // slot of plane q + d
        if (act) strow(&Ph[q & 1][li], ph[sl(0)]);
        __syncthreads();
        // ---- update of plane q: cut, tolerance sums, grad, b += d, shrink, b -= d, d -= b
        if (q >= 0 && q <= X - 1 && qunit) {
            const T* pl = Ph[q & 1];
            // x rows: the ring's planes q+1, q-1 inside; the first and last planes read their
            // one-sided rows from memory (the ring's slots are only ever indexed by
            // constants, so it stays in registers)
            DRows<T> rx;
            rx.c = plane_class(q, X);
            rx.r[0] = ph[sl(1)];
            rx.r[1] = ph[sl(-1)];
            if (rx.c == 0) {
                rx.r[0] = ld(a.phi, 0);
                rx.r[1] = ld(a.phi, 1);
                rx.r[2] = ld(a.phi, 2);
            } else if (rx.c == 2) {
                rx.r[0] = ld(a.phi, X - 1);
                rx.r[1] = ld(a.phi, X - 2);
                rx.r[2] = ld(a.phi, X - 3);
            }
            const DRows<T> ry = yrows(pl);
            const Row<T> cc = ph[sl(0)];
            const T zl = k0 > 0 ? pl[li - 1] : (T)0, zr = k0 + V < Z ? pl[li + V] : (T)0;
            T d[3][V];
            drow_eval<T, P2>(rx, a.hx, cl, d[0]);
            drow_eval<T, P2>(ry, a.hy, cl, d[1]);
            T c[V + 2];
            c[0] = k0 > 0 ? cl(zl) : (T)0;
            c[V + 1] = k0 + V < Z ? cl(zr) : (T)0;
#pragma unroll
            for (int e = 0; e < V; ++e) c[e + 1] = cl(cc.v[e]);
#pragma unroll
            for (int e = 0; e < V; ++e) d[2][e] = dz_win<T, V, P2>(c, e, k0 + e, Z, a.hz);
            const bool mine = own && q >= ia && q < ib;
            const long v0 = (long)q * sX + (long)j * Z + k0;
            if (mine) {
                const Row<T> po = rpo[sl(0)];
                Row<T> pc;
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    pc.v[e] = c[e + 1];
                    const double dd = (double)pc.v[e] - (double)po.v[e];
                    s1 += dd * dd;
                    s2 += (double)po.v[e] * (double)po.v[e];
                }
                strow(a.phi + v0, pc);
            }
            const Row<T> gv = rg[sl(0)];
            Row<T> bb[3] = {rb0[sl(0)], rb1[sl(0)], rb2[sl(0)]};
            Row<T> qq[3];
#pragma unroll
            for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    T b = bb[ax].v[e];
                    b += d[ax][e];
                    const T m = fabs(b) - gv.v[e];
                    const T z = 0;
                    T dn = (T)((int)(b > 0) - (int)(b < 0)) * ((m < z) ? z : m);
                    b -= dn;
                    dn -= b;
                    bb[ax].v[e] = b;
                    qq[ax].v[e] = dn;
                }
                if (mine) strow(a.bout[ax] + v0, bb[ax]);
            }
            qx[sl(0)] = qq[0];
            strow(&Qy[q & 1][li], qq[1]);
            strow(&Qz[q & 1][li], qq[2]);
        }
        // ---- div of plane p = q - 1 (image3d_imp.hxx:947-1383, both quirks)
        const int p = q - 1;
        if (own && p >= ia && p < ib) {
            const T* py = Qy[p & 1];
            const T* pz = Qz[p & 1];
            const DRows<T> ry = yrows(py);
            T ty[V], tz[V];
            drow_eval<T, P2>(ry, a.hy, id, ty);
            const int ci = plane_class(p, X);
            T c[V + 2];
            c[0] = k0 > 0 ? pz[li - 1] : (T)0;
            c[V + 1] = k0 + V < Z ? pz[li + V] : (T)0;
            const Row<T> zc = ldrow(pz + li);
#pragma unroll
            for (int e = 0; e < V; ++e) c[e + 1] = zc.v[e];
#pragma unroll
            for (int e = 0; e < V; ++e) {
                const int k = k0 + e;
                if (ci == 1 && cj == 0 && k == Z - 1)
                    // edge (i inside, 0, Z-1): Z-2 used twice (image3d_imp.hxx:1192-1194)
                    tz[e] = q2h<T, P2>(3 * c[e + 1] - 4 * c[e] + c[e], a.hz);
                else
                    tz[e] = dz_win<T, V, P2>(c, e, k, Z, a.hz);
            }
            if (p == 0) {
                // x term needs q_x of plane 2, formed at the next step
#pragma unroll
                for (int e = 0; e < V; ++e) {
                    ty0.v[e] = ty[e];
                    tz0.v[e] = tz[e];
                }
            } else {
                // x rows: planes p+1, p-1 inside; p = X-1: X-1, X-2, X-3 (the ring holds q_x of
                // planes q-3 .. q; each case reads its own constant slots)
                DRows<T> rx;
                rx.c = ci;
                if (ci == 1) {
                    rx.r[0] = qx[sl(0)];
                    rx.r[1] = qx[sl(-2)];
                } else {
                    const Row<T> x1 = qx[sl(-1)], x2 = qx[sl(-2)], x3 = qx[sl(-3)];
                    rx.r[0] = x1;
                    rx.r[1] = x2;
                    rx.r[2] = x3;
                }
                T tx[V];
                drow_eval<T, P2>(rx, a.hx, id, tx);
                if (ci == 2 && cj == 1) {
                    // face i = X-1 (j, k inside): "/ 2. * (hx)" (image3d_imp.hxx:1092-1095),
                    // "/ 2." as a product with 0.5: the same number for every operand
#pragma unroll
                    for (int e = 0; e < V; ++e) {
                        const int k = k0 + e;
                        if (k >= 1 && k <= Z - 2)
                            tx[e] = (3 * rx.r[0].v[e] - 4 * rx.r[1].v[e] + rx.r[2].v[e]) *
                                    (T)0.5 * (T)a.hx;
                    }
                }
                Row<T> o;
#pragma unroll
                for (int e = 0; e < V; ++e) o.v[e] = tx[e] + ty[e] + tz[e];
                strow(a.divdb + (long)p * sX + (long)j * Z + k0, o);
            }
        }
        // plane 0's x term: q_x of planes 0, 1, 2 (the step that formed plane 2)
        if (own && q == 2 && ia == 0) {
            DRows<T> rx;
            rx.c = 0;
            rx.r[0] = qx[sl(-2)];
            rx.r[1] = qx[sl(-1)];
            rx.r[2] = qx[sl(0)];
            T tx[V];
            drow_eval<T, P2>(rx, a.hx, id, tx);
            Row<T> o;
#pragma unroll
            for (int e = 0; e < V; ++e) o.v[e] = tx[e] + ty0.v[e] + tz0.v[e];
            strow(a.divdb + (long)j * Z + k0, o);
        }
        // refill the slots this step was the last to use (phi q-1, the fields of q)
        ph[sl(3)] = ld(a.phi, q + 3);
        rb0[sl(2)] = ldo(a.bin[0], q + 2, offq);
        rb1[sl(2)] = ldo(a.bin[1], q + 2, offq);
        rb2[sl(2)] = ldo(a.bin[2], q + 2, offq);
        rg[sl(2)] = ldo(a.g, q + 2, offq);
        rpo[sl(2)] = ldo(a.phiold, q + 2, offw);
    };
    int q = qs;
    for (; q + K <= qe + 1; q += K)
        static_unroll<K>([&](auto u_c) { step(u_c, q + decltype(u_c)::value); });
    static_unroll<K - 1>([&](auto u_c) {
        if (q + decltype(u_c)::value <= qe) step(u_c, q + decltype(u_c)::value);
    });
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        const int b = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        a.part[2 * b] = s1;
        a.part[2 * b + 1] = s2;
    }
}

// the update and div of one volume in one pass (b_in -> b_out); *nblocks: the partial sums
// written to part.  Needs 16-byte rows (nz % V == 0), every edge >= 3.
template <class T>
int launch_sb_div(tk_ctx ctx, tk_vol phi, tk_vol phiold, tk_vol g, tk_vol* b_in, tk_vol* b_out,
                  tk_vol divdb, double a0, double b0, int* nblocks, double* part)
{
    // 8-row tiles: 16 rows measured 189 us, 12 rows 283 us and 4 rows 226 us against 178 us
    // at 256^3 FP32
    constexpr int TJ = 8;
    constexpr int NT = sd_threads<T, TJ>();
    const int X = phi->nx, Y = phi->ny, Z = phi->nz;
    static int resident = 0;
    if (resident <= 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sb_div<T, true, TJ>, NT, 0) !=
                hipSuccess ||
            nb < 1)
            nb = 1;
        resident = nb;
    }
    const int kt = (Z + kSdTK - 1) / kSdTK, jt = (Y + TJ - 1) / TJ;
    // one round of resident workgroups; chunks of at least 4 planes (each recomputes one
    // plane on each side), and no chunk of one plane at the end
    long chunks = ((long)ctx->num_cu * resident + (long)kt * jt / 2) / ((long)kt * jt);
    chunks = std::max(1L, std::min(chunks, (long)X / 4));
    chunks = std::min(chunks, (long)kMaxPart / ((long)kt * jt));
    int CI = (int)((X + chunks - 1) / chunks);
    if (X % CI == 1) ++CI;
    const int nz = (X + CI - 1) / CI;
    TK_CHECK((long)kt * jt * nz <= kMaxPart, TK_ERR_UNSUPPORTED,
             "update and div: %ld workgroups exceed the partial-sum slots", (long)kt * jt * nz);
    SdArgs<T> args;
    args.phi = (T*)phi->d;
    args.phiold = (const T*)phiold->d;
    args.g = (const T*)g->d;
    for (int c = 0; c < 3; ++c) {
        args.bin[c] = (const T*)b_in[c]->d;
        args.bout[c] = (T*)b_out[c]->d;
    }
    args.divdb = (T*)divdb->d;
    args.X = X;
    args.Y = Y;
    args.Z = Z;
    args.CI = CI;
    args.hx = phi->hx;
    args.hy = phi->hy;
    args.hz = phi->hz;
    args.lo = (T)a0;
    args.hi = (T)b0;
    args.part = part;
    const dim3 grid(kt, jt, nz);
    if (pow2_2h3(phi->hx, phi->hy, phi->hz))
        hipLaunchKernelGGL((k_sb_div<T, true, TJ>), grid, dim3(NT), 0, ctx->stream, args);
    else
        hipLaunchKernelGGL((k_sb_div<T, false, TJ>), grid, dim3(NT), 0, ctx->stream, args);
    TK_LAUNCHED();
    *nblocks = (int)(grid.x * grid.y * grid.z);
    return TK_OK;
}
template int launch_sb_div<float>(tk_ctx, tk_vol, tk_vol, tk_vol, tk_vol*, tk_vol*, tk_vol, double,
                                  double, int*, double*);
template int launch_sb_div<double>(tk_ctx, tk_vol, tk_vol, tk_vol, tk_vol*, tk_vol*, tk_vol,
                                   double, double, int*, double*);

int launch_sb_batch(tk_ctx ctx, tk_dtype dt, const SbJob* jobs, int n, unsigned nblocks)
{
    TimedLaunch tl(ctx, TK_K_SB);
    if (dt == TK_F64)
        hipLaunchKernelGGL(k_sb_update_batch<double>, dim3(nblocks), dim3(256), 0, ctx->stream,
                           jobs, n);
    else
        hipLaunchKernelGGL(k_sb_update_batch<float>, dim3(nblocks), dim3(256), 0, ctx->stream,
                           jobs, n);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_sum_pairs_batch, dim3((unsigned)n), dim3(1024), 0, ctx->stream, jobs);
    TK_LAUNCHED();
    return TK_OK;
}

int launch_div_batch(tk_ctx ctx, tk_dtype dt, const DivJob* jobs, int n, unsigned nblocks)
{
    TimedLaunch tl(ctx, TK_K_DIV);
    if (dt == TK_F64)
        hipLaunchKernelGGL(k_div_batch<double>, dim3(nblocks), dim3(256), 0, ctx->stream, jobs, n);
    else
        hipLaunchKernelGGL(k_div_batch<float>, dim3(nblocks), dim3(256), 0, ctx->stream, jobs, n);
    TK_LAUNCHED();
    return TK_OK;
}

template int launch_div<float>(tk_ctx, tk_vol, tk_vol, tk_vol, tk_vol);
template int launch_div<double>(tk_ctx, tk_vol, tk_vol, tk_vol, tk_vol);
template int launch_edge<float>(tk_ctx, tk_vol, double, double, double, tk_vol);
template int launch_edge<double>(tk_ctx, tk_vol, double, double, double, tk_vol);
template int launch_sb_update<float>(tk_ctx, tk_vol, tk_vol, tk_vol*, tk_vol, tk_vol*, double,
                                     double, int*, double*);
template int launch_sb_update<double>(tk_ctx, tk_vol, tk_vol, tk_vol*, tk_vol, tk_vol*, double,
                                      double, int*, double*);

}  // namespace tk

using namespace tk;

static int check3(tk_vol a, tk_vol b)
{
    TK_CHECK(a && b && a->d && b->d, TK_ERR_ARG, "stencil: null volume");
    TK_CHECK(same_dims(a, b), TK_ERR_SHAPE, "stencil: dims differ");
    TK_CHECK(a->dtype == b->dtype, TK_ERR_DTYPE, "stencil: dtypes differ");
    TK_CHECK(a->dtype != TK_U8, TK_ERR_DTYPE, "stencil: U8 volume");
    TK_CHECK(a->nx >= 3 && a->ny >= 3 && a->nz >= 3, TK_ERR_UNSUPPORTED,
             "stencil: 3-D volumes with every edge >= 3 only");
    return TK_OK;
}

extern "C" {

int tk_grad(tk_ctx ctx, tk_vol f, tk_vol g0, tk_vol g1, tk_vol g2)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_grad: null ctx");
    TK_TRY(check3(f, g0));
    TK_TRY(check3(f, g1));
    TK_TRY(check3(f, g2));
    return f->dtype == TK_F64 ? launch_grad<double>(ctx, f, g0, g1, g2)
                              : launch_grad<float>(ctx, f, g0, g1, g2);
}

int tk_div(tk_ctx ctx, tk_vol f0, tk_vol f1, tk_vol f2, tk_vol out)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_div: null ctx");
    TK_TRY(check3(out, f0));
    TK_TRY(check3(out, f1));
    TK_TRY(check3(out, f2));
    return out->dtype == TK_F64 ? launch_div<double>(ctx, f0, f1, f2, out)
                                : launch_div<float>(ctx, f0, f1, f2, out);
}

int tk_edge_detect(tk_ctx ctx, tk_vol u0, double beta, double lamda, double nu, tk_vol g)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_edge_detect: null ctx");
    TK_TRY(check3(u0, g));
    return u0->dtype == TK_F64 ? launch_edge<double>(ctx, u0, beta, lamda, nu, g)
                               : launch_edge<float>(ctx, u0, beta, lamda, nu, g);
}

}  // extern "C"
