// tk_stencil.hip -- image3d::grad, im3d::div, the GRAD edge detector and the fused
// Split-Bregman update (rsfe_splitbregman_imp.hxx:1868-1936).
#include <hip/hip_runtime.h>

#include "tk_device.hpp"
#include "tk_internal.hpp"
#include "tk_stencil.hpp"

namespace tk {

template <class T>
__global__ void k_grad(const T* __restrict__ f, T* __restrict__ g0, T* __restrict__ g1,
                       T* __restrict__ g2, int X, int Y, int Z, double hx, double hy, double hz)
{
    const long n = (long)X * Y * Z, sX = (long)Y * Z, sY = Z;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % Z), j = (int)((v / Z) % Y), i = (int)(v / sX);
        g0[v] = dterm(f, v, sX, plane_class(i, X), X, i, hx);
        g1[v] = dterm(f, v, sY, plane_class(j, Y), Y, j, hy);
        g2[v] = dterm(f, v, 1, plane_class(k, Z), Z, k, hz);
    }
}

template <class T>
__global__ void k_div(const T* __restrict__ f0, const T* __restrict__ f1, const T* __restrict__ f2,
                      T* __restrict__ out, int X, int Y, int Z, double hx, double hy, double hz)
{
    const long n = (long)X * Y * Z, sX = (long)Y * Z;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % Z), j = (int)((v / Z) % Y), i = (int)(v / sX);
        out[v] = div_at(f0, f1, f2, i, j, k, X, Y, Z, hx, hy, hz);
    }
}

// rsfe_splitbregman_imp.hxx:233-242 and 308-313 (GRAD, edge_detector_sigma == 0)
template <class T>
__global__ void k_edge(const T* __restrict__ f, T* __restrict__ g, int X, int Y, int Z, double hx,
                       double hy, double hz, T beta, T lamda, T nu)
{
    const long n = (long)X * Y * Z, sX = (long)Y * Z, sY = Z;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % Z), j = (int)((v / Z) % Y), i = (int)(v / sX);
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
    const long n = (long)X * Y * Z, sX = (long)Y * Z, sY = Z;
    double s1 = 0, s2 = 0;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % Z), j = (int)((v / Z) % Y), i = (int)(v / sX);
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
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

template <class T>
int launch_grad(tk_ctx ctx, tk_vol f, tk_vol g0, tk_vol g1, tk_vol g2)
{
    const long n = f->n();
    hipLaunchKernelGGL(k_grad<T>, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, ctx->stream,
                       (const T*)f->d, (T*)g0->d, (T*)g1->d, (T*)g2->d, f->nx, f->ny, f->nz, f->hx,
                       f->hy, f->hz);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
int launch_div(tk_ctx ctx, tk_vol f0, tk_vol f1, tk_vol f2, tk_vol out)
{
    const long n = out->n();
    hipLaunchKernelGGL(k_div<T>, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, ctx->stream,
                       (const T*)f0->d, (const T*)f1->d, (const T*)f2->d, (T*)out->d, out->nx,
                       out->ny, out->nz, out->hx, out->hy, out->hz);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
int launch_edge(tk_ctx ctx, tk_vol u0, double beta, double lamda, double nu, tk_vol g)
{
    const long n = u0->n();
    hipLaunchKernelGGL(k_edge<T>, dim3(blocks_for(n, 256, 8192)), dim3(256), 0, ctx->stream,
                       (const T*)u0->d, (T*)g->d, u0->nx, u0->ny, u0->nz, u0->hx, u0->hy, u0->hz,
                       (T)beta, (T)lamda, (T)nu);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
int launch_sb_update(tk_ctx ctx, tk_vol phi, tk_vol phiold, tk_vol* b, tk_vol g, tk_vol* q,
                     double a0, double b0, int* nblocks, double* part)
{
    const long n = phi->n();
    const unsigned nb = blocks_for(n, 256, 1024);
    hipLaunchKernelGGL(k_sb_update<T>, dim3(nb), dim3(256), 0, ctx->stream, (T*)phi->d,
                       (const T*)phiold->d, (T*)b[0]->d, (T*)b[1]->d, (T*)b[2]->d, (const T*)g->d,
                       (T*)q[0]->d, (T*)q[1]->d, (T*)q[2]->d, phi->nx, phi->ny, phi->nz, phi->hx,
                       phi->hy, phi->hz, (T)a0, (T)b0, part);
    TK_LAUNCHED();
    *nblocks = (int)nb;
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
