// tk_gs.hip -- Dirichlet Gauss-Seidel sweeps for (1/dt - lap_h) phi = rhs + phi/dt
// (lapl::DirGaussSeidel, poisson_imp.hxx:340-429) and the sweep loop of
// rsfe_splitbregman::apply (rsfe_splitbregman_imp.hxx:1762-1797).
//
// TK_GS_LEX -- bit-exact lexicographic order on a GPU.  In the reference's i-j-k loop
// every update reads new values at the minus neighbours and old values at the plus
// neighbours.  The interior is cut into 8x8xTZ tiles; tile (I,J,K) performs sweep s
// at pipeline step tau = I+J+K+2s.  At that step its minus-neighbour tiles finished
// sweep s (step tau-1), its plus-neighbour tiles finished sweep s-1 (step tau-1) and
// have not started sweep s (step tau+1), and no two face-adjacent tiles are active
// together, so each tile sees exactly the values the serial loop would.  Inside a
// tile, one 64-lane wave walks the skewed schedule k = t - a - b (lane (a,b) owns
// column (a,b)), which orders the updates the same way.  C sweeps take
// (nTx+nTy+nTz-3) + 2C - 1 steps: about two launches per sweep, each over half the
// tiles.  Expression order and -ffp-contract=off match the reference bit for bit.
//
// The relative-update stop rule needs ||phi_s - phi_{s-1}|| for every sweep: each
// tile-step belongs to one sweep, so its wave reduces (new-old)^2 and old^2 in FP64
// and adds them to that sweep's slot.  Sweeps run in chunks; when the rule fires
// inside a chunk the chunk is replayed from a saved copy up to the stopping sweep
// (deterministic, hence identical).
//
// Neumann (lapl::NeuGaussSeidel, poisson_imp.hxx:115-334, 3-D branch): the same
// pipelined wavefront over every voxel, the boundary ones included.  The reference's loop
// runs k outer and i fastest, the Dirichlet one i outer and k fastest; for a 7-point stencil
// every lexicographic order reads new values at all three minus neighbours and old values
// at the plus neighbours, so both give the same result and share the schedule.
//
// TK_GS_REDBLACK -- the FP32 fast mode: red (i+j+k even) then black, each pass
// in place; FP64 per-sweep update norms; the black pass's last workgroup evaluates
// the stop rule on the device and raises a flag that makes later sweeps return at once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {

struct GsCoef {
    double hx2, hy2, hz2, dt, htilde;
};

inline GsCoef gs_coef(const tk_vol phi, double dt)
{
    GsCoef c;
    c.hx2 = phi->hx * phi->hx;
    c.hy2 = phi->hy * phi->hy;
    c.hz2 = phi->hz * phi->hz;
    c.dt = dt;
    // poisson_imp.hxx:414
    c.htilde = 1. / (1. / dt + 2. / (phi->hx * phi->hx) + 2. / (phi->hy * phi->hy) +
                     2. / (phi->hz * phi->hz));
    return c;
}

// Neumann boundary coefficients: htilde per boundary mask (bit 0: i on a face, 1: j, 2: k;
// 1/h^2 instead of 2/h^2 on a boundary axis, poisson_imp.hxx:175-197) and the derivative
// term (n bc) h per number of boundary planes (n = 1, 1/sqrt 2, 1/sqrt 3, 201-205) and axis
struct NeuCoef {
    double ht[8];
    double q[4][3];
};

inline NeuCoef neu_coef(const tk_vol phi, double dt, double bc)
{
    NeuCoef nc;
    const double hx = phi->hx, hy = phi->hy, hz = phi->hz;
    for (int m = 0; m < 8; ++m) {
        const double cx = (m & 1) ? 1. : 2., cy = (m & 2) ? 1. : 2., cz = (m & 4) ? 1. : 2.;
        nc.ht[m] = 1. / (1. / dt + cx / (hx * hx) + cy / (hy * hy) + cz / (hz * hz));
    }
    const double n_of[4] = {0., 1., 1. / std::sqrt(2.), 1. / std::sqrt(3.)};
    for (int c = 0; c < 4; ++c) {
        nc.q[c][0] = n_of[c] * bc * hx;
        nc.q[c][1] = n_of[c] * bc * hy;
        nc.q[c][2] = n_of[c] * bc * hz;
    }
    return nc;
}

// poisson_imp.hxx:420-424, same operand order; UNIT skips the exact divisions by 1
template <class T, bool UNIT>
__device__ __forceinline__ T gs_update(T xp, T xm, T yp, T ym, T zp, T zm, T b, T self,
                                       const GsCoef& c)
{
    if constexpr (std::is_same<T, double>::value) {
        if constexpr (UNIT) return ((xp + xm) + (yp + ym) + (zp + zm) + b + self) * c.htilde;
        else
            return ((xp + xm) / c.hx2 + (yp + ym) / c.hy2 + (zp + zm) / c.hz2 + b + self / c.dt) *
                   c.htilde;
    } else {
        const float ht = (float)c.htilde;
        if constexpr (UNIT) return ((xp + xm) + (yp + ym) + (zp + zm) + b + self) * ht;
        else
            return ((xp + xm) / (float)c.hx2 + (yp + ym) / (float)c.hy2 + (zp + zm) / (float)c.hz2 +
                    b + self / (float)c.dt) *
                   ht;
    }
}

// ---- boundary: all six faces := bc, with the sweep-0 update norm of the faces -----
template <class T>
__global__ void k_set_boundary(T* __restrict__ phi, int X, int Y, int Z, T bc, double* __restrict__ acc)
{
    // enumerate the voxels of the boundary shell once each
    const long nxy = (long)X * Y;
    double d2 = 0, o2 = 0;
    const long total = 2 * nxy + 2L * X * (Z - 2 > 0 ? Z - 2 : 0) +
                       2L * (Y - 2 > 0 ? Y - 2 : 0) * (Z - 2 > 0 ? Z - 2 : 0);
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total;
         q += (long)gridDim.x * blockDim.x) {
        int i, j, k;
        long r = q;
        if (r < 2 * nxy) {  // k = 0 and k = Z-1 planes
            k = (r < nxy) ? 0 : Z - 1;
            r %= nxy;
            i = (int)(r / Y);
            j = (int)(r % Y);
        } else {
            r -= 2 * nxy;
            const int zi = Z - 2;
            const long nxz = (long)X * zi;
            if (r < 2 * nxz) {  // j = 0 and j = Y-1, k interior
                j = (r < nxz) ? 0 : Y - 1;
                r %= nxz;
                i = (int)(r / zi);
                k = 1 + (int)(r % zi);
            } else {            // i = 0 and i = X-1, j and k interior
                r -= 2 * nxz;
                const long nyz = (long)(Y - 2) * zi;
                i = (r < nyz) ? 0 : X - 1;
                r %= nyz;
                j = 1 + (int)(r / zi);
                k = 1 + (int)(r % zi);
            }
        }
        const long v = ((long)i * Y + j) * Z + k;
        const double old = (double)phi[v];
        const double d = (double)bc - old;
        d2 += d * d;
        o2 += old * old;
        phi[v] = bc;
    }
    if (acc) {
        block_sum2(d2, o2);
        if (threadIdx.x == 0) {
            atomicAdd(&acc[0], d2);
            atomicAdd(&acc[1], o2);
        }
    }
}

inline long boundary_count(int X, int Y, int Z)
{
    const long zi = Z - 2 > 0 ? Z - 2 : 0, yi = Y - 2 > 0 ? Y - 2 : 0;
    return 2L * X * Y + 2L * X * zi + 2L * yi * zi;
}

// ---- lexicographic pipelined tile wavefront --------------------------------------
constexpr int kLexT = 8;    // tile edge in x and y (one lane per column)
constexpr int kLexTZ = 32;  // tile depth in z
constexpr int kLexBatch = 16;  // tile loads per lane in flight at once
constexpr int kLexNT = 256;    // threads per tile: staging and stores use all four waves, the
                               // in-tile wavefront the first (one lane per (i, j) column)
// the staging loop records which of a lane's kLexBatch loads fell outside the volume as one
// bit each of a 32-bit mask
static_assert(kLexBatch <= 32, "kLexBatch loads per lane must fit the 32-bit out-of-volume mask");

// one Neumann update of a voxel on the boundary (mask m != 0): per axis (p+ - q) on the
// low face, (p- + q) on the high face, (p+ + p-) inside; then the interior expression with
// the boundary htilde (poisson_imp.hxx:200-331)
template <class T, bool UNIT>
__device__ __forceinline__ T neu_update(T xp, T xm, T yp, T ym, T zp, T zm, T b, T self,
                                        const GsCoef& c, const double* ht, const double* q,
                                        int m, bool lx, bool ly, bool lz)
{
    const T tx = (m & 1) ? (lx ? xp - (T)q[0] : xm + (T)q[0]) : xp + xm;
    const T ty = (m & 2) ? (ly ? yp - (T)q[1] : ym + (T)q[1]) : yp + ym;
    const T tz = (m & 4) ? (lz ? zp - (T)q[2] : zm + (T)q[2]) : zp + zm;
    if constexpr (UNIT) return (tx + ty + tz + b + self) * (T)ht[m];
    else
        return (tx / (T)c.hx2 + ty / (T)c.hy2 + tz / (T)c.hz2 + b + self / (T)c.dt) * (T)ht[m];
}

// Tiles (I, J, K) of an a x b x c tile grid on the diagonal plane I + J + K = d: how many,
// and the n-th of them in the order (I, J) increasing.  Step tau of a chunk of C sweeps
// runs sweep s on the plane d = tau - 2 s, so a launch holds only those planes' tiles.
__host__ __device__ inline int lex_plane_count(int d, int a, int b, int c)
{
    int n = 0;
    const int i1 = d < a - 1 ? d : a - 1;
    for (int I = d - (b - 1) - (c - 1) > 0 ? d - (b - 1) - (c - 1) : 0; I <= i1; ++I) {
        const int m = d - I, lo = m - (c - 1) > 0 ? m - (c - 1) : 0, hi = m < b - 1 ? m : b - 1;
        if (hi >= lo) n += hi - lo + 1;
    }
    return n;
}
__device__ inline void lex_plane_tile(int d, int n, int a, int b, int c, int& I, int& J, int& K)
{
    for (I = max(0, d - (b - 1) - (c - 1)); I < a; ++I) {
        const int m = d - I, lo = max(0, m - (c - 1)), hi = min(b - 1, m);
        const int cnt = hi >= lo ? hi - lo + 1 : 0;
        if (n < cnt) {
            J = lo + n;
            K = m - J;
            return;
        }
        n -= cnt;
    }
}
// the sweeps of a chunk whose planes are on the grid at step tau
__host__ __device__ inline void lex_step_sweeps(int tau, int C, int dmax, int& s_lo, int& s_hi)
{
    s_lo = tau - dmax > 0 ? (tau - dmax + 1) / 2 : 0;
    s_hi = tau / 2 < C - 1 ? tau / 2 : C - 1;
}

// NEU: tiles cover every voxel (Neumann sweep), else the interior (Dirichlet sweep).  One
// workgroup per tile on the active planes of step tau (blockIdx.x enumerates them, sweep by
// sweep).
template <class T, bool UNIT, bool NEU>
__global__ __launch_bounds__(kLexNT) void k_gs_lex(T* __restrict__ phi, const T* __restrict__ rhs,
                                               int X, int Y, int Z, int nTx, int nTy, int nTz,
                                               int tau, int C, GsCoef c, NeuCoef nc,
                                               double* __restrict__ acc)
{
    constexpr int TX = kLexT, TY = kLexT, TZ = kLexTZ;
    constexpr int HX = TX + 2, HY = TY + 2, HZ = TZ + 2;
    constexpr int O = NEU ? 0 : 1;  // first updated index along each axis
    __shared__ T ph[HX * HY * HZ];
    __shared__ T rb[TX * TY * TZ];
    __shared__ double nct[8 + 12];
    int s_lo, s_hi, s = -1, I = 0, J = 0, K = 0;
    lex_step_sweeps(tau, C, nTx + nTy + nTz - 3, s_lo, s_hi);
    {
        int n = (int)blockIdx.x;
        for (int q = s_lo; q <= s_hi; ++q) {
            const int d = tau - 2 * q, cnt = lex_plane_count(d, nTx, nTy, nTz);
            if (n < cnt) {
                s = q;
                lex_plane_tile(d, n, nTx, nTy, nTz, I, J, K);
                break;
            }
            n -= cnt;
        }
    }
    if (s < 0) return;
    const int i0 = O + I * TX, j0 = O + J * TY, k0 = O + K * TZ;
    const int tx = min(TX, X - O - i0), ty = min(TY, Y - O - j0), tz = min(TZ, Z - O - k0);
    const long sY = Z, sX = (long)Y * Z;
    const int lane = threadIdx.x;
    if constexpr (NEU) {
        if (lane < 8) nct[lane] = nc.ht[lane];
        else if (lane < 20) nct[lane] = nc.q[(lane - 8) / 3][(lane - 8) % 3];
    }
    // load phi tile + halo and the rhs tile (rows along z are contiguous); the Neumann
    // tiles' halo positions outside the volume are never read and hold 0
    // kLexBatch loads per lane are issued before any is written to LDS (a load written in
    // the iteration that issues it is waited for before the next one goes out), and every
    // lane issues all of them: an index past the end is clamped to the last element and its
    // write skipped, so the compiler need not wait between conditional loads
    // element -> (ii, jj, kk) by exact float-reciprocal quotients (nh <= 3400, divisors <= 34):
    // four 32-bit integer divisions per element made the staging the tile's longest phase
    const int hz = tz + 2;
    const int nh = (tx + 2) * (ty + 2) * hz, nr = tx * ty * tz;
    const float inv_hz = 1.0f / (float)hz, inv_hy = 1.0f / (float)(ty + 2);
    const float inv_tz = 1.0f / (float)tz, inv_ty = 1.0f / (float)ty;
    for (int e0 = lane; e0 < nh; e0 += kLexNT * kLexBatch) {
        T v[kLexBatch];
        int at[kLexBatch];
        unsigned zero = 0;  // Neumann halo positions outside the volume hold 0
#pragma unroll
        for (int u = 0; u < kLexBatch; ++u) {
            const int e = min(e0 + kLexNT * u, nh - 1);
            const int r = qdiv(e, inv_hz), kk = e - r * hz;
            const int ii = qdiv(r, inv_hy), jj = r - ii * (ty + 2);
            const int gi = i0 - 1 + ii, gj = j0 - 1 + jj, gk = k0 - 1 + kk;
            const bool out =
                NEU && (gi < 0 || gi >= X || gj < 0 || gj >= Y || gk < 0 || gk >= Z);
            v[u] = phi[out ? 0 : gi * sX + gj * sY + gk];
            zero |= (unsigned)out << u;
            at[u] = e0 + kLexNT * u < nh ? (ii * HY + jj) * HZ + kk : -1;
        }
#pragma unroll
        for (int u = 0; u < kLexBatch; ++u)
            if (at[u] >= 0) ph[at[u]] = (zero >> u) & 1u ? (T)0 : v[u];
    }
    for (int e0 = lane; e0 < nr; e0 += kLexNT * kLexBatch) {
        T v[kLexBatch];
        int at[kLexBatch];
#pragma unroll
        for (int u = 0; u < kLexBatch; ++u) {
            const int e = min(e0 + kLexNT * u, nr - 1);
            const int r = qdiv(e, inv_tz), kk = e - r * tz;
            const int ii = qdiv(r, inv_ty), jj = r - ii * ty;
            v[u] = rhs[(i0 + ii) * sX + (j0 + jj) * sY + (k0 + kk)];
            at[u] = e0 + kLexNT * u < nr ? (ii * TY + jj) * TZ + kk : -1;
        }
#pragma unroll
        for (int u = 0; u < kLexBatch; ++u)
            if (at[u] >= 0) rb[at[u]] = v[u];
    }
    __syncthreads();
    // lanes of the first wave own the (a, b) columns; a >= TX for the other waves, so they
    // only meet the barriers
    const int a = lane / TY, b = lane % TY;
    double d2 = 0, o2 = 0;
    const int steps = tz + (tx - 1) + (ty - 1);
    for (int t = 0; t < steps; ++t) {
        const int k = t - a - b;
        if (a < tx && b < ty && k >= 0 && k < tz) {
            T* p = &ph[((a + 1) * HY + (b + 1)) * HZ + (k + 1)];
            const T old = p[0];
            T v;
            if constexpr (NEU) {
                const int gi = i0 + a, gj = j0 + b, gk = k0 + k;
                const bool lx = gi == 0, ly = gj == 0, lz = gk == 0;
                const int bx = lx || gi == X - 1, by = ly || gj == Y - 1, bz = lz || gk == Z - 1;
                const int m = bx | (by << 1) | (bz << 2);
                if (m == 0)
                    v = gs_update<T, UNIT>(p[HY * HZ], p[-HY * HZ], p[HZ], p[-HZ], p[1], p[-1],
                                           rb[(a * TY + b) * TZ + k], old, c);
                else
                    v = neu_update<T, UNIT>(p[HY * HZ], p[-HY * HZ], p[HZ], p[-HZ], p[1], p[-1],
                                            rb[(a * TY + b) * TZ + k], old, c, nct,
                                            nct + 8 + 3 * (bx + by + bz), m, lx, ly, lz);
            } else {
                v = gs_update<T, UNIT>(p[HY * HZ], p[-HY * HZ], p[HZ], p[-HZ], p[1], p[-1],
                                       rb[(a * TY + b) * TZ + k], old, c);
            }
            p[0] = v;
            const double d = (double)v - (double)old;
            d2 += d * d;
            o2 += (double)old * (double)old;
        }
        __syncthreads();
    }
    for (int e = lane; e < tx * ty * tz; e += kLexNT) {
        const int r = qdiv(e, inv_tz), kk = e - r * tz;
        const int ii = qdiv(r, inv_ty), jj = r - ii * ty;
        phi[(i0 + ii) * sX + (j0 + jj) * sY + (k0 + kk)] =
            ph[((ii + 1) * HY + (jj + 1)) * HZ + (kk + 1)];
    }
    if (acc && lane < 64) {
        d2 = wave_sum(d2);
        o2 = wave_sum(o2);
        if (lane == 0) {
            atomicAdd(&acc[2 * s], d2);
            atomicAdd(&acc[2 * s + 1], o2);
        }
    }
}

// C lexicographic sweeps; neu: the Neumann sweep over every voxel with its coefficients
template <class T>
static int lex_chunk(tk_ctx ctx, tk_vol phi, tk_vol rhs, const GsCoef& c, int C, double* acc,
                     const NeuCoef* neu = nullptr)
{
    const int X = phi->nx, Y = phi->ny, Z = phi->nz;
    const int e = neu ? 0 : 2;  // voxels left out along each axis
    const int nTx = (X - e + kLexT - 1) / kLexT, nTy = (Y - e + kLexT - 1) / kLexT,
              nTz = (Z - e + kLexTZ - 1) / kLexTZ;
    const int last = (nTx - 1) + (nTy - 1) + (nTz - 1) + 2 * (C - 1);
    const bool unit = c.hx2 == 1. && c.hy2 == 1. && c.hz2 == 1. && c.dt == 1.;
    const NeuCoef nc = neu ? *neu : NeuCoef{};
#define TK_LEX(U, N)                                                                           \
    hipLaunchKernelGGL((k_gs_lex<T, U, N>), dim3(nb), dim3(kLexNT), 0, ctx->stream, (T*)phi->d, \
                       (const T*)rhs->d, X, Y, Z, nTx, nTy, nTz, tau, C, c, nc, acc)
    const int dmax = nTx + nTy + nTz - 3;
    for (int tau = 0; tau <= last; ++tau) {
        int s_lo, s_hi, nb = 0;
        lex_step_sweeps(tau, C, dmax, s_lo, s_hi);
        for (int q = s_lo; q <= s_hi; ++q) nb += lex_plane_count(tau - 2 * q, nTx, nTy, nTz);
        if (nb == 0) continue;
        TimedLaunch tl(ctx, TK_K_GS);
        if (neu) {
            if (unit) TK_LEX(true, true);
            else TK_LEX(false, true);
        } else {
            if (unit) TK_LEX(true, false);
            else TK_LEX(false, false);
        }
    }
#undef TK_LEX
    TK_LAUNCHED();
    return TK_OK;
}

// ---- red-black -----------------------------------------------------------------------
// Device stop protocol (slots in ctx->cnt): cnt[0] = 1 + stopping sweep (0: running);
// cnt[1 + slot] = arrivals of the black pass of the sweep using acc slot `slot`.
// one volume of a batched sweep launch (k_gs_rb_pv_batch)
struct RbJob {
    const void* src;
    void* dst;
    const void* rhs;
    int X, Y, Z, CI;
    int gx, gy, gz;  // tiles in k, j, i
    int first, nblk; // first block of the volume (a multiple of kNumXcd) and its block count
};

struct RbGate {
    double* acc;        // pair per sweep slot: sum (new-old)^2, sum old^2 (interior); the
                        // column kernels use acc[0..1] only: the faces' terms of sweep 0
    double* err;        // per slot: relative update error
    unsigned* stop;     // cnt[0]
    unsigned* arrive;   // cnt + 1
    int ls_steps;
    double ls_tol;
    double h3;          // hx*hy*hz
    double bnd_o2;      // boundary contribution to sum old^2 for sweeps >= 1
    // column kernels: per-block sum pairs of sweep s at part + (s & 3) * 2 * kRbGateBlk,
    // and the previous launch of the batch (its first sweep, sweep count and block count;
    // prev_n = 0: none), whose stop test the launch runs before its own sweeps
    double* part;
    int base;           // sweep index of error slot 0
    int prev_first, prev_n;
    unsigned prev_nblk;
};

// The update norm sums of sweep sp: the per-block pairs its launch stored (nblk blocks),
// added in one fixed order (lane-strided sums, then the DPP wave sum), so every workgroup
// that tests a sweep gets the same values.  Called by a whole wave; uniform result.
__device__ double2 rb_gate_sums(const RbGate& g, int sp, unsigned nblk)
{
    const double2* pp = (const double2*)(g.part + (size_t)(sp & 3) * (2 * kRbGateBlk));
    const unsigned l = threadIdx.x & 63;
    double sd = 0, so = 0;
    // the first 256 pairs in straight-line code: four loads per lane issued together, with
    // clamped indices (a loop, or loads under a branch, would make the wave wait first for
    // every load in flight, the march's prologue included)
    {
        double2 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = pp[min(64u * u + l, nblk - 1)];
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (64u * u + l < nblk) {
                sd += v[u].x;
                so += v[u].y;
            }
    }
    for (unsigned b = 256 + l; b < nblk; b += 64) {
        sd += pp[b].x;
        so += pp[b].y;
    }
    return double2{wave_sum_dpp(sd), wave_sum_dpp(so)};
}

// The stop test of sweep sp from its sums (plus the faces' terms: those of sweep 0 in
// acc[0..1], the constant face value's for later sweeps); the writer's lane 0 records the
// error and, on a stop, the sweep count.
__device__ bool rb_gate_decide(const RbGate& g, int sp, double2 s, bool writer)
{
    double sd = s.x, so = s.y;
    if (sp == 0) {
        sd += g.acc[0];
        so += g.acc[1];
    } else {
        so += g.bnd_o2;
    }
    const double err = sqrt(g.h3 * sd) / sqrt(g.h3 * so);
    const bool stop = sp >= g.ls_steps - 1 && !(err > g.ls_tol);
    if (writer && (threadIdx.x & 63) == 0) {
        g.err[sp - g.base] = err;
        if (stop) *(volatile unsigned*)g.stop = (unsigned)(sp + 1);
    }
    return stop;
}

// Whether a gated column launch runs: the stop flag is clear and no sweep of the previous
// launch met the stop condition (tested in sweep order).  Wave 0 decides, the workgroup
// learns it through LDS; block 0 is the writer.  The flag is read with a vector load, so it
// is in flight together with the sums' loads (and the march's prologue loads issued
// before): the flag that matters was set by an earlier launch; one set by this launch's
// writer only repeats the decision that every workgroup reaches by itself.
__device__ bool rb_gate_live(const RbGate& g, bool writer)
{
    __shared__ int live;
    if (threadIdx.x < 64) {
        const unsigned stopped = __hip_atomic_load(g.stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool go = true;
        if (g.prev_n > 0) {
            const double2 s0 = rb_gate_sums(g, g.prev_first, g.prev_nblk);
            go = !rb_gate_decide(g, g.prev_first, s0, writer && stopped == 0);
            if (go && g.prev_n > 1) {
                const double2 s1 = rb_gate_sums(g, g.prev_first + 1, g.prev_nblk);
                go = !rb_gate_decide(g, g.prev_first + 1, s1, writer && stopped == 0);
            }
        }
        if (threadIdx.x == 0) live = go && stopped == 0 ? 1 : 0;
    }
    __syncthreads();
    return live != 0;
}

// the stop test of the batch's last launch (its successor would have run it)
__global__ __launch_bounds__(64) void k_rb_gate_final(RbGate g)
{
    if (*(volatile const unsigned*)g.stop) return;
    for (int t = 0; t < g.prev_n; ++t)
        if (rb_gate_decide(g, g.prev_first + t, rb_gate_sums(g, g.prev_first + t, g.prev_nblk),
                           true))
            return;
}

// Fused red-black sweep, one launch per sweep, dst = sweep(src) (ping-pong buffers).
// A workgroup owns a TJ x TK tile of (j, k) and marches over CI planes in i.  For each
// plane p it forms R(p), the state after the red half-sweep, on the tile plus a one-voxel
// ring (from old values on a two-voxel ring), then the black half-sweep of the tile from
// R(p-1), R(p), R(p+1).  Neighbouring tiles recompute the same ring values from the same
// old values, so they agree bit for bit and no workgroup waits for another.  The result
// equals the in-place red pass followed by the black pass.
constexpr int kRbTJ = 8, kRbTK = 64, kRbCI = 16;

template <class T, bool UNIT>
__global__ __launch_bounds__(256) void k_gs_rb_pp(const T* __restrict__ src, T* __restrict__ dst,
                                                  const T* __restrict__ rhs, int X, int Y, int Z,
                                                  GsCoef c, int s, int slot, int gated, RbGate g)
{
    if (gated && *(volatile unsigned*)g.stop) return;
    constexpr int TJ = kRbTJ, TK = kRbTK;
    constexpr int OJ = TJ + 4, OK = TK + 4;  // old planes, two-voxel ring
    constexpr int RJ = TJ + 2, RK = TK + 2;  // red planes, one-voxel ring
    __shared__ T O[4][OJ * OK];
    __shared__ T Rp[3][RJ * RK];
    const uint3 bl = xcd_block3();  // tiles sharing halo rows run on one XCD
    const int k0 = 1 + bl.x * TK, j0 = 1 + bl.y * TJ;
    const int ia = 1 + bl.z * kRbCI;
    const int ib = min(ia + kRbCI, X - 1);
    const long sX = (long)Y * Z;
    const int tid = threadIdx.x;

    auto load_O = [&](int p) {
        T* o = O[p & 3];
        const T* plane = src + (long)p * sX;
        for (int e = tid; e < OJ * OK; e += 256) {
            const int jj = e / OK, kk = e - jj * OK;
            const int j = j0 - 2 + jj, k = k0 - 2 + kk;
            o[e] = (j >= 0 && j < Y && k >= 0 && k < Z) ? plane[(long)j * Z + k] : (T)0;
        }
    };
    auto red = [&](int p) {
        T* r = Rp[p % 3];
        const T* o = O[p & 3];
        const T* om = O[(p + 3) & 3];
        const T* op = O[(p + 1) & 3];
        const bool inner_p = p >= 1 && p <= X - 2;
        for (int e = tid; e < RJ * RK; e += 256) {
            const int jj = e / RK, kk = e - jj * RK;
            const int j = j0 - 1 + jj, k = k0 - 1 + kk;
            const int oi = (jj + 1) * OK + (kk + 1);
            T v = o[oi];
            if (inner_p && j >= 1 && j <= Y - 2 && k >= 1 && k <= Z - 2 && ((p + j + k) & 1) == 0)
                v = gs_update<T, UNIT>(op[oi], om[oi], o[oi + OK], o[oi - OK], o[oi + 1], o[oi - 1],
                                       rhs[(long)p * sX + (long)j * Z + k], v, c);
            r[e] = v;
        }
    };
    double d2 = 0, o2 = 0;
    auto black = [&](int p) {
        const T* r = Rp[p % 3];
        const T* rm = Rp[(p + 2) % 3];
        const T* rp = Rp[(p + 1) % 3];
        const T* o = O[p & 3];
        for (int e = tid; e < TJ * TK; e += 256) {
            const int jj = e / TK, kk = e - jj * TK;
            const int j = j0 + jj, k = k0 + kk;
            if (j > Y - 2 || k > Z - 2) continue;
            const int ri = (jj + 1) * RK + (kk + 1);
            T v = r[ri];
            if (((p + j + k) & 1) == 1)
                v = gs_update<T, UNIT>(rp[ri], rm[ri], r[ri + RK], r[ri - RK], r[ri + 1], r[ri - 1],
                                       rhs[(long)p * sX + (long)j * Z + k], v, c);
            dst[(long)p * sX + (long)j * Z + k] = v;
            if (gated) {
                const double old = (double)o[(jj + 2) * OK + (kk + 2)];
                const double d = (double)v - old;
                d2 += d * d;
                o2 += old * old;
            }
        }
    };
    if (ia - 2 >= 0) load_O(ia - 2);
    load_O(ia - 1);
    load_O(ia);
    if (ia + 1 <= X - 1) load_O(ia + 1);
    __syncthreads();
    red(ia - 1);
    red(ia);
    // the first load of the loop reuses the slot of O(ia - 2), read by red(ia - 1)
    __syncthreads();
    for (int p = ia; p < ib; ++p) {
        if (p + 2 <= X - 1) load_O(p + 2);
        __syncthreads();
        red(p + 1);
        __syncthreads();
        black(p);
    }
    if (!gated) return;
    block_sum2(d2, o2);
    if (tid == 0) {
        atomicAdd(&g.acc[2 * slot], d2);
        atomicAdd(&g.acc[2 * slot + 1], o2);
        __threadfence();
        const unsigned nblk = gridDim.x * gridDim.y * gridDim.z;
        const unsigned ticket = atomicAdd(&g.arrive[slot], 1u);
        if (ticket == nblk - 1) {
            const double sd = atomicAdd(&g.acc[2 * slot], 0.0);
            double so = atomicAdd(&g.acc[2 * slot + 1], 0.0);
            if (s > 0) so += g.bnd_o2;
            const double err = sqrt(g.h3 * sd) / sqrt(g.h3 * so);
            g.err[slot] = err;
            g.arrive[slot] = 0;
            if (s >= g.ls_steps - 1 && !(err > g.ls_tol)) atomicExch(g.stop, (unsigned)(s + 1));
        }
    }
}

// The same sweep for rows of 16 bytes (nz a multiple of V): the tile's k columns are
// aligned to multiples of 64, so every plane row is moved with 16-byte accesses.  TJ (rows
// per tile) is 8 or 16, picked per volume shape and type by timing both.
constexpr int kRbW = kRbTK + 8;   // LDS row: k from 64*bx - 4 to 64*bx + 67

// Red-black sweeps with the tile's columns in registers.  Thread u owns one 16-byte row
// segment (row r, quad c4) of the tile and its ring, and keeps in registers the planes of
// its own column that the half-sweeps need along x; only the neighbours along y and z go
// through LDS.  NS sweeps (1 or 2) are 2 NS half-sweep stages (0: red of sweep 1, 1: black
// of sweep 1, 2: red of sweep 2, 3: black of sweep 2); stage s updates plane q - s at
// step q, over rows [s + 1, OJ - 1 - s) and columns [4 - (L - s), 67 + (L - s)] of the
// extended tile (L = 2 NS - 1, the last stage, whose range is the tile's own TJ rows and 64
// columns), from the output of stage s - 1 (the old field for stage 0), so the rings
// recompute what the neighbouring tiles compute from the same values, bit for bit.  One
// barrier per step: stage s reads its input plane's y and z neighbours from the LDS slot
// that stage s - 1 filled one step earlier (stage 0: the old plane written at the start of
// the step), and writes its own output to the other slot.  Every voxel goes through the
// updates of NS one-sweep launches, in the same order with the same operands.
template <class T, bool UNIT, int NS, int TJ, int NT>
__device__ __forceinline__ void rb_col_tile(const T* __restrict__ src, T* __restrict__ dst,
                                            const T* __restrict__ rhs, int X, int Y, int Z, int CI,
                                            const GsCoef& c, int s_idx, int gated,
                                            const RbGate& g, uint3 bl, unsigned fb)
{
    constexpr int V = Vec16<T>::n;
    constexpr int W = kRbW, WV = W / V;
    constexpr int NST = 2 * NS, LS = NST - 1;
    constexpr int OJ = TJ + 2 * NST;  // rows j0-NST .. j0+TJ+NST-1
    static_assert(OJ * WV <= NT, "one row segment per thread");
    __shared__ __align__(16) T P[NST][2][OJ * W];  // each stage's input plane, two slots
    const int kb = bl.x * kRbTK, kbase = kb - 4;
    const int j0 = 1 + bl.y * TJ;
    const int ia = 1 + bl.z * CI;
    const int ib = min(ia + CI, X - 1);
    const long sX = (long)Y * Z;
    const int tid = threadIdx.x;
    const bool act = tid < OJ * WV;
    const int r = act ? tid / WV : 0, c4 = act ? tid - r * WV : 0;
    const int j = j0 - NST + r, k0 = kbase + c4 * V;
    // every lane loads on every step (a clamped row where its segment is outside the
    // volume), so the wait before a use counts only the loads issued after it.  Values
    // outside the volume are never updated and are neighbours only of boundary voxels,
    // which are not updated either, so what the clamped rows hold does not matter.
    const long off = (long)min(max(j, 0), Y - 1) * Z + min(max(k0, 0), Z - V);
    // rhs is used on the first stage's rows 1 .. OJ-2; the two outer rows read a row that
    // another lane reads (no extra bytes from memory, the load count stays fixed)
    const int jr = j0 - NST + min(max(r, 1), OJ - 2);
    const long offr = (long)min(max(jr, 0), Y - 1) * Z + min(max(k0, 0), Z - V);
    const int li = r * W + c4 * V;
    auto ld_plane = [&](const T* f, int p) {
        return ldrow(f + (long)min(max(p, 0), X - 1) * sX + off);
    };
    auto ld_rhs = [&](int p) { return ldrow(rhs + (long)min(max(p, 0), X - 1) * sX + offr); };
    // Per stage, what does not change along the march: whether this segment's row is in
    // the stage's rows and inside the volume's interior in j (rowok), and which of its V
    // voxels lie in the stage's columns and in the interior in k (colm, bit u); the colour
    // of voxel u on plane p is (p + j + k0 + u) & 1.
    unsigned colm[NST];
    bool rowok[NST];
#pragma unroll
    for (int s = 0; s < NST; ++s) {
        const int cA = 4 - (LS - s), cB = 67 + (LS - s);
        rowok[s] = act && r >= s + 1 && r < OJ - 1 - s && j >= 1 && j <= Y - 2;
        unsigned m = 0;
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int col = c4 * V + u, k = k0 + u;
            if (col >= cA && col <= cB && k >= 1 && k <= Z - 2) m |= 1u << u;
        }
        colm[s] = m;
    }
    // y and z neighbours of one stage, read from its LDS plane before any stage of the step
    // runs (the planes a step reads were written in earlier steps)
    struct Nbr {
        Row<T> yp, ym;
        T zl, zr;
    };
    auto nbr = [&](const T* pl) {
        Nbr n;
        n.yp = ldrow(pl + li + W);
        n.ym = ldrow(pl + li - W);
        n.zl = pl[li - 1];
        n.zr = pl[li + V];
        return n;
    };
    // stage s on plane p: centre ctr, x neighbours xm (p - 1) and xp (p + 1), y and z
    // neighbours n, rhs b
    auto half = [&](int s, int p, const Row<T>& xm, const Row<T>& ctr, const Row<T>& xp,
                    const Nbr& n, const Row<T>& b) {
        Row<T> o = ctr;
        if (rowok[s] && p >= 1 && p <= X - 2) {
            // voxels of the stage's colour: bit u set where (p + j + k0 + u) & 1 == s & 1
            const unsigned par = ((p + j + k0 + s) & 1) ? 0xAu : 0x5u;
            const unsigned m = colm[s] & par;
#pragma unroll
            for (int u = 0; u < V; ++u) {
                const T zm = u > 0 ? ctr.v[u - 1] : n.zl;
                const T zp = u < V - 1 ? ctr.v[u + 1] : n.zr;
                const T v = gs_update<T, UNIT>(xp.v[u], xm.v[u], n.yp.v[u], n.ym.v[u], zp, zm,
                                               b.v[u], ctr.v[u], c);
                if ((m >> u) & 1u) o.v[u] = v;
            }
        }
        return o;
    };
    const int qa = ia - LS, qe = ib + LS;  // steps; stage s runs on planes q - s
    // Registers, as rings of K slots with plane p in slot (p - qa) mod K; the march is
    // unrolled by K steps, so every slot index is a constant and no value is copied from
    // one register to another (a copy would wait for the load that fills it):
    //   old planes q-1 .. q+1 in use and q+2 .. q+1+KL in flight (loaded KL steps ahead),
    //   rhs planes q-2NS+1 .. q in use (stage s takes plane q-s) and q+1 .. q+KB in flight,
    //   the last three outputs of every stage but the last.
    constexpr int K = 6, KL = K - 3, KB = K - NST;
    static_assert(KB >= 1, "rhs ring");
    Row<T> ro[K], rb[K], ra[NST > 1 ? NST - 1 : 1][K];
    // prologue: old planes qa-1 .. qa+1+KL, rhs planes qa-NST+1 .. qa+KB
#pragma unroll
    for (int d = -1; d <= 1 + KL; ++d) ro[(d + K) % K] = ld_plane(src, qa + d);
#pragma unroll
    for (int d = 1 - NST; d <= KB; ++d) rb[(d + K) % K] = ld_rhs(qa + d);
#pragma unroll
    for (int s = 0; s < NST - 1; ++s)
#pragma unroll
        for (int m = 0; m < K; ++m) ra[s][m] = Row<T>{};
    // gated: the stop test of the previous launch, while the prologue's loads are in flight
    if (gated && !rb_gate_live(g, fb == 0)) return;
    // update norm sums of each sweep over the voxels this workgroup owns
    double d2[NS], o2[NS];
#pragma unroll
    for (int n = 0; n < NS; ++n) d2[n] = o2[n] = 0;
    auto acc_norm = [&](int n, const Row<T>& nw, const Row<T>& old) {
#pragma unroll
        for (int e = 0; e < V; ++e) {
            const int k = k0 + e;
            if (k >= 1 && k <= Z - 2) {
                const double d = (double)nw.v[e] - (double)old.v[e];
                d2[n] += d * d;
                o2[n] += (double)old.v[e] * (double)old.v[e];
            }
        }
    };
    const bool own_seg = act && r >= NST && r < NST + TJ && j <= Y - 2;
    const int owned_c0 = 4 / V, owned_c1 = (4 + kRbTK) / V;
    // step q = q0 + u (q0 - qa a multiple of K)
    auto step = [&](auto u_c, int q) {
        constexpr int u = decltype(u_c)::value;
        auto sl = [](int d) { return ((u + d) % K + K) % K; };  // slot of plane q + d
        if (act) strow(&P[0][q & 1][li], ro[sl(0)]);
        __syncthreads();
        Nbr nb[NST];
        static_unroll<NST>([&](auto s_c) {
            constexpr int s = decltype(s_c)::value;
            nb[s] = nbr(P[s][(q - s) & 1]);
        });
        // stage 0 inputs: old planes q-1, q, q+1
        Row<T> x = ro[sl(-1)], y = ro[sl(0)], z = ro[sl(1)];
        static_unroll<NST>([&](auto s_c) {
            constexpr int s = decltype(s_c)::value;
            const int p = q - s;
            const bool run = p >= ia - (LS - s) && p < ib + (LS - s) && p >= 0 && p <= X - 1;
            Row<T> out = y;
            if (run) out = half(s, p, x, y, z, nb[s], rb[sl(-s)]);
            const bool own = run && own_seg && p >= ia && p < ib && c4 >= owned_c0 &&
                             c4 < owned_c1 && k0 < Z;
            if constexpr (s < LS) {
                // plane p of this stage; the next stage (plane p - 1) takes p - 2, p - 1, p
                if (act && run) strow(&P[s + 1][p & 1][li], out);
                ra[s][sl(-s)] = out;
                x = ra[s][sl(-s - 2)];
                y = ra[s][sl(-s - 1)];
                z = out;
                // stage 1 of two sweeps ends the first: its old values are plane q - 1
                if (s == 1 && gated && own) acc_norm(0, out, ro[sl(-1)]);
            } else {
                if (own) {
                    strow(dst + (long)p * sX + (long)j * Z + k0, out);
                    // old values of plane q - LS: the field (one sweep) or the first
                    // sweep's output (two)
                    if (gated) acc_norm(NS - 1, out, NS == 1 ? ro[sl(-1)] : ra[NS - 1][sl(-3)]);
                }
            }
        });
        // refill: old plane q+2+KL into the slot of q-1, rhs plane q+1+KB into the slot of
        // q+1-NST (the planes this step was the last to use)
        ro[sl(2 + KL)] = ld_plane(src, q + 2 + KL);
        rb[sl(1 + KB)] = ld_rhs(q + 1 + KB);
    };
    int q = qa;
    for (; q + K <= qe; q += K)
        static_unroll<K>([&](auto u_c) { step(u_c, q + decltype(u_c)::value); });
    static_unroll<K - 1>([&](auto u_c) {
        if (q + decltype(u_c)::value < qe) step(u_c, q + decltype(u_c)::value);
    });
    if (!gated) return;
    // this block's sums of each sweep, for the stop test that the next launch runs
#pragma unroll
    for (int n = 0; n < NS; ++n) {
        block_sum2_dpp(d2[n], o2[n]);
        if (tid == 0) {
            double* pp = g.part + (size_t)((s_idx + n) & 3) * (2 * kRbGateBlk) + 2 * (size_t)fb;
            pp[0] = d2[n];
            pp[1] = o2[n];
        }
    }
}

// threads of a column-march workgroup: one per row segment of the extended tile
template <class T, int NS, int TJ>
constexpr int rbc_threads()
{
    return ((TJ + 4 * NS) * (kRbW / Vec16<T>::n) + 63) / 64 * 64;
}

template <class T, bool UNIT, int NS, int TJ>
__global__ __launch_bounds__((rbc_threads<T, NS, TJ>())) void k_gs_rbc(
    const T* __restrict__ src, T* __restrict__ dst, const T* __restrict__ rhs, int X, int Y, int Z,
    int CI, GsCoef c, int s, int gated, RbGate g)
{
    rb_col_tile<T, UNIT, NS, TJ, rbc_threads<T, NS, TJ>()>(
        src, dst, rhs, X, Y, Z, CI, c, s, gated, g, xcd_block3(),
        blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// Many volumes, one sweep each, in one launch (the batch's fixed-sweep path): the job table
// gives each volume a block range starting at a multiple of the XCD count and padded to
// one, so the XCD-aware order inside a volume is that of its own launch; padding blocks
// exit at once.  The sweep of each voxel does not depend on the tiling, so every volume
// gets the values of its own one-volume sweep.
template <class T, bool UNIT, int TJ>
__global__ __launch_bounds__((rbc_threads<T, 1, TJ>())) void k_gs_rb_pv_batch(
    const RbJob* __restrict__ jobs, int njobs, GsCoef c)
{
    int lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= (int)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    const RbJob jb = jobs[lo];
    const unsigned n = (unsigned)jb.gx * jb.gy * jb.gz;
    unsigned w = xcd_remap(blockIdx.x - (unsigned)jb.first, (unsigned)jb.nblk);
    if (w >= n) return;
    uint3 bl;
    bl.x = w % jb.gx;
    w /= jb.gx;
    bl.y = w % jb.gy;
    bl.z = w / jb.gy;
    rb_col_tile<T, UNIT, 1, TJ, rbc_threads<T, 1, TJ>()>((const T*)jb.src, (T*)jb.dst,
                                                         (const T*)jb.rhs, jb.X, jb.Y, jb.Z,
                                                         jb.CI, c, 0, 0, RbGate{}, bl, 0);
}

__global__ void k_zero_d(double* p, int n)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0;
}
__global__ void k_zero_u(unsigned* p, int n)
{
    for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < n; i += blockDim.x * gridDim.x) p[i] = 0;
}

inline long long rb_key(int X, int Y, int Z, int w)
{
    return (((long long)X * 65536 + Y) * 65536 + Z) * 16 + w;
}

// LDS bytes of a column-march workgroup: two slots per stage of (TJ + 4 NS) rows
template <class T, int NS, int TJ>
constexpr long rbc_lds_bytes()
{
    return (long)sizeof(T) * 2L * 2 * NS * (TJ + 4L * NS) * kRbW;
}

// grid of NS aligned red-black sweeps with TJ-row tiles: the chunk length in x is chosen so
// the grid is one round of as many workgroups per CU as the kernel's registers and LDS
// allow (asked once per instance); a chunk recomputes 2 NS - 1 planes on each side for the
// first stage, so chunks keep at least 2 planes for one sweep (small volumes then spread
// over more workgroups: a 64^3 sweep is a few march steps long) and 8 for two
template <class T, int NS, int TJ>
static void rbc_grid(tk_ctx ctx, int X, int Y, int Z, dim3* grid, int* CI)
{
    static int resident = 0;
    if (resident <= 0) {
        int nb = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gs_rbc<T, true, NS, TJ>,
                                                         rbc_threads<T, NS, TJ>(), 0) !=
                hipSuccess ||
            nb < 1)
            nb = 1;
        resident = nb;
    }
    const long per_cu = resident;
    const int kt = (Z + kRbTK - 1) / kRbTK, jt = (Y - 2 + TJ - 1) / TJ;
    long chunks = ((long)ctx->num_cu * per_cu + (long)kt * jt / 2) / ((long)kt * jt);
    if (chunks < 1) chunks = 1;
    constexpr int min_ci = 2;
    if (chunks > (X - 2) / min_ci) chunks = std::max(1, (X - 2) / min_ci);
    *CI = (int)((X - 2 + chunks - 1) / chunks);
    *grid = dim3(kt, jt, (X - 2 + *CI - 1) / *CI);
}

// NS aligned red-black sweeps src -> dst (sweep indices s, s + 1) with TJ-row tiles
template <class T, int NS, int TJ>
static int launch_rbc(tk_ctx ctx, const void* src, void* dst, tk_vol rhs, const GsCoef& c, int s,
                      int gated, const RbGate& g, bool unit)
{
    constexpr int NT = rbc_threads<T, NS, TJ>();
    const int X = rhs->nx, Y = rhs->ny, Z = rhs->nz;
    dim3 grid;
    int CI = 0;
    rbc_grid<T, NS, TJ>(ctx, X, Y, Z, &grid, &CI);
    if (unit)
        hipLaunchKernelGGL((k_gs_rbc<T, true, NS, TJ>), grid, dim3(NT), 0, ctx->stream,
                           (const T*)src, (T*)dst, (const T*)rhs->d, X, Y, Z, CI, c, s, gated, g);
    else
        hipLaunchKernelGGL((k_gs_rbc<T, false, NS, TJ>), grid, dim3(NT), 0, ctx->stream,
                           (const T*)src, (T*)dst, (const T*)rhs->d, X, Y, Z, CI, c, s, gated, g);
    TK_LAUNCHED();
    return TK_OK;
}

// the autotuner's key for the two-sweep launch of a shape and type (next to the one-sweep
// key, whose last field is the element size)
template <class T>
inline long long rb2_key(int X, int Y, int Z)
{
    return rb_key(X, Y, Z, sizeof(T) == 4 ? 15 : 14);
}

// tile rows of the column kernel for n (1 or 2) sweeps per launch on rhs's shape: 8 or 16
// (one sweep in FP32 also 32), or 0 where it does not serve: rows not 16-byte aligned; for
// two sweeps also fewer than 3
// interior planes, two-sweep launches switched off, an untuned shape, or one where the
// autotuner found them slower than two one-sweep launches
template <class T>
static int rbc_rows(tk_ctx ctx, tk_vol rhs, int n)
{
    const int X = rhs->nx, Y = rhs->ny, Z = rhs->nz;
    if (Z % Vec16<T>::n != 0) return 0;
    if (n == 1) {
        const auto it = ctx->rb_rows.find(rb_key(X, Y, Z, (int)sizeof(T)));
        if (it == ctx->rb_rows.end()) return 8;
        return it->second == 16 || (sizeof(T) == 4 && it->second == 32) ? it->second : 8;
    }
    if (!ctx->rb_pair || X < 5 || Y < 3) return 0;
    const auto it = ctx->rb_rows.find(rb2_key<T>(X, Y, Z));
    return it == ctx->rb_rows.end() ? 0 : it->second;
}

// FP64 two-sweep tiles: 6 rows, so the 14-row extended tile is 504 threads (8 waves, up to
// 256 registers each; the 8-row tile's 9 waves would allow 168, and its rings spill)
constexpr int kRb2TJ64 = 6;

// blocks of the column launch of n sweeps with `rows` tile rows
template <class T>
static unsigned rbc_blocks(tk_ctx ctx, tk_vol rhs, int n, int rows)
{
    dim3 grid;
    int CI = 0;
    const int X = rhs->nx, Y = rhs->ny, Z = rhs->nz;
    if (n == 2) {
        if constexpr (sizeof(T) == 8) rbc_grid<T, 2, kRb2TJ64>(ctx, X, Y, Z, &grid, &CI);
        else if (rows == 16) rbc_grid<T, 2, 16>(ctx, X, Y, Z, &grid, &CI);
        else rbc_grid<T, 2, 8>(ctx, X, Y, Z, &grid, &CI);
    } else {
        if constexpr (sizeof(T) == 4)
            if (rows == 32) {
                rbc_grid<T, 1, 32>(ctx, X, Y, Z, &grid, &CI);
                return grid.x * grid.y * grid.z;
            }
        if (rows == 16) rbc_grid<T, 1, 16>(ctx, X, Y, Z, &grid, &CI);
        else rbc_grid<T, 1, 8>(ctx, X, Y, Z, &grid, &CI);
    }
    return grid.x * grid.y * grid.z;
}

// n (1 or 2) sweeps src -> dst with the column kernel of `rows` tile rows
template <class T>
static int rbc_launch(tk_ctx ctx, const void* src, void* dst, tk_vol rhs, const GsCoef& c, int n,
                      int rows, int s, int gated, const RbGate& g)
{
    const bool unit = c.hx2 == 1. && c.hy2 == 1. && c.hz2 == 1. && c.dt == 1.;
    if (n == 2) {
        TimedLaunch tl(ctx, TK_K_GS2);
        if constexpr (sizeof(T) == 8)
            return launch_rbc<T, 2, kRb2TJ64>(ctx, src, dst, rhs, c, s, gated, g, unit);
        else
            return rows == 16 ? launch_rbc<T, 2, 16>(ctx, src, dst, rhs, c, s, gated, g, unit)
                              : launch_rbc<T, 2, 8>(ctx, src, dst, rhs, c, s, gated, g, unit);
    }
    TimedLaunch tl(ctx, TK_K_GS);
    if constexpr (sizeof(T) == 4)
        if (rows == 32) return launch_rbc<T, 1, 32>(ctx, src, dst, rhs, c, s, gated, g, unit);
    return rows == 16 ? launch_rbc<T, 1, 16>(ctx, src, dst, rhs, c, s, gated, g, unit)
                      : launch_rbc<T, 1, 8>(ctx, src, dst, rhs, c, s, gated, g, unit);
}

// the batched sweep: 8-row tiles of 256 threads (the per-volume autotuner's choice at
// every size measured), the chunk length in x common to all volumes so that the whole
// launch is about one round of resident workgroups
constexpr int kRbBatchTJ = 8;
constexpr int kRbBatchRounds = 6;

template <class T>
static int rb_batch_table_t(tk_ctx ctx, const RbBatchVol* vols, int n, void** table, int* nblocks)
{
    constexpr int TJ = kRbBatchTJ;
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, k_gs_rb_pv_batch<T, true, TJ>,
                                                     rbc_threads<T, 1, TJ>(), 0) != hipSuccess ||
        resident < 1)
        resident = 1;
    const long per_cu = resident;
    long planes = 0;  // tile-planes of the whole batch
    for (int v = 0; v < n; ++v) {
        const int kt = (vols[v].Z + kRbTK - 1) / kRbTK, jt = (vols[v].Y - 2 + TJ - 1) / TJ;
        planes += (long)kt * jt * (vols[v].X - 2);
    }
    // several rounds of resident workgroups: chunks of different volumes differ in length,
    // and one round of whole-length chunks left a long tail (382 us per sweep of config 3
    // against 180 us at the 256^3 rate)
    const long target = std::max(1L, (long)ctx->num_cu * per_cu * kRbBatchRounds);
    const int CI0 = (int)std::max(4L, (planes + target - 1) / target);
    std::vector<RbJob> jobs((size_t)n);
    int first = 0;
    for (int v = 0; v < n; ++v) {
        RbJob& j = jobs[(size_t)v];
        j.src = vols[v].src;
        j.dst = vols[v].dst;
        j.rhs = vols[v].rhs;
        j.X = vols[v].X;
        j.Y = vols[v].Y;
        j.Z = vols[v].Z;
        j.CI = std::min(CI0, j.X - 2);
        j.gx = (j.Z + kRbTK - 1) / kRbTK;
        j.gy = (j.Y - 2 + TJ - 1) / TJ;
        j.gz = (j.X - 2 + j.CI - 1) / j.CI;
        const int nb = j.gx * j.gy * j.gz;
        j.nblk = (nb + kNumXcd - 1) / kNumXcd * kNumXcd;
        j.first = first;
        first += j.nblk;
    }
    void* d = nullptr;
    TK_CHECK(hipMalloc(&d, sizeof(RbJob) * (size_t)n) == hipSuccess, TK_ERR_NOMEM,
             "batched sweep: job table allocation failed");
    if (hipMemcpy(d, jobs.data(), sizeof(RbJob) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        set_error("batched sweep: job table upload failed");
        return TK_ERR_HIP;
    }
    *table = d;
    *nblocks = first;
    return TK_OK;
}

// one red-black sweep with the per-voxel kernel (rows not 16-byte aligned, or a gated grid
// with more blocks than the column kernels' stop test holds)
template <class T>
static int rb_sweep_pp(tk_ctx ctx, const void* src, void* dst, tk_vol rhs, const GsCoef& c, int s,
                       int slot, int gated, const RbGate& g)
{
    const int X = rhs->nx, Y = rhs->ny, Z = rhs->nz;
    const bool unit = c.hx2 == 1. && c.hy2 == 1. && c.hz2 == 1. && c.dt == 1.;
    TimedLaunch tl(ctx, TK_K_GS);
    const dim3 grid((Z - 2 + kRbTK - 1) / kRbTK, (Y - 2 + kRbTJ - 1) / kRbTJ,
                    (X - 2 + kRbCI - 1) / kRbCI);
    if (unit)
        hipLaunchKernelGGL((k_gs_rb_pp<T, true>), grid, dim3(256), 0, ctx->stream, (const T*)src,
                           (T*)dst, (const T*)rhs->d, X, Y, Z, c, s, slot, gated, g);
    else
        hipLaunchKernelGGL((k_gs_rb_pp<T, false>), grid, dim3(256), 0, ctx->stream, (const T*)src,
                           (T*)dst, (const T*)rhs->d, X, Y, Z, c, s, slot, gated, g);
    TK_LAUNCHED();
    return TK_OK;
}

// time the 8- and 16-row tile variants of the aligned sweep on rhs's shape, and 32 rows in
// FP32 (ungated sweeps of a zero field into a scratch volume); keep the fastest for this
// context
template <class T>
static int rb_autotune(tk_ctx ctx, tk_vol rhs, int* rows_out, double* ms_out)
{
    const int X = rhs->nx, Y = rhs->ny, Z = rhs->nz;
    if (Z % Vec16<T>::n != 0) {
        if (rows_out) *rows_out = 0;  // unaligned rows: single scalar variant
        return TK_OK;
    }
    void *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double ms[3] = {0, 0, 0};
    int rc = TK_OK;
    const GsCoef c = gs_coef(rhs, 1.0);
    const bool unit = c.hx2 == 1. && c.hy2 == 1. && c.hz2 == 1.;
    RbGate g = {};
    auto body = [&]() -> int {
        TK_HIP(hipMalloc(&a, rhs->bytes()));
        TK_HIP(hipMalloc(&b, rhs->bytes()));
        TK_HIP(hipMemsetAsync(a, 0, rhs->bytes(), ctx->stream));
        TK_HIP(hipEventCreate(&e0));
        TK_HIP(hipEventCreate(&e1));
        for (int v = 0; v < (sizeof(T) == 4 ? 3 : 2); ++v) {
            auto launch = [&]() {
                if constexpr (sizeof(T) == 4)
                    if (v == 2) return launch_rbc<T, 1, 32>(ctx, a, b, rhs, c, 0, 0, g, unit);
                return v == 0 ? launch_rbc<T, 1, 8>(ctx, a, b, rhs, c, 0, 0, g, unit)
                              : launch_rbc<T, 1, 16>(ctx, a, b, rhs, c, 0, 0, g, unit);
            };
            TK_TRY(launch());
            TK_HIP(hipEventRecord(e0, ctx->stream));
            for (int r = 0; r < 3; ++r) TK_TRY(launch());
            TK_HIP(hipEventRecord(e1, ctx->stream));
            TK_HIP(hipEventSynchronize(e1));
            float t = 0;
            TK_HIP(hipEventElapsedTime(&t, e0, e1));
            ms[v] = t / 3.0;
        }
        return TK_OK;
    };
    rc = body();
    // the two-sweep launch: the same two tile heights (FP64: one, of 6 rows, in ms2[0])
    double ms2[2] = {0, 0};
    if (rc == TK_OK && X >= 5 && Y >= 3) {
        rc = [&]() -> int {
            for (int v = 0; v < (sizeof(T) == 8 ? 1 : 2); ++v) {
                auto launch = [&]() {
                    if constexpr (sizeof(T) == 8)
                        return launch_rbc<T, 2, kRb2TJ64>(ctx, a, b, rhs, c, 0, 0, g, unit);
                    else
                        return v == 0 ? launch_rbc<T, 2, 8>(ctx, a, b, rhs, c, 0, 0, g, unit)
                                      : launch_rbc<T, 2, 16>(ctx, a, b, rhs, c, 0, 0, g, unit);
                };
                TK_TRY(launch());
                TK_HIP(hipEventRecord(e0, ctx->stream));
                for (int r = 0; r < 3; ++r) TK_TRY(launch());
                TK_HIP(hipEventRecord(e1, ctx->stream));
                TK_HIP(hipEventSynchronize(e1));
                float t = 0;
                TK_HIP(hipEventElapsedTime(&t, e0, e1));
                ms2[v] = t / 3.0;
            }
            // kept only where one launch of two sweeps beats two launches of one
            const double best1 = sizeof(T) == 4 ? std::min(std::min(ms[0], ms[1]), ms[2])
                                                : std::min(ms[0], ms[1]);
            const double best2 = sizeof(T) == 8 ? ms2[0] : std::min(ms2[0], ms2[1]);
            ctx->rb_rows[rb2_key<T>(X, Y, Z)] =
                best2 < 2 * best1 ? (sizeof(T) == 4 && ms2[1] < ms2[0] ? 16 : 8) : 0;
            return TK_OK;
        }();
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (rc != TK_OK) return rc;
    int rows = ms[1] < ms[0] ? 16 : 8;
    if (sizeof(T) == 4 && ms[2] < std::min(ms[0], ms[1])) rows = 32;
    ctx->rb_rows[rb_key(X, Y, Z, (int)sizeof(T))] = rows;
    if (rows_out) *rows_out = rows;
    if (ms_out) {
        ms_out[0] = ms[0];
        ms_out[1] = ms[1];
        ms_out[2] = ms[2];
        ms_out[3] = ms2[0];
        ms_out[4] = ms2[1];
    }
    return TK_OK;
}

template <class T>
static int set_boundary(tk_ctx ctx, tk_vol phi, double bc, double* acc)
{
    const long nb_vox = boundary_count(phi->nx, phi->ny, phi->nz);
    const unsigned nb = blocks_for(nb_vox, 256, 1024);
    hipLaunchKernelGGL(k_set_boundary<T>, dim3(nb), dim3(256), 0, ctx->stream, (T*)phi->d, phi->nx,
                       phi->ny, phi->nz, (T)bc, acc);
    TK_LAUNCHED();
    return TK_OK;
}

// sweep loop of rsfe_splitbregman_imp.hxx:1762-1797
template <class T>
// faces_set: phi's (and *alt's) six faces already hold bc and no sweep, cut or copy since
// has written them, so the fixed-sweep paths skip setting them again.  neumann: the sweep is
// NeuGaussSeidel (lexicographic; every voxel updated, bc is the boundary derivative)
static int solve(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order, int ls_steps,
                 double ls_tol, int max_sweeps, int* sweeps_out, double* err_out,
                 void** alt = nullptr, bool faces_set = false, bool neumann = false)
{
    const GsCoef c = gs_coef(phi, dt);
    const int X = phi->nx, Y = phi->ny, Z = phi->nz;
    const double h3 = phi->hx * phi->hy * phi->hz;
    NeuCoef ncoef{};
    const NeuCoef* neu = nullptr;
    if (neumann) {
        TK_CHECK(order == TK_GS_LEX, TK_ERR_UNSUPPORTED,
                 "poisson: the Neumann sweep runs in the reference's lexicographic order only");
        ncoef = neu_coef(phi, dt, bc);
        neu = &ncoef;
    }
    // Dirichlet: the faces (bc) enter the update norms; Neumann: the kernel covers them
    const double bnd_o2 =
        neu ? 0. : (double)boundary_count(X, Y, Z) * (double)(T)bc * (double)(T)bc;
    if (ls_steps < 1) ls_steps = 1;
    if (order == TK_GS_REDBLACK) {
        // ping-pong between phi and a scratch volume (the caller's *alt, else the context's
        // slot 4); both carry the boundary value
        void* other;
        if (alt) other = *alt;
        else TK_TRY(ctx_work(ctx, 4, phi->bytes(), &other));
        tk_vol_s tmp = *phi;
        tmp.d = other;
        tmp.owned = false;
        if (!(faces_set && ls_tol == 0)) TK_TRY(set_boundary<T>(ctx, &tmp, bc, nullptr));
        void* buf[2] = {phi->d, other};
        int sweeps = 0;
        double err = -1;
        int r = TK_OK;
        int launches = 0;  // the result is in buf[launches & 1]
        // column kernels: tile rows of the one- and two-sweep launches (0: not served)
        const int rows1 = rbc_rows<T>(ctx, rhs, 1), rows2 = rows1 ? rbc_rows<T>(ctx, rhs, 2) : 0;
        if (ls_tol == 0) {
            if (!faces_set) TK_TRY(set_boundary<T>(ctx, phi, bc, nullptr));
            RbGate g{};
            for (int q = 0; q < ls_steps; ++launches) {
                void* from = buf[launches & 1];
                void* to = buf[(launches + 1) & 1];
                if (rows2 && ls_steps - q >= 2) {
                    TK_TRY(rbc_launch<T>(ctx, from, to, rhs, c, 2, rows2, q, 0, g));
                    q += 2;
                } else {
                    if (rows1) TK_TRY(rbc_launch<T>(ctx, from, to, rhs, c, 1, rows1, q, 0, g));
                    else TK_TRY(rb_sweep_pp<T>(ctx, from, to, rhs, c, q, 0, 0, g));
                    q += 1;
                }
            }
            sweeps = ls_steps;
        } else {
            if (max_sweeps <= 0) max_sweeps = 1 << 30;
            // batches of B sweeps; the device stops itself, the host polls between batches
            const int B = 256;
            RbGate g{};
            g.acc = ctx->acc + kAccSweep;
            g.err = ctx->acc + kAccErr;  // B error slots
            g.stop = ctx->cnt;
            g.arrive = ctx->cnt + 1;
            g.ls_steps = ls_steps;
            g.ls_tol = ls_tol;
            g.h3 = h3;
            g.bnd_o2 = bnd_o2;
            g.part = ctx->acc + kAccGate;
            // Column kernels where their grids fit the per-block sums of the stop test: each
            // launch first tests the sweeps of the launch before it (all its workgroups add
            // the same sums in the same order), and one more kernel tests the batch's last
            // launch.  The per-voxel kernel tests its own sweep (atomic sums).
            int r1 = rows1, r2 = rows2;
            const unsigned nb1 = r1 ? rbc_blocks<T>(ctx, rhs, 1, r1) : 0;
            const unsigned nb2 = r2 ? rbc_blocks<T>(ctx, rhs, 2, r2) : 0;
            if (nb1 > (unsigned)kRbGateBlk) r1 = r2 = 0;
            if (nb2 > (unsigned)kRbGateBlk) r2 = 0;
            struct Launch {
                int first, n, src;  // first sweep, sweeps, source buffer
            };
            std::vector<Launch> ls;
            ls.reserve(B);
            hipLaunchKernelGGL(k_zero_u, dim3(1), dim3(256), 0, ctx->stream, ctx->cnt, B + 1);
            int done = 0, cur = 0;  // cur: the buffer that holds the latest sweep's output
            r = TK_ERR_SWEEPS;
            while (done < max_sweeps) {
                const int nb = std::min(B, max_sweeps - done);
                hipLaunchKernelGGL(k_zero_d, dim3(4), dim3(256), 0, ctx->stream, g.acc, 2 * B);
                if (done == 0) TK_TRY(set_boundary<T>(ctx, phi, bc, g.acc));
                g.base = done;
                g.prev_n = 0;
                ls.clear();
                for (int q = 0; q < nb;) {
                    const int sw = done + q;
                    const int n = r2 && nb - q >= 2 ? 2 : 1;
                    ls.push_back({sw, n, cur});
                    if (r1) {
                        TK_TRY(rbc_launch<T>(ctx, buf[cur], buf[cur ^ 1], rhs, c, n,
                                             n == 2 ? r2 : r1, sw, 1, g));
                        g.prev_first = sw;
                        g.prev_n = n;
                        g.prev_nblk = n == 2 ? nb2 : nb1;
                    } else {
                        TK_TRY(rb_sweep_pp<T>(ctx, buf[cur], buf[cur ^ 1], rhs, c, sw, q, 1, g));
                    }
                    cur ^= 1;
                    q += n;
                }
                if (r1) {
                    hipLaunchKernelGGL(k_rb_gate_final, dim3(1), dim3(64), 0, ctx->stream, g);
                    TK_LAUNCHED();
                }
                unsigned stop = 0;
                TK_HIP(hipMemcpyAsync(ctx->h_acc, g.err, sizeof(double) * B, hipMemcpyDeviceToHost,
                                      ctx->stream));
                TK_HIP(hipMemcpyAsync(&ctx->h_acc[B], ctx->cnt, sizeof(unsigned),
                                      hipMemcpyDeviceToHost, ctx->stream));
                TK_HIP(hipStreamSynchronize(ctx->stream));
                std::memcpy(&stop, &ctx->h_acc[B], sizeof(unsigned));
                if (stop) {
                    sweeps = (int)stop;
                    err = ctx->h_acc[sweeps - 1 - done];
                    r = TK_OK;
                    // the launches after the stop sweep's ran nothing; where that sweep was
                    // the first of two in one launch, run it again by itself over the
                    // second's output (its source is untouched)
                    const int sp = sweeps - 1;
                    size_t i = 0;
                    while (i + 1 < ls.size() && sp >= ls[i].first + ls[i].n) ++i;
                    const Launch& L = ls[i];
                    if (L.n == 2 && sp == L.first)
                        TK_TRY(rbc_launch<T>(ctx, buf[L.src], buf[L.src ^ 1], rhs, c, 1, r1, sp, 0,
                                             g));
                    cur = L.src ^ 1;
                    break;
                }
                done += nb;
                sweeps = done;
                err = ctx->h_acc[nb - 1];
            }
            if (r != TK_OK)
                set_error("tk_poisson_dir_solve: %d sweeps without reaching ls_tol", sweeps);
            launches = cur;
        }
        if (launches & 1) {
            // the result is in the scratch volume: hand it to phi
            if (alt) {
                *alt = phi->d;
                phi->d = other;
            } else if (phi->owned) {
                ctx->work[4] = phi->d;
                ctx->work_bytes[4] = phi->bytes();
                phi->d = other;
            } else {
                TK_HIP(hipMemcpyAsync(phi->d, other, phi->bytes(), hipMemcpyDeviceToDevice,
                                      ctx->stream));
            }
        }
        if (sweeps_out) *sweeps_out = sweeps;
        if (err_out) *err_out = err;
        return r;
    }
    if (ls_tol == 0) {
        // exactly ls_steps sweeps, no norms (the reference computes none either)
        if (!faces_set && !neu) TK_TRY(set_boundary<T>(ctx, phi, bc, nullptr));
        TK_TRY(lex_chunk<T>(ctx, phi, rhs, c, ls_steps, nullptr, neu));
        if (sweeps_out) *sweeps_out = ls_steps;
        if (err_out) *err_out = -1;
        return TK_OK;
    }
    if (max_sweeps <= 0) max_sweeps = 1 << 30;
    double* acc = ctx->acc + kAccSweep;  // per-sweep pairs (up to 2048 sweeps per chunk)
    const int kMaxChunk = 2048;

    // lexicographic: chunks with replay
    void* save_p;
    TK_TRY(ctx_work(ctx, 2, phi->bytes(), &save_p));
    int done = 0;
    int C = std::max(ls_steps, 16);
    while (done < max_sweeps) {
        C = std::min(C, std::min(kMaxChunk, max_sweeps - done));
        TK_HIP(hipMemcpyAsync(save_p, phi->d, phi->bytes(), hipMemcpyDeviceToDevice, ctx->stream));
        hipLaunchKernelGGL(k_zero_d, dim3(16), dim3(256), 0, ctx->stream, acc, 2 * C);
        if (done == 0 && !neu) TK_TRY(set_boundary<T>(ctx, phi, bc, acc));
        TK_TRY(lex_chunk<T>(ctx, phi, rhs, c, C, acc, neu));
        TK_HIP(hipMemcpyAsync(ctx->h_acc, acc, sizeof(double) * 2 * C, hipMemcpyDeviceToHost,
                              ctx->stream));
        TK_HIP(hipStreamSynchronize(ctx->stream));
        int stop_at = -1;
        double err = -1;
        for (int q = 0; q < C; ++q) {
            const int s = done + q;
            double so = ctx->h_acc[2 * q + 1];
            if (s > 0) so += bnd_o2;
            err = std::sqrt(h3 * ctx->h_acc[2 * q]) / std::sqrt(h3 * so);
            if (s >= ls_steps - 1 && !(err > ls_tol)) {
                stop_at = q;
                break;
            }
        }
        if (stop_at >= 0) {
            if (stop_at < C - 1) {
                // replay the chunk up to the stopping sweep from the saved state
                TK_HIP(hipMemcpyAsync(phi->d, save_p, phi->bytes(), hipMemcpyDeviceToDevice,
                                      ctx->stream));
                if (done == 0 && !neu) TK_TRY(set_boundary<T>(ctx, phi, bc, nullptr));
                TK_TRY(lex_chunk<T>(ctx, phi, rhs, c, stop_at + 1, nullptr, neu));
            }
            if (sweeps_out) *sweeps_out = done + stop_at + 1;
            if (err_out) *err_out = err;
            return TK_OK;
        }
        done += C;
        if (err_out) *err_out = err;
        C *= 2;
    }
    if (sweeps_out) *sweeps_out = done;
    set_error("tk_poisson_dir_solve: %d sweeps without reaching ls_tol", done);
    return TK_ERR_SWEEPS;
}

// residual r = rhs + lap_h phi on the interior; sums r^2 and rhs^2
template <class T>
__global__ __launch_bounds__(256) void k_residual(const T* __restrict__ phi, const T* __restrict__ rhs,
                                                  int X, int Y, int Z, double ihx2, double ihy2,
                                                  double ihz2, double* __restrict__ part)
{
    const long total = (long)(X - 2) * (Y - 2) * (Z - 2);
    const long sY = Z, sX = (long)Y * Z;
    double r2 = 0, b2 = 0;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total;
         q += (long)gridDim.x * blockDim.x) {
        const int k = 1 + (int)(q % (Z - 2));
        const long rr = q / (Z - 2);
        const int j = 1 + (int)(rr % (Y - 2)), i = 1 + (int)(rr / (Y - 2));
        const long v = i * sX + j * sY + k;
        const double c = phi[v];
        const double lap = ((double)phi[v + sX] + (double)phi[v - sX] - 2 * c) * ihx2 +
                           ((double)phi[v + sY] + (double)phi[v - sY] - 2 * c) * ihy2 +
                           ((double)phi[v + 1] + (double)phi[v - 1] - 2 * c) * ihz2;
        const double b = rhs[v];
        const double r = b + lap;
        r2 += r * r;
        b2 += b * b;
    }
    block_sum2(r2, b2);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = r2;
        part[2 * blockIdx.x + 1] = b2;
    }
}

// r = rhs + lap_h phi on the interior, 0 on the faces (the Dirichlet problem's residual)
template <class T>
__global__ void k_residual_field(const T* __restrict__ phi, const T* __restrict__ rhs, int X,
                                 int Y, int Z, double ihx2, double ihy2, double ihz2,
                                 T* __restrict__ out)
{
    const long n = (long)X * Y * Z, sY = Z, sX = (long)Y * Z;
    for (long v = blockIdx.x * (long)blockDim.x + threadIdx.x; v < n;
         v += (long)gridDim.x * blockDim.x) {
        const int k = (int)(v % Z);
        const long rr = v / Z;
        const int j = (int)(rr % Y), i = (int)(rr / Y);
        T r = 0;
        if (i > 0 && i < X - 1 && j > 0 && j < Y - 1 && k > 0 && k < Z - 1) {
            const double c = phi[v];
            const double lap = ((double)phi[v + sX] + (double)phi[v - sX] - 2 * c) * ihx2 +
                               ((double)phi[v + sY] + (double)phi[v - sY] - 2 * c) * ihy2 +
                               ((double)phi[v + 1] + (double)phi[v - 1] - 2 * c) * ihz2;
            r = (T)((double)rhs[v] + lap);
        }
        out[v] = r;
    }
}

__global__ void k_sum_pairs_gs(const double* __restrict__ part, int nb, double* __restrict__ out)
{
    double s1 = 0, s2 = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        s1 += part[2 * i];
        s2 += part[2 * i + 1];
    }
    block_sum2(s1, s2);
    if (threadIdx.x == 0) {
        out[0] = s1;
        out[1] = s2;
    }
}

}  // namespace tk

using namespace tk;

namespace tk {
int rb_batch_table(tk_ctx ctx, tk_dtype dt, const RbBatchVol* vols, int n, void** table,
                   int* nblocks)
{
    TK_CHECK(ctx && vols && n > 0 && table && nblocks, TK_ERR_ARG, "batched sweep: argument");
    for (int v = 0; v < n; ++v)
        TK_CHECK(vols[v].X >= 3 && vols[v].Y >= 3 && vols[v].Z >= 3 &&
                     vols[v].Z % (dt == TK_F64 ? 2 : 4) == 0,
                 TK_ERR_UNSUPPORTED, "batched sweep: volume %d: 16-byte rows needed", v);
    return dt == TK_F64 ? rb_batch_table_t<double>(ctx, vols, n, table, nblocks)
                        : rb_batch_table_t<float>(ctx, vols, n, table, nblocks);
}

int rb_batch_sweep(tk_ctx ctx, tk_dtype dt, const void* table, int n, int nblocks, double dt_step,
                   double hx, double hy, double hz)
{
    tk_vol_s like{};
    like.hx = hx;
    like.hy = hy;
    like.hz = hz;
    const GsCoef c = gs_coef(&like, dt_step);
    const bool unit = c.hx2 == 1. && c.hy2 == 1. && c.hz2 == 1. && c.dt == 1.;
    constexpr int TJ = kRbBatchTJ;
    constexpr int NT64 = rbc_threads<double, 1, TJ>(), NT32 = rbc_threads<float, 1, TJ>();
    TimedLaunch tl(ctx, TK_K_GS);
    const RbJob* jobs = (const RbJob*)table;
    if (dt == TK_F64) {
        if (unit)
            hipLaunchKernelGGL((k_gs_rb_pv_batch<double, true, TJ>), dim3(nblocks), dim3(NT64), 0,
                               ctx->stream, jobs, n, c);
        else
            hipLaunchKernelGGL((k_gs_rb_pv_batch<double, false, TJ>), dim3(nblocks), dim3(NT64),
                               0, ctx->stream, jobs, n, c);
    } else {
        if (unit)
            hipLaunchKernelGGL((k_gs_rb_pv_batch<float, true, TJ>), dim3(nblocks), dim3(NT32), 0,
                               ctx->stream, jobs, n, c);
        else
            hipLaunchKernelGGL((k_gs_rb_pv_batch<float, false, TJ>), dim3(nblocks), dim3(NT32),
                               0, ctx->stream, jobs, n, c);
    }
    TK_LAUNCHED();
    return TK_OK;
}
}  // namespace tk

// min_edge: 3 for the Dirichlet sweep (an interior to update), 2 for the Neumann sweep
// (its 2-D branch, Z == 1, is not built)
static int check_gs(tk_vol phi, tk_vol rhs, int min_edge = 3)
{
    TK_CHECK(phi && rhs && phi->d && rhs->d, TK_ERR_ARG, "poisson: null volume");
    TK_CHECK(same_dims(phi, rhs), TK_ERR_SHAPE, "poisson: phi and rhs dims differ");
    TK_CHECK(phi->dtype == rhs->dtype, TK_ERR_DTYPE, "poisson: dtypes differ");
    TK_CHECK(phi->dtype == TK_F32 || phi->dtype == TK_F64, TK_ERR_DTYPE, "poisson: not FP");
    TK_CHECK(phi->nx >= min_edge && phi->ny >= min_edge && phi->nz >= min_edge,
             TK_ERR_UNSUPPORTED, "poisson: 3-D volumes with every edge >= %d only", min_edge);
    return TK_OK;
}

namespace tk {
int poisson_solve_residual_alt(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc,
                               int order, double res_tol, int check_every, int max_sweeps,
                               void** alt, int* sweeps_out, double* res_out)
{
    TK_TRY(check_gs(phi, rhs));
    TK_CHECK(res_tol > 0 && check_every > 0, TK_ERR_ARG,
             "poisson residual solve: res_tol and check_every must be > 0");
    if (max_sweeps <= 0) max_sweeps = 1 << 30;
    if (phi->dtype == TK_F64) TK_TRY(set_boundary<double>(ctx, phi, bc, nullptr));
    else TK_TRY(set_boundary<float>(ctx, phi, bc, nullptr));
    double res = 0;
    TK_TRY(tk_poisson_residual(ctx, phi, rhs, &res));
    int sweeps = 0;
    while (!(res <= res_tol) && sweeps < max_sweeps) {
        const int k = std::min(check_every, max_sweeps - sweeps);
        if (phi->dtype == TK_F64)
            TK_TRY(solve<double>(ctx, phi, rhs, dt, bc, order, k, 0, 0, nullptr, nullptr, alt));
        else
            TK_TRY(solve<float>(ctx, phi, rhs, dt, bc, order, k, 0, 0, nullptr, nullptr, alt));
        sweeps += k;
        TK_TRY(tk_poisson_residual(ctx, phi, rhs, &res));
    }
    if (sweeps_out) *sweeps_out = sweeps;
    if (res_out) *res_out = res;
    if (!(res <= res_tol)) {
        set_error("poisson residual solve: %d sweeps, residual %.3e > %.3e", sweeps, res,
                  res_tol);
        return TK_ERR_SWEEPS;
    }
    return TK_OK;
}

int poisson_set_boundary(tk_ctx ctx, tk_vol phi, double bc)
{
    if (phi->dtype == TK_F64) return set_boundary<double>(ctx, phi, bc, nullptr);
    return set_boundary<float>(ctx, phi, bc, nullptr);
}

int poisson_dir_solve_alt(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int ls_steps, double ls_tol, int max_sweeps, void** alt,
                          int* sweeps_out, double* err_out, bool faces_set)
{
    TK_TRY(check_gs(phi, rhs));
    if (phi->dtype == TK_F64)
        return solve<double>(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps,
                             sweeps_out, err_out, alt, faces_set);
    return solve<float>(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps, sweeps_out,
                        err_out, alt, faces_set);
}

int poisson_neu_solve_alt(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int ls_steps, double ls_tol, int max_sweeps, int* sweeps_out,
                          double* err_out)
{
    TK_TRY(check_gs(phi, rhs, 2));
    if (phi->dtype == TK_F64)
        return solve<double>(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps,
                             sweeps_out, err_out, nullptr, false, true);
    return solve<float>(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps, sweeps_out,
                        err_out, nullptr, false, true);
}
}  // namespace tk

extern "C" {

int tk_poisson_dir_sweeps(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int n_sweeps)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_dir_sweeps: null ctx");
    TK_TRY(check_gs(phi, rhs));
    TK_CHECK(order == TK_GS_LEX || order == TK_GS_REDBLACK, TK_ERR_ARG, "bad order");
    TK_CHECK(dt > 0, TK_ERR_ARG, "tk_poisson_dir_sweeps: dt <= 0");
    if (n_sweeps <= 0) return TK_OK;
    if (phi->dtype == TK_F64)
        return solve<double>(ctx, phi, rhs, dt, bc, order, n_sweeps, 0, 0, nullptr, nullptr);
    return solve<float>(ctx, phi, rhs, dt, bc, order, n_sweeps, 0, 0, nullptr, nullptr);
}

int tk_poisson_dir_solve(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                         int ls_steps, double ls_tol, int max_sweeps, int* sweeps_out,
                         double* err_out)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_dir_solve: null ctx");
    TK_TRY(check_gs(phi, rhs));
    TK_CHECK(order == TK_GS_LEX || order == TK_GS_REDBLACK, TK_ERR_ARG, "bad order");
    TK_CHECK(dt > 0 && ls_tol >= 0, TK_ERR_ARG, "tk_poisson_dir_solve: dt <= 0 or ls_tol < 0");
    if (phi->dtype == TK_F64)
        return solve<double>(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps, sweeps_out,
                             err_out);
    return solve<float>(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps, sweeps_out,
                        err_out);
}

int tk_poisson_neu_sweeps(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int n_sweeps)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_neu_sweeps: null ctx");
    TK_TRY(check_gs(phi, rhs, 2));
    TK_CHECK(order == TK_GS_LEX || order == TK_GS_REDBLACK, TK_ERR_ARG, "bad order");
    TK_CHECK(dt > 0, TK_ERR_ARG, "tk_poisson_neu_sweeps: dt <= 0");
    if (n_sweeps <= 0) return TK_OK;
    return poisson_neu_solve_alt(ctx, phi, rhs, dt, bc, order, n_sweeps, 0, 0, nullptr, nullptr);
}

int tk_poisson_neu_solve(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                         int ls_steps, double ls_tol, int max_sweeps, int* sweeps_out,
                         double* err_out)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_neu_solve: null ctx");
    TK_TRY(check_gs(phi, rhs, 2));
    TK_CHECK(order == TK_GS_LEX || order == TK_GS_REDBLACK, TK_ERR_ARG, "bad order");
    TK_CHECK(dt > 0 && ls_tol >= 0, TK_ERR_ARG, "tk_poisson_neu_solve: dt <= 0 or ls_tol < 0");
    return poisson_neu_solve_alt(ctx, phi, rhs, dt, bc, order, ls_steps, ls_tol, max_sweeps,
                                 sweeps_out, err_out);
}

int tk_poisson_residual_field(tk_ctx ctx, tk_vol phi, tk_vol rhs, tk_vol out)
{
    TK_CHECK(ctx && out && out->d, TK_ERR_ARG, "tk_poisson_residual_field: argument");
    TK_TRY(check_gs(phi, rhs));
    TK_CHECK(same_dims(out, phi) && out->dtype == phi->dtype, TK_ERR_SHAPE,
             "tk_poisson_residual_field: out must match phi");
    const unsigned nb = blocks_for(phi->n(), 256, 8192);
    const double ihx2 = 1. / (phi->hx * phi->hx), ihy2 = 1. / (phi->hy * phi->hy),
                 ihz2 = 1. / (phi->hz * phi->hz);
    if (phi->dtype == TK_F64)
        hipLaunchKernelGGL(k_residual_field<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double*)phi->d, (const double*)rhs->d, phi->nx, phi->ny, phi->nz,
                           ihx2, ihy2, ihz2, (double*)out->d);
    else
        hipLaunchKernelGGL(k_residual_field<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float*)phi->d, (const float*)rhs->d, phi->nx, phi->ny, phi->nz,
                           ihx2, ihy2, ihz2, (float*)out->d);
    TK_LAUNCHED();
    return TK_OK;
}

int tk_poisson_solve_residual(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc,
                              int order, double res_tol, int check_every, int max_sweeps,
                              int* sweeps_out, double* res_out)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_solve_residual: null context");
    return poisson_solve_residual_alt(ctx, phi, rhs, dt, bc, order, res_tol, check_every,
                                      max_sweeps, nullptr, sweeps_out, res_out);
}

int tk_poisson_rb_autotune(tk_ctx ctx, tk_vol rhs, int* tile_rows, double* ms)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_rb_autotune: null context");
    TK_TRY(check_gs(rhs, rhs));
    return rhs->dtype == TK_F64 ? rb_autotune<double>(ctx, rhs, tile_rows, ms)
                                : rb_autotune<float>(ctx, rhs, tile_rows, ms);
}

int tk_poisson_rb_set_pair(tk_ctx ctx, int enable)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_rb_set_pair: null context");
    ctx->rb_pair = enable ? 1 : 0;
    return TK_OK;
}

int tk_poisson_rb_set_tile(tk_ctx ctx, tk_vol rhs, int tile_rows)
{
    TK_CHECK(ctx, TK_ERR_ARG, "tk_poisson_rb_set_tile: null context");
    TK_TRY(check_gs(rhs, rhs));
    TK_CHECK(tile_rows == 8 || tile_rows == 16 || (tile_rows == 32 && rhs->dtype == TK_F32),
             TK_ERR_ARG, "tk_poisson_rb_set_tile: tile_rows must be 8 or 16 (or 32 in FP32)");
    ctx->rb_rows[rb_key(rhs->nx, rhs->ny, rhs->nz, (int)dtype_size(rhs->dtype))] = tile_rows;
    // the two-sweep launch has 8- and 16-row tiles
    ctx->rb_rows[rhs->dtype == TK_F32 ? rb2_key<float>(rhs->nx, rhs->ny, rhs->nz)
                                      : rb2_key<double>(rhs->nx, rhs->ny, rhs->nz)] =
        tile_rows > 16 ? 16 : tile_rows;
    return TK_OK;
}

int tk_poisson_residual(tk_ctx ctx, tk_vol phi, tk_vol rhs, double* rel)
{
    TK_CHECK(ctx && rel, TK_ERR_ARG, "tk_poisson_residual: null argument");
    TK_TRY(check_gs(phi, rhs));
    const long total = (long)(phi->nx - 2) * (phi->ny - 2) * (phi->nz - 2);
    const int nb = (int)blocks_for(total, 256, 1024);
    double* part = ctx->acc + 64;
    const double ihx2 = 1. / (phi->hx * phi->hx), ihy2 = 1. / (phi->hy * phi->hy),
                 ihz2 = 1. / (phi->hz * phi->hz);
    if (phi->dtype == TK_F64)
        hipLaunchKernelGGL(k_residual<double>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const double*)phi->d, (const double*)rhs->d, phi->nx, phi->ny, phi->nz,
                           ihx2, ihy2, ihz2, part);
    else
        hipLaunchKernelGGL(k_residual<float>, dim3(nb), dim3(256), 0, ctx->stream,
                           (const float*)phi->d, (const float*)rhs->d, phi->nx, phi->ny, phi->nz,
                           ihx2, ihy2, ihz2, part);
    TK_LAUNCHED();
    hipLaunchKernelGGL(k_sum_pairs_gs, dim3(1), dim3(1024), 0, ctx->stream, part, nb, ctx->acc);
    TK_LAUNCHED();
    TK_HIP(hipMemcpyAsync(ctx->h_acc, ctx->acc, 2 * sizeof(double), hipMemcpyDeviceToHost,
                          ctx->stream));
    TK_HIP(hipStreamSynchronize(ctx->stream));
    *rel = std::sqrt(ctx->h_acc[0]) / std::sqrt(ctx->h_acc[1]);
    return TK_OK;
}

}  // extern "C"
