// tk_gen.cpp -- deterministic synthetic volumes (host code).
//
// The reference reads its images from disk through VTK (interface_imp.hxx:33-88);
// there are no datasets here, so every workload is generated from a seed by the
// functions below.  One generator serves the GPU path, the benchmark and the
// tests, so every consumer sees identical inputs for a given seed.
//
// Random numbers: SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom
// number generators", OOPSLA 2014); normals by the Box-Muller transform (cosine
// branch, one normal per two uniforms).  Layout: index (i*ny + j)*nz + k, z fastest.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/tk.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }  // [0,1)
    double normal()
    {
        double u1 = 1.0 - uniform();  // (0,1]
        double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586476925 * u2);
    }
    int uniform_int(int lo, int hi)  // inclusive
    {
        if (hi < lo) return lo;
        return lo + (int)(next() % (uint64_t)(hi - lo + 1));
    }
};

inline long vidx(int i, int j, int k, int ny, int nz) { return ((long)i * ny + j) * nz + k; }

}  // namespace

extern "C" {

int tk_gen_ellipsoids(int nx, int ny, int nz, uint64_t seed, int n_ell, const double* semi_axes,
                      int margin, double sphere_radius, double* img, double* phi0)
{
    if (nx < 3 || ny < 3 || nz < 1 || n_ell < 0 || !img) return TK_ERR_ARG;
    Rng rng(seed);
    std::vector<int> c(3 * (size_t)n_ell);
    const int dims[3] = {nx, ny, nz};
    for (int e = 0; e < n_ell; ++e)
        for (int a = 0; a < 3; ++a) {
            int m = margin;
            if (2 * m > dims[a] - 1) m = (dims[a] - 1) / 2;
            c[3 * e + a] = rng.uniform_int(m, dims[a] - 1 - m);
        }
    const long n = (long)nx * ny * nz;
    for (long v = 0; v < n; ++v) img[v] = 0.2 + 0.05 * rng.normal();
    for (int e = 0; e < n_ell; ++e) {
        const double ax = semi_axes[3 * e], ay = semi_axes[3 * e + 1], az = semi_axes[3 * e + 2];
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    double dx = (i - c[3 * e]) / ax, dy = (j - c[3 * e + 1]) / ay,
                           dz = (k - c[3 * e + 2]) / az;
                    if (dx * dx + dy * dy + dz * dz <= 1.0) img[vidx(i, j, k, ny, nz)] = 0.8;
                }
    }
    for (int i = 0; i < nx; ++i) {
        const double bias = 0.7 + 0.6 * (double)i / (double)(nx - 1);
        for (long v = (long)i * ny * nz; v < (long)(i + 1) * ny * nz; ++v) img[v] *= bias;
    }
    for (long v = 0; v < n; ++v) img[v] += 0.05 * rng.normal();
    if (phi0) {
        const double cx = (nx - 1) / 2.0, cy = (ny - 1) / 2.0, cz = (nz - 1) / 2.0;
        const double r2 = sphere_radius * sphere_radius;
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) {
                    double dx = i - cx, dy = j - cy, dz = k - cz;
                    phi0[vidx(i, j, k, ny, nz)] = (dx * dx + dy * dy + dz * dz <= r2) ? 1.0 : 0.0;
                }
    }
    return TK_OK;
}

int tk_gen_standard(int nx, int ny, int nz, uint64_t seed, int scale, int n_ell, double* img,
                    double* phi0)
{
    // ellipsoid semi-axes cycle through (12,9,7), (8,8,10), scaled; margin 15*scale;
    // phi0 sphere radius 20*scale, clipped to 0.3125 of the smallest edge
    if (scale < 1 || n_ell < 1 || n_ell > 16) return TK_ERR_ARG;
    double ax[48];
    const double base[2][3] = {{12, 9, 7}, {8, 8, 10}};
    for (int e = 0; e < n_ell; ++e)
        for (int a = 0; a < 3; ++a) ax[3 * e + a] = base[e % 2][a] * scale;
    int mn = nx < ny ? nx : ny;
    mn = mn < nz ? mn : nz;
    double r = 20.0 * scale;
    if (r > 0.3125 * mn) r = 0.3125 * mn;
    return tk_gen_ellipsoids(nx, ny, nz, seed, n_ell, ax, 15 * scale, r, img, phi0);
}

int tk_gen_batch_dims(uint64_t seed, int n_vol, int* dims3, int* n_ell)
{
    // edge lengths drawn independently per axis from {48,64,80,96,128,160};
    // 1-4 ellipsoids per volume
    static const int edges[6] = {48, 64, 80, 96, 128, 160};
    if (n_vol < 0 || !dims3) return TK_ERR_ARG;
    Rng rng(seed);
    for (int v = 0; v < n_vol; ++v) {
        for (int a = 0; a < 3; ++a) dims3[3 * v + a] = edges[rng.uniform_int(0, 5)];
        if (n_ell) n_ell[v] = rng.uniform_int(1, 4);
    }
    return TK_OK;
}

int tk_gen_ct(int nx, int ny, int nz, uint64_t seed, int n_vessels, double* img, int* seed_ijk)
{
    // CT-like volume: background -1000+-20, tissue slab 40+-15 over the central 60% of y,
    // tubular vessels 300+-30 of radius U[3,8] along Catmull-Rom splines through 4-6
    // random control points; min-max rescale to [0,1]; bias 0.8 -> 1.2 along z.
    if (nx < 8 || ny < 8 || nz < 8 || !img || n_vessels < 1) return TK_ERR_ARG;
    Rng rng(seed);
    const long n = (long)nx * ny * nz;
    std::vector<uint8_t> cls(n, 0);
    const int y0 = (int)(0.2 * ny), y1 = (int)(0.8 * ny);
    for (int i = 0; i < nx; ++i)
        for (int j = y0; j < y1; ++j)
            std::memset(&cls[vidx(i, j, 0, ny, nz)], 1, nz);
    for (int v = 0; v < n_vessels; ++v) {
        const double rad = 3.0 + 5.0 * rng.uniform();
        const int np = rng.uniform_int(4, 6);
        double P[8][3];
        const int mg = 10;
        for (int p = 0; p < np; ++p) {
            P[p + 1][0] = rng.uniform_int(mg, nx - 1 - mg);
            P[p + 1][1] = rng.uniform_int(mg, ny - 1 - mg);
            P[p + 1][2] = rng.uniform_int(mg, nz - 1 - mg);
        }
        if (v == 0 && seed_ijk) {
            seed_ijk[0] = (int)P[1][0]; seed_ijk[1] = (int)P[1][1]; seed_ijk[2] = (int)P[1][2];
        }
        for (int a = 0; a < 3; ++a) {  // phantom end points for the end segments
            P[0][a] = 2 * P[1][a] - P[2][a];
            P[np + 1][a] = 2 * P[np][a] - P[np - 1][a];
        }
        const int ir = (int)std::ceil(rad);
        for (int s = 1; s < np; ++s) {
            double seg = 0;
            for (int a = 0; a < 3; ++a) seg += (P[s + 1][a] - P[s][a]) * (P[s + 1][a] - P[s][a]);
            int ns = (int)std::ceil(std::sqrt(seg) / 0.5) + 1;
            for (int q = 0; q <= ns; ++q) {
                double t = (double)q / ns, t2 = t * t, t3 = t2 * t;
                double c[3];
                for (int a = 0; a < 3; ++a)
                    c[a] = 0.5 * ((2 * P[s][a]) + (-P[s - 1][a] + P[s + 1][a]) * t +
                                  (2 * P[s - 1][a] - 5 * P[s][a] + 4 * P[s + 1][a] - P[s + 2][a]) * t2 +
                                  (-P[s - 1][a] + 3 * P[s][a] - 3 * P[s + 1][a] + P[s + 2][a]) * t3);
                int ci = (int)std::lround(c[0]), cj = (int)std::lround(c[1]), ck = (int)std::lround(c[2]);
                for (int i = ci - ir; i <= ci + ir; ++i) {
                    if (i < 0 || i >= nx) continue;
                    for (int j = cj - ir; j <= cj + ir; ++j) {
                        if (j < 0 || j >= ny) continue;
                        for (int k = ck - ir; k <= ck + ir; ++k) {
                            if (k < 0 || k >= nz) continue;
                            double dx = i - c[0], dy = j - c[1], dz = k - c[2];
                            if (dx * dx + dy * dy + dz * dz <= rad * rad) cls[vidx(i, j, k, ny, nz)] = 2;
                        }
                    }
                }
            }
        }
    }
    double mn = 1e300, mx = -1e300;
    for (long t = 0; t < n; ++t) {
        double g = rng.normal();
        double val = cls[t] == 0 ? -1000.0 + 20.0 * g : (cls[t] == 1 ? 40.0 + 15.0 * g : 300.0 + 30.0 * g);
        img[t] = val;
        mn = val < mn ? val : mn;
        mx = val > mx ? val : mx;
    }
    const double inv = 1.0 / (mx - mn);
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int k = 0; k < nz; ++k) {
                long t = vidx(i, j, k, ny, nz);
                img[t] = (img[t] - mn) * inv * (0.8 + 0.4 * (double)k / (double)(nz - 1));
            }
    return TK_OK;
}

}  // extern "C"
