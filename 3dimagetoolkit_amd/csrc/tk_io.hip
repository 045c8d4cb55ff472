// tk_io.hip -- MetaImage (.mhd header + .raw voxels) writer and reader for volumes: the
// format of the reference's dumps, im3d::interface<S>::write(name, ".mhd")
// (interface_imp.hxx:672-690, vtkMetaImageWriter, uncompressed) of the image converted by
// convertfromimage3d (interface_imp.hxx:104-122: VTK_DOUBLE scalars, extent dims, spacing
// h), and its reader, interface<S>(name.mhd) (interface_imp.hxx:33-47, vtkMetaImageReader)
// followed by convert2image3d (interface_imp.hxx:125-150: component 0 cast to S).
//
// The file order is VTK's, x (i) fastest: raw index i + X (j + Y k); volumes here keep z
// (k) fastest, so both directions run a tiled i <-> k transpose on the device (32 x 32 tiles
// through LDS, one j plane per tile row, reads and writes coalesced).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "tk_device.hpp"
#include "tk_internal.hpp"

namespace tk {

constexpr int kTile = 32;

// dst (x fastest, double) <- src (z fastest, T)
template <class T>
__global__ __launch_bounds__(256) void k_to_xfast(const T* __restrict__ src, double* __restrict__ dst,
                                                  int X, int Y, int Z, int tz)
{
    __shared__ double tile[kTile][kTile + 1];
    const int j = blockIdx.x / tz, k0 = (blockIdx.x - j * tz) * kTile, i0 = blockIdx.y * kTile;
    for (int r = threadIdx.y; r < kTile; r += 8) {
        const int i = i0 + r, k = k0 + threadIdx.x;
        if (i < X && k < Z) tile[r][threadIdx.x] = (double)src[((long)i * Y + j) * Z + k];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < kTile; r += 8) {
        const int k = k0 + r, i = i0 + threadIdx.x;
        if (i < X && k < Z) dst[i + (long)X * (j + (long)Y * k)] = tile[threadIdx.x][r];
    }
}

// dst (z fastest, T) <- src (x fastest, double), cast as convert2image3d does
template <class T>
__global__ __launch_bounds__(256) void k_from_xfast(const double* __restrict__ src, T* __restrict__ dst,
                                                    int X, int Y, int Z, int tz)
{
    __shared__ double tile[kTile][kTile + 1];
    const int j = blockIdx.x / tz, k0 = (blockIdx.x - j * tz) * kTile, i0 = blockIdx.y * kTile;
    for (int r = threadIdx.y; r < kTile; r += 8) {
        const int k = k0 + r, i = i0 + threadIdx.x;
        if (i < X && k < Z) tile[threadIdx.x][r] = src[i + (long)X * (j + (long)Y * k)];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < kTile; r += 8) {
        const int i = i0 + r, k = k0 + threadIdx.x;
        if (i < X && k < Z) dst[((long)i * Y + j) * Z + k] = (T)tile[r][threadIdx.x];
    }
}

static dim3 transpose_grid(int X, int Y, int Z, int* tz)
{
    *tz = (Z + kTile - 1) / kTile;
    return dim3((unsigned)(*tz) * (unsigned)Y, (unsigned)((X + kTile - 1) / kTile));
}

namespace {

struct Header {
    int ndims = 0;
    int dims[3] = {1, 1, 1};
    double h[3] = {1, 1, 1};
    bool have_spacing = false;
    std::string type, data_file;
    bool msb = false, compressed = false;
    long header_size = 0;
    int channels = 1;
};

std::string trim(const std::string& s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

std::string lower(std::string s)
{
    for (char& c : s) c = (char)std::tolower((unsigned char)c);
    return s;
}

bool truthy(const std::string& v)
{
    const std::string l = lower(v);
    return l == "true" || l == "1" || l == "yes";
}

// MetaIO element types and sizes
int type_size(const std::string& t)
{
    if (t == "MET_CHAR" || t == "MET_UCHAR") return 1;
    if (t == "MET_SHORT" || t == "MET_USHORT") return 2;
    if (t == "MET_INT" || t == "MET_UINT" || t == "MET_LONG" || t == "MET_ULONG" ||
        t == "MET_FLOAT")
        return 4;
    if (t == "MET_LONG_LONG" || t == "MET_ULONG_LONG" || t == "MET_DOUBLE") return 8;
    return 0;
}

double load_elem(const unsigned char* p, const std::string& t, int size, bool swap)
{
    unsigned char b[8];
    for (int q = 0; q < size; ++q) b[q] = swap ? p[size - 1 - q] : p[q];
    if (t == "MET_CHAR") return (double)(int8_t)b[0];
    if (t == "MET_UCHAR") return (double)b[0];
    if (t == "MET_SHORT") { int16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "MET_USHORT") { uint16_t v; std::memcpy(&v, b, 2); return v; }
    if (t == "MET_INT" || t == "MET_LONG") { int32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "MET_UINT" || t == "MET_ULONG") { uint32_t v; std::memcpy(&v, b, 4); return v; }
    if (t == "MET_FLOAT") { float v; std::memcpy(&v, b, 4); return v; }
    if (t == "MET_LONG_LONG") { int64_t v; std::memcpy(&v, b, 8); return (double)v; }
    if (t == "MET_ULONG_LONG") { uint64_t v; std::memcpy(&v, b, 8); return (double)v; }
    double v;
    std::memcpy(&v, b, 8);
    return v;
}

// parse "Key = Value" lines up to ElementDataFile (the last field of a MetaImage header);
// *data_offset: where the header ends in the file (for ElementDataFile = LOCAL)
int parse_header(FILE* f, Header* hd, long* data_offset)
{
    char line[4096];
    bool done = false;
    while (!done && std::fgets(line, sizeof line, f)) {
        const std::string s(line);
        const size_t eq = s.find('=');
        if (eq == std::string::npos) continue;
        const std::string key = lower(trim(s.substr(0, eq))), val = trim(s.substr(eq + 1));
        if (key == "ndims") {
            hd->ndims = std::atoi(val.c_str());
        } else if (key == "dimsize") {
            std::sscanf(val.c_str(), "%d %d %d", &hd->dims[0], &hd->dims[1], &hd->dims[2]);
        } else if (key == "elementspacing") {
            std::sscanf(val.c_str(), "%lf %lf %lf", &hd->h[0], &hd->h[1], &hd->h[2]);
            hd->have_spacing = true;
        } else if (key == "elementsize" && !hd->have_spacing) {
            std::sscanf(val.c_str(), "%lf %lf %lf", &hd->h[0], &hd->h[1], &hd->h[2]);
        } else if (key == "elementtype") {
            hd->type = val;
        } else if (key == "binarydatabyteordermsb" || key == "elementbyteordermsb") {
            hd->msb = truthy(val);
        } else if (key == "compresseddata") {
            hd->compressed = truthy(val);
        } else if (key == "headersize") {
            hd->header_size = std::atol(val.c_str());
        } else if (key == "elementnumberofchannels") {
            hd->channels = std::atoi(val.c_str());
        } else if (key == "elementdatafile") {
            hd->data_file = val;
            done = true;
        }
    }
    *data_offset = std::ftell(f);
    return done ? TK_OK : TK_ERR_IO;
}

}  // namespace

template <class T>
static int to_xfast(tk_vol v, double* dev)
{
    int tz;
    const dim3 g = transpose_grid(v->nx, v->ny, v->nz, &tz);
    hipLaunchKernelGGL(k_to_xfast<T>, g, dim3(kTile, 8), 0, v->ctx->stream, (const T*)v->d, dev,
                       v->nx, v->ny, v->nz, tz);
    TK_LAUNCHED();
    return TK_OK;
}

template <class T>
static int from_xfast(const double* dev, tk_vol v)
{
    int tz;
    const dim3 g = transpose_grid(v->nx, v->ny, v->nz, &tz);
    hipLaunchKernelGGL(k_from_xfast<T>, g, dim3(kTile, 8), 0, v->ctx->stream, dev, (T*)v->d,
                       v->nx, v->ny, v->nz, tz);
    TK_LAUNCHED();
    return TK_OK;
}

int write_mhd(tk_vol v, const char* base)
{
    const std::string b(base), hdr = b + ".mhd", raw = b + ".raw";
    const size_t slash = raw.find_last_of('/');
    const std::string raw_name = slash == std::string::npos ? raw : raw.substr(slash + 1);
    const long n = v->n();
    double* dev = nullptr;
    TK_CHECK(hipMalloc(&dev, sizeof(double) * (size_t)n) == hipSuccess, TK_ERR_NOMEM,
             "tk_vol_write_mhd: device staging of %ld voxels", n);
    int r = TK_OK;
    if (v->dtype == TK_F64) r = to_xfast<double>(v, dev);
    else if (v->dtype == TK_F32) r = to_xfast<float>(v, dev);
    else r = to_xfast<uint8_t>(v, dev);
    std::vector<double> host((size_t)n);
    if (r == TK_OK &&
        hipMemcpyAsync(host.data(), dev, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost,
                       v->ctx->stream) != hipSuccess)
        r = TK_ERR_HIP;
    if (r == TK_OK && hipStreamSynchronize(v->ctx->stream) != hipSuccess) r = TK_ERR_HIP;
    (void)hipFree(dev);
    if (r != TK_OK) {
        set_error("tk_vol_write_mhd: device transpose or copy failed");
        return r;
    }
    FILE* f = std::fopen(hdr.c_str(), "w");
    TK_CHECK(f, TK_ERR_IO, "tk_vol_write_mhd: cannot open %s", hdr.c_str());
    std::fprintf(f,
                 "ObjectType = Image\nNDims = 3\nBinaryData = True\nBinaryDataByteOrderMSB = False\n"
                 "CompressedData = False\nTransformMatrix = 1 0 0 0 1 0 0 0 1\nOffset = 0 0 0\n"
                 "CenterOfRotation = 0 0 0\nAnatomicalOrientation = RAI\n"
                 "ElementSpacing = %.17g %.17g %.17g\nDimSize = %d %d %d\n"
                 "ElementType = MET_DOUBLE\nElementDataFile = %s\n",
                 v->hx, v->hy, v->hz, v->nx, v->ny, v->nz, raw_name.c_str());
    const bool ok_h = std::fclose(f) == 0;
    TK_CHECK(ok_h, TK_ERR_IO, "tk_vol_write_mhd: writing %s failed", hdr.c_str());
    f = std::fopen(raw.c_str(), "wb");
    TK_CHECK(f, TK_ERR_IO, "tk_vol_write_mhd: cannot open %s", raw.c_str());
    const size_t w = std::fwrite(host.data(), sizeof(double), (size_t)n, f);
    const bool ok_r = std::fclose(f) == 0 && w == (size_t)n;
    TK_CHECK(ok_r, TK_ERR_IO, "tk_vol_write_mhd: writing %s failed", raw.c_str());
    return TK_OK;
}

int read_mhd(tk_ctx ctx, const char* path, tk_dtype dtype, tk_vol* out)
{
    FILE* f = std::fopen(path, "rb");
    TK_CHECK(f, TK_ERR_IO, "tk_vol_read_mhd: cannot open %s", path);
    Header hd;
    long off = 0;
    const int pr = parse_header(f, &hd, &off);
    if (pr != TK_OK) {
        std::fclose(f);
        set_error("tk_vol_read_mhd: %s has no ElementDataFile field", path);
        return TK_ERR_IO;
    }
    const int es = type_size(hd.type);
    bool ok = (hd.ndims == 2 || hd.ndims == 3) && es > 0 && !hd.compressed && hd.channels >= 1;
    if (hd.ndims == 2) hd.dims[2] = 1;
    ok = ok && hd.dims[0] > 0 && hd.dims[1] > 0 && hd.dims[2] > 0;
    if (!ok) {
        std::fclose(f);
        set_error("tk_vol_read_mhd: %s: unsupported header (NDims %d, ElementType '%s', "
                  "compressed %d)", path, hd.ndims, hd.type.c_str(), (int)hd.compressed);
        return TK_ERR_UNSUPPORTED;
    }
    const long n = (long)hd.dims[0] * hd.dims[1] * hd.dims[2];
    const size_t bytes = (size_t)n * es * hd.channels;
    FILE* df = f;
    long start = off;
    if (hd.data_file != "LOCAL") {
        std::fclose(f);
        std::string dp = hd.data_file;
        const std::string p(path);
        const size_t slash = p.find_last_of('/');
        if (dp.empty() || dp[0] != '/') dp = (slash == std::string::npos ? "" : p.substr(0, slash + 1)) + dp;
        df = std::fopen(dp.c_str(), "rb");
        TK_CHECK(df, TK_ERR_IO, "tk_vol_read_mhd: cannot open data file %s", dp.c_str());
        start = 0;
    }
    if (hd.header_size > 0) start += hd.header_size;
    if (hd.header_size == -1) {
        std::fseek(df, 0, SEEK_END);
        start = std::ftell(df) - (long)bytes;
    }
    std::vector<unsigned char> raw(bytes);
    const bool rd = std::fseek(df, start, SEEK_SET) == 0 && std::fread(raw.data(), 1, bytes, df) == bytes;
    std::fclose(df);
    TK_CHECK(rd, TK_ERR_IO, "tk_vol_read_mhd: %s: short data (%zu bytes expected)", path, bytes);
    // component 0 of each voxel, as double (convert2image3d reads it with
    // GetScalarComponentAsDouble); the file's byte order
    const uint16_t probe = 1;
    const bool little = *(const unsigned char*)&probe == 1;
    const bool swap = hd.msb == little;
    std::vector<double> host((size_t)n);
    for (long q = 0; q < n; ++q)
        host[(size_t)q] = load_elem(&raw[(size_t)q * es * hd.channels], hd.type, es, swap);
    tk_vol v = nullptr;
    TK_TRY(tk_vol_alloc(ctx, hd.dims[0], hd.dims[1], hd.dims[2], hd.h[0], hd.h[1], hd.h[2], dtype, &v));
    double* dev = nullptr;
    int r = hipMalloc(&dev, sizeof(double) * (size_t)n) == hipSuccess ? TK_OK : TK_ERR_NOMEM;
    if (r == TK_OK &&
        hipMemcpyAsync(dev, host.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice,
                       ctx->stream) != hipSuccess)
        r = TK_ERR_HIP;
    if (r == TK_OK) {
        if (dtype == TK_F64) r = from_xfast<double>(dev, v);
        else if (dtype == TK_F32) r = from_xfast<float>(dev, v);
        else r = from_xfast<uint8_t>(dev, v);
    }
    if (r == TK_OK && hipStreamSynchronize(ctx->stream) != hipSuccess) r = TK_ERR_HIP;
    if (dev) (void)hipFree(dev);
    if (r != TK_OK) {
        tk_vol_free(v);
        set_error("tk_vol_read_mhd: device staging failed for %s", path);
        return r;
    }
    *out = v;
    return TK_OK;
}

}  // namespace tk

extern "C" {

int tk_vol_write_mhd(tk_vol v, const char* basename)
{
    TK_CHECK(v && v->d && basename && basename[0], TK_ERR_ARG, "tk_vol_write_mhd: argument");
    return tk::write_mhd(v, basename);
}

int tk_vol_read_mhd(tk_ctx ctx, const char* path, tk_dtype dtype, tk_vol* out)
{
    TK_CHECK(ctx && path && out, TK_ERR_ARG, "tk_vol_read_mhd: argument");
    TK_CHECK(dtype == TK_F32 || dtype == TK_F64 || dtype == TK_U8, TK_ERR_DTYPE,
             "tk_vol_read_mhd: dtype");
    return tk::read_mhd(ctx, path, dtype, out);
}

}  // extern "C"
