// tk_internal.hpp -- shared definitions of libtk.so (not installed).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/tk.h"

namespace tk {

void set_error(const char* fmt, ...);

#define TK_HIP(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::tk::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,                \
                            hipGetErrorString(e_));                                     \
            return TK_ERR_HIP;                                                          \
        }                                                                               \
    } while (0)

#define TK_CHECK(cond, code, ...)                                                       \
    do {                                                                                \
        if (!(cond)) {                                                                  \
            ::tk::set_error(__VA_ARGS__);                                               \
            return (code);                                                              \
        }                                                                               \
    } while (0)

#define TK_TRY(expr)                                                                    \
    do {                                                                                \
        int r_ = (expr);                                                                \
        if (r_ != TK_OK) return r_;                                                     \
    } while (0)

// after a kernel launch
#define TK_LAUNCHED() TK_HIP(hipGetLastError())

inline size_t dtype_size(tk_dtype t) { return t == TK_F64 ? 8 : (t == TK_F32 ? 4 : 1); }

// Device scratch owned by a context: FP64 accumulators, counters and flags used by
// the reductions and the device-side Poisson stop test.  Sizes are fixed.
constexpr int kAccSlots = 163840;  // doubles
constexpr int kCntSlots = 4096;    // unsigned
// layout of ctx->acc: [0, 64) results; [kAccPart, +2*kMaxPart) per-block partial pairs;
// [kAccErr, +512) per-sweep relative errors; [kAccSweep, +4096) per-sweep sum pairs;
// [kAccGate, +4*2*kRbGateBlk) per-block sum pairs of the red-black sweeps' stop test, four
// sweeps deep (a launch reads those of the one or two sweeps before it and writes its own)
constexpr int kAccPart = 64;
constexpr int kMaxPart = 8192;
constexpr int kAccErr = kAccPart + 2 * kMaxPart;
constexpr int kAccSweep = kAccErr + 512;
constexpr int kAccGate = kAccSweep + 4096;
constexpr int kRbGateBlk = 16384;
static_assert(kAccGate + 4 * 2 * kRbGateBlk <= kAccSlots, "acc layout");

}  // namespace tk

struct tk_ctx_s {
    int device = 0;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    double* acc = nullptr;        // device, kAccSlots
    unsigned* cnt = nullptr;      // device, kCntSlots
    double* h_acc = nullptr;      // pinned host mirror, kAccSlots
    int num_cu = 256;
    // cached complex work buffers (bytes); reused across calls of the same size
    void* work[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t work_bytes[6] = {0, 0, 0, 0, 0, 0};
    // per-kernel event timing
    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    struct Pending {
        hipEvent_t a, b;
        int tag;
    };
    std::vector<Pending> pending;
    double tag_ms[TK_K_COUNT] = {0};
    long tag_n[TK_K_COUNT] = {0};
    // red-black sweep tile rows (8 or 16) picked by timing, per volume shape and type
    std::map<long long, int> rb_rows;
    int rb_pair = 1;  // fixed red-black sweep counts run two sweeps per launch where possible
    int sb_fused = 1;  // single-volume runs: update and div in one pass
};

struct tk_vol_s {
    tk_ctx ctx = nullptr;
    int nx = 0, ny = 0, nz = 0;
    double hx = 1, hy = 1, hz = 1;
    tk_dtype dtype = TK_F64;
    void* d = nullptr;
    bool owned = true;
    long n() const { return (long)nx * ny * nz; }
    size_t bytes() const { return (size_t)n() * tk::dtype_size(dtype); }
};

namespace tk {

// the address of tk_noalloc.hip's scratch array on `device` (current), looked up once
int noalloc_setup(int device, double** out);

// get (grow if needed) context work buffer `slot` of at least `bytes`
int ctx_work(tk_ctx ctx, int slot, size_t bytes, void** out);

// batch support: one outer iteration that enqueues kernels only (no host wait; capturable),
// the tollerance of the last iteration read back (waits on the run's stream), the
// iteration counter
int rsfe_step_async(tk_rsfe s);
// the run's red-black buffers (faces set to the boundary value if not yet) and rhs; the
// exchange of phi and its partner after an odd number of sweeps
int rsfe_rb_prepare(tk_rsfe s, void** phi, void** alt, const void** rhs);
void rsfe_rb_exchange(tk_rsfe s);
// the five convolution passes of many runs, one launch per pass and variant; tables for
// both roles of phi and its red-black partner (role 0: as when built), built outside any
// capture
struct RsfeBatchConv;
struct RsfeBatchUpdate;
// the Split-Bregman update, tolerance sums and div of many runs, one launch each (tables for
// both roles of phi, as the convolution tables); then each run's iteration count
int rsfe_batch_update_build(tk_rsfe* runs, int n, RsfeBatchUpdate** out);
int rsfe_batch_update_run(tk_ctx ctx, const RsfeBatchUpdate* t, tk_rsfe* runs, int n, int role);
void rsfe_batch_update_free(RsfeBatchUpdate* t);
int rsfe_batch_conv_build(tk_rsfe* runs, int n, RsfeBatchConv** out);
int rsfe_batch_conv_run(tk_ctx ctx, const RsfeBatchConv* t, int role);
void rsfe_batch_conv_free(RsfeBatchConv* t);
const tk_rsfe_params& rsfe_params(tk_rsfe s);
int rsfe_read_tollerance(tk_rsfe s, double* tol);
int rsfe_iterations(tk_rsfe s);
void rsfe_set_iterations(tk_rsfe s, int iter);

// tk_poisson_solve_residual with the caller-owned ping-pong buffer *alt (nullptr: slot)
int poisson_solve_residual_alt(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc,
                               int order, double res_tol, int check_every, int max_sweeps,
                               void** alt, int* sweeps_out, double* res_out);
// the context and dims a network was created for
tk_ctx unet_ctx(tk_unet n);
void unet_dims(tk_unet n, int d[3]);

// set phi's six faces to bc (the first step of DirGaussSeidel, poisson_imp.hxx:370-390)
int poisson_set_boundary(tk_ctx ctx, tk_vol phi, double bc);

// tk_poisson_dir_solve with a caller-owned red-black ping-pong buffer *alt of phi's size
// (swapped with phi->d when the sweep count is odd); alt == nullptr uses the context's slot
// faces_set: the faces of phi and *alt already hold bc (the fixed-sweep paths skip them)
int poisson_dir_solve_alt(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int ls_steps, double ls_tol, int max_sweeps, void** alt,
                          int* sweeps_out, double* err_out, bool faces_set = false);
// red-black sweeps of many volumes in one launch (fixed sweep count): the job table of one
// sweep direction (src -> dst per volume; device memory, free with hipFree) and the launch
struct RbBatchVol {
    const void* src;
    void* dst;
    const void* rhs;
    int X, Y, Z;
};
int rb_batch_table(tk_ctx ctx, tk_dtype dt, const RbBatchVol* vols, int n, void** table,
                   int* nblocks);
int rb_batch_sweep(tk_ctx ctx, tk_dtype dt, const void* table, int n, int nblocks, double dt_step,
                   double hx, double hy, double hz);
// NeuGaussSeidel sweeps (lexicographic only), the same loop rule
int poisson_neu_solve_alt(tk_ctx ctx, tk_vol phi, tk_vol rhs, double dt, double bc, int order,
                          int ls_steps, double ls_tol, int max_sweeps, int* sweeps_out,
                          double* err_out);

inline bool same_dims(const tk_vol a, const tk_vol b)
{
    return a->nx == b->nx && a->ny == b->ny && a->nz == b->nz;
}

// 1-D grid helper
inline unsigned blocks_for(long n, int threads, long cap = 1 << 20)
{
    long b = (n + threads - 1) / threads;
    if (b < 1) b = 1;
    if (b > cap) b = cap;
    return (unsigned)b;
}

// FP64 sums of two quantities over a volume; results copied to host.
// kinds implemented in tk_core.hip
int reduce_sumsq(tk_ctx ctx, tk_vol a, tk_vol b_or_null, double* sum_d2, double* sum_b2);

// internal stage launchers shared between translation units
int launch_cut(tk_ctx ctx, tk_vol phi, double a0, double b0);

// event timing around one launch: begin returns a slot (or -1 when timing is off)
int timing_begin(tk_ctx ctx, int tag);
void timing_end(tk_ctx ctx, int slot);
int timing_flush(tk_ctx ctx);

struct TimedLaunch {
    tk_ctx ctx;
    int slot;
    TimedLaunch(tk_ctx c, int tag) : ctx(c), slot(timing_begin(c, tag)) {}
    ~TimedLaunch() { timing_end(ctx, slot); }
};

}  // namespace tk
