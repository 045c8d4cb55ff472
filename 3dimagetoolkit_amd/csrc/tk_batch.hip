// tk_batch.hip -- many independent Split-Bregman runs over volumes of uneven sizes.
//
// Each volume keeps its own filter spectrum, state, red-black ping-pong buffer and
// tolerance slot (the reference's filter state is static and class-wide,
// convolution.hxx:307-314, so it could not serve two sizes at once).  The volumes are spread
// over up to four lanes; a lane is a context of its own (one HIP stream, its own scratch and
// reduction slots), so lanes run concurrently and share nothing.  Volumes are assigned to
// lanes by longest-processing-time on their voxel counts.
//
// With a fixed sweep count (ls_tol == 0) and no early stop (tol == 0) an outer iteration
// enqueues kernels only, so:
//  - eager mode enqueues iteration after iteration of every volume without a host wait;
//  - graph mode captures two outer iterations of all volumes (forked over the lanes and
//    joined back) into one hipGraph and replays it.  Two, because an odd sweep count swaps
//    phi with its ping-pong partner every iteration: after two the buffers are back where
//    the capture found them.  After an odd number of eager iterations the graph is captured
//    again.  The graph holds kernel nodes only (no memset nodes).
// Otherwise (ls_tol > 0 gates the Poisson loop on the host, tol > 0 stops a run early) each
// volume runs the reference's loop with its host checks, lane after lane.
//
// Red-black sweeps with a fixed count run staged, one launch per stage over all volumes,
// on lane 0: each convolution pass of every volume (k_conv_axis_batch, one launch per pass
// and filter variant), each sweep (k_gs_rb_pv_batch), the Split-Bregman update with the
// tolerance sums (k_sb_update_batch, k_sum_pairs_batch) and div (k_div_batch): about 13
// launches per outer iteration whatever the number of volumes.  The job tables are built
// once, for both roles of phi and its red-black partner (an odd sweep count swaps the roles
// each iteration).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "tk_internal.hpp"

using namespace tk;

struct tk_batch_s {
    tk_ctx ctx = nullptr;
    tk_dtype dtype = TK_F32;
    tk_rsfe_params p{};
    std::vector<int> dims;
    std::vector<tk_rsfe> runs;
    std::vector<int> lane;       // lane of each volume
    std::vector<tk_ctx> lanes;   // lane contexts
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_nodes = 0;
    int graph_parity = 0;  // iteration parity at capture (phi and its partner swap roles)
    // staged path: job tables phi -> partner (fwd) and partner -> phi (bwd) as built, and
    // whether phi and its partner have swapped roles since
    void* rb_fwd = nullptr;
    void* rb_bwd = nullptr;
    int rb_blocks = 0;
    bool rb_swapped = false;
    RsfeBatchConv* conv = nullptr;
    RsfeBatchUpdate* upd = nullptr;
};

namespace {

void drop_graph(tk_batch b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void drop_tables(tk_batch b)
{
    if (b->rb_fwd) (void)hipFree(b->rb_fwd);
    if (b->rb_bwd) (void)hipFree(b->rb_bwd);
    b->rb_fwd = b->rb_bwd = nullptr;
    b->rb_blocks = 0;
    b->rb_swapped = false;
    if (b->conv) rsfe_batch_conv_free(b->conv);
    if (b->upd) rsfe_batch_update_free(b->upd);
    b->conv = nullptr;
    b->upd = nullptr;
}

// the staged path applies: red-black, a fixed sweep count, no host gate, 16-byte rows
bool staged(tk_batch b)
{
    const tk_rsfe_params& p = b->p;
    if (p.gs_order != TK_GS_REDBLACK || p.ls_tol > 0 || p.tol > 0 || p.neumann ||
        p.warm_start > 0 || p.ls_res_tol > 0)
        return false;
    // the faces keep the boundary value from the first sweep on (cut_phi leaves it)
    const double bc = p.bc_dir ? p.a0 : 0.;
    if (!(bc >= p.a0 && bc <= p.b0)) return false;
    const int V = b->dtype == TK_F64 ? 2 : 4;
    for (size_t v = 0; v < b->runs.size(); ++v)
        if (b->dims[3 * v + 2] % V != 0) return false;
    return true;
}

// volumes in an order that alternates lanes, so consecutive launches go to different streams
std::vector<int> lane_order(tk_batch b)
{
    std::vector<std::vector<int>> per(b->lanes.size());
    for (size_t v = 0; v < b->runs.size(); ++v) per[b->lane[v]].push_back((int)v);
    std::vector<int> order;
    for (size_t r = 0;; ++r) {
        bool any = false;
        for (auto& l : per)
            if (r < l.size()) {
                order.push_back(l[r]);
                any = true;
            }
        if (!any) break;
    }
    return order;
}

int join_into_lane0(tk_batch b)
{
    for (size_t k = 1; k < b->lanes.size(); ++k) {
        TK_HIP(hipEventRecord(b->join[k], b->lanes[k]->stream));
        TK_HIP(hipStreamWaitEvent(b->lanes[0]->stream, b->join[k], 0));
    }
    return TK_OK;
}

// the staged path's job tables, built for the buffers' present roles: allocation and a
// host copy, so never inside a capture (tk_batch_iterate builds them first).  Each run's
// faces get the boundary value here if they have not yet.
int ensure_tables(tk_batch b)
{
    if (b->rb_fwd) return TK_OK;
    std::vector<RbBatchVol> vols(b->runs.size());
    for (size_t v = 0; v < b->runs.size(); ++v) {
        RbBatchVol& q = vols[v];
        void *phi, *alt;
        const void* rhs;
        TK_TRY(rsfe_rb_prepare(b->runs[v], &phi, &alt, &rhs));
        q.src = phi;
        q.dst = alt;
        q.rhs = rhs;
        q.X = b->dims[3 * v];
        q.Y = b->dims[3 * v + 1];
        q.Z = b->dims[3 * v + 2];
    }
    std::vector<RbBatchVol> back(vols);
    for (RbBatchVol& q : back) {
        void* from = const_cast<void*>(q.src);
        q.src = q.dst;
        q.dst = from;
    }
    for (tk_ctx c : b->lanes) TK_HIP(hipStreamSynchronize(c->stream));
    int n1 = 0, n2 = 0;
    TK_TRY(rb_batch_table(b->lanes[0], b->dtype, vols.data(), (int)vols.size(), &b->rb_fwd, &n1));
    TK_TRY(rb_batch_table(b->lanes[0], b->dtype, back.data(), (int)back.size(), &b->rb_bwd, &n2));
    b->rb_blocks = n1;
    b->rb_swapped = false;
    TK_TRY(rsfe_batch_conv_build(b->runs.data(), (int)b->runs.size(), &b->conv));
    TK_TRY(rsfe_batch_update_build(b->runs.data(), (int)b->runs.size(), &b->upd));
    return TK_OK;
}

// one outer iteration of every volume, staged (see the file comment); the tables exist
int staged_step(tk_batch b)
{
    const tk_rsfe_params& p = b->p;
    const int S = p.ls_steps < 1 ? 1 : p.ls_steps;
    // the previous iteration's updates, on every lane, come first
    TK_TRY(join_into_lane0(b));
    TK_TRY(rsfe_batch_conv_run(b->lanes[0], b->conv, b->rb_swapped ? 1 : 0));
    for (int q = 0; q < S; ++q) {
        const bool fwd = ((q & 1) == 0) != b->rb_swapped;
        TK_TRY(rb_batch_sweep(b->lanes[0], b->dtype, fwd ? b->rb_fwd : b->rb_bwd,
                              (int)b->runs.size(), b->rb_blocks, p.dt, 1., 1., 1.));
    }
    if (S & 1) {
        for (tk_rsfe run : b->runs) rsfe_rb_exchange(run);
        b->rb_swapped = !b->rb_swapped;
    }
    return rsfe_batch_update_run(b->lanes[0], b->upd, b->runs.data(), (int)b->runs.size(),
                                 b->rb_swapped ? 1 : 0);
}

int step_all(tk_batch b, const std::vector<int>& order)
{
    if (staged(b)) {
        for (tk_rsfe run : b->runs)
            if (rsfe_iterations(run) >= rsfe_params(run).maxiter) return TK_OK;
        return staged_step(b);
    }
    for (int v : order) TK_TRY(rsfe_step_async(b->runs[v]));
    return TK_OK;
}

int capture(tk_batch b, const std::vector<int>& order)
{
    drop_graph(b);
    hipStream_t s0 = b->lanes[0]->stream;
    std::vector<int> it0(b->runs.size());
    for (size_t v = 0; v < b->runs.size(); ++v) it0[v] = rsfe_iterations(b->runs[v]);
    b->graph_parity = it0[0] & 1;
    TK_HIP(hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal));
    int r = TK_OK;
    if (hipEventRecord(b->fork, s0) != hipSuccess) r = TK_ERR_HIP;
    for (size_t k = 1; r == TK_OK && k < b->lanes.size(); ++k)
        if (hipStreamWaitEvent(b->lanes[k]->stream, b->fork, 0) != hipSuccess) r = TK_ERR_HIP;
    for (int rep = 0; r == TK_OK && rep < 2; ++rep) r = step_all(b, order);
    for (size_t k = 1; r == TK_OK && k < b->lanes.size(); ++k) {
        if (hipEventRecord(b->join[k], b->lanes[k]->stream) != hipSuccess ||
            hipStreamWaitEvent(s0, b->join[k], 0) != hipSuccess)
            r = TK_ERR_HIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(s0, &g);
    // the host side ran two iterations' bookkeeping; the device has run none
    for (size_t v = 0; v < b->runs.size(); ++v) rsfe_set_iterations(b->runs[v], it0[v]);
    if (r != TK_OK || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_error("tk_batch_iterate: graph capture failed (%s)", hipGetErrorString(e));
        return TK_ERR_HIP;
    }
    b->graph = g;
    size_t nn = 0;
    TK_HIP(hipGraphGetNodes(g, nullptr, &nn));
    b->graph_nodes = (int)nn;
    TK_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
    return TK_OK;
}

int sync_lanes(tk_batch b)
{
    for (tk_ctx c : b->lanes) TK_HIP(hipStreamSynchronize(c->stream));
    return TK_OK;
}

}  // namespace

extern "C" {

int tk_batch_create(tk_ctx ctx, int n_vol, const int* dims3, tk_dtype dtype,
                    const tk_rsfe_params* p, tk_batch* out)
{
    TK_CHECK(ctx && dims3 && p && out && n_vol > 0, TK_ERR_ARG, "tk_batch_create: argument");
    TK_CHECK(dtype == TK_F32 || dtype == TK_F64, TK_ERR_DTYPE, "tk_batch_create: dtype");
    // the batch enqueues whole iterations without host waits (and captures them): the
    // automatic extraction reads its seed test back and dumps write files, per run
    TK_CHECK(!p->auto_extract_conn_comp && p->dumpfrequency == 0 && !p->verbosity,
             TK_ERR_UNSUPPORTED,
             "tk_batch_create: auto_extract_conn_comp, dumpfrequency and verbosity are per-run "
             "options (tk_rsfe_*)");
    for (int v = 0; v < n_vol; ++v)
        TK_CHECK(dims3[3 * v] >= 3 && dims3[3 * v + 1] >= 3 && dims3[3 * v + 2] >= 3,
                 TK_ERR_SHAPE, "tk_batch_create: volume %d: every edge must be >= 3", v);
    tk_batch b = new tk_batch_s();
    b->ctx = ctx;
    b->dtype = dtype;
    b->p = *p;
    b->dims.assign(dims3, dims3 + 3 * (size_t)n_vol);
    b->runs.assign(n_vol, nullptr);
    const int n_lanes = std::min(4, n_vol);
    int r = TK_OK;
    for (int k = 0; k < n_lanes && r == TK_OK; ++k) {
        tk_ctx c = nullptr;
        r = tk_ctx_create(ctx->device, &c);
        if (r == TK_OK) b->lanes.push_back(c);
    }
    if (r == TK_OK && hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        r = TK_ERR_HIP;
    b->join.assign(b->lanes.size(), nullptr);
    for (size_t k = 1; r == TK_OK && k < b->lanes.size(); ++k)
        if (hipEventCreateWithFlags(&b->join[k], hipEventDisableTiming) != hipSuccess)
            r = TK_ERR_HIP;
    if (r != TK_OK) {
        tk_batch_destroy(b);
        return r;
    }
    // longest processing time first: biggest volume to the least loaded lane
    std::vector<int> idx(n_vol);
    for (int v = 0; v < n_vol; ++v) idx[v] = v;
    auto vox = [&](int v) { return (long)dims3[3 * v] * dims3[3 * v + 1] * dims3[3 * v + 2]; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int c) { return vox(a) > vox(c); });
    std::vector<long> load(b->lanes.size(), 0);
    b->lane.assign(n_vol, 0);
    for (int v : idx) {
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        b->lane[v] = k;
        load[k] += vox(v);
    }
    *out = b;
    return TK_OK;
}

int tk_batch_set_volume(tk_batch b, int v, const void* u0, const void* init, tk_dtype hd)
{
    TK_CHECK(b && u0 && v >= 0 && v < (int)b->runs.size(), TK_ERR_ARG,
             "tk_batch_set_volume: argument");
    const int* d = &b->dims[3 * (size_t)v];
    tk_ctx c = b->lanes[b->lane[v]];
    tk_vol vu = nullptr, vi = nullptr;
    int r = tk_vol_alloc(c, d[0], d[1], d[2], 1, 1, 1, b->dtype, &vu);
    if (r == TK_OK && init) r = tk_vol_alloc(c, d[0], d[1], d[2], 1, 1, 1, b->dtype, &vi);
    if (r == TK_OK) r = tk_vol_upload(vu, u0, hd);
    if (r == TK_OK && init) r = tk_vol_upload(vi, init, hd);
    if (r == TK_OK) {
        drop_graph(b);
        drop_tables(b);
        if (b->runs[v]) tk_rsfe_destroy(b->runs[v]);
        b->runs[v] = nullptr;
        r = tk_rsfe_create(c, vu, vi, &b->p, &b->runs[v]);
    }
    tk_vol_free(vu);
    tk_vol_free(vi);
    return r;
}

int tk_batch_iterate(tk_batch b, int n_iters, int use_graph)
{
    TK_CHECK(b && n_iters >= 0, TK_ERR_ARG, "tk_batch_iterate: argument");
    for (size_t v = 0; v < b->runs.size(); ++v)
        TK_CHECK(b->runs[v], TK_ERR_ARG, "tk_batch_iterate: volume %zu not set", v);
    if (b->p.ls_tol > 0 || b->p.tol > 0) {
        // host-gated runs: the reference's loop, volume by volume
        for (tk_rsfe run : b->runs) {
            int done = 0;
            TK_TRY(tk_rsfe_iterate(run, n_iters, nullptr, &done));
        }
        return TK_OK;
    }
    // every run advances together up to maxiter
    int it = rsfe_iterations(b->runs[0]);
    for (tk_rsfe run : b->runs)
        TK_CHECK(rsfe_iterations(run) == it, TK_ERR_ARG,
                 "tk_batch_iterate: volumes are at different iterations");
    int n = std::min(n_iters, std::max(0, b->p.maxiter - it));
    const std::vector<int> order = lane_order(b);
    if (staged(b) && n > 0) TK_TRY(ensure_tables(b));
    if (use_graph && n >= 2) {
        // the graph holds the buffers as the capture found them: with an odd sweep count an
        // odd number of iterations since then has swapped them, so capture again
        if (!b->exec || (it & 1) != b->graph_parity) {
            TK_TRY(sync_lanes(b));
            TK_TRY(capture(b, order));
        }
        hipStream_t s0 = b->lanes[0]->stream;
        // the graph's first nodes wait for nothing on the other lanes: drain them first
        TK_TRY(sync_lanes(b));
        for (int q = 0; q + 2 <= n; q += 2) TK_HIP(hipGraphLaunch(b->exec, s0));
        for (tk_rsfe run : b->runs) rsfe_set_iterations(run, it + (n / 2) * 2);
        if (n & 1) {
            TK_HIP(hipStreamSynchronize(s0));
            TK_TRY(step_all(b, order));
        }
    } else {
        for (int q = 0; q < n; ++q) TK_TRY(step_all(b, order));
    }
    TK_TRY(sync_lanes(b));
    for (tk_rsfe run : b->runs) TK_TRY(rsfe_read_tollerance(run, nullptr));
    return TK_OK;
}

int tk_batch_get_phi(tk_batch b, int v, void* host, tk_dtype hd)
{
    TK_CHECK(b && host && v >= 0 && v < (int)b->runs.size() && b->runs[v], TK_ERR_ARG,
             "tk_batch_get_phi: argument");
    return tk_vol_download(tk_rsfe_phi(b->runs[v]), host, hd);
}

int tk_batch_info(tk_batch b, int* n_lanes, int* graph_nodes, int* iterations)
{
    TK_CHECK(b, TK_ERR_ARG, "tk_batch_info: null batch");
    if (n_lanes) *n_lanes = (int)b->lanes.size();
    if (graph_nodes) *graph_nodes = b->graph_nodes;
    if (iterations) *iterations = b->runs.empty() || !b->runs[0] ? 0 : rsfe_iterations(b->runs[0]);
    return TK_OK;
}

int tk_batch_destroy(tk_batch b)
{
    if (!b) return TK_OK;
    for (tk_ctx c : b->lanes) (void)hipStreamSynchronize(c->stream);
    drop_graph(b);
    drop_tables(b);
    for (tk_rsfe r : b->runs)
        if (r) tk_rsfe_destroy(r);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (hipEvent_t e : b->join)
        if (e) (void)hipEventDestroy(e);
    for (tk_ctx c : b->lanes) tk_ctx_destroy(c);
    delete b;
    return TK_OK;
}

}  // extern "C"
