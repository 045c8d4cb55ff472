// tk_batch.hip -- many independent Split-Bregman runs over volumes of uneven sizes.
//
// Each volume keeps its own filter spectrum and state (the reference's filter state
// is static and class-wide, convolution.hxx:307-314, so it cannot serve two sizes at
// once).  This first version advances the volumes one after another on the context's
// stream; every stage of a volume is already a single launch.
#include <hip/hip_runtime.h>

#include <vector>

#include "tk_internal.hpp"

struct tk_batch_s {
    tk_ctx ctx = nullptr;
    tk_dtype dtype = TK_F32;
    tk_rsfe_params p{};
    std::vector<int> dims;
    std::vector<tk_rsfe> runs;
};

extern "C" {

int tk_batch_create(tk_ctx ctx, int n_vol, const int* dims3, tk_dtype dtype,
                    const tk_rsfe_params* p, tk_batch* out)
{
    TK_CHECK(ctx && dims3 && p && out && n_vol > 0, TK_ERR_ARG, "tk_batch_create: argument");
    TK_CHECK(dtype == TK_F32 || dtype == TK_F64, TK_ERR_DTYPE, "tk_batch_create: dtype");
    tk_batch b = new tk_batch_s();
    b->ctx = ctx;
    b->dtype = dtype;
    b->p = *p;
    b->dims.assign(dims3, dims3 + 3 * (size_t)n_vol);
    b->runs.assign(n_vol, nullptr);
    *out = b;
    return TK_OK;
}

int tk_batch_set_volume(tk_batch b, int v, const void* u0, const void* init, tk_dtype hd)
{
    TK_CHECK(b && u0 && init && v >= 0 && v < (int)b->runs.size(), TK_ERR_ARG,
             "tk_batch_set_volume: argument");
    const int* d = &b->dims[3 * (size_t)v];
    tk_vol vu = nullptr, vi = nullptr;
    int r = tk_vol_alloc(b->ctx, d[0], d[1], d[2], 1, 1, 1, b->dtype, &vu);
    if (r == TK_OK) r = tk_vol_alloc(b->ctx, d[0], d[1], d[2], 1, 1, 1, b->dtype, &vi);
    if (r == TK_OK) r = tk_vol_upload(vu, u0, hd);
    if (r == TK_OK) r = tk_vol_upload(vi, init, hd);
    if (r == TK_OK) {
        if (b->runs[v]) tk_rsfe_destroy(b->runs[v]);
        b->runs[v] = nullptr;
        r = tk_rsfe_create(b->ctx, vu, vi, &b->p, &b->runs[v]);
    }
    tk_vol_free(vu);
    tk_vol_free(vi);
    return r;
}

int tk_batch_iterate(tk_batch b, int n_iters, int use_graph)
{
    TK_CHECK(b, TK_ERR_ARG, "tk_batch_iterate: null batch");
    (void)use_graph;
    for (size_t v = 0; v < b->runs.size(); ++v) {
        TK_CHECK(b->runs[v], TK_ERR_ARG, "tk_batch_iterate: volume %zu not set", v);
        int done = 0;
        TK_TRY(tk_rsfe_iterate(b->runs[v], n_iters, nullptr, &done));
    }
    return TK_OK;
}

int tk_batch_get_phi(tk_batch b, int v, void* host, tk_dtype hd)
{
    TK_CHECK(b && host && v >= 0 && v < (int)b->runs.size() && b->runs[v], TK_ERR_ARG,
             "tk_batch_get_phi: argument");
    return tk_vol_download(tk_rsfe_phi(b->runs[v]), host, hd);
}

int tk_batch_destroy(tk_batch b)
{
    if (!b) return TK_OK;
    for (tk_rsfe r : b->runs)
        if (r) tk_rsfe_destroy(r);
    delete b;
    return TK_OK;
}

}  // extern "C"
