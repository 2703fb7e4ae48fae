// mrlbm_batch.hip -- many independent runs of uneven size advanced together.
//
// Each run is an ordinary handle (own state, geometry and boundary tables).  The batch
// keeps the runs' kernel argument blocks in device memory, one array per state parity and
// recording mode, and a flattened tile table (run, bulk block) for the bulk kernel.  A step
// is four launches for all runs: the bulk over the tile table on one stream and, on a
// second stream, the edges, the stencil moments and the boundary workgroups (one per run),
// joined as in the single-run step loop.  Runs of graph_steps steps are captured once per
// (start parity, recording mode) and replayed.  The autotuner times the candidate bulk
// variants and grid modes on a saved copy of the states and keeps the fastest.
#include "mrlbm_handle.h"
#include "mrlbm_lattice.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

namespace {
thread_local std::string g_batch_create_error;
}

struct mrlbm_batch {
    std::vector<mrlbm_handle *> runs;
    bool f32 = false;
    int kernel = 0, persistent = 1, grid = 0, ntiles = 0, graph_steps = 50, tune_steps = 20;
    int max_ring = 0, max_nU = 0;
    int64_t nodes = 0;
    hipStream_t st = nullptr, st2 = nullptr;
    hipEvent_t evA = nullptr, evB = nullptr, e0 = nullptr, e1 = nullptr;
    void *d_args[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}}; // [parity][record]
    int2 *d_tiles = nullptr;
    double *d_facc = nullptr; // 4 per run
    hipGraphExec_t exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    std::vector<int> tune_kernel, tune_persistent;
    std::vector<double> tune_ms;
    std::string err;
};

#define BFAIL(code, msg)           \
    do {                           \
        b->err = (msg);            \
        return (code);             \
    } while (0)
#define BHIP(...)                                                                        \
    do {                                                                                 \
        hipError_t e_ = (__VA_ARGS__);                                                   \
        if (e_ != hipSuccess) BFAIL(MRLBM_E_HIP, std::string(#__VA_ARGS__) + ": " + hipGetErrorString(e_)); \
    } while (0)

static void drop_graphs(mrlbm_batch *b)
{
    for (auto &row : b->exec)
        for (auto &g : row)
            if (g) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
}

// Argument blocks and tile table for the runs' current bulk shape.
template <class S, class T> static int build_tables(mrlbm_batch *b)
{
    drop_graphs(b);
    const int R = (int)b->runs.size();
    std::vector<StepArgs<S, T>> args[2][2];
    std::vector<int2> tiles;
    b->max_ring = 0;
    b->max_nU = 0;
    for (int r = 0; r < R; r++) {
        mrlbm_handle *h = b->runs[r];
        for (int par = 0; par < 2; par++)
            for (int rec = 0; rec < 2; rec++) {
                auto a = base_args<S, T>(h);
                a.src = (const S *)h->S[par];
                a.dst = (S *)h->S[par ^ 1];
                a.pops = nullptr;
                a.rec_slot = rec ? 0 : -1;
                a.facc = rec ? b->d_facc + 4 * r : nullptr;
                args[par][rec].push_back(a);
            }
        const int nb = h->ncb * h->nstrips;
        for (int t = 0; t < nb; t++) tiles.push_back(make_int2(r, t));
        b->max_ring = std::max(b->max_ring, 2 * h->nx + 2 * (h->ny - 2));
        if (h->has_b) b->max_nU = std::max(b->max_nU, h->bt.nU);
    }
    for (int par = 0; par < 2; par++)
        for (int rec = 0; rec < 2; rec++) {
            if (!b->d_args[par][rec]) BHIP(hipMalloc(&b->d_args[par][rec], (size_t)R * sizeof(StepArgs<S, T>)));
            BHIP(hipMemcpy(b->d_args[par][rec], args[par][rec].data(), (size_t)R * sizeof(StepArgs<S, T>),
                           hipMemcpyHostToDevice));
        }
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    b->d_tiles = nullptr;
    BHIP(hipMalloc(&b->d_tiles, tiles.size() * sizeof(int2)));
    BHIP(hipMemcpy(b->d_tiles, tiles.data(), tiles.size() * sizeof(int2), hipMemcpyHostToDevice));
    b->ntiles = (int)tiles.size();
    const int resident = bulk_batch_resident_blocks<S, T>(b->runs[0]->variant);
    b->grid = (b->persistent && resident > 0) ? std::min(resident, b->ntiles) : b->ntiles;
    return 0;
}

// One step of every run, parity par (runs read S[par]): the single-run stream pattern.
template <class S, class T> static int launch_step(mrlbm_batch *b, int par, int rec)
{
    const auto *args = (const StepArgs<S, T> *)b->d_args[par][rec];
    BHIP(hipStreamWaitEvent(b->st, b->evB, 0));
    BHIP(launch_bulk_batch<S, T>(args, b->d_tiles, b->ntiles, b->runs[0]->variant, b->grid, b->st));
    BHIP(hipStreamWaitEvent(b->st2, b->evA, 0));
    BHIP(launch_side_batch<S, T>(args, (int)b->runs.size(), b->max_ring, b->max_nU, b->st2));
    BHIP(hipEventRecord(b->evB, b->st2));
    BHIP(hipEventRecord(b->evA, b->st));
    return 0;
}

template <class S, class T> static int capture(mrlbm_batch *b, int par, int rec)
{
    hipGraph_t g = nullptr;
    BHIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    // bring the side stream into the capture, and seed the two events inside it
    hipError_t e = hipEventRecord(b->evA, b->st);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->st2, b->evA, 0);
    if (e == hipSuccess) e = hipEventRecord(b->evB, b->st2);
    int rc = e == hipSuccess ? 0 : MRLBM_E_HIP;
    for (int s = 0; s < b->graph_steps && !rc; s++) rc = launch_step<S, T>(b, par ^ (s & 1), rec);
    if (!rc && hipStreamWaitEvent(b->st, b->evB, 0) != hipSuccess) rc = MRLBM_E_HIP;
    const hipError_t ee = hipStreamEndCapture(b->st, &g);
    if (rc || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        BFAIL(MRLBM_E_HIP, "graph capture of the batched step failed");
    }
    e = hipGraphInstantiate(&b->exec[par][rec], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BHIP(e);
    return 0;
}

template <class S, class T> static int step_typed(mrlbm_batch *b, int nsteps, int rec)
{
    const int par0 = b->runs[0]->cur;
    for (mrlbm_handle *h : b->runs) {
        if (h->cur != par0) BFAIL(MRLBM_E_STATE, "a run was stepped outside the batch");
        if (h->from_pops) BFAIL(MRLBM_E_STATE, "a run holds populations from mrlbm_set_populations: step it once alone first");
        if (h->chi) BFAIL(MRLBM_E_ARG, "batched runs take no Brinkman field");
        BHIP(hipStreamSynchronize(h->st)); // state uploads on the run's own stream
    }
    BHIP(hipEventRecord(b->evA, b->st));
    BHIP(hipEventRecord(b->evB, b->st));
    int par = par0, s = 0;
    while (s < nsteps) {
        if (nsteps - s >= b->graph_steps) {
            if (!b->exec[par][rec]) {
                const int rc = capture<S, T>(b, par, rec);
                if (rc) return rc;
            }
            BHIP(hipGraphLaunch(b->exec[par][rec], b->st));
            BHIP(hipEventRecord(b->evA, b->st));
            BHIP(hipEventRecord(b->evB, b->st));
            s += b->graph_steps; // graph_steps is even: the parity is unchanged
        } else {
            const int rc = launch_step<S, T>(b, par, rec);
            if (rc) return rc;
            par ^= 1;
            s++;
        }
    }
    BHIP(hipStreamWaitEvent(b->st, b->evB, 0));
    BHIP(hipStreamSynchronize(b->st));
    for (mrlbm_handle *h : b->runs) h->cur = par;
    return 0;
}

// Times every candidate on a saved copy of the states, keeps the fastest.
template <class S, class T> static int autotune_typed(mrlbm_batch *b, int fixed_kernel, int fixed_persistent)
{
    std::vector<int> kernels;
    if (fixed_kernel) kernels = {fixed_kernel};
    else if (b->f32) kernels = {304, 308, 8, 16, 4};
    else kernels = {8, 304, 16, 4, 308};
    std::vector<int> modes;
    if (fixed_persistent >= 0) modes = {fixed_persistent};
    else modes = {1, 0};
    const bool tune = kernels.size() * modes.size() > 1;
    std::vector<void *> saved;
    const int par0 = b->runs[0]->cur;
    if (tune)
        for (mrlbm_handle *h : b->runs) {
            const size_t bytes = (size_t)6 * h->plane * sizeof(S);
            void *p = nullptr;
            if (hipMalloc(&p, bytes) != hipSuccess) {
                for (void *q : saved) (void)hipFree(q);
                BFAIL(MRLBM_E_NOMEM, "no device memory for the autotuner's state copy");
            }
            saved.push_back(p);
            BHIP(hipMemcpy(p, h->S[h->cur], bytes, hipMemcpyDeviceToDevice));
        }
    b->tune_kernel.clear();
    b->tune_persistent.clear();
    b->tune_ms.clear();
    BHIP(hipEventRecord(b->evA, b->st));
    BHIP(hipEventRecord(b->evB, b->st));
    int best_k = kernels[0], best_p = modes[0];
    double best = 1e300;
    int rc = 0;
    for (int k : kernels) {
        for (int pm : modes) {
            std::string why;
            for (mrlbm_handle *h : b->runs)
                if (!rc) rc = set_bulk_shape(h, k, 0, why);
            if (rc) break;
            b->persistent = pm;
            rc = build_tables<S, T>(b);
            if (rc || !tune) break;
            const int par = b->runs[0]->cur;
            for (int s = 0; s < 2 && !rc; s++) rc = launch_step<S, T>(b, par ^ (s & 1), 0);
            if (!rc && hipEventRecord(b->e0, b->st) != hipSuccess) rc = MRLBM_E_HIP;
            for (int s = 0; s < b->tune_steps && !rc; s++) rc = launch_step<S, T>(b, par ^ (s & 1), 0);
            if (!rc && hipStreamWaitEvent(b->st, b->evB, 0) != hipSuccess) rc = MRLBM_E_HIP;
            if (!rc && hipEventRecord(b->e1, b->st) != hipSuccess) rc = MRLBM_E_HIP;
            if (!rc && hipEventSynchronize(b->e1) != hipSuccess) rc = MRLBM_E_HIP;
            float ms = 0;
            if (!rc && hipEventElapsedTime(&ms, b->e0, b->e1) != hipSuccess) rc = MRLBM_E_HIP;
            if (rc) break;
            const double per = ms / b->tune_steps;
            b->tune_kernel.push_back(k);
            b->tune_persistent.push_back(pm);
            b->tune_ms.push_back(per);
            if (per < best) {
                best = per;
                best_k = k;
                best_p = pm;
            }
            // restore the states and parity for the next candidate
            for (size_t r = 0; r < b->runs.size(); r++) {
                mrlbm_handle *h = b->runs[r];
                h->cur = par0;
                if (hipMemcpy(h->S[par0], saved[r], (size_t)6 * h->plane * sizeof(S), hipMemcpyDeviceToDevice) != hipSuccess)
                    rc = MRLBM_E_HIP;
            }
        }
        if (rc || !tune) break;
    }
    for (void *q : saved) (void)hipFree(q);
    if (rc) return rc;
    b->kernel = best_k;
    b->persistent = best_p;
    std::string why;
    for (mrlbm_handle *h : b->runs) {
        const int r2 = set_bulk_shape(h, best_k, 0, why);
        if (r2) BFAIL(r2, why);
    }
    return build_tables<S, T>(b);
}

static void batch_free(mrlbm_batch *b)
{
    drop_graphs(b);
    for (auto &row : b->d_args)
        for (void *p : row)
            if (p) (void)hipFree(p);
    if (b->d_tiles) (void)hipFree(b->d_tiles);
    if (b->d_facc) (void)hipFree(b->d_facc);
    for (hipEvent_t e : {b->evA, b->evB, b->e0, b->e1})
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t s : {b->st, b->st2})
        if (s) (void)hipStreamDestroy(s);
    for (mrlbm_handle *h : b->runs) mrlbm_destroy(h);
    b->runs.clear();
}

extern "C" void mrlbm_default_batch_opts(mrlbm_batch_opts *o)
{
    std::memset(o, 0, sizeof *o);
    o->kernel = 0;
    o->persistent = -1;
    o->graph_steps = 50;
    o->tune_steps = 20;
}

extern "C" int mrlbm_batch_create(const mrlbm_params *runs, int32_t nruns, const mrlbm_batch_opts *o,
                                  mrlbm_batch **out)
{
    g_batch_create_error.clear();
    if (!runs || nruns < 1 || !out) {
        g_batch_create_error = "need runs, nruns >= 1 and out";
        return MRLBM_E_ARG;
    }
    *out = nullptr;
    mrlbm_batch_opts opts;
    if (o) opts = *o;
    else mrlbm_default_batch_opts(&opts);
    if (opts.graph_steps <= 0) opts.graph_steps = 50;
    if (opts.tune_steps <= 0) opts.tune_steps = 20;
    auto *b = new mrlbm_batch();
    auto fail = [&](int code, const std::string &m) {
        g_batch_create_error = m.empty() ? b->err : m;
        batch_free(b);
        delete b;
        return code;
    };
    if (opts.graph_steps % 2) return fail(MRLBM_E_ARG, "graph_steps must be even");
    if (opts.kernel == 1) return fail(MRLBM_E_ARG, "the batched bulk has no one-node-per-thread kernel");
    b->graph_steps = opts.graph_steps;
    b->tune_steps = opts.tune_steps;
    b->f32 = runs[0].precision == MRLBM_FP32_DEV;
    for (int r = 0; r < nruns; r++) {
        if ((runs[r].precision == MRLBM_FP32_DEV) != b->f32) return fail(MRLBM_E_ARG, "every run needs the same precision");
        mrlbm_handle *h = nullptr;
        mrlbm_params pr = runs[r];
        pr.time_blocking = 0; // the batch steps its runs with its own launches
        const int rc = mrlbm_create(&pr, &h);
        if (rc) return fail(rc, "run " + std::to_string(r) + ": " + mrlbm_last_error(nullptr));
        b->runs.push_back(h);
        b->nodes += (int64_t)h->nx * h->ny;
        if (!h->d_rec) { // force reduction target of the boundary workgroup
            h->rec_stride = 14 + 2 * h->bt.nbct;
            h->rec_cap = 256;
            if (hipMalloc(&h->d_rec, (size_t)h->rec_cap * h->rec_stride * sizeof(double)) != hipSuccess)
                return fail(MRLBM_E_NOMEM, "force record buffer");
        }
    }
    if (hipStreamCreate(&b->st) != hipSuccess || hipStreamCreate(&b->st2) != hipSuccess ||
        hipEventCreateWithFlags(&b->evA, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->evB, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess)
        return fail(MRLBM_E_HIP, "stream or event creation failed");
    if (hipMalloc(&b->d_facc, (size_t)4 * nruns * sizeof(double)) != hipSuccess ||
        hipMemset(b->d_facc, 0, (size_t)4 * nruns * sizeof(double)) != hipSuccess)
        return fail(MRLBM_E_HIP, "force sums");
    const int rc = b->f32 ? autotune_typed<float, float>(b, opts.kernel, opts.persistent)
                          : autotune_typed<double, double>(b, opts.kernel, opts.persistent);
    if (rc) return fail(rc, "");
    *out = b;
    return MRLBM_OK;
}

extern "C" int mrlbm_batch_destroy(mrlbm_batch *b)
{
    if (!b) return MRLBM_E_ARG;
    batch_free(b);
    delete b;
    return MRLBM_OK;
}

extern "C" const char *mrlbm_batch_last_error(const mrlbm_batch *b)
{
    return b ? b->err.c_str() : g_batch_create_error.c_str();
}

extern "C" mrlbm_handle *mrlbm_batch_run(mrlbm_batch *b, int32_t r)
{
    if (!b || r < 0 || r >= (int)b->runs.size()) return nullptr;
    return b->runs[r];
}

extern "C" int mrlbm_batch_autotune(mrlbm_batch *b)
{
    if (!b) return MRLBM_E_ARG;
    return b->f32 ? autotune_typed<float, float>(b, 0, -1) : autotune_typed<double, double>(b, 0, -1);
}

extern "C" int mrlbm_batch_step(mrlbm_batch *b, int32_t nsteps, int32_t record_forces)
{
    if (!b || nsteps < 0) return MRLBM_E_ARG;
    const int rec = record_forces ? 1 : 0;
    return b->f32 ? step_typed<float, float>(b, nsteps, rec) : step_typed<double, double>(b, nsteps, rec);
}

extern "C" int mrlbm_batch_forces(mrlbm_batch *b, double *fd_sum, double *fl_sum, double *fl2_sum, int64_t *count)
{
    if (!b) return MRLBM_E_ARG;
    const size_t R = b->runs.size();
    std::vector<double> acc(4 * R);
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemcpy(acc.data(), b->d_facc, acc.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < R; r++) {
        if (fd_sum) fd_sum[r] = acc[4 * r];
        if (fl_sum) fl_sum[r] = acc[4 * r + 1];
        if (fl2_sum) fl2_sum[r] = acc[4 * r + 2];
        if (count) count[r] = (int64_t)acc[4 * r + 3];
    }
    return MRLBM_OK;
}

extern "C" int mrlbm_batch_reset_forces(mrlbm_batch *b)
{
    if (!b) return MRLBM_E_ARG;
    BHIP(hipStreamSynchronize(b->st));
    BHIP(hipMemset(b->d_facc, 0, 4 * b->runs.size() * sizeof(double)));
    return MRLBM_OK;
}

extern "C" int mrlbm_batch_get_info(const mrlbm_batch *b, mrlbm_batch_info *info)
{
    if (!b || !info) return MRLBM_E_ARG;
    std::memset(info, 0, sizeof *info);
    info->nruns = (int32_t)b->runs.size();
    info->kernel = b->kernel;
    info->persistent = b->persistent;
    info->grid = b->grid;
    info->ntiles = b->ntiles;
    info->graph_steps = b->graph_steps;
    info->nodes = b->nodes;
    info->ntune = (int32_t)std::min<size_t>(16, b->tune_ms.size());
    for (int i = 0; i < info->ntune; i++) {
        info->tune_kernel[i] = b->tune_kernel[i];
        info->tune_persistent[i] = b->tune_persistent[i];
        info->tune_ms[i] = b->tune_ms[i];
    }
    return MRLBM_OK;
}
