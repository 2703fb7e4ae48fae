// mrlbm_batch.hip -- many independent runs of uneven size advanced together.
//
// Each run is an ordinary handle (own state, geometry and boundary tables).  The batch
// keeps the runs' kernel argument blocks in device memory, one array per state parity and
// recording mode, and a flattened tile table (run, bulk block) for the bulk kernel.  A step
// is four launches for all runs: the bulk over the tile table on one stream and, on a
// second stream, the edges, the stencil moments and the boundary workgroups (one per run),
// joined as in the single-run step loop.  Runs of graph_steps steps are captured once per
// (start parity, recording mode) and replayed.  With two steps per pass, a pair of steps is
// the two-step bulk over its own tile table plus two band sub-steps on the second stream,
// as in the single-run loop; with four steps per pass, the four-step bulk over its tile table
// plus four band sub-steps through the runs' two scratch states.  The autotuner times the
// candidate bulk variants, grid modes and two- and four-step strip heights on a saved copy
// of the states and keeps the fastest.
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
    int tb = 0, ntiles2 = 0;      // two steps per bulk pass: strip height (0 off), tiles
    bool tb_possible = true;      // every run has the two-step scratch state
    int qh = 0, ntiles4 = 0;      // four steps per bulk pass: strip height (0 off), tiles
    bool q_possible = true;       // every run has the four-step scratch states and rectangles
    int spp = 0;                  // steps_per_pass option: 0 autotune, 2 at most two, 4 four
    int2 *d_tiles4 = nullptr;
    int2 *d_qtiles[4] = {nullptr, nullptr, nullptr, nullptr}; // side blocks of sub-steps 1..4
    int nqtiles[4] = {0, 0, 0, 0};
    void *d_b4[2] = {nullptr, nullptr};                        // four-step bulk [parity]
    void *d_q[4][2][2] = {};                                   // sub-steps 1..4 [parity][record]
    int2 *d_stiles[3] = {nullptr, nullptr, nullptr}; // side blocks (run, block): single step, sub-step 1, 2
    int nstiles[3] = {0, 0, 0};
    int2 *d_wtiles = nullptr;     // wall blocks (run, block of 64 boundary nodes) of the runs
    int nwtiles = 0;              // whose walls are not in wall groups
    bool wall_levels = false;     // some wall has more than one dependency level
    bool walls = false;           // some run has a curved wall
    int64_t nodes = 0;
    hipStream_t st = nullptr, st2 = nullptr;
    hipEvent_t evA = nullptr, evB = nullptr, e0 = nullptr, e1 = nullptr;
    void *d_args[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}}; // single step [parity][record]
    void *d_b2[2] = {nullptr, nullptr};                            // two-step bulk [parity]
    void *d_s1[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // sub-step 1 [parity][record]
    void *d_s2[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};   // sub-step 2 [parity][record]
    int2 *d_tiles = nullptr, *d_tiles2 = nullptr;
    double *d_facc = nullptr; // 4 per run
    hipGraphExec_t exec[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    int exec_flip = 0;        // parity change of one graph replay
    std::vector<int> tune_kernel, tune_persistent, tune_tb, tune_quad;
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

template <class S, class T> static int upload_args(mrlbm_batch *b, void *&dst, const std::vector<StepArgs<S, T>> &v)
{
    if (!dst) BHIP(hipMalloc(&dst, v.size() * sizeof(StepArgs<S, T>)));
    BHIP(hipMemcpy(dst, v.data(), v.size() * sizeof(StepArgs<S, T>), hipMemcpyHostToDevice));
    return 0;
}

template <class T> static int upload_tiles(mrlbm_batch *b, int2 *&dst, const std::vector<int2> &v)
{
    if (dst) (void)hipFree(dst);
    dst = nullptr;
    BHIP(hipMalloc(&dst, std::max<size_t>(v.size(), 1) * sizeof(int2)));
    if (!v.empty()) BHIP(hipMemcpy(dst, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice));
    return 0;
}

// Argument blocks and tile tables for the runs' current bulk shape and b->tb.
template <class S, class T> static int build_tables(mrlbm_batch *b)
{
    drop_graphs(b);
    const int R = (int)b->runs.size();
    std::vector<StepArgs<S, T>> args[2][2], b2[2], s1[2][2], s2[2][2], b4[2], q[4][2][2];
    std::vector<int2> tiles, tiles2, tiles4;
    std::vector<int2> stiles[3], wtiles, qtiles[4];
    b->wall_levels = b->walls = false;
    for (int r = 0; r < R; r++) {
        b->walls = b->walls || b->runs[r]->has_b;
        b->wall_levels = b->wall_levels || (b->runs[r]->has_b && b->runs[r]->bt.nlevels > 1);
    }
    auto add_blocks = [](std::vector<int2> &v, int r, int n) {
        for (int k = 0; k < (n + 63) / 64; k++) v.push_back(make_int2(r, k));
    };
    for (int r = 0; r < R; r++) {
        mrlbm_handle *h = b->runs[r];
        // A wall of one level in wall groups runs inside the side blocks, as in a single run
        // (then no stencil items and no wall blocks); with some wall of several levels every
        // wall goes to the per-run boundary workgroups.
        const int nwg = (h->has_b && !b->wall_levels) ? h->bd.nwg : 0;
        // side blocks: the wall groups, the marched band rows, the ring's column tiles, then 64
        // items each (ring rows or the whole ring, the other bands, stencil nodes), as
        // side_block counts them
        const int nct = ring_column_tiles(h->nx, h->ny, false);
        const int ring = 64 * (nwg + nct) + (nct ? 2 * h->nx : 2 * h->nx + 2 * (h->ny - 2));
        const int nU = h->has_b && nwg == 0 ? h->bt.nU : 0;
        auto band = [&](const int rc[5][4], int n) {
            const int bm = band_march_rects(rc, n);
            return (int)(64 * band_march_blocks(rc, n, bm) + band_items(rc, n, bm));
        };
        add_blocks(stiles[0], r, ring + nU);
        if (h->has_b && nwg == 0) add_blocks(wtiles, r, h->bt.nbct);
        for (int par = 0; par < 2; par++)
            for (int rec = 0; rec < 2; rec++) {
                auto a = base_args<S, T>(h);
                a.pops = nullptr;
                a.wall_groups = nwg;
                a.rec_slot = rec ? 0 : -1;
                a.facc = rec ? b->d_facc + 4 * r : nullptr;
                auto one = a;
                one.src = (const S *)h->S[par];
                one.dst = (S *)h->S[par ^ 1];
                args[par][rec].push_back(one);
                if (b->tb) {
                    auto x = a; // sub-step 1: state t -> scratch
                    x.src = (const S *)h->S[par];
                    x.dst = (S *)h->S2;
                    x.nrect = h->nrect1;
                    std::memcpy(x.rect, h->rect1, sizeof h->rect1);
                    x.band_march = band_march_rects(x.rect, x.nrect);
                    s1[par][rec].push_back(x);
                    auto y = a; // sub-step 2: scratch -> state t+2
                    y.src = (const S *)h->S2;
                    y.dst = (S *)h->S[par ^ 1];
                    y.nrect = h->nrect2;
                    std::memcpy(y.rect, h->rect2, sizeof h->rect2);
                    y.band_march = band_march_rects(y.rect, y.nrect);
                    s2[par][rec].push_back(y);
                    if (rec == 0) {
                        auto z = one; // two-step bulk: state t -> state t+2
                        z.nstrips = (h->ny - 4 + b->tb - 1) / b->tb;
                        z.ncg = h->ncg2;
                        z.ncb = h->ncb2;
                        b2[par].push_back(z);
                    }
                }
                if (b->qh) { // sub-steps 1..4: state t -> S2 -> S3 -> S2 -> state t+4
                    const void *src[4] = {h->S[par], h->S2, h->S3, h->S2};
                    void *dst[4] = {h->S2, h->S3, h->S2, h->S[par ^ 1]};
                    for (int k = 0; k < 4; k++) {
                        auto x = a;
                        x.src = (const S *)src[k];
                        x.dst = (S *)dst[k];
                        x.nrect = h->nrectq[k];
                        std::memcpy(x.rect, h->rectq[k], sizeof h->rectq[k]);
                        x.band_march = band_march_rects(x.rect, x.nrect);
                        q[k][par][rec].push_back(x);
                    }
                    if (rec == 0) {
                        auto z = one; // four-step bulk: state t -> state t+4
                        z.nstrips = (h->ny - 8 + b->qh - 1) / b->qh;
                        z.ncg = h->ncgN;
                        z.ncb = h->ncbN;
                        b4[par].push_back(z);
                    }
                }
            }
        const int nb = h->ncb * h->nstrips;
        for (int t = 0; t < nb; t++) tiles.push_back(make_int2(r, t));
        if (b->tb) {
            add_blocks(stiles[1], r, ring + band(h->rect1, h->nrect1) + nU);
            add_blocks(stiles[2], r, ring + band(h->rect2, h->nrect2) + nU);
            const int nb2 = h->ncb2 * ((h->ny - 4 + b->tb - 1) / b->tb);
            for (int t = 0; t < nb2; t++) tiles2.push_back(make_int2(r, t));
        }
        if (b->qh) {
            for (int k = 0; k < 4; k++) add_blocks(qtiles[k], r, ring + band(h->rectq[k], h->nrectq[k]) + nU);
            const int nb4 = h->ncbN * ((h->ny - 8 + b->qh - 1) / b->qh);
            for (int t = 0; t < nb4; t++) tiles4.push_back(make_int2(r, t));
        }
    }
    int rc = 0;
    for (int par = 0; par < 2 && !rc; par++) {
        for (int rec = 0; rec < 2 && !rc; rec++) {
            rc = upload_args<S, T>(b, b->d_args[par][rec], args[par][rec]);
            if (!rc && b->tb) rc = upload_args<S, T>(b, b->d_s1[par][rec], s1[par][rec]);
            if (!rc && b->tb) rc = upload_args<S, T>(b, b->d_s2[par][rec], s2[par][rec]);
            for (int k = 0; k < 4 && !rc && b->qh; k++) rc = upload_args<S, T>(b, b->d_q[k][par][rec], q[k][par][rec]);
        }
        if (!rc && b->tb) rc = upload_args<S, T>(b, b->d_b2[par], b2[par]);
        if (!rc && b->qh) rc = upload_args<S, T>(b, b->d_b4[par], b4[par]);
    }
    if (!rc) rc = upload_tiles<T>(b, b->d_tiles, tiles);
    if (!rc) rc = upload_tiles<T>(b, b->d_tiles2, tiles2);
    if (!rc) rc = upload_tiles<T>(b, b->d_tiles4, tiles4);
    for (int k = 0; k < 4 && !rc; k++) rc = upload_tiles<T>(b, b->d_qtiles[k], qtiles[k]);
    for (int k = 0; k < 3 && !rc; k++) rc = upload_tiles<T>(b, b->d_stiles[k], stiles[k]);
    if (!rc) rc = upload_tiles<T>(b, b->d_wtiles, wtiles);
    if (rc) return rc;
    for (int k = 0; k < 3; k++) b->nstiles[k] = (int)stiles[k].size();
    b->nwtiles = (int)wtiles.size();
    b->ntiles = (int)tiles.size();
    b->ntiles2 = (int)tiles2.size();
    b->ntiles4 = (int)tiles4.size();
    for (int k = 0; k < 4; k++) b->nqtiles[k] = (int)qtiles[k].size();
    const int resident = bulk_batch_resident_blocks<S, T>(b->runs[0]->variant);
    b->grid = (b->persistent && resident > 0) ? std::min(resident, b->ntiles) : b->ntiles;
    return 0;
}

// One step of every run, parity par (runs read S[par]): the single-run stream pattern.
template <class S, class T> static int launch_step(mrlbm_batch *b, int par, int rec)
{
    const auto *args = (const StepArgs<S, T> *)b->d_args[par][rec];
    const int R = (int)b->runs.size();
    BHIP(hipStreamWaitEvent(b->st, b->evB, 0));
    BHIP(launch_bulk_batch<S, T>(args, b->d_tiles, b->ntiles, b->runs[0]->variant, b->grid, b->st));
    BHIP(hipStreamWaitEvent(b->st2, b->evA, 0));
    BHIP(launch_side_tiles<S, T>(args, b->d_stiles[0], b->nstiles[0], b->d_wtiles, b->nwtiles, R, b->wall_levels,
                                 b->walls, rec != 0, b->st2));
    BHIP(hipEventRecord(b->evB, b->st2));
    BHIP(hipEventRecord(b->evA, b->st));
    return 0;
}

// Two steps of every run from parity par: the two-step bulk on stream A, the two band
// sub-steps on stream B; the runs end at parity par ^ 1.
template <class S, class T> static int launch_pair(mrlbm_batch *b, int par, int rec)
{
    const int R = (int)b->runs.size();
    BHIP(hipStreamWaitEvent(b->st, b->evB, 0));
    BHIP(launch_bulk2_batch<S, T>((const StepArgs<S, T> *)b->d_b2[par], b->d_tiles2, b->ntiles2, b->tb, b->st));
    BHIP(hipStreamWaitEvent(b->st2, b->evA, 0));
    BHIP(launch_side_tiles<S, T>((const StepArgs<S, T> *)b->d_s1[par][rec], b->d_stiles[1], b->nstiles[1], b->d_wtiles,
                                 b->nwtiles, R, b->wall_levels, b->walls, rec != 0, b->st2));
    BHIP(launch_side_tiles<S, T>((const StepArgs<S, T> *)b->d_s2[par][rec], b->d_stiles[2], b->nstiles[2], b->d_wtiles,
                                 b->nwtiles, R, b->wall_levels, b->walls, rec != 0, b->st2));
    BHIP(hipEventRecord(b->evB, b->st2));
    BHIP(hipEventRecord(b->evA, b->st));
    return 0;
}

// Four steps of every run from parity par: the four-step bulk on stream A, the four band
// sub-steps on stream B; the runs end at parity par ^ 1.
template <class S, class T> static int launch_quad(mrlbm_batch *b, int par, int rec)
{
    const int R = (int)b->runs.size();
    BHIP(hipStreamWaitEvent(b->st, b->evB, 0));
    BHIP(launch_bulkn_batch<S, T>((const StepArgs<S, T> *)b->d_b4[par], b->d_tiles4, b->ntiles4, b->qh, b->st));
    BHIP(hipStreamWaitEvent(b->st2, b->evA, 0));
    for (int k = 0; k < 4; k++)
        BHIP(launch_side_tiles<S, T>((const StepArgs<S, T> *)b->d_q[k][par][rec], b->d_qtiles[k], b->nqtiles[k],
                                     b->d_wtiles, b->nwtiles, R, b->wall_levels, b->walls, rec != 0, b->st2));
    BHIP(hipEventRecord(b->evB, b->st2));
    BHIP(hipEventRecord(b->evA, b->st));
    return 0;
}

// Steps remaining in a run of n from s, taken by the next pass: four, two or one.
static int pass_steps(const mrlbm_batch *b, int left)
{
    if (b->qh && left >= 4) return 4;
    return (b->tb && left >= 2) ? 2 : 1;
}

// n steps from parity par: four-step passes, pairs or single steps
template <class S, class T> static int launch_steps(mrlbm_batch *b, int par, int rec, int n, int *par_out)
{
    int rc = 0, s = 0;
    while (s < n && !rc) {
        if (pass_steps(b, n - s) == 4) {
            rc = launch_quad<S, T>(b, par, rec);
            s += 4;
        } else if (pass_steps(b, n - s) == 2) {
            rc = launch_pair<S, T>(b, par, rec);
            s += 2;
        } else {
            rc = launch_step<S, T>(b, par, rec);
            s += 1;
        }
        par ^= 1;
    }
    if (par_out) *par_out = par;
    return rc;
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
    int par_end = par;
    if (!rc) rc = launch_steps<S, T>(b, par, rec, b->graph_steps, &par_end);
    if (!rc && hipStreamWaitEvent(b->st, b->evB, 0) != hipSuccess) rc = MRLBM_E_HIP;
    const hipError_t ee = hipStreamEndCapture(b->st, &g);
    if (rc || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        BFAIL(MRLBM_E_HIP, "graph capture of the batched step failed");
    }
    e = hipGraphInstantiate(&b->exec[par][rec], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    BHIP(e);
    b->exec_flip = par_end ^ par;
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
            par ^= b->exec_flip;
            s += b->graph_steps;
        } else {
            const int n = pass_steps(b, nsteps - s);
            const int rc = launch_steps<S, T>(b, par, rec, n, &par);
            if (rc) return rc;
            s += n;
        }
    }
    BHIP(hipStreamWaitEvent(b->st, b->evB, 0));
    BHIP(hipStreamSynchronize(b->st));
    for (mrlbm_handle *h : b->runs) h->cur = par;
    return 0;
}

// Times every candidate on a saved copy of the states, keeps the fastest: the single-step
// bulk variants x grid modes, then two steps per pass with the strip heights in `tbs`, then
// four steps per pass with the strip heights in `qhs` (the rest of a graph by the best pair).
template <class S, class T>
static int autotune_typed(mrlbm_batch *b, int fixed_kernel, int fixed_persistent, int fixed_tb)
{
    std::vector<int> kernels;
    if (fixed_kernel) kernels = {fixed_kernel};
    else if (b->f32) kernels = {304, 308, 8, 16, 4};
    else kernels = {8, 304, 16, 4, 308};
    std::vector<int> modes;
    if (fixed_persistent >= 0) modes = {fixed_persistent};
    else modes = {1, 0};
    std::vector<int> tbs;
    if (fixed_tb > 0 && b->tb_possible) tbs = {fixed_tb};
    else if (fixed_tb < 0 && b->tb_possible) tbs = b->f32 ? std::vector<int>{6, 10} : std::vector<int>{6, 8};
    std::vector<int> qhs;
    if (fixed_tb != 0 && b->q_possible && b->spp != 2)
        qhs = b->f32 ? std::vector<int>{20, 24, 12} : std::vector<int>{20, 16, 8};
    // steps_per_pass 4: only the four-step candidates (the single-step variant and pair strip
    // height of the remainders are the first ones, untimed)
    const bool only_q = b->spp == 4 && !qhs.empty();
    const bool only_tb = !only_q && fixed_tb > 0 && b->tb_possible;
    const bool time_single = !only_q && !only_tb, time_pairs = !only_q;
    const size_t ncand = (time_single ? kernels.size() * modes.size() : 0) + (time_pairs ? tbs.size() : 0) + qhs.size();
    const bool tune = ncand > 1;
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
    b->tune_tb.clear();
    b->tune_quad.clear();
    b->tune_ms.clear();
    BHIP(hipEventRecord(b->evA, b->st));
    BHIP(hipEventRecord(b->evB, b->st));
    int best_k = kernels[0], best_p = modes[0], best_tb = (only_tb || only_q) && !tbs.empty() ? tbs[0] : 0;
    int best_q = only_q ? qhs[0] : 0;
    double best = 1e300;
    int rc = 0;
    auto try_candidate = [&](int k, int pm, int tbc, int qhc) -> int {
        std::string why;
        for (mrlbm_handle *h : b->runs) {
            const int r2 = set_bulk_shape(h, k, 0, why);
            if (r2) BFAIL(r2, why);
        }
        b->persistent = pm;
        b->tb = tbc;
        b->qh = qhc;
        int r3 = build_tables<S, T>(b);
        if (r3 || !tune) return r3;
        const int par = b->runs[0]->cur;
        int pe = par;
        r3 = launch_steps<S, T>(b, par, 0, qhc ? 4 : 2, &pe);
        if (!r3 && hipEventRecord(b->e0, b->st) != hipSuccess) r3 = MRLBM_E_HIP;
        if (!r3) r3 = launch_steps<S, T>(b, pe, 0, b->tune_steps, &pe);
        if (!r3 && hipStreamWaitEvent(b->st, b->evB, 0) != hipSuccess) r3 = MRLBM_E_HIP;
        if (!r3 && hipEventRecord(b->e1, b->st) != hipSuccess) r3 = MRLBM_E_HIP;
        if (!r3 && hipEventSynchronize(b->e1) != hipSuccess) r3 = MRLBM_E_HIP;
        float ms = 0;
        if (!r3 && hipEventElapsedTime(&ms, b->e0, b->e1) != hipSuccess) r3 = MRLBM_E_HIP;
        if (r3) return r3;
        const double per = ms / b->tune_steps;
        b->tune_kernel.push_back(k);
        b->tune_persistent.push_back(pm);
        b->tune_tb.push_back(tbc);
        b->tune_quad.push_back(qhc);
        b->tune_ms.push_back(per);
        if (per < best) {
            best = per;
            best_k = k;
            best_p = pm;
            best_tb = tbc;
            best_q = qhc;
        }
        // restore the states and parity for the next candidate
        for (size_t r = 0; r < b->runs.size(); r++) {
            mrlbm_handle *h = b->runs[r];
            h->cur = par0;
            if (hipMemcpy(h->S[par0], saved[r], (size_t)6 * h->plane * sizeof(S), hipMemcpyDeviceToDevice) != hipSuccess)
                return MRLBM_E_HIP;
        }
        return 0;
    };
    if (time_single)
        for (int k : kernels) {
            for (int pm : modes) {
                rc = try_candidate(k, pm, 0, 0);
                if (rc || !tune) break;
            }
            if (rc || !tune) break;
        }
    const int single_k = best_k, single_p = best_p; // the remainder steps of a multi-step run
    if (!rc && time_pairs && (tune || only_tb))
        for (int tbc : tbs) {
            rc = try_candidate(single_k, only_tb ? modes[0] : single_p, tbc, 0);
            if (rc || !tune) break;
        }
    // the pair strip height of the four-step candidates' remainders: the best pair timed
    int pair_tb = only_q ? best_tb : 0;
    if (!only_q) {
        double bp = 1e300;
        for (size_t i = 0; i < b->tune_ms.size(); i++)
            if (b->tune_tb[i] > 0 && b->tune_ms[i] < bp) {
                bp = b->tune_ms[i];
                pair_tb = b->tune_tb[i];
            }
        if (!pair_tb && !tbs.empty()) pair_tb = tbs[0];
    }
    if (!rc && (tune || only_q))
        for (int qhc : qhs) {
            rc = try_candidate(single_k, single_p, pair_tb, qhc);
            if (rc || !tune) break;
        }
    for (void *q : saved) (void)hipFree(q);
    if (rc) return rc;
    b->kernel = best_k;
    std::string why;
    for (mrlbm_handle *h : b->runs) {
        const int r2 = set_bulk_shape(h, best_k, 0, why);
        if (r2) BFAIL(r2, why);
    }
    b->persistent = best_p;
    b->tb = best_q ? pair_tb : best_tb;
    b->qh = best_q;
    return build_tables<S, T>(b);
}

static void batch_free(mrlbm_batch *b)
{
    drop_graphs(b);
    for (int par = 0; par < 2; par++) {
        for (int rec = 0; rec < 2; rec++)
            for (void *p : {b->d_args[par][rec], b->d_s1[par][rec], b->d_s2[par][rec]})
                if (p) (void)hipFree(p);
        if (b->d_b2[par]) (void)hipFree(b->d_b2[par]);
        if (b->d_b4[par]) (void)hipFree(b->d_b4[par]);
        for (int k = 0; k < 4; k++)
            for (int rec = 0; rec < 2; rec++)
                if (b->d_q[k][par][rec]) (void)hipFree(b->d_q[k][par][rec]);
    }
    for (int2 *p : {b->d_tiles, b->d_tiles2, b->d_tiles4, b->d_stiles[0], b->d_stiles[1], b->d_stiles[2], b->d_wtiles,
                    b->d_qtiles[0], b->d_qtiles[1], b->d_qtiles[2], b->d_qtiles[3]})
        if (p) (void)hipFree(p);
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
    o->time_blocking = -1;
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
    if (opts.time_blocking < -1 || opts.time_blocking > 10 || (opts.time_blocking > 0 && opts.time_blocking % 2))
        return fail(MRLBM_E_ARG, "time_blocking must be -1, 0, 2, 4, 6, 8 or 10");
    if (opts.kernel == 1) return fail(MRLBM_E_ARG, "the batched bulk has no one-node-per-thread kernel");
    if (opts.steps_per_pass != 0 && opts.steps_per_pass != 2 && opts.steps_per_pass != 4)
        return fail(MRLBM_E_ARG, "steps_per_pass must be 0, 2 or 4");
    b->spp = opts.steps_per_pass;
    b->graph_steps = opts.graph_steps;
    b->tune_steps = opts.tune_steps;
    b->f32 = runs[0].precision == MRLBM_FP32_DEV;
    for (int r = 0; r < nruns; r++) {
        if ((runs[r].precision == MRLBM_FP32_DEV) != b->f32) return fail(MRLBM_E_ARG, "every run needs the same precision");
        mrlbm_handle *h = nullptr;
        mrlbm_params pr = runs[r];
        // any strip height: it allocates the scratch state and the band rectangles of the
        // two-step pass and, unless the batch stays at two steps per pass, the second scratch
        // state and the rectangles of the four-step pass (the batch picks the strip heights)
        pr.time_blocking = opts.time_blocking == 0 ? 0 : 2;
        pr.steps_per_pass = opts.steps_per_pass == 2 ? 2 : 0;
        const int rc = mrlbm_create(&pr, &h);
        if (rc) return fail(rc, "run " + std::to_string(r) + ": " + mrlbm_last_error(nullptr));
        b->runs.push_back(h);
        b->nodes += (int64_t)h->nx * h->ny;
        if (!h->S2) b->tb_possible = false;
        if (!h->S3 || !h->ms_H) b->q_possible = false;
        if (!h->d_rec) { // force reduction target of the boundary workgroup
            h->rec_stride = 14 + 2 * h->bt.nbct;
            h->rec_cap = 256;
            if (hipMalloc(&h->d_rec, (size_t)h->rec_cap * h->rec_stride * sizeof(double)) != hipSuccess)
                return fail(MRLBM_E_NOMEM, "force record buffer");
        }
    }
    if (hipStreamCreate(&b->st) != hipSuccess || hipStreamCreate(&b->st2) != hipSuccess ||
        hipEventCreateWithFlags(&b->evA, hipEventDisableTiming | hipEventReleaseToDevice) != hipSuccess ||
        hipEventCreateWithFlags(&b->evB, hipEventDisableTiming | hipEventReleaseToDevice) != hipSuccess ||
        hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess)
        return fail(MRLBM_E_HIP, "stream or event creation failed");
    if (hipMalloc(&b->d_facc, (size_t)4 * nruns * sizeof(double)) != hipSuccess ||
        hipMemset(b->d_facc, 0, (size_t)4 * nruns * sizeof(double)) != hipSuccess)
        return fail(MRLBM_E_HIP, "force sums");
    const int rc = b->f32 ? autotune_typed<float, float>(b, opts.kernel, opts.persistent, opts.time_blocking)
                          : autotune_typed<double, double>(b, opts.kernel, opts.persistent, opts.time_blocking);
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
    const int tbc = b->tb_possible ? -1 : 0;
    return b->f32 ? autotune_typed<float, float>(b, 0, -1, tbc) : autotune_typed<double, double>(b, 0, -1, tbc);
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
    info->time_blocking = b->tb;
    info->quad_rows = b->qh;
    info->ntune = (int32_t)std::min<size_t>(16, b->tune_ms.size());
    for (int i = 0; i < info->ntune; i++) {
        info->tune_kernel[i] = b->tune_kernel[i];
        info->tune_persistent[i] = b->tune_persistent[i];
        info->tune_tb[i] = b->tune_tb[i];
        info->tune_ms[i] = b->tune_ms[i];
        info->tune_quad[i] = b->tune_quad[i];
    }
    return MRLBM_OK;
}
