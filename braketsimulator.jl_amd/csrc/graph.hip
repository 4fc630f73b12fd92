// Captured evolutions (hipGraph) and the tile autotuner.
//
// A captured evolution is the fused plan of an instruction list, its device
// tables uploaded once, and the launches -- optionally preceded by the reset
// to |0...0> -- captured from the state's stream into one hipGraph.
// Replaying it issues the whole evolution as one graph launch, which is what
// small, launch-bound circuits need (the reference's per-gate loop,
// src/sv_simulator.jl:140-148, pays one dispatch per gate).  Only plans made
// of tile passes are captured: single-op launches stage matrices through
// host copies.
//
// The autotuner times the fused evolution of one instruction list for every
// candidate tile width (tile bits), contiguous-low-bit count and sweep width
// (register bits) on the handle's own state, then the fusion width (the
// FP64 budget of one pass) at the best of those, and keeps the fastest
// configuration.
#include <float.h>

#include <algorithm>
#include <vector>

#include "fused.h"

struct bsim_graph {
    bsim_state* h = nullptr;   // null once the state has been destroyed
    uint64_t gen = 0;          // generation of h at capture
    double2* d = nullptr;      // the amplitudes the captured kernels write
    char* tables = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

using namespace bsim;

namespace bsim {

// Called by bsim_state_destroy: the state's graphs lose their handle, so a
// later launch fails cleanly instead of replaying kernels on freed memory.
void detach_graphs(bsim_state* h) {
    for (bsim_graph* g : h->graphs) g->h = nullptr;
    h->graphs.clear();
}

}  // namespace bsim

extern "C" {

int bsim_graph_create(bsim_state* h, const bsim_op* ops, int n_ops, int reset, bsim_graph** out) {
    if (int rc = check_state(h)) return rc;
    if (!out || n_ops < 0 || (n_ops > 0 && !ops)) return fail(BSIM_EINVAL, "bad graph arguments");
    *out = nullptr;
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, ops, n_ops, eops, false)) return rc;
    Plan plan;
    if (int rc = make_plan(h->vq, eops, tile_config(), plan)) return rc;
    for (const Plan::Step& st : plan.steps)
        if (!st.tile) return fail(BSIM_ENOTSUP, "graph capture takes fusable ops only (dense <= 2 targets, diagonal <= 4)");
    if (int rc = prepare_tile_kernel()) return rc;
    bsim_graph* g = new bsim_graph();
    g->h = h;
    g->gen = h->gen;
    g->d = h->d;
    auto cleanup = [&](int rc) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        if (g->tables) (void)hipFree(g->tables);
        delete g;
        return rc;
    };
    hipError_t e = hipMalloc((void**)&g->tables, plan_table_bytes(plan));
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipMalloc(graph tables)"));
    if (int rc = upload_plan(plan, g->tables, h->stream)) return cleanup(rc);
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return cleanup(hip_fail(e, "hipStreamSynchronize"));
    if ((e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return cleanup(hip_fail(e, "hipStreamBeginCapture"));
    int rc = BSIM_OK;
    if (reset) rc = bsim_state_reset(h);  // a fill kernel, no memset node
    if (rc == BSIM_OK) rc = launch_plan(h, h->d, eops, plan, g->tables, false);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(h->stream, &graph);
    g->graph = graph;
    if (rc != BSIM_OK) return cleanup(rc);
    if (e != hipSuccess) return cleanup(hip_fail(e, "hipStreamEndCapture"));
    if ((e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0)) != hipSuccess)
        return cleanup(hip_fail(e, "hipGraphInstantiate"));
    h->graphs.push_back(g);
    *out = g;
    return BSIM_OK;
}

int bsim_graph_launch(bsim_graph* g) {
    if (!g || !g->exec) return fail(BSIM_EINVAL, "null graph");
    if (!g->h) return fail(BSIM_EINVAL, "the state this graph was captured on has been destroyed");
    if (int rc = check_state(g->h)) return rc;
    if (g->h->gen != g->gen || g->h->d != g->d) return fail(BSIM_EINVAL, "the graph's state has been re-created");
    BSIM_HIP(hipGraphLaunch(g->exec, g->h->stream));
    g->h->alias_ready = false;
    return BSIM_OK;
}

int bsim_graph_destroy(bsim_graph* g) {
    if (!g) return BSIM_OK;
    if (g->h) {
        std::vector<bsim_graph*>& v = g->h->graphs;
        v.erase(std::remove(v.begin(), v.end(), g), v.end());
        if (g->h->stream) (void)hipStreamSynchronize(g->h->stream);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    if (g->tables) (void)hipFree(g->tables);
    delete g;
    return BSIM_OK;
}

// Time one configuration: the faster of two runs (reset + fused evolve)
// after a warm-up run.
static int time_config(bsim_state* h, const std::vector<EOp>& eops, const TileConfig& cfg, hipEvent_t e0,
                       hipEvent_t e1, double& best) {
    best = DBL_MAX;
    Plan plan;
    if (int rc = make_plan(h->vq, eops, cfg, plan)) return rc;
    // one warm-up run, then the best of two timed runs -- or of one when a
    // run takes over 20 ms (large states: run-to-run spread is well under
    // the differences between candidates there)
    for (int rep = 0; rep < 3; ++rep) {
        if (int rc = bsim_state_reset(h)) return rc;
        BSIM_HIP(hipEventRecord(e0, h->stream));
        if (int rc = run_plan(h, h->d, eops, plan)) return rc;
        BSIM_HIP(hipEventRecord(e1, h->stream));
        BSIM_HIP(hipEventSynchronize(e1));
        float ms = 0;
        BSIM_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0 && ms < best) best = ms;
        if (rep == 1 && ms > 20.f) break;
    }
    return BSIM_OK;
}

int bsim_autotune(bsim_state* h, const bsim_op* ops, int n_ops, int* tile_bits, int* low_bits, int* sweep_bits,
                  int* fusion_limit, double* best_ms) {
    if (int rc = check_state(h)) return rc;
    if (n_ops < 0 || (n_ops > 0 && !ops)) return fail(BSIM_EINVAL, "bad autotune arguments");
    std::vector<EOp> eops;
    if (int rc = lower_ops(h->nq, h->kind, ops, n_ops, eops, false)) return rc;
    if (int rc = prepare_tile_kernel()) return rc;
    hipEvent_t e0, e1;
    BSIM_HIP(hipEventCreate(&e0));
    BSIM_HIP(hipEventCreate(&e1));
    struct EvGuard {
        hipEvent_t a, b;
        ~EvGuard() {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    } evg{ e0, e1 };
    // Candidates are planned with configurations of their own; the library
    // configuration changes once, to the winner, at the end.
    const TileConfig keep = tile_config();
    TileConfig best_cfg = keep;
    double best = DBL_MAX;
    const int Tmax = h->vq < 13 ? h->vq : 13;
    // A coordinate search: stage 1, the tile width and sweep width at the
    // current contiguous low bits; stage 2, the low bits at the best shape;
    // stage 3, the fusion width.
    auto trial = [&](const TileConfig& c) -> int {
        double t = DBL_MAX;
        if (int rc = time_config(h, eops, c, e0, e1, t)) return rc;
        if (t < best) {
            best = t;
            best_cfg = c;
        }
        return BSIM_OK;
    };
    for (int RB = 4; RB >= 3; --RB)
        for (int T = 10; T <= Tmax; ++T) {
            TileConfig c = keep;
            c.tile_bits = T;
            c.low_bits = std::min(keep.low_bits, T - 2);
            c.reg_bits = RB;
            if (int rc = trial(c)) return rc;
        }
    if (best < DBL_MAX) {
        const TileConfig shape = best_cfg;
        for (int L = 2; L <= 6 && L <= shape.tile_bits - 2; ++L) {
            if (L == shape.low_bits) continue;
            TileConfig c = shape;
            c.low_bits = L;
            if (int rc = trial(c)) return rc;
        }
    }
    // stage 3: the fusion width of the tile passes (FP64 operations per
    // amplitude one pass may spend; 0 = unlimited) at the best tile
    // configuration; stage 4: the width of the merged dense blocks (0 = no
    // merging, 4, 5, 6 bits), only when the plan has merged steps
    if (best < DBL_MAX) {
        static const int limits[] = { 0, 48, 96, 192, 384 };
        const TileConfig tiled = best_cfg;
        for (int lim : limits) {
            if (lim == tiled.max_cost) continue;
            TileConfig c = tiled;
            c.max_cost = lim;
            if (int rc = trial(c)) return rc;
        }
        Plan probe;
        if (int rc = make_plan(h->vq, eops, best_cfg, probe)) return rc;
        if (!probe.merged.empty()) {
            const TileConfig fused = best_cfg;
            for (int mb : { 0, 4, 5, 6 }) {
                if (mb == fused.merge_bits) continue;
                TileConfig c = fused;
                c.merge_bits = mb;
                if (int rc = trial(c)) return rc;
            }
        }
    }
    if (best == DBL_MAX) {  // no candidate fits this state: keep the current configuration
        best = 0;
        best_cfg = keep;
    } else {
        set_tile_config(best_cfg);
    }
    h->alias_ready = false;
    if (tile_bits) *tile_bits = best_cfg.tile_bits;
    if (low_bits) *low_bits = best_cfg.low_bits;
    if (sweep_bits) *sweep_bits = best_cfg.reg_bits;
    if (fusion_limit) *fusion_limit = best_cfg.max_cost;
    if (best_ms) *best_ms = best;
    return BSIM_OK;
}

}  // extern "C"
