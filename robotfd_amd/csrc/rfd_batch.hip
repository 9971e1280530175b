// Batch entry points of include/rfd.h: many pairs resident on the device,
// every round of every pair's computation on the context's stream, captured
// once as a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>
#include <string>
#include <vector>

#include "batch_dec.hpp"
#include "batch_kernels.hpp"
#include "mlp_internal.hpp"
#include "rfd.h"
#include "rfd_internal.hpp"

namespace {

using rfd_host::set_err;

// Decision rounds captured in the graph.  A pair converges within
// log2((ub - lb) / tolerance) rounds; rounds without active pairs cost one
// empty plan/sweep/update.  Pairs still active afterwards (only possible for
// an absolute tolerance far below the distances' scale) continue in
// replays of a graph holding only decision rounds.
constexpr int kDecisionRounds = 48;
constexpr int kContinueRounds = 16;

struct GraphKey {
    int rpl, stop_mode, prune, gate, rounds, with_bounds, depth;
    double eps;
    const void *net;   // learned runs: the network whose forward pass is captured
    double tau, delta;
    int bcfg;          // fused bound round configuration (0: separate DFD and VE-FD rounds)
    int dwin;          // decision rounds on the windowed sweep (configuration; 0: band_sweep)
    bool operator<(const GraphKey &o) const {
        return std::tie(rpl, stop_mode, prune, gate, rounds, with_bounds, depth, eps, net, tau, delta, bcfg, dwin) <
               std::tie(o.rpl, o.stop_mode, o.prune, o.gate, o.rounds, o.with_bounds, o.depth, o.eps,
                        o.net, o.tau, o.delta, o.bcfg, o.dwin);
    }
};

// Process-wide cache of the autotuner's choice, keyed by a batch signature.
std::mutex g_tune_mu;
std::map<std::string, int> g_tune_cache;

}  // namespace

struct rfd_batch {
    rfd_ctx *ctx = nullptr;
    int64_t npairs = 0;
    std::vector<rfd::BatchPair> hpairs;
    std::vector<int> order;
    rfd::BatchPair *pairs = nullptr;
    rfd::BatchState *state = nullptr;
    int *d_order = nullptr;
    double *pts = nullptr, *qseg = nullptr, *slots = nullptr;
    int *progress = nullptr;
    unsigned long long *flags64 = nullptr;  // band words of windowed decision rounds
    int2 *items = nullptr;
    int2 *bitems = nullptr;    // the fused bound round's fixed work list
    int *count = nullptr;
    unsigned long long *evals = nullptr, *t0 = nullptr;
    int *abort = nullptr;      // raised by a hand-off wait that gave up (rfd::Watch)
    float *feat = nullptr;     // [npairs x 386] FP32 features of the last bounds
    uint16_t *xin = nullptr;   // [npairs x 400] BF16 network input
    float *pred = nullptr;     // [npairs] network output
    int64_t total_items = 0;
    int tuned_rpl = 0;
    std::string signature;
    std::map<GraphKey, hipGraphExec_t> graphs;
    std::map<GraphKey, hipGraph_t> graph_defs;
};

namespace {

template <int OP, bool PRUNE>
void launch_batch_sweep_p(int rpl, int grid, hipStream_t st, const rfd::BatchArgs &a) {
    if (rpl == 4)
        hipLaunchKernelGGL((rfd::batch_sweep<OP, 4, PRUNE>), dim3(grid), dim3(64), 0, st, a);
    else if (rpl == 3)
        hipLaunchKernelGGL((rfd::batch_sweep<OP, 3, PRUNE>), dim3(grid), dim3(64), 0, st, a);
    else if (rpl == 2)
        hipLaunchKernelGGL((rfd::batch_sweep<OP, 2, PRUNE>), dim3(grid), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL((rfd::batch_sweep<OP, 1, PRUNE>), dim3(grid), dim3(64), 0, st, a);
}

// The pruning instantiation serves the first VE-FD round when pruning is on
// (its per-pair thresholds are then finite); the redo round and every other
// round run without the per-cell test.
template <int OP>
void launch_batch_sweep(int rpl, int grid, hipStream_t st, const rfd::BatchArgs &a, bool prune = false) {
    if (OP == rfd::OP_VEFD && prune) launch_batch_sweep_p<OP, true>(rpl, grid, st, a);
    else launch_batch_sweep_p<OP, false>(rpl, grid, st, a);
}

template <int R, int NP>
int batch_dec_grid_cfg(rfd_ctx *ctx) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, rfd::batch_dec<R, NP>, 64 * (NP + 2), 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (ctx->batch_dec_wgcu > 0) per_cu = std::min(per_cu, ctx->batch_dec_wgcu);
    return std::max(1, std::min(ctx->cu_count * per_cu, ctx->max_waves));
}

// Workgroups of a windowed batch decision round at R rows per lane (six
// producer waves).
int rfd_batch_dec_grid(rfd_ctx *ctx, int rpl) {
    switch (rpl) {
        case 1: return batch_dec_grid_cfg<1, 6>(ctx);
        case 3: return batch_dec_grid_cfg<3, 6>(ctx);
        case 4: return batch_dec_grid_cfg<4, 6>(ctx);
        default: return batch_dec_grid_cfg<2, 6>(ctx);
    }
}

void launch_batch_dec(rfd_ctx *ctx, hipStream_t st, const rfd::BatchArgs &a) {
    const int rpl = a.dec_win / 10;
    const int grid = rfd_batch_dec_grid(ctx, rpl);
    switch (rpl) {
        case 1: hipLaunchKernelGGL((rfd::batch_dec<1, 6>), dim3(grid), dim3(64 * 8), 0, st, a); break;
        case 3: hipLaunchKernelGGL((rfd::batch_dec<3, 6>), dim3(grid), dim3(64 * 8), 0, st, a); break;
        case 4: hipLaunchKernelGGL((rfd::batch_dec<4, 6>), dim3(grid), dim3(64 * 8), 0, st, a); break;
        default: hipLaunchKernelGGL((rfd::batch_dec<2, 6>), dim3(grid), dim3(64 * 8), 0, st, a); break;
    }
}

// A decision or certification round's sweep.
void launch_decisions(rfd_ctx *ctx, int rpl, hipStream_t st, const rfd::BatchArgs &a) {
    if (a.dec_win) launch_batch_dec(ctx, st, a);
    else launch_batch_sweep<rfd::OP_DEC>(rpl, ctx->max_waves, st, a);
}

rfd::BatchArgs make_args(rfd_batch *b, int rpl, const rfd_opts &o, int prune) {
    rfd::BatchArgs a{};
    a.bounds_rows = b->ctx->batch_bounds >= 1000 ? b->ctx->batch_bounds / 100 : b->ctx->batch_bounds / 10;
    a.pairs = b->pairs;
    a.state = b->state;
    a.order = b->d_order;
    a.npairs = (int)b->npairs;
    a.pts = b->pts;
    a.qseg = b->qseg;
    a.slots = b->slots;
    a.progress = b->progress;
    a.flags64 = b->flags64;
    a.items = b->items;
    a.bitems = b->bitems;
    a.count = b->count;
    a.evals = b->evals;
    a.t0 = b->t0;
    a.rows_per_lane = rpl;
    a.stop_mode = o.stop_mode;
    a.eps = o.eps;
    a.prune = prune;
    a.gate = o.vefd_gate;
    a.max_waves = b->ctx->max_waves;
    // Decision rounds on the windowed sweep at the batch's rows per lane, six
    // producer waves.  Speculation depth: they count workgroups (several
    // waves each, one band at a time) instead of single-wave bands.
    a.dec_win = b->ctx->batch_dec ? 10 * rpl + 6 : 0;
    if (a.dec_win) a.max_waves = rfd_batch_dec_grid(b->ctx, rpl) * b->ctx->batch_dec_budget;
    a.max_depth = o.spec_depth >= 1 ? std::min(o.spec_depth, rfd::kMaxDepth) : rfd::kMaxDepth;
    a.watch = rfd::Watch{b->abort, b->ctx->spin_limit};
    return a;
}

void enqueue_rounds(rfd_batch *b, const rfd::BatchArgs &a, bool with_bounds, int rounds,
                    rfd_mlp *net = nullptr, double tau = 0.0, double delta = 0.0);

// Enqueues the rounds on the stream (captured or not).  With a network: after
// the bounds, features -> forward pass -> predicted bracket -> one round of
// certifying decisions, then the bisection rounds.
void enqueue_rounds(rfd_batch *b, const rfd::BatchArgs &a, bool with_bounds, int rounds,
                    rfd_mlp *net, double tau, double delta) {
    rfd_ctx *ctx = b->ctx;
    hipStream_t st = ctx->stream;
    const unsigned pg = (unsigned)((b->npairs + 255) / 256);
    const int grid = ctx->max_waves;
    const int rpl = a.rows_per_lane;
    if (with_bounds) {
        hipLaunchKernelGGL(rfd::batch_reset, dim3(pg), dim3(256), 0, st, a);
        hipLaunchKernelGGL(rfd::batch_qseg, dim3(8, (unsigned)b->npairs), dim3(256), 0, st, a);
        if (ctx->batch_bounds) {
            // DFD and VE-FD of every pair in one fused round.
            hipLaunchKernelGGL(rfd::batch_plan, dim3(1), dim3(1024), 0, st, a, (int)rfd::PH_BOUNDS);
            rfd_host::rfd_launch_batch_bounds(ctx, ctx->batch_bounds, st, a);
            hipLaunchKernelGGL(rfd::batch_update, dim3(pg), dim3(256), 0, st, a, (int)rfd::PH_BOUNDS);
        } else {
            const int phases[3] = {rfd::PH_DFD, rfd::PH_VEFD, rfd::PH_VEFD_REDO};
            for (int ph : phases) {
                hipLaunchKernelGGL(rfd::batch_plan, dim3(1), dim3(1024), 0, st, a, ph);
                if (ph == rfd::PH_DFD) launch_batch_sweep<rfd::OP_DFD>(rpl, grid, st, a);
                else launch_batch_sweep<rfd::OP_VEFD>(rpl, grid, st, a, ph == rfd::PH_VEFD && a.prune);
                hipLaunchKernelGGL(rfd::batch_update, dim3(pg), dim3(256), 0, st, a, ph);
            }
        }
        if (net) {
            hipLaunchKernelGGL(rfd::batch_features, dim3((unsigned)b->npairs), dim3(128), 0, st, a,
                               b->feat, b->xin);
            rfd_host::mlp_forward_dev(net, b->xin, (int)b->npairs, b->pred);
            hipLaunchKernelGGL(rfd::batch_learned_prepare, dim3(pg), dim3(256), 0, st, a,
                               (const float *)b->pred, tau, delta);
            hipLaunchKernelGGL(rfd::batch_plan, dim3(1), dim3(1024), 0, st, a, (int)rfd::PH_CERT);
            launch_decisions(ctx, rpl, st, a);
            hipLaunchKernelGGL(rfd::batch_update, dim3(pg), dim3(256), 0, st, a, (int)rfd::PH_CERT);
        }
    }
    for (int r = 0; r < rounds; ++r) {
        hipLaunchKernelGGL(rfd::batch_plan, dim3(1), dim3(1024), 0, st, a, (int)rfd::PH_DEC);
        launch_decisions(ctx, rpl, st, a);
        hipLaunchKernelGGL(rfd::batch_update, dim3(pg), dim3(256), 0, st, a, (int)rfd::PH_DEC);
    }
}

rfd_status get_graph(rfd_batch *b, const GraphKey &key, const rfd::BatchArgs &a, hipGraphExec_t *out,
                     rfd_mlp *net = nullptr) {
    rfd_ctx *ctx = b->ctx;
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return RFD_OK;
    }
    hipGraph_t g = nullptr;
    RFD_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    enqueue_rounds(b, a, key.with_bounds != 0, key.rounds, net, key.tau, key.delta);
    RFD_HIP(ctx, hipStreamEndCapture(ctx->stream, &g));
    hipGraphExec_t ge = nullptr;
    RFD_HIP(ctx, hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
    b->graphs[key] = ge;
    b->graph_defs[key] = g;
    *out = ge;
    return RFD_OK;
}

// One full run of the batch with a given tile height; fills host states.
rfd_status run_once(rfd_batch *b, const rfd_opts &o, int rpl, int prune,
                    std::vector<rfd::BatchState> &hst, unsigned long long &t0, float &ms,
                    rfd_mlp *net = nullptr, double tau = 0.0, double delta = 0.0) {
    rfd_ctx *ctx = b->ctx;
    rfd::BatchArgs a = make_args(b, rpl, o, prune);
    if (net) {
        rfd_status s = rfd_host::mlp_reserve(net, (int)b->npairs);
        if (s != RFD_OK) return s;
    }
    RFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    if (o.use_graph) {
        hipGraphExec_t ge;
        GraphKey key{rpl, o.stop_mode, prune, o.vefd_gate, kDecisionRounds, 1, a.max_depth, o.eps,
                     net, tau, delta, ctx->batch_bounds, ctx->batch_dec};
        rfd_status s = get_graph(b, key, a, &ge, net);
        if (s != RFD_OK) return s;
        RFD_HIP(ctx, hipGraphLaunch(ge, ctx->stream));
    } else {
        enqueue_rounds(b, a, true, kDecisionRounds, net, tau, delta);
        RFD_HIP(ctx, hipGetLastError());
    }
    hst.resize((size_t)b->npairs);
    for (int guard = 0;; ++guard) {
        int habort = 0;
        RFD_HIP(ctx, hipMemcpyAsync(hst.data(), b->state, sizeof(rfd::BatchState) * b->npairs,
                                    hipMemcpyDeviceToHost, ctx->stream));
        RFD_HIP(ctx, hipMemcpyAsync(&habort, b->abort, sizeof(habort), hipMemcpyDeviceToHost, ctx->stream));
        RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        if (habort) {
            set_err(ctx, "batch: a hand-off wait exceeded %u polls and gave up", ctx->spin_limit);
            return RFD_ERR_STATE;
        }
        bool any = false;
        for (const auto &s : hst) any = any || s.active;
        if (!any) break;
        if (guard > 64) {
            set_err(ctx, "batch bisection did not converge");
            return RFD_ERR_STATE;
        }
        if (o.use_graph) {
            hipGraphExec_t ge;
            GraphKey key{rpl, o.stop_mode, prune, o.vefd_gate, kContinueRounds, 0, a.max_depth, o.eps,
                         nullptr, 0.0, 0.0, ctx->batch_bounds, ctx->batch_dec};
            rfd_status s = get_graph(b, key, a, &ge);
            if (s != RFD_OK) return s;
            RFD_HIP(ctx, hipGraphLaunch(ge, ctx->stream));
        } else {
            enqueue_rounds(b, a, false, kContinueRounds);
        }
    }
    RFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    RFD_HIP(ctx, hipEventSynchronize(ctx->ev1));
    RFD_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    RFD_HIP(ctx, hipMemcpy(&t0, b->t0, sizeof(t0), hipMemcpyDeviceToHost));
    return RFD_OK;
}

int resolve_prune(const rfd_opts &o) { return o.prune < 0 ? 1 : (o.prune ? 1 : 0); }

}  // namespace

extern "C" {

rfd_status rfd_batch_destroy(rfd_batch *b) {
    if (!b) return RFD_ERR_ARG;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    for (auto &kv : b->graphs) (void)hipGraphExecDestroy(kv.second);
    for (auto &kv : b->graph_defs) (void)hipGraphDestroy(kv.second);
    (void)hipFree(b->pairs);
    (void)hipFree(b->state);
    (void)hipFree(b->d_order);
    (void)hipFree(b->pts);
    (void)hipFree(b->qseg);
    (void)hipFree(b->slots);
    (void)hipFree(b->progress);
    (void)hipFree(b->flags64);
    (void)hipFree(b->items);
    (void)hipFree(b->bitems);
    (void)hipFree(b->count);
    (void)hipFree(b->evals);
    (void)hipFree(b->t0);
    (void)hipFree(b->abort);
    (void)hipFree(b->feat);
    (void)hipFree(b->xin);
    (void)hipFree(b->pred);
    delete b;
    return RFD_OK;
}

rfd_status rfd_batch_create(rfd_ctx *ctx, int64_t npairs, const double *const *P, const int64_t *m,
                            const double *const *Q, const int64_t *n, const int32_t *dims,
                            rfd_batch **out) {
    if (!ctx || !P || !Q || !m || !n || !dims || !out || npairs <= 0) return RFD_ERR_ARG;
    *out = nullptr;
    if (npairs > (1 << 24)) {
        set_err(ctx, "at most 2^24 pairs per batch");
        return RFD_ERR_ARG;
    }
    for (int64_t k = 0; k < npairs; ++k) {
        if (!P[k] || !Q[k] || m[k] < 1 || n[k] < 1 || (dims[k] != 2 && dims[k] != 3) ||
            m[k] > (1 << 30) || n[k] > (1 << 30)) {
            set_err(ctx, "pair %lld: need non-empty curves of dims 2 or 3", (long long)k);
            return RFD_ERR_ARG;
        }
        if (rfd_status fs = rfd_host::check_finite(ctx, P[k], m[k], Q[k], n[k], dims[k], k); fs != RFD_OK)
            return fs;
    }
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    rfd_batch *b = new rfd_batch();
    b->ctx = ctx;
    b->npairs = npairs;
    b->hpairs.resize((size_t)npairs);
    long long pofs = 0, qsofs = 0, slofs = 0, bofs = 0;
    double cells_total = 0;
    for (int64_t k = 0; k < npairs; ++k) {
        rfd::BatchPair &p = b->hpairs[(size_t)k];
        p.m = m[k];
        p.n = n[k];
        p.poff = pofs;
        pofs += 3 * m[k];
        p.qoff = pofs;
        pofs += 3 * n[k];
        p.qsoff = qsofs;
        qsofs += 2 * std::max<int64_t>(n[k] - 1, 1);
        p.sloff = slofs;
        slofs += (long long)rfd::kSpec * 2 * n[k];
        p.bandoff = bofs;
        p.bandstride = (m[k] + 63) / 64 + 1;
        bofs += (long long)rfd::kSpec * p.bandstride;
        p.endpoint_lb = rfd_host::endpoint_bound(P[k], m[k], Q[k], n[k], dims[k]);
        p.big = (m[k] * n[k] > 10000) ? 1 : 0;
        p.dec_form = rfd::DEC_FUSED;
        p.dims = dims[k];
        cells_total += (double)m[k] * (double)n[k];
    }
    b->total_items = bofs;
    // Longest pairs first (their chains of bands are the critical path).
    b->order.resize((size_t)npairs);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int x, int y) {
        return (double)m[x] * n[x] > (double)m[y] * n[y];
    });
    // The fused bound round's work list.  Every pair takes part, so the list is
    // fixed: pair-major in `order` (the plan's order), then spread so that a
    // pair's band b sits kBandSpread places after its band b - 1.  Listed back
    // to back, band b would be taken while band b - 1 is a few steps in, and
    // its workgroup would wait ~64 steps per band of lag for its input rows:
    // the wait of a tall pair's last band (25 bands of 832 rows) exceeds a
    // narrow pair's whole band.  Spread by about one pick per workgroup, band
    // b - 1 is a band's length ahead.  Keys are increasing within a pair, so
    // every band still follows its predecessor in the list (no deadlock).
    std::vector<int2> hbitems;
    if (ctx->batch_bounds) {
        constexpr long long kBandSpread = 256;
        const int br = ctx->batch_bounds >= 1000 ? ctx->batch_bounds / 100 : ctx->batch_bounds / 10;
        std::vector<std::pair<long long, int2>> keyed;
        long long pos = 0;
        for (int k : b->order) {
            const long long nb = (m[k] + 64LL * br - 1) / (64LL * br);
            for (long long bb = 0; bb < nb; ++bb)
                keyed.push_back({pos + kBandSpread * bb, make_int2(k * rfd::kSpec, (int)bb)});
            pos += nb;
        }
        std::stable_sort(keyed.begin(), keyed.end(),
                         [](const auto &x, const auto &y) { return x.first < y.first; });
        hbitems.reserve(keyed.size());
        for (const auto &kv : keyed) hbitems.push_back(kv.second);
    }
    char sig[128];
    std::snprintf(sig, sizeof(sig), "n%lld_c%.0f", (long long)npairs, std::log2(cells_total + 1) * 4);
    b->signature = sig;

    std::vector<double> hpts((size_t)pofs, 0.0);
    for (int64_t k = 0; k < npairs; ++k) {
        const rfd::BatchPair &p = b->hpairs[(size_t)k];
        const int d = dims[k];
        for (int64_t i = 0; i < m[k]; ++i)
            for (int c = 0; c < d; ++c) hpts[(size_t)(p.poff + 3 * i + c)] = P[k][i * d + c];
        for (int64_t i = 0; i < n[k]; ++i)
            for (int c = 0; c < d; ++c) hpts[(size_t)(p.qoff + 3 * i + c)] = Q[k][i * d + c];
    }
    auto fail = [&](hipError_t e, const char *what) {
        set_err(ctx, "%s failed: %s", what, hipGetErrorString(e));
        rfd_batch_destroy(b);
        return e == hipErrorOutOfMemory ? RFD_ERR_NOMEM : RFD_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&b->pairs, sizeof(rfd::BatchPair) * npairs)) != hipSuccess) return fail(e, "hipMalloc(pairs)");
    if ((e = hipMalloc(&b->state, sizeof(rfd::BatchState) * npairs)) != hipSuccess) return fail(e, "hipMalloc(state)");
    if ((e = hipMalloc(&b->d_order, sizeof(int) * npairs)) != hipSuccess) return fail(e, "hipMalloc(order)");
    if ((e = hipMalloc(&b->pts, sizeof(double) * pofs)) != hipSuccess) return fail(e, "hipMalloc(points)");
    if ((e = hipMalloc(&b->qseg, sizeof(double) * qsofs)) != hipSuccess) return fail(e, "hipMalloc(qseg)");
    if ((e = hipMalloc(&b->slots, sizeof(double) * slofs)) != hipSuccess) return fail(e, "hipMalloc(slots)");
    if ((e = hipMalloc(&b->progress, sizeof(int) * bofs)) != hipSuccess) return fail(e, "hipMalloc(progress)");
    if ((e = hipMalloc(&b->flags64, sizeof(unsigned long long) * bofs)) != hipSuccess) return fail(e, "hipMalloc(flags64)");
    if ((e = hipMalloc(&b->items, sizeof(int2) * bofs)) != hipSuccess) return fail(e, "hipMalloc(items)");
    if (!hbitems.empty()) {
        if ((e = hipMalloc(&b->bitems, sizeof(int2) * hbitems.size())) != hipSuccess) return fail(e, "hipMalloc(bound items)");
        if ((e = hipMemcpy(b->bitems, hbitems.data(), sizeof(int2) * hbitems.size(), hipMemcpyHostToDevice)) != hipSuccess)
            return fail(e, "hipMemcpy(bound items)");
    }
    if ((e = hipMalloc(&b->count, sizeof(int) * 2)) != hipSuccess) return fail(e, "hipMalloc(count)");
    if ((e = hipMalloc(&b->evals, sizeof(unsigned long long))) != hipSuccess) return fail(e, "hipMalloc(evals)");
    if ((e = hipMalloc(&b->t0, sizeof(unsigned long long))) != hipSuccess) return fail(e, "hipMalloc(t0)");
    if ((e = hipMalloc(&b->abort, sizeof(int))) != hipSuccess) return fail(e, "hipMalloc(abort)");
    if ((e = hipMemset(b->abort, 0, sizeof(int))) != hipSuccess) return fail(e, "hipMemset(abort)");
    if ((e = hipMalloc(&b->feat, sizeof(float) * npairs * rfd::MLP_FEAT_ROW)) != hipSuccess) return fail(e, "hipMalloc(features)");
    if ((e = hipMalloc(&b->xin, sizeof(uint16_t) * npairs * rfd::MLP_IN_ROW)) != hipSuccess) return fail(e, "hipMalloc(net input)");
    if ((e = hipMalloc(&b->pred, sizeof(float) * npairs)) != hipSuccess) return fail(e, "hipMalloc(predictions)");
    if ((e = hipMemcpy(b->pairs, b->hpairs.data(), sizeof(rfd::BatchPair) * npairs, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy(pairs)");
    if ((e = hipMemcpy(b->d_order, b->order.data(), sizeof(int) * npairs, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy(order)");
    if ((e = hipMemcpy(b->pts, hpts.data(), sizeof(double) * pofs, hipMemcpyHostToDevice)) != hipSuccess)
        return fail(e, "hipMemcpy(points)");
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        auto it = g_tune_cache.find(b->signature);
        if (it != g_tune_cache.end()) b->tuned_rpl = it->second;
    }
    *out = b;
    return RFD_OK;
}

rfd_status rfd_batch_tune(rfd_batch *b, const rfd_opts *opts, int32_t *rows_per_lane, double *ms) {
    if (!b) return RFD_ERR_ARG;
    rfd_opts o;
    if (opts) o = *opts; else rfd_default_opts(&o);
    rfd_ctx *ctx = b->ctx;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    const int cand[4] = {1, 2, 3, 4};
    double best = INFINITY;
    int best_r = 4;
    std::vector<rfd::BatchState> hst;
    for (int c = 0; c < 4; ++c) {
        unsigned long long t0 = 0;
        float t = 0;
        // Warm (graph instantiation, first-touch) then time.
        rfd_status s = run_once(b, o, cand[c], resolve_prune(o), hst, t0, t);
        if (s != RFD_OK) return s;
        s = run_once(b, o, cand[c], resolve_prune(o), hst, t0, t);
        if (s != RFD_OK) return s;
        if (ms) ms[c] = t;
        if (t < best) {
            best = t;
            best_r = cand[c];
        }
    }
    b->tuned_rpl = best_r;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune_cache[b->signature] = best_r;
    }
    if (rows_per_lane) *rows_per_lane = best_r;
    return RFD_OK;
}

// Each pair's decision form from opts->decide_form (RFD_DECIDE_REFERENCE:
// the reference's free-memory rule for that pair's size); the pair records
// are uploaded again when a form changes (the captured graph reads them from
// device memory).
static rfd_status apply_forms(rfd_batch *b, const rfd_opts &o) {
    rfd_ctx *ctx = b->ctx;
    if (o.decide_form < RFD_DECIDE_FUSED || o.decide_form > RFD_DECIDE_REFERENCE) return RFD_ERR_ARG;
    bool changed = false;
    for (auto &p : b->hpairs) {
        int f = rfd::DEC_FUSED;
        if (o.decide_form == RFD_DECIDE_TWO_PHASE) f = rfd::DEC_TWO_PHASE;
        else if (o.decide_form == RFD_DECIDE_REFERENCE) f = rfd_host::reference_form(ctx, p.m, p.n);
        if (p.dec_form != f) {
            p.dec_form = f;
            changed = true;
        }
    }
    if (changed)
        RFD_HIP(ctx, hipMemcpy(b->pairs, b->hpairs.data(), sizeof(rfd::BatchPair) * b->npairs,
                               hipMemcpyHostToDevice));
    return RFD_OK;
}

static rfd_status batch_run_impl(rfd_batch *b, const rfd_opts *opts, rfd_result *results,
                                 double *latency_ms, rfd_mlp *net, double tau, double delta,
                                 int32_t *certified) {
    if (!b || !results) return RFD_ERR_ARG;
    rfd_opts o;
    if (opts) o = *opts; else rfd_default_opts(&o);
    rfd_ctx *ctx = b->ctx;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    if (rfd_status fs = apply_forms(b, o); fs != RFD_OK) return fs;
    int rpl = o.rows_per_lane;
    if (rpl < 1 || rpl > 4) {
        if (!b->tuned_rpl) {
            rfd_status s = rfd_batch_tune(b, &o, nullptr, nullptr);
            if (s != RFD_OK) return s;
        }
        rpl = b->tuned_rpl;
    }
    std::vector<rfd::BatchState> hst;
    unsigned long long t0 = 0;
    float ms = 0;
    unsigned long long ev0 = 0;
    rfd_status s = run_once(b, o, rpl, resolve_prune(o), hst, t0, ms, net, tau, delta);
    if (s != RFD_OK) return s;
    RFD_HIP(ctx, hipMemcpy(&ev0, b->evals, sizeof(ev0), hipMemcpyDeviceToHost));
    double cells = 0;
    for (int64_t k = 0; k < b->npairs; ++k) {
        const rfd::BatchState &st = hst[(size_t)k];
        const rfd::BatchPair &p = b->hpairs[(size_t)k];
        rfd_result &r = results[k];
        std::memset(&r, 0, sizeof(r));
        r.distance = st.distance;
        r.endpoint_lb = p.endpoint_lb;
        r.dfd = st.dfd;
        r.vefd = st.vefd;
        r.lb = st.lb;
        r.ub = st.ub;
        r.err_bound = rfd_host::error_bound(st.lb, st.ub, st.distance);
        r.iterations = st.iterations;
        r.ms_total = st.t_done >= t0 ? (double)(st.t_done - t0) / ctx->wall_khz : 0.0;
        if (latency_ms) latency_ms[k] = r.ms_total;
        if (certified) certified[k] = st.certified;
        const double inner = (double)(p.m - 1) * (double)(p.n - 1);
        const bool vefd_ran = p.m >= 2 && p.n >= 2 && (!o.vefd_gate || p.big);
        cells += (double)p.m * (double)p.n + (vefd_ran ? inner : 0.0) + inner * (double)st.iterations;
    }
    ctx->stats.launches[RFD_K_BATCH] += 1;
    ctx->stats.ms[RFD_K_BATCH] += ms;
    ctx->stats.cells[RFD_K_BATCH] += cells;
    ctx->stats.evaluated[RFD_K_BATCH] += (double)ev0;
    return RFD_OK;
}

rfd_status rfd_batch_run(rfd_batch *b, const rfd_opts *opts, rfd_result *results, double *latency_ms) {
    return batch_run_impl(b, opts, results, latency_ms, nullptr, 0.0, 0.0, nullptr);
}

rfd_status rfd_batch_run_learned(rfd_batch *b, rfd_mlp *net, const rfd_opts *opts,
                                 rfd_result *results, double *latency_ms, int32_t *certified) {
    if (!b || !net || net->ctx != b->ctx) return RFD_ERR_ARG;
    return batch_run_impl(b, opts, results, latency_ms, net, net->tau, net->delta, certified);
}

rfd_status rfd_batch_features(rfd_batch *b, float *out) {
    if (!b || !out) return RFD_ERR_ARG;
    rfd_ctx *ctx = b->ctx;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    rfd_opts o;
    rfd_default_opts(&o);
    rfd::BatchArgs a = make_args(b, 1, o, 0);
    hipLaunchKernelGGL(rfd::batch_features, dim3((unsigned)b->npairs), dim3(128), 0, ctx->stream, a,
                       b->feat, b->xin);
    RFD_HIP(ctx, hipGetLastError());
    RFD_HIP(ctx, hipMemcpyAsync(out, b->feat, sizeof(float) * b->npairs * rfd::MLP_FEAT_ROW,
                                hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RFD_OK;
}

}  // extern "C"
