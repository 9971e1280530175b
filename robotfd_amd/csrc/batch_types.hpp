// Batch types shared by the batch kernels and the batch bound round: pair
// records, per-pair bisection state, round arguments and their device helpers.
// (The batched rounds themselves are described in batch_kernels.hpp.)
#pragma once

#include "sweep_kernels.hpp"
namespace rfd {
constexpr int MLP_FEAT_ROW = 386;  // features per pair (see mlp_kernels.hpp)
constexpr int MLP_IN_ROW = 400;    // padded network input
__device__ inline uint16_t f2bf_dev(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
}  // namespace rfd

namespace rfd {

// PH_BOUNDS: DFD and VE-FD of every pair in one fused round (bounds_kernels.hpp).
enum BatchPhase { PH_DFD = 0, PH_VEFD = 1, PH_VEFD_REDO = 2, PH_DEC = 3, PH_CERT = 4, PH_BOUNDS = 5 };

// Speculative bisection.  A decision round may evaluate the next d levels of a
// pair's bisection tree at once: node 0 splits (lb, ub); node i's children
// 2i+1 and 2i+2 split the lower and upper half it leaves.  Every node's radius
// is computed exactly as the sequential loop would compute it on reaching that
// node, so walking the tree with the decisions reproduces the sequential
// trajectory bit for bit, in d times fewer rounds.  d is chosen per round
// from how many bands the active pairs have: speculation fills waves that
// would otherwise idle when few pairs remain.
constexpr int kSpec = 8;      // tree nodes per pair per round (depth <= 3 uses 7)
constexpr int kMaxDepth = 3;

struct BatchPair {
    long long m, n;          // points
    long long poff, qoff;    // offsets (in doubles) of P and Q in the point pool (3 per point)
    long long qsoff;         // offset of the decision's per-column terms (2 per column)
    long long sloff;         // offset of the two boundary columns (2 * n)
    long long bandoff;       // offset of this pair's progress flags (kSpec * bandstride)
    long long bandstride;    // progress flags per tree node
    double endpoint_lb;      // frechet.cpp:30, computed on the host
    int big;                 // m*n > 10000 (.cu:1343, in 64 bits)
    int dec_form;            // decision interval form of this pair: DEC_FUSED or DEC_TWO_PHASE
    int dims;                // dimensions of the pair's points (stored 3D either way)
};

struct BatchState {
    double lb, ub, split, dfd, vefd, F, distance;
    double result[kSpec][2]; // terminal value of each tree node's last sweep (node 0: bounds)
    double radius[kSpec];    // decision radius of each tree node this round
    double thr;              // VE-FD prune threshold used
    long long iterations;
    unsigned long long t_done;  // wall clock when the pair's result became final
    int active;
    int need_redo;
    unsigned cert_mask;      // learned run: nodes 0 (lower radius) and 1 (upper radius) to certify
    int certified;           // learned run: the predicted bracket held
};

struct BatchArgs {
    const BatchPair *pairs;
    BatchState *state;
    const int *order;        // pair indices, longest first
    int npairs;
    const double *pts;
    const double *qseg;
    double *slots;
    int *progress;
    unsigned long long *flags64;  // windowed decision rounds: band words, indexed as progress
    int2 *items;             // (pair, band) work list of the current round
    const int2 *bitems;      // the fused bound round's work list, fixed per batch (or null)
    int *count;              // [0] items, [1] queue, [2] speculation depth of the round
    unsigned long long *evals;
    unsigned long long *t0;  // wall clock at batch start
    int rows_per_lane;
    int bounds_rows;         // rows per lane of the fused bound round
    int stop_mode;
    double eps;
    int prune;
    int gate;                // 1: VE-FD only when m*n > 10000
    int max_waves;           // waves of a sweep launch (sets the speculation depth)
    int dec_win;             // decision rounds on the windowed sweep: 10 * rows per lane + producer waves; 0: band_sweep
    int max_depth;           // 1 = plain sequential bisection
    Watch watch;             // bound on the hand-off waits (abort word cleared by batch_reset)
};

__device__ inline bool batch_keep(double lb, double ub, int stop_mode, double eps) {
    return stop_mode == 0 ? (ub - lb > eps) : (ub - lb > eps * ub);
}

__device__ inline long long batch_rows(const BatchPair &p, int phase) {
    return (phase == PH_DFD || phase == PH_BOUNDS) ? p.m : p.m - 1;
}

__device__ inline bool batch_runs_vefd(const BatchPair &p, int gate) {
    return p.m >= 2 && p.n >= 2 && (!gate || p.big);
}

// Whether pair k takes part in a round of `phase`.
__device__ inline bool batch_in_round(const BatchPair &p, const BatchState &st, int phase, int gate) {
    if (phase == PH_DFD || phase == PH_BOUNDS) return true;
    if (p.m < 2 || p.n < 2) return false;
    if (phase == PH_VEFD) return batch_runs_vefd(p, gate);
    if (phase == PH_VEFD_REDO) return st.need_redo != 0;
    if (phase == PH_CERT) return st.active != 0 && st.cert_mask != 0;
    return st.active != 0;
}

}  // namespace rfd
