// Arguments and band hand-off word of the windowed decision sweep
// (dec_kernels.hpp), shared with the host code that launches it.
#pragma once

#include "sweep_kernels.hpp"

namespace rfd {

struct DecArgs {
    const void *P, *Q;          // AoS points of the row and column curves (T)
    const void *qseg;           // per column j: {4 |q_{j+1}-q_j|^2, 0.5 / |q_{j+1}-q_j|^2} (T)
    void *slots;                // 2 * cols values of T: band-boundary columns
    unsigned long long *flags;  // per band: final columns (low 32 bits), lo (high 32 bits)
    int *queue;                 // next band to hand out
    double *result;             // terminal rightUB
    long long rows, cols;       // m - 1, n - 1
    int nbands;
    double r;                   // radius
    unsigned long long *evals;  // [0] cells with a reachable input, [1] steps not run / 16
    unsigned long long *tlog;   // optional, 12 per band: courier start, chain start, end, hw id,
                                //  first step, last step, ticks waiting on producers, on input,
                                //  on output room, courier publishes, courier input batches
    Watch watch;                // bound on the hand-off waits
};

constexpr unsigned kLoUnknown = 0x7fffffffu;
constexpr int kDecNone = 0x3fffffff;

__device__ inline unsigned long long flag64_load(const unsigned long long *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void flag64_store(unsigned long long *p, unsigned long long v) {
    __builtin_amdgcn_s_waitcnt(0);  // the column stores of every lane have completed
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline unsigned long long flag64(unsigned count, unsigned lo) {
    return ((unsigned long long)lo << 32) | (unsigned long long)count;
}

}  // namespace rfd
