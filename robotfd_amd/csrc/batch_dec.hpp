// Batch decision rounds on the windowed decision sweep (dec_kernels.hpp):
// workgroups take (pair x tree node, band) items from the round's work list in
// order and run the band with the chain, producer and courier waves of the
// single-pair kernel.  A pair's bands are listed consecutively, so every band
// a workgroup waits on has been taken already (no deadlock for any grid).
// Each (pair, node) keeps its own band words (flags64, indexed as the
// progress flags), boundary columns and result slot.  Points are stored 3D;
// a 2D pair has z = 0, which leaves every free interval bit-identical.
#pragma once

#include "batch_types.hpp"
#include "dec_kernels.hpp"

namespace rfd {

template <int R, int NP>
__global__ __launch_bounds__(64 * (NP + 2)) void batch_dec(BatchArgs ba) {
    __shared__ DecShared<3, double, R, NP> sh;
    __shared__ int2 item_s;
    unsigned long long evals = 0, skips = 0;
    const int total = __hip_atomic_load(&ba.count[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (threadIdx.x == 0) {
            const int idx = atomicAdd(&ba.count[1], 1);
            sh.item = idx;
            item_s = idx < total ? ba.items[idx] : make_int2(0, 0);
        }
        dec_clear(sh);
        __syncthreads();
        if (sh.item >= total) break;
        const int kn = item_s.x, band = item_s.y;
        const int k = kn / kSpec, node = kn % kSpec;
        const BatchPair p = ba.pairs[k];
        DecArgs a{};
        a.P = ba.pts + p.poff;
        a.Q = ba.pts + p.qoff;
        a.qseg = ba.qseg + p.qsoff;
        a.slots = ba.slots + p.sloff + (long long)node * 2 * p.n;
        a.flags = ba.flags64 + p.bandoff + (long long)node * p.bandstride;
        a.queue = nullptr;
        a.result = ba.state[k].result[node];
        a.rows = p.m - 1;
        a.cols = p.n - 1;
        a.nbands = (int)((a.rows + 64LL * R - 1) / (64LL * R));
        a.r = ba.state[k].radius[node];
        a.evals = nullptr;
        a.tlog = nullptr;
        a.watch = ba.watch;
        // The pair's interval form (uniform over the workgroup).
        if (p.dec_form == DEC_TWO_PHASE)
            dec_roles<3, double, R, NP, true>(a, band, sh, evals, skips);
        else
            dec_roles<3, double, R, NP, false>(a, band, sh, evals, skips);
        __syncthreads();
    }
    if (threadIdx.x < 64 && ba.evals) atomicAdd(ba.evals, evals);
}

}  // namespace rfd
