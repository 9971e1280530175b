// The batch bound round's kernel: the single-pair fused bound sweep run per
// (pair, band) item.  Instantiated in rfd_bounds.hip only.
#pragma once

#include "batch_types.hpp"
#include "bounds_kernels.hpp"

namespace rfd {

// Fused bound round: workgroups take (pair, band) items in order and run the
// band with the single-pair fused code; a pair's DFD lands in its node-0
// result, its VE-FD in node 1's, and the band hand-off uses nodes 0 and 1's
// boundary columns (4n values).  The items come from the batch's fixed list
// when it has one (every pair takes part in this round), else from the plan.
// Pairs of at most two dimensions run the two-dimensional producers over the
// stored 3D points: their third coordinate is 0, and each term it would add is
// fma(0, 0, acc) = acc, so the results are the same bits for a third less
// point arithmetic.
template <int R, int NPW>
__global__ __launch_bounds__(64 * (NPW + 2)) void batch_bounds(BatchArgs a) {
    __shared__ BoundsShared<R, NPW> sh;
    const int total = __hip_atomic_load(&a.count[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        if (threadIdx.x == 0) {
            sh.item = atomicAdd(&a.count[1], 1);
            bounds_clear(sh);
        }
        __syncthreads();
        const int idx = sh.item;
        if (idx >= total) break;
        const int2 it = a.bitems ? a.bitems[idx] : a.items[idx];
        const int k = it.x / kSpec;
        const BatchPair p = a.pairs[k];
        BoundsArgs b;
        b.P = a.pts + p.poff;
        b.Q = a.pts + p.qoff;
        b.slots = a.slots + p.sloff;
        b.progress = a.progress + p.bandoff;
        b.queue = nullptr;
        b.result = &a.state[k].result[0][0];
        b.rows = p.m;
        b.cols = p.n;
        b.nbands = (int)((p.m + 64LL * R - 1) / (64LL * R));
        b.r2 = p.endpoint_lb * p.endpoint_lb;
        b.tlog = nullptr;
        b.watch = a.watch;
        if (p.dims <= 2) {
            bounds_band<2, R, NPW, 1, 3>(b, it.y, sh);
        } else {
            bounds_band<3, R, NPW>(b, it.y, sh);
        }
        __syncthreads();
    }
}

}  // namespace rfd
