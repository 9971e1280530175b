// Launches of the windowed decision sweep (one pair, FP64).  Configuration
// codes: 10 * rows per lane + producer waves (one chain and one courier wave
// besides): 12, 16, 23, 26, 36, 46.
#include "rfd_internal.hpp"
#include "dec_kernels.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace rfd {

// Flags, queue, result and counters of one windowed decision sweep.
__global__ void dec_init(unsigned long long *flags, int nbands, int *queue, double *result, double sentinel,
                         unsigned long long *evals, int *abort) {
    const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int k = tid; k < nbands; k += stride) flags[k] = flag64(0u, kLoUnknown);
    if (tid == 0) {
        *queue = 0;
        result[0] = sentinel;
        evals[0] = 0;
        evals[1] = 0;
        *abort = 0;
    }
}

}  // namespace rfd

namespace {

template <int D, int R, int NP, bool VAR>
hipError_t launch_dec_cfg(rfd_ctx *ctx, int64_t nb, const rfd::DecArgs &a, int *grid_out) {
    auto kern = rfd::dec_sweep<D, double, R, NP, VAR>;
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * (NP + 2), 0);
    if (e != hipSuccess) return e;
    per_cu = std::max(per_cu, 1);
    // Workgroups per CU: a band whose CU is shared steps slower, and the
    // bands above it wait on it (RFD_DEC_WGCU caps them; default one).
    if (ctx->dec_wg_per_cu > 0) per_cu = std::min(per_cu, ctx->dec_wg_per_cu);
    int64_t grid = std::min<int64_t>(nb, (int64_t)ctx->cu_count * per_cu);
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, ctx->max_waves));  // RFD_MAX_WAVES caps workgroups too
    *grid_out = (int)grid;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * (NP + 2)), 0, ctx->stream, a);
    return hipGetLastError();
}

template <int D, bool VAR>
hipError_t launch_dec_d(rfd_ctx *ctx, int cfg, int64_t nb, const rfd::DecArgs &a, int *grid) {
    switch (cfg) {
        case 12: return launch_dec_cfg<D, 1, 2, VAR>(ctx, nb, a, grid);
        case 16: return launch_dec_cfg<D, 1, 6, VAR>(ctx, nb, a, grid);
        case 23: return launch_dec_cfg<D, 2, 3, VAR>(ctx, nb, a, grid);
        case 36: return launch_dec_cfg<D, 3, 6, VAR>(ctx, nb, a, grid);
        case 46: return launch_dec_cfg<D, 4, 6, VAR>(ctx, nb, a, grid);
        default: return launch_dec_cfg<D, 2, 6, VAR>(ctx, nb, a, grid);
    }
}

}  // namespace

namespace rfd_host {

bool rfd_valid_dec_cfg(int cfg) {
    return cfg == 12 || cfg == 16 || cfg == 23 || cfg == 26 || cfg == 36 || cfg == 46;
}

hipError_t rfd_launch_dec(rfd_ctx *ctx, int dims, int cfg, bool two_phase, int64_t nb, const rfd::DecArgs &a,
                          int *grid) {
    if (dims == 2)
        return two_phase ? launch_dec_d<2, true>(ctx, cfg, nb, a, grid) : launch_dec_d<2, false>(ctx, cfg, nb, a, grid);
    return two_phase ? launch_dec_d<3, true>(ctx, cfg, nb, a, grid) : launch_dec_d<3, false>(ctx, cfg, nb, a, grid);
}

void rfd_launch_dec_init(rfd_ctx *ctx, const rfd::DecArgs &a, double sentinel, int *abort) {
    hipLaunchKernelGGL(rfd::dec_init, dim3((unsigned)std::min<int64_t>((a.nbands + 255) / 256, 1024)), dim3(256), 0,
                       ctx->stream, a.flags, a.nbands, a.queue, a.result, sentinel, a.evals, abort);
}

}  // namespace rfd_host
