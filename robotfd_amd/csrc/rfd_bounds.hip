// Launches of the fused DFD + VE-FD bound sweep (one pair, and the batch
// bound round).  The kernels are instantiated in this translation unit only,
// which is compiled with -fno-honor-nans: no value on the fused bound path is
// NaN (every weight is a sum of squares or a clamped projection whose divide
// runs only for 0 < apab < |ab|^2, and the recursion takes min / max of those
// and of +-FLT_MAX), so fmin / fmax lower to bare v_min_f64 / v_max_f64 whose
// latency the scheduler knows.  The decision and free-interval code, which
// relies on fmin / fmax of NaN, stays in the other translation units.
#define RFD_BOUNDS_NO_NANS 1

#include "rfd_internal.hpp"
#include "bounds_kernels.hpp"
#include "batch_bounds.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace {

// Configurations: rows per lane R and producer waves NPW (one chain wave and
// one courier wave besides), encoded as 10 * R + NPW (100 * R + NPW from 10
// producer waves on), plus 20000 for two columns per step.
template <int D, int R, int NPW, int C = 1, int MODE = rfd::kModeF64>
hipError_t launch_bounds_cfg(rfd_ctx *ctx, int64_t nb, const rfd::BoundsArgs &a, int *grid_out) {
    auto kern = rfd::bounds_sweep<D, R, NPW, C, MODE>;
    // Optional cap on workgroups per CU (RFD_WGCU): unused dynamic LDS sized so
    // that one more workgroup does not fit.  Bands on a CU shared by more
    // workgroups run slower, and every later band waits on the slowest.
    size_t dyn = 0;
    if (ctx->wg_per_cu > 0 && ctx->lds_per_cu > 0) {
        hipFuncAttributes fa;
        hipError_t e0 = hipFuncGetAttributes(&fa, reinterpret_cast<const void *>(kern));
        if (e0 != hipSuccess) return e0;
        const size_t want = (size_t)ctx->lds_per_cu / (size_t)(ctx->wg_per_cu + 1) + 1;
        if (want > fa.sharedSizeBytes) dyn = want - fa.sharedSizeBytes;
    }
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * (NPW + 2), dyn);
    if (e != hipSuccess) return e;
    int64_t grid = std::min<int64_t>(nb, (int64_t)ctx->cu_count * std::max(per_cu, 1));
    grid = std::max<int64_t>(1, std::min<int64_t>(grid, ctx->max_waves));  // RFD_MAX_WAVES caps workgroups too
    *grid_out = (int)grid;
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * (NPW + 2)), dyn, ctx->stream, a);
    return hipGetLastError();
}

template <int D>
hipError_t launch_bounds_d(rfd_ctx *ctx, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid) {
    switch (cfg) {
        case 21: return launch_bounds_cfg<D, 2, 1>(ctx, nb, a, grid);
        case 62: return launch_bounds_cfg<D, 6, 2>(ctx, nb, a, grid);
        case 63: return launch_bounds_cfg<D, 6, 3>(ctx, nb, a, grid);
        case 72: return launch_bounds_cfg<D, 7, 2>(ctx, nb, a, grid);
        case 73: return launch_bounds_cfg<D, 7, 3>(ctx, nb, a, grid);
        case 83: return launch_bounds_cfg<D, 8, 3>(ctx, nb, a, grid);
        case 84: return launch_bounds_cfg<D, 8, 4>(ctx, nb, a, grid);
        case 136: return launch_bounds_cfg<D, 13, 6>(ctx, nb, a, grid);
        case 137: return launch_bounds_cfg<D, 13, 7>(ctx, nb, a, grid);
        case 77: return launch_bounds_cfg<D, 7, 7>(ctx, nb, a, grid);
        case 107: return launch_bounds_cfg<D, 10, 7>(ctx, nb, a, grid);
        case 1310: return launch_bounds_cfg<D, 13, 10>(ctx, nb, a, grid);
        case 1313: return launch_bounds_cfg<D, 13, 13>(ctx, nb, a, grid);
        case 1012: return launch_bounds_cfg<D, 10, 10>(ctx, nb, a, grid);
        case 20077: return launch_bounds_cfg<D, 7, 7, 2>(ctx, nb, a, grid);
        case 20107: return launch_bounds_cfg<D, 10, 7, 2>(ctx, nb, a, grid);
        case 20137: return launch_bounds_cfg<D, 13, 7, 2>(ctx, nb, a, grid);
        case 21310: return launch_bounds_cfg<D, 13, 10, 2>(ctx, nb, a, grid);
        default: return launch_bounds_cfg<D, 4, 2>(ctx, nb, a, grid);
    }
}

// The DFD-only sweep (one weight per cell) in the default configurations.
template <int D>
hipError_t launch_dfd_d(rfd_ctx *ctx, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid) {
    switch (cfg) {
        case 42: return launch_bounds_cfg<D, 4, 2, 1, rfd::kModeDfd>(ctx, nb, a, grid);
        case 77: return launch_bounds_cfg<D, 7, 7, 1, rfd::kModeDfd>(ctx, nb, a, grid);
        case 107: return launch_bounds_cfg<D, 10, 7, 1, rfd::kModeDfd>(ctx, nb, a, grid);
        case 1310: return launch_bounds_cfg<D, 13, 10, 1, rfd::kModeDfd>(ctx, nb, a, grid);
        default: return hipErrorInvalidValue;
    }
}

// DFD in FP64 and VE-FD in FP32, in the configurations with staged column
// terms.
template <int D>
hipError_t launch_mixed_d(rfd_ctx *ctx, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid) {
    switch (cfg) {
        case 77: return launch_bounds_cfg<D, 7, 7, 1, rfd::kModeV32>(ctx, nb, a, grid);
        case 107: return launch_bounds_cfg<D, 10, 7, 1, rfd::kModeV32>(ctx, nb, a, grid);
        case 1310: return launch_bounds_cfg<D, 13, 10, 1, rfd::kModeV32>(ctx, nb, a, grid);
        default: return hipErrorInvalidValue;
    }
}

template <int R, int NPW>
void launch_batch_bounds_cfg(rfd_ctx *ctx, hipStream_t st, const rfd::BatchArgs &a) {
    auto kern = rfd::batch_bounds<R, NPW>;
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * (NPW + 2), 0) != hipSuccess || per_cu < 1)
        per_cu = 1;
    const int grid = std::max(1, std::min(ctx->cu_count * per_cu, ctx->max_waves));
    hipLaunchKernelGGL(kern, dim3((unsigned)grid), dim3(64 * (NPW + 2)), 0, st, a);
}

}  // namespace

namespace rfd_host {

bool rfd_valid_bounds_cfg(int cfg) {
    return cfg == 42 || cfg == 62 || cfg == 63 || cfg == 72 || cfg == 73 || cfg == 83 || cfg == 84 || cfg == 21 ||
           cfg == 136 || cfg == 137 || cfg == 77 || cfg == 107 || cfg == 1310 || cfg == 1313 || cfg == 1012 ||
           cfg == 20077 || cfg == 20107 || cfg == 20137 || cfg == 21310;
}

hipError_t rfd_launch_bounds(rfd_ctx *ctx, int dims, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid) {
    return dims == 2 ? launch_bounds_d<2>(ctx, cfg, nb, a, grid) : launch_bounds_d<3>(ctx, cfg, nb, a, grid);
}

bool rfd_valid_dfd_cfg(int cfg) { return cfg == 42 || cfg == 77 || cfg == 107 || cfg == 1310; }

bool rfd_valid_mixed_cfg(int cfg) { return cfg == 77 || cfg == 107 || cfg == 1310; }

hipError_t rfd_launch_mixed(rfd_ctx *ctx, int dims, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid) {
    return dims == 2 ? launch_mixed_d<2>(ctx, cfg, nb, a, grid) : launch_mixed_d<3>(ctx, cfg, nb, a, grid);
}

hipError_t rfd_launch_dfd(rfd_ctx *ctx, int dims, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid) {
    return dims == 2 ? launch_dfd_d<2>(ctx, cfg, nb, a, grid) : launch_dfd_d<3>(ctx, cfg, nb, a, grid);
}

bool rfd_valid_batch_bounds_cfg(int cfg) {
    return cfg == 21 || cfg == 42 || cfg == 63 || cfg == 77 || cfg == 137 || cfg == 1310;
}

void rfd_launch_batch_bounds(rfd_ctx *ctx, int cfg, hipStream_t st, const rfd::BatchArgs &a) {
    switch (cfg) {
        case 21: launch_batch_bounds_cfg<2, 1>(ctx, st, a); break;
        case 63: launch_batch_bounds_cfg<6, 3>(ctx, st, a); break;
        case 137: launch_batch_bounds_cfg<13, 7>(ctx, st, a); break;
        case 77: launch_batch_bounds_cfg<7, 7>(ctx, st, a); break;
        case 1310: launch_batch_bounds_cfg<13, 10>(ctx, st, a); break;
        default: launch_batch_bounds_cfg<4, 2>(ctx, st, a); break;
    }
}

}  // namespace rfd_host
