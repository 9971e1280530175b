// Host-side internals shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>

#include "rfd.h"
#include "sweep_kernels.hpp"

struct rfd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    int cu_count = 0;
    int max_waves = 0;
    double wall_khz = 100000.0;   // device wall clock (wall_clock64) rate
    int sync = rfd::SYNC_SCOPED;  // band hand-off protocol (RFD_SYNC=0 selects acquire/release)
    int rows_per_lane = 0;        // R of band_sweep, 1-4 (RFD_ROWS); 0: chosen per sweep
    bool prune = false;           // VE-FD skips cells above DFD^2 (RFD_PRUNE=1 enables)
    bool fused = true;            // FP64 bounds by the fused DFD + VE-FD sweep (RFD_FUSED=0 disables)
    int bounds_cfg = 0;           // fused sweep: 10 * rows per lane + producer waves, 0 = by size (RFD_BOUNDS)
    int wg_per_cu = 0;            // fused sweep: cap on workgroups per CU, 0 = none (RFD_WGCU)
    unsigned spin_limit = 1u << 22;  // polls one hand-off wait may take before it gives up (RFD_SPIN_LIMIT)
    int batch_bounds = 1310;      // batches: fused bound round configuration, 0 = separate rounds (RFD_BATCH_BOUNDS)
    int dec_cfg = -1;             // FP64 decisions: windowed sweep, 10 * rows per lane + producer waves; -1 = by size,
                                  // 0 = band_sweep (RFD_DEC)
    int dec_wg_per_cu = 1;        // windowed decision sweep: workgroups per CU, 0 = as many as fit (RFD_DEC_WGCU)
    int batch_dec = 1;            // batch decision rounds on the windowed sweep at the batch's rows per lane; 0 = band_sweep (RFD_BATCH_DEC)
    int batch_dec_wgcu = 0;       // their workgroups per CU, 0 = as many as fit (RFD_BATCH_DEC_WGCU)
    int batch_dec_budget = 2;     // speculation: bands x nodes up to this many per workgroup (RFD_BATCH_DEC_BUDGET)
    int lds_per_cu = 0;           // bytes of LDS per CU
    const char *tlog_path = nullptr;  // RFD_TLOG: append per-band timestamps of every sweep
    void *scratch = nullptr;      // device address of the library's fixed working array on this device
    rfd_stats stats{};
    char err[512] = {0};
};

namespace rfd_host {

inline void set_err(rfd_ctx *ctx, const char *fmt, ...) {
    if (!ctx) return;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
}

}  // namespace rfd_host

namespace rfd {
struct BoundsArgs;
struct BatchArgs;
struct DecArgs;
}  // namespace rfd

namespace rfd_host {

// The fused bound sweep's launches (rfd_bounds.hip).  Configuration codes:
// 10 * rows per lane + producer waves (100 * rows + waves from 10 waves on),
// plus 20000 for two columns per step.
bool rfd_valid_bounds_cfg(int cfg);
hipError_t rfd_launch_bounds(rfd_ctx *ctx, int dims, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid);
// The fused sweep's DFD-only form (configurations 42, 77, 107, 1310).
bool rfd_valid_dfd_cfg(int cfg);
hipError_t rfd_launch_dfd(rfd_ctx *ctx, int dims, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid);
// The fused sweep with VE-FD in FP32 (configurations 77, 107, 1310).
bool rfd_valid_mixed_cfg(int cfg);
hipError_t rfd_launch_mixed(rfd_ctx *ctx, int dims, int cfg, int64_t nb, const rfd::BoundsArgs &a, int *grid);
bool rfd_valid_batch_bounds_cfg(int cfg);
void rfd_launch_batch_bounds(rfd_ctx *ctx, int cfg, hipStream_t st, const rfd::BatchArgs &a);

// The windowed decision sweep's launches (rfd_dec.hip).  Configuration codes:
// 10 * rows per lane + producer waves.
bool rfd_valid_dec_cfg(int cfg);
hipError_t rfd_launch_dec(rfd_ctx *ctx, int dims, int cfg, bool two_phase, int64_t nb, const rfd::DecArgs &a,
                          int *grid);
void rfd_launch_dec_init(rfd_ctx *ctx, const rfd::DecArgs &a, double sentinel, int *abort);

// The reference's choice of decision path for an m x n pair from free device
// memory (rfd_capi.hip): rfd::DEC_TWO_PHASE or rfd::DEC_FUSED.
int reference_form(rfd_ctx *ctx, int64_t m, int64_t n);

}  // namespace rfd_host

#define RFD_HIP(ctx, call)                                                          \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess) {                                                     \
            rfd_host::set_err((ctx), "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                    __FILE__, __LINE__);                                            \
            return RFD_ERR_HIP;                                                     \
        }                                                                           \
    } while (0)

namespace rfd_host {

// frechet.cpp:30 with Point::dist_sqr (geometry_basics.hpp:72-81): host
// arithmetic, each product rounded before the sum.
inline double endpoint_bound(const double *P, int64_t m, const double *Q, int64_t n, int dims) {
    auto d2 = [&](const double *a, const double *b) {
        double result = 0.0;
        for (int k = 0; k < dims; ++k) {
            volatile double ds = a[k] - b[k];
            volatile double sq = ds * ds;
            result = result + sq;
        }
        return result;
    };
    double a = d2(P, Q);
    double b = d2(P + (m - 1) * dims, Q + (n - 1) * dims);
    return std::sqrt(std::max(a, b));
}


// Relative error bound of a result v = (lb + ub) / 2 for a distance FD in
// [lb, ub]: |v - FD| / FD <= max(ub - v, v - lb) / lb.  The differences are
// exact (Sterbenz) and account for the rounding of v itself; 0 when the
// bracket is a point.
inline double error_bound(double lb, double ub, double v) {
    const double w = std::max(ub - v, v - lb);
    return w > 0.0 ? (lb > 0.0 ? w / lb : INFINITY) : 0.0;
}

// Rejects curves holding a NaN or infinite coordinate: the sweeps' min / max
// recursions and the fused bound sweep (compiled without NaN semantics) assume
// finite points.  pair >= 0 names a batch pair in the message.
inline rfd_status check_finite(rfd_ctx *ctx, const double *P, int64_t m, const double *Q, int64_t n,
                               int dims, int64_t pair) {
    for (int c = 0; c < 2; ++c) {
        const double *X = c ? Q : P;
        const int64_t len = (c ? n : m) * dims;
        for (int64_t k = 0; k < len; ++k) {
            if (!std::isfinite(X[k])) {
                if (pair >= 0)
                    set_err(ctx, "pair %lld: curve %c point %lld holds a non-finite coordinate", (long long)pair,
                            c ? 'Q' : 'P', (long long)(k / dims));
                else
                    set_err(ctx, "curve %c point %lld holds a non-finite coordinate", c ? 'Q' : 'P',
                            (long long)(k / dims));
                return RFD_ERR_ARG;
            }
        }
    }
    return RFD_OK;
}

}  // namespace rfd_host
