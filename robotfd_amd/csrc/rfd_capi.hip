// C ABI of the gfx950 continuous Frechet distance library (include/rfd.h).
//
// Host side of the path: sessions hold both curves resident in HBM; the FP64
// bound sweeps (DFD, VE-FD) run as one fused sweep (bounds_kernels.hpp), the
// decision sweep and the FP32 sweeps as band-wavefront kernels
// (sweep_kernels.hpp); the bisection of frechet.cpp:18-75 runs on the host
// between decision sweeps.
#include "rfd.h"
#include "rfd_internal.hpp"
#include "sweep_kernels.hpp"
#include "bounds_kernels.hpp"
#include "dec_types.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

struct rfd_session {
    rfd_ctx *ctx = nullptr;
    int64_t m = 0, n = 0;
    int dims = 0;
    double *dP = nullptr, *dQ = nullptr;
    float *fP = nullptr, *fQ = nullptr;
    double *qseg64 = nullptr;
    float *qseg32 = nullptr;
    void *slots = nullptr;
    int *progress = nullptr;
    int *queue = nullptr;
    double *result = nullptr;            // [0] terminal value, [1] its square; fused bounds: [0..3]
    unsigned long long *evals = nullptr; // cells evaluated by the last sweep
    int *abort = nullptr;                // raised by a hand-off wait that gave up (rfd::Watch)
    double endpoint_lb = 0.0;
    double dfd_sq = -1.0;                // DFD^2 of this pair once known (prunes VE-FD)
    int dec_form = 0;                    // rfd::DEC_FUSED or rfd::DEC_TWO_PHASE
    bool noalloc = false;                // buffers in the library's fixed array: allocate nothing
};

namespace rfd {

__global__ void sweep_init(int *progress, int nbands, int *queue, double *result, int nres,
                           double sentinel, unsigned long long *evals, int *abort) {
    const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int k = tid; k < nbands; k += stride) progress[k] = 0;
    if (tid == 0) {
        *queue = 0;
        for (int k = 0; k < nres; ++k) result[k] = sentinel;
        evals[0] = 0;
        evals[1] = 0;
        *abort = 0;
    }
}

__global__ void to_f32(const double *src, float *dst, long long count) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) dst[k] = (float)src[k];
}

// One level of midpoint refinement: dst gets src's points with a + 0.5 (b - a)
// between consecutive points a, b (2 len - 1 points).
__global__ void refine_level(const double *src, long long len, int dims, double *dst) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= (2 * len - 1) * dims) return;
    const long long pt = k / dims, c = k % dims, i = pt >> 1;
    const double a = src[i * dims + c];
    dst[k] = (pt & 1) ? a + 0.5 * (src[(i + 1) * dims + c] - a) : a;
}

// flag[0] <- 0 (a kernel, not a memset, so that the call stays capturable).
__global__ void clear_flag(int *flag) { *flag = 0; }

// flag[0] <- 1 if any of x[0 .. count) is NaN or infinite.
__global__ void scan_finite(const double *x, long long count, int *flag) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (long long)gridDim.x * blockDim.x)
        if (!__builtin_isfinite(x[k])) *flag = 1;
}

}  // namespace rfd

// ---- Fixed working array of rfd_distance_noalloc ----
//
// Every buffer a distance computation needs beyond its inputs and outputs
// (band boundary columns, the decision's per-column terms, progress flags, the
// band queue, the terminal values, the evaluated-cell count, the abort word
// and a flag) lives in this array, so that the entry point allocates nothing.
// Each GPU the library runs on holds its own copy until the process ends; its
// address on a device comes from hipGetSymbolAddress with that device current
// (rfd_create), never from the array's name, which host code sees only as a
// stand-in.  Layout (bytes): [0, 256) words and terminal values, then the
// progress flags, the boundary columns (4 n doubles) and the per-column terms
// (2 (n - 1) doubles) for n <= RFD_NOALLOC_MAX_POINTS.
namespace {
constexpr long long kNaMax = RFD_NOALLOC_MAX_POINTS;
constexpr long long kNaProgressOff = 256;
constexpr long long kNaProgressBytes = ((kNaMax / 64 + 2) * 4 + 255) / 256 * 256;
constexpr long long kNaSlotsOff = kNaProgressOff + kNaProgressBytes;
constexpr long long kNaQsegOff = kNaSlotsOff + 4 * kNaMax * 8;
constexpr long long kNaBytes = kNaQsegOff + 2 * kNaMax * 8;
static_assert(kNaBytes <= RFD_NOALLOC_BYTES, "fixed working array too small");
}  // namespace

static __device__ __align__(256) double scratch[RFD_NOALLOC_BYTES / 8];

namespace {
std::mutex g_scratch_mu;
void *g_scratch_ptr[64] = {};  // per device: the array's device address

// The array's address on `device` (current), looked up once per device.
hipError_t scratch_address(int device, void **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    if (!g_scratch_ptr[device]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return hipSuccess;
}
}  // namespace

namespace {

using rfd_host::set_err;
using rfd_host::endpoint_bound;
using rfd_host::rfd_launch_bounds;
using rfd_host::rfd_launch_dfd;
using rfd_host::rfd_valid_dfd_cfg;
using rfd_host::rfd_launch_mixed;
using rfd_host::rfd_valid_mixed_cfg;
using rfd_host::rfd_valid_batch_bounds_cfg;
using rfd_host::rfd_valid_bounds_cfg;

int64_t nbands_for(int64_t rows) { return (rows + 63) / 64; }

int grid_for(rfd_ctx *ctx, int64_t nb) {
    int64_t g = std::min<int64_t>(nb, ctx->max_waves);
    return (int)std::max<int64_t>(g, 1);
}

template <int OP, typename T, int SYNC, int R, bool PRUNE>
void launch_sweep_rp(int dims, int grid, hipStream_t st, const rfd::SweepArgs &a) {
    if (dims == 2)
        hipLaunchKernelGGL((rfd::band_sweep<OP, 2, T, SYNC, R, PRUNE>), dim3(grid), dim3(64), 0, st, a);
    else
        hipLaunchKernelGGL((rfd::band_sweep<OP, 3, T, SYNC, R, PRUNE>), dim3(grid), dim3(64), 0, st, a);
}

// VE-FD with a finite threshold runs the pruning instantiation, and the
// decision in the two-phase form its variant; every other sweep runs the plain
// instantiation.
template <int OP, typename T, int SYNC, int R>
void launch_sweep_r(int dims, int grid, hipStream_t st, const rfd::SweepArgs &a) {
    if ((OP == rfd::OP_VEFD && std::isfinite(a.thr)) || (OP == rfd::OP_DEC && a.dec_form == rfd::DEC_TWO_PHASE))
        launch_sweep_rp<OP, T, SYNC, R, true>(dims, grid, st, a);
    else
        launch_sweep_rp<OP, T, SYNC, R, false>(dims, grid, st, a);
}

// Rows per lane the kernels are instantiated for.
bool valid_rows_per_lane(int r) { return r >= 1 && r <= 4; }

template <int OP, typename T>
hipError_t launch_sweep(int dims, int grid, int sync, int rpl, hipStream_t st,
                        const rfd::SweepArgs &a) {
    if (sync == rfd::SYNC_ACQREL) {
        launch_sweep_r<OP, T, rfd::SYNC_ACQREL, 1>(dims, grid, st, a);
    } else if (rpl == 4) {
        launch_sweep_r<OP, T, rfd::SYNC_SCOPED, 4>(dims, grid, st, a);
    } else if (rpl == 3) {
        launch_sweep_r<OP, T, rfd::SYNC_SCOPED, 3>(dims, grid, st, a);
    } else if (rpl == 2) {
        launch_sweep_r<OP, T, rfd::SYNC_SCOPED, 2>(dims, grid, st, a);
    } else {
        launch_sweep_r<OP, T, rfd::SYNC_SCOPED, 1>(dims, grid, st, a);
    }
    return hipGetLastError();
}

rfd_status ensure_f32(rfd_session *s) {
    rfd_ctx *ctx = s->ctx;
    if (s->fP) return RFD_OK;
    RFD_HIP(ctx, hipMalloc(&s->fP, sizeof(float) * s->m * s->dims));
    RFD_HIP(ctx, hipMalloc(&s->fQ, sizeof(float) * s->n * s->dims));
    long long cp = s->m * s->dims, cq = s->n * s->dims;
    hipLaunchKernelGGL(rfd::to_f32, dim3((unsigned)((cp + 255) / 256)), dim3(256), 0, ctx->stream,
                       s->dP, s->fP, cp);
    hipLaunchKernelGGL(rfd::to_f32, dim3((unsigned)((cq + 255) / 256)), dim3(256), 0, ctx->stream,
                       s->dQ, s->fQ, cq);
    RFD_HIP(ctx, hipGetLastError());
    return RFD_OK;
}

template <typename T>
rfd_status ensure_qseg(rfd_session *s, T **out, const T *Qd) {
    rfd_ctx *ctx = s->ctx;
    if (*out) return RFD_OK;
    long long cols = s->n - 1;
    RFD_HIP(ctx, hipMalloc(out, sizeof(T) * 2 * cols));
    unsigned g = (unsigned)((cols + 255) / 256);
    if (s->dims == 2)
        hipLaunchKernelGGL((rfd::qseg_terms<2, T>), dim3(g), dim3(256), 0, ctx->stream, Qd, cols, *out);
    else
        hipLaunchKernelGGL((rfd::qseg_terms<3, T>), dim3(g), dim3(256), 0, ctx->stream, Qd, cols, *out);
    RFD_HIP(ctx, hipGetLastError());
    return RFD_OK;
}

// Rows per lane of a band_sweep when the context does not fix them.  A band of
// 64 R rows trails its predecessor by about 80 steps, so fewer rows per lane
// shorten the pipeline's fill, until bands outnumber what the SIMDs hold.
// Measured at 20k x 20k, 100k x 120k and 200k x 200k: DFD and VE-FD sweeps are
// fastest at R = 2 everywhere (200k: 397 ms against 406 at R = 4 and 462 at
// R = 1).  Decisions, whose dead cells are skipped, are fastest at the fewest
// rows whose bands number at most 1.25 per SIMD, and at no more than 3 rows
// (200k: R = 3, 181 ms against 198 at R = 2, 225 at R = 4 and 263 at R = 1;
// 20k: R = 1).
int auto_rows_per_lane(const rfd_ctx *ctx, int op, int64_t rows) {
    if (op != rfd::OP_DEC) return 2;
    const int64_t simds = 4 * (int64_t)std::max(ctx->cu_count, 1);
    for (int r = 1; r < 3; ++r)
        if ((rows + 64 * r - 1) / (64 * r) * 4 <= 5 * simds) return r;
    return 3;
}

// One sweep: init kernel + band kernel, timed with HIP events on the
// context's stream, result copied back.
template <int OP, typename T>
rfd_status run_sweep(rfd_session *s, const T *P, const T *Q, const T *qseg, double param,
                     rfd_kernel_kind kind, double *result, double thr = INFINITY,
                     double *result_sq = nullptr) {
    rfd_ctx *ctx = s->ctx;
    const int64_t rows = (OP == rfd::OP_DFD) ? s->m : s->m - 1;
    const int64_t cols = (OP == rfd::OP_DFD) ? s->n : s->n - 1;
    const int rpl = ctx->sync == rfd::SYNC_ACQREL ? 1
                    : ctx->rows_per_lane > 0 ? ctx->rows_per_lane
                                             : auto_rows_per_lane(ctx, OP, rows);
    const int64_t nb = (rows + 64 * rpl - 1) / (64 * rpl);
    rfd::SweepArgs a{};
    a.P = P;
    a.Q = Q;
    a.qseg = qseg;
    a.slots = s->slots;
    a.progress = s->progress;
    a.queue = s->queue;
    a.result = s->result;
    a.rows = rows;
    a.cols = cols;
    a.nbands = (int)nb;
    a.param = param;
    a.thr = thr;
    a.evals = s->evals;
    a.dec_form = (OP == rfd::OP_DEC) ? s->dec_form : rfd::DEC_FUSED;
    a.watch = rfd::Watch{s->abort, ctx->spin_limit};
    const double sentinel = -3.0;  // .cu:193 (result = -3 until the terminal cell writes it)
    unsigned long long *tlog = nullptr;
    if (ctx->tlog_path && !s->noalloc) {
        RFD_HIP(ctx, hipMalloc(&tlog, sizeof(unsigned long long) * 4 * nb));
        RFD_HIP(ctx, hipMemsetAsync(tlog, 0, sizeof(unsigned long long) * 4 * nb, ctx->stream));
    }
    a.tlog = tlog;
    const int grid = grid_for(ctx, nb);
    RFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(rfd::sweep_init, dim3((unsigned)std::min<int64_t>((nb + 255) / 256, 1024)),
                       dim3(256), 0, ctx->stream, s->progress, (int)nb, s->queue, s->result, 2,
                       sentinel, s->evals, s->abort);
    RFD_HIP(ctx, hipGetLastError());
    RFD_HIP(ctx, (launch_sweep<OP, T>(s->dims, grid, ctx->sync, rpl, ctx->stream, a)));
    RFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    double hres[2] = {0, 0};
    unsigned long long hev[2] = {0, 0};
    int habort = 0;
    RFD_HIP(ctx, hipMemcpyAsync(hres, s->result, 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipMemcpyAsync(hev, s->evals, sizeof(hev), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipMemcpyAsync(&habort, s->abort, sizeof(habort), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const double h = hres[0];
    if (result_sq) *result_sq = hres[1];
    float ms = 0;
    RFD_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.launches[kind] += 1;
    ctx->stats.ms[kind] += ms;
    ctx->stats.cells[kind] += (double)rows * (double)cols;
    ctx->stats.evaluated[kind] += (double)hev[0];
    ctx->stats.skipped_blocks[kind] += (double)hev[1];
    if (tlog) {
        // Debug instrumentation: header {kind, rows, cols, nbands, grid, sync} + 4 u64 per band.
        std::vector<unsigned long long> h_t((size_t)(4 * nb));
        RFD_HIP(ctx, hipMemcpy(h_t.data(), tlog, sizeof(unsigned long long) * 4 * nb, hipMemcpyDeviceToHost));
        RFD_HIP(ctx, hipFree(tlog));
        if (FILE *f = std::fopen(ctx->tlog_path, "ab")) {
            long long hdr[6] = {(long long)kind, (long long)rows, (long long)cols, (long long)nb,
                                (long long)grid, (long long)(ctx->sync + 10 * rpl)};
            std::fwrite(hdr, sizeof(hdr), 1, f);
            std::fwrite(h_t.data(), sizeof(unsigned long long), h_t.size(), f);
            std::fclose(f);
        }
    }
    if (habort) {
        set_err(ctx, "sweep kernel %d: a hand-off wait exceeded %u polls and gave up", (int)kind,
                ctx->spin_limit);
        return RFD_ERR_STATE;
    }
    if (h == sentinel) {
        set_err(ctx, "sweep kernel %d wrote no terminal value", (int)kind);
        return RFD_ERR_STATE;
    }
    *result = h;
    return RFD_OK;
}

// One fused bound sweep of a session (m, n >= 2): DFD, DFD^2 and VE-FD
// (mode kModeF64), VE-FD in FP32 (kModeV32; `init` is then rounded to FP32
// and squared in FP32), or DFD and DFD^2 alone (kModeDfd).  *ran is false
// when the configuration has no such form, and the caller runs band_sweep.
rfd_status run_bounds(rfd_session *s, double init, double *dfd, double *dfd_sq, double *vefd,
                      int mode = rfd::kModeF64, bool *ran = nullptr) {
    const bool ve = rfd::mode_ve(mode);
    rfd_ctx *ctx = s->ctx;
    // Configuration: RFD_BOUNDS if set; otherwise one band per CU with the
    // fewest rows per lane (7, 10 or 13) that cover the pair (7 producer
    // waves, 10 at 13 rows per lane), or 4 rows per lane with 2 producer waves
    // (several bands per CU) for pairs too short to give every CU a band of 7
    // rows per lane.
    // Two columns per step (20000 + the one-column code, opt-in) need even n;
    // for odd n the one-column kernel runs.
    if (s->n >= (int64_t)0x7fffffff - 64) {
        set_err(ctx, "the fused bound sweep needs n < 2^31 - 64 (n=%lld)", (long long)s->n);
        return RFD_ERR_ARG;
    }
    int cfg = ctx->bounds_cfg;
    if (cfg == 0) {
        const int64_t per = 64 * (int64_t)ctx->cu_count;
        const int64_t need = (s->m + per - 1) / per;  // rows per lane for one band per CU
        cfg = need <= 4 ? 42 : need <= 7 ? 77 : need <= 10 ? 107 : 1310;
    }
    if (cfg >= 20000 && (s->n & 1)) cfg -= 20000;
    if (ran) *ran = true;
    if ((mode == rfd::kModeDfd && !rfd_valid_dfd_cfg(cfg)) || (mode == rfd::kModeV32 && !rfd_valid_mixed_cfg(cfg))) {
        if (ran) *ran = false;
        return RFD_OK;
    }
    const int base = cfg % 10000;
    const int R = base >= 1000 ? base / 100 : base / 10;  // 10 R + NPW, or 100 R + NPW for NPW >= 10
    const int64_t rows = s->m, cols = s->n;
    const int64_t nb = (rows + 64 * R - 1) / (64 * R);
    rfd::BoundsArgs a{};
    a.P = s->dP;
    a.Q = s->dQ;
    a.slots = (double *)s->slots;
    a.progress = s->progress;
    a.queue = s->queue;
    a.result = s->result;
    a.rows = rows;
    a.cols = cols;
    a.nbands = (int)nb;
    a.r2 = (mode == rfd::kModeV32) ? (double)((float)init * (float)init) : init * init;
    a.watch = rfd::Watch{s->abort, ctx->spin_limit};
    const double sentinel = -3.0;
    unsigned long long *tlog = nullptr;
    if (ctx->tlog_path && !s->noalloc) {
        RFD_HIP(ctx, hipMalloc(&tlog, sizeof(unsigned long long) * 4 * nb));
        RFD_HIP(ctx, hipMemsetAsync(tlog, 0, sizeof(unsigned long long) * 4 * nb, ctx->stream));
    }
    a.tlog = tlog;
    RFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    hipLaunchKernelGGL(rfd::sweep_init, dim3((unsigned)std::min<int64_t>((nb + 255) / 256, 1024)),
                       dim3(256), 0, ctx->stream, s->progress, (int)nb, s->queue, s->result, 4,
                       sentinel, s->evals, s->abort);
    RFD_HIP(ctx, hipGetLastError());
    int grid = 0;
    if (mode == rfd::kModeF64) {
        RFD_HIP(ctx, rfd_launch_bounds(ctx, s->dims, cfg, nb, a, &grid));
    } else if (mode == rfd::kModeV32) {
        RFD_HIP(ctx, rfd_launch_mixed(ctx, s->dims, cfg, nb, a, &grid));
    } else {
        RFD_HIP(ctx, rfd_launch_dfd(ctx, s->dims, cfg, nb, a, &grid));
    }
    RFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    double hres[4] = {0, 0, 0, 0};
    int habort = 0;
    RFD_HIP(ctx, hipMemcpyAsync(hres, s->result, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipMemcpyAsync(&habort, s->abort, sizeof(habort), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    RFD_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    // Cells of both bound sweeps: m * n DFD cells and (m-1)(n-1) VE-FD cells;
    // the DFD-only form counts as a DFD sweep.
    const double cells = (double)rows * (double)cols + (ve ? (double)(rows - 1) * (double)(cols - 1) : 0.0);
    const int kind = mode == rfd::kModeF64 ? RFD_K_BOUNDS64 : mode == rfd::kModeV32 ? RFD_K_BOUNDS_MIXED : RFD_K_DFD;
    ctx->stats.launches[kind] += 1;
    ctx->stats.ms[kind] += ms;
    ctx->stats.cells[kind] += cells;
    ctx->stats.evaluated[kind] += cells;
    if (tlog) {
        std::vector<unsigned long long> h_t((size_t)(4 * nb));
        RFD_HIP(ctx, hipMemcpy(h_t.data(), tlog, sizeof(unsigned long long) * 4 * nb, hipMemcpyDeviceToHost));
        RFD_HIP(ctx, hipFree(tlog));
        if (FILE *f = std::fopen(ctx->tlog_path, "ab")) {
            long long hdr[6] = {(long long)kind, (long long)rows, (long long)cols, (long long)nb,
                                (long long)grid, (long long)cfg};
            std::fwrite(hdr, sizeof(hdr), 1, f);
            std::fwrite(h_t.data(), sizeof(unsigned long long), h_t.size(), f);
            std::fclose(f);
        }
    }
    if (habort) {
        set_err(ctx, "fused bound sweep: a hand-off wait exceeded %u polls and gave up", ctx->spin_limit);
        return RFD_ERR_STATE;
    }
    if (hres[0] == sentinel || (ve && hres[2] == sentinel)) {
        set_err(ctx, "fused bound sweep wrote no terminal value");
        return RFD_ERR_STATE;
    }
    *dfd = hres[0];
    *dfd_sq = hres[1];
    if (ve) *vefd = hres[2];
    return RFD_OK;
}

// Flags of the windowed decision sweep a session's progress array holds (one
// 64-bit word per band).
int64_t dec_flag_capacity(const rfd_session *s) {
    return s->noalloc ? kNaProgressBytes / 8 : nbands_for(s->m) + 2;
}

// One windowed decision sweep (FP64) of a session (m, n >= 2): the terminal
// rightUB at `radius`.  false in *ran when the configuration does not fit the
// session's flag array (the caller then runs band_sweep).
rfd_status run_dec(rfd_session *s, double radius, double *result, bool *ran) {
    rfd_ctx *ctx = s->ctx;
    *ran = false;
    const int64_t rows = s->m - 1, cols = s->n - 1;
    // By size: six producer waves and the rows per lane that give about three
    // bands per CU.  Measured near the critical radius: 100k x 120k fastest at
    // 2 rows per lane (26.8 ms; 36 ms at 4), 200k x 200k at 4 (67.6 ms; 72.6
    // at 2); one workgroup per CU (two: 75 ms at 200k).
    int cfg = ctx->dec_cfg;
    if (cfg < 0) {
        const int64_t per = 192 * (int64_t)std::max(ctx->cu_count, 1);  // rows of three 1-row bands per CU
        const int64_t r = (rows + per / 2) / per;
        cfg = 10 * (int)std::max<int64_t>(1, std::min<int64_t>(4, r)) + 6;
    }
    const int R = cfg / 10;
    const int64_t nb = (rows + 64 * R - 1) / (64 * R);
    if (nb > dec_flag_capacity(s) || cols >= ((int64_t)1 << 30)) return RFD_OK;
    *ran = true;
    rfd::DecArgs a{};
    a.P = s->dP;
    a.Q = s->dQ;
    a.qseg = s->qseg64;
    a.slots = s->slots;
    a.flags = reinterpret_cast<unsigned long long *>(s->progress);
    a.queue = s->queue;
    a.result = s->result;
    a.rows = rows;
    a.cols = cols;
    a.nbands = (int)nb;
    a.r = radius;
    a.evals = s->evals;
    a.watch = rfd::Watch{s->abort, ctx->spin_limit};
    const double sentinel = -3.0;  // .cu:193 (result = -3 until the terminal cell writes it)
    unsigned long long *tlog = nullptr;
    if (ctx->tlog_path && !s->noalloc) {
        RFD_HIP(ctx, hipMalloc(&tlog, sizeof(unsigned long long) * 12 * nb));
        RFD_HIP(ctx, hipMemsetAsync(tlog, 0, sizeof(unsigned long long) * 12 * nb, ctx->stream));
    }
    a.tlog = tlog;
    RFD_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    rfd_host::rfd_launch_dec_init(ctx, a, sentinel, s->abort);
    RFD_HIP(ctx, hipGetLastError());
    int grid = 0;
    RFD_HIP(ctx, rfd_host::rfd_launch_dec(ctx, s->dims, cfg, s->dec_form == rfd::DEC_TWO_PHASE, nb, a, &grid));
    RFD_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    double h = 0;
    unsigned long long hev[2] = {0, 0};
    int habort = 0;
    RFD_HIP(ctx, hipMemcpyAsync(&h, s->result, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipMemcpyAsync(hev, s->evals, sizeof(hev), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipMemcpyAsync(&habort, s->abort, sizeof(habort), hipMemcpyDeviceToHost, ctx->stream));
    RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
    float ms = 0;
    RFD_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.launches[RFD_K_DEC64] += 1;
    ctx->stats.ms[RFD_K_DEC64] += ms;
    ctx->stats.cells[RFD_K_DEC64] += (double)rows * (double)cols;
    ctx->stats.evaluated[RFD_K_DEC64] += (double)hev[0];
    ctx->stats.skipped_blocks[RFD_K_DEC64] += (double)hev[1];
    if (tlog) {
        std::vector<unsigned long long> h_t((size_t)(12 * nb));
        RFD_HIP(ctx, hipMemcpy(h_t.data(), tlog, sizeof(unsigned long long) * 12 * nb, hipMemcpyDeviceToHost));
        RFD_HIP(ctx, hipFree(tlog));
        if (FILE *f = std::fopen(ctx->tlog_path, "ab")) {
            long long hdr[6] = {(long long)RFD_K_DEC64, (long long)rows, (long long)cols, (long long)nb,
                                (long long)grid, (long long)(100000 + cfg)};
            std::fwrite(hdr, sizeof(hdr), 1, f);
            std::fwrite(h_t.data(), sizeof(unsigned long long), h_t.size(), f);
            std::fclose(f);
        }
    }
    if (habort) {
        set_err(ctx, "windowed decision sweep: a hand-off wait exceeded %u polls and gave up", ctx->spin_limit);
        return RFD_ERR_STATE;
    }
    if (h == sentinel) {
        set_err(ctx, "windowed decision sweep wrote no terminal value");
        return RFD_ERR_STATE;
    }
    *result = h;
    return RFD_OK;
}

// Midpoint refinement of a device curve, `levels` times, into a new device
// array (*out, *len_out); the caller frees it.
rfd_status refine_curve(rfd_ctx *ctx, const double *src, int64_t len, int dims, int levels, double **out,
                        int64_t *len_out) {
    int64_t cur_len = len;
    double *cur = nullptr;
    RFD_HIP(ctx, hipMalloc(&cur, sizeof(double) * len * dims));
    RFD_HIP(ctx, hipMemcpyAsync(cur, src, sizeof(double) * len * dims, hipMemcpyDeviceToDevice, ctx->stream));
    for (int l = 0; l < levels && cur_len >= 2; ++l) {
        const int64_t nl = 2 * cur_len - 1;
        double *nxt = nullptr;
        if (hipMalloc(&nxt, sizeof(double) * nl * dims) != hipSuccess) {
            (void)hipFree(cur);
            set_err(ctx, "hipMalloc(refined curve) failed");
            return RFD_ERR_NOMEM;
        }
        const long long cnt = nl * dims;
        hipLaunchKernelGGL(rfd::refine_level, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, ctx->stream,
                           (const double *)cur, (long long)cur_len, dims, nxt);
        RFD_HIP(ctx, hipGetLastError());
        RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        (void)hipFree(cur);
        cur = nxt;
        cur_len = nl;
    }
    *out = cur;
    *len_out = cur_len;
    return RFD_OK;
}

// DFD of both curves of a session refined `levels` times (device buffers of
// the refined sizes for the sweep; the session's small words are reused).
rfd_status refined_ub(rfd_session *s, int levels, double *ub) {
    rfd_ctx *ctx = s->ctx;
    rfd_session t;
    t.ctx = ctx;
    t.dims = s->dims;
    t.queue = s->queue;
    t.result = s->result;
    t.evals = s->evals;
    t.abort = s->abort;
    rfd_status st = refine_curve(ctx, s->dP, s->m, s->dims, levels, &t.dP, &t.m);
    if (st == RFD_OK) st = refine_curve(ctx, s->dQ, s->n, s->dims, levels, &t.dQ, &t.n);
    // Boundary columns for the fused sweep's layout ([2 buffers][2 kinds][n]).
    if (st == RFD_OK && hipMalloc(&t.slots, sizeof(double) * 4 * t.n) != hipSuccess) st = RFD_ERR_NOMEM;
    if (st == RFD_OK && hipMalloc(&t.progress, sizeof(int) * (nbands_for(t.m) + 1)) != hipSuccess) st = RFD_ERR_NOMEM;
    if (st == RFD_OK) {
        // The fused sweep's DFD-only form where it has one, else the band sweep.
        double F = 0;
        bool ran = false;
        if (ctx->fused && t.m >= 2 && t.n >= 2) st = run_bounds(&t, 0.0, ub, &F, nullptr, rfd::kModeDfd, &ran);
        if (st == RFD_OK && !ran)
            st = run_sweep<rfd::OP_DFD, double>(&t, t.dP, t.dQ, nullptr, 0.0, RFD_K_DFD, ub, INFINITY, &F);
    } else if (st == RFD_ERR_NOMEM) {
        set_err(ctx, "no device memory for the refined curves (%lld x %lld points)", (long long)t.m,
                (long long)t.n);
    }
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(t.dP);
    (void)hipFree(t.dQ);
    (void)hipFree(t.slots);
    (void)hipFree(t.progress);
    return st;
}

bool keep_bisecting(double lb, double ub, const rfd_opts &o, double eps) {
    return o.stop_mode == RFD_STOP_ABS ? (ub - lb > eps) : (ub - lb > eps * ub);
}

}  // namespace

namespace rfd_host {

// The reference's choice between its decision paths (.cu:87-96): two-phase
// when its four interval arrays of ((m-1)(n+m-4) + m) doubles, plus
// n_block * n_thread_per_block bytes, fit in free device memory; it reads
// n_block = multiProcessorCount and n_thread_per_block = maxThreadsPerBlock
// (.cu:34-36).
int reference_form(rfd_ctx *ctx, int64_t m, int64_t n) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return rfd::DEC_FUSED;
    hipDeviceProp_t prop;
    int tpb = 1024;
    if (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess) tpb = prop.maxThreadsPerBlock;
    const long double need = (long double)ctx->cu_count * tpb +
                             ((long double)(m - 1) * (long double)(n - 2 + m - 2) + (long double)m) * 8.0L * 4.0L;
    return (long double)free_b < need ? rfd::DEC_FUSED : rfd::DEC_TWO_PHASE;
}

int resolve_form(rfd_session *s, int form) {
    if (form == RFD_DECIDE_TWO_PHASE) return rfd::DEC_TWO_PHASE;
    if (form == RFD_DECIDE_REFERENCE) return reference_form(s->ctx, s->m, s->n);
    return rfd::DEC_FUSED;
}

}  // namespace rfd_host

extern "C" {

const char *rfd_status_string(rfd_status s) {
    switch (s) {
        case RFD_OK: return "ok";
        case RFD_ERR_ARG: return "invalid argument";
        case RFD_ERR_HIP: return "HIP runtime error";
        case RFD_ERR_NO_DEVICE: return "no GPU device";
        case RFD_ERR_STATE: return "inconsistent sweep result";
        case RFD_ERR_NOMEM: return "device allocation failed";
    }
    return "unknown status";
}

const char *rfd_last_error(rfd_ctx *ctx) { return ctx ? ctx->err : "null context"; }

void rfd_default_opts(rfd_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->stop_mode = RFD_STOP_ABS;
    o->precision = RFD_PREC_FP64;
    o->eps = (double)FLT_EPSILON;  // compute_distance_wrapper.cpp:6
    o->vefd_gate = 1;
    o->verbose = 0;
    o->mixed_switch = 1e-4;
    o->rows_per_lane = 0;
    o->use_graph = 1;
    o->prune = -1;
    o->spec_depth = 0;
    o->refine_ub = 0;
}

rfd_status rfd_create(rfd_ctx **out, int device) {
    if (!out) return RFD_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return RFD_ERR_NO_DEVICE;
    if (device < 0 || device >= count) return RFD_ERR_ARG;
    rfd_ctx *ctx = new rfd_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return RFD_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        delete ctx;
        return RFD_ERR_HIP;
    }
    ctx->cu_count = prop.multiProcessorCount;
    if (scratch_address(device, &ctx->scratch) != hipSuccess) {
        (void)hipEventDestroy(ctx->ev0);
        (void)hipEventDestroy(ctx->ev1);
        (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return RFD_ERR_HIP;
    }
    {
        int khz = 0;
        if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
            ctx->wall_khz = (double)khz;
    }
    int waves_per_cu = 16;
    if (const char *e = std::getenv("RFD_WAVES_PER_CU")) waves_per_cu = std::max(1, std::atoi(e));
    ctx->max_waves = ctx->cu_count * waves_per_cu;
    // Absolute cap, used by tests to make waves take several bands each.
    if (const char *e = std::getenv("RFD_MAX_WAVES")) ctx->max_waves = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("RFD_SYNC")) ctx->sync = std::atoi(e) == 0 ? rfd::SYNC_ACQREL : rfd::SYNC_SCOPED;
    ctx->tlog_path = std::getenv("RFD_TLOG");
    if (const char *e = std::getenv("RFD_PRUNE")) ctx->prune = std::atoi(e) != 0;
    if (const char *e = std::getenv("RFD_FUSED")) ctx->fused = std::atoi(e) != 0;
    if (const char *e = std::getenv("RFD_BOUNDS")) {
        int c = std::atoi(e);
        if (rfd_valid_bounds_cfg(c)) ctx->bounds_cfg = c;
    }
    if (const char *e = std::getenv("RFD_WGCU")) ctx->wg_per_cu = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("RFD_SPIN_LIMIT")) {
        const long long v = std::atoll(e);
        ctx->spin_limit = (unsigned)std::max(1LL, std::min(v, 0xffffffffLL));
    }
    if (const char *e = std::getenv("RFD_BATCH_BOUNDS")) {
        const int c = std::atoi(e);
        if (c == 0 || rfd_valid_batch_bounds_cfg(c)) ctx->batch_bounds = c;
    }
    {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) == hipSuccess)
            ctx->lds_per_cu = lds;
    }
    if (const char *e = std::getenv("RFD_DEC")) {
        const int c = std::atoi(e);
        if (c == 0 || rfd_host::rfd_valid_dec_cfg(c)) ctx->dec_cfg = c;
    }
    if (const char *e = std::getenv("RFD_DEC_WGCU")) ctx->dec_wg_per_cu = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("RFD_BATCH_DEC")) ctx->batch_dec = std::atoi(e) != 0 ? 1 : 0;
    if (const char *e = std::getenv("RFD_BATCH_DEC_WGCU")) ctx->batch_dec_wgcu = std::max(0, std::atoi(e));
    if (const char *e = std::getenv("RFD_BATCH_DEC_BUDGET")) ctx->batch_dec_budget = std::max(1, std::atoi(e));
    if (const char *e = std::getenv("RFD_ROWS")) {
        int r = std::atoi(e);
        if (valid_rows_per_lane(r)) ctx->rows_per_lane = r;
    }
    *out = ctx;
    return RFD_OK;
}

rfd_status rfd_destroy(rfd_ctx *ctx) {
    if (!ctx) return RFD_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RFD_OK;
}

rfd_status rfd_get_stats(rfd_ctx *ctx, rfd_stats *out) {
    if (!ctx || !out) return RFD_ERR_ARG;
    *out = ctx->stats;
    return RFD_OK;
}

rfd_status rfd_reset_stats(rfd_ctx *ctx) {
    if (!ctx) return RFD_ERR_ARG;
    std::memset(&ctx->stats, 0, sizeof(ctx->stats));
    return RFD_OK;
}

rfd_status rfd_session_destroy(rfd_session *s) {
    if (!s) return RFD_ERR_ARG;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    (void)hipFree(s->dP);
    (void)hipFree(s->dQ);
    (void)hipFree(s->fP);
    (void)hipFree(s->fQ);
    (void)hipFree(s->qseg64);
    (void)hipFree(s->qseg32);
    (void)hipFree(s->slots);
    (void)hipFree(s->progress);
    (void)hipFree(s->queue);
    (void)hipFree(s->result);
    (void)hipFree(s->evals);
    (void)hipFree(s->abort);
    delete s;
    return RFD_OK;
}

rfd_status rfd_session_create(rfd_ctx *ctx, const double *P, int64_t m, const double *Q,
                              int64_t n, int dims, rfd_session **out) {
    if (!ctx || !P || !Q || !out) return RFD_ERR_ARG;
    *out = nullptr;
    if (m < 1 || n < 1 || (dims != 2 && dims != 3)) {
        set_err(ctx, "need non-empty curves and dims in {2,3} (m=%lld n=%lld dims=%d)",
                (long long)m, (long long)n, dims);
        return RFD_ERR_ARG;
    }
    if (m > (int64_t)1 << 30 || n > (int64_t)1 << 30) {
        set_err(ctx, "curves longer than 2^30 points are not supported");
        return RFD_ERR_ARG;
    }
    if (rfd_status fs = rfd_host::check_finite(ctx, P, m, Q, n, dims, -1); fs != RFD_OK) return fs;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    rfd_session *s = new rfd_session();
    s->ctx = ctx;
    s->m = m;
    s->n = n;
    s->dims = dims;
    s->endpoint_lb = endpoint_bound(P, m, Q, n, dims);
    auto fail = [&](hipError_t e, const char *what) {
        set_err(ctx, "%s failed: %s", what, hipGetErrorString(e));
        rfd_session_destroy(s);
        return e == hipErrorOutOfMemory ? RFD_ERR_NOMEM : RFD_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc(&s->dP, sizeof(double) * m * dims)) != hipSuccess) return fail(e, "hipMalloc(P)");
    if ((e = hipMalloc(&s->dQ, sizeof(double) * n * dims)) != hipSuccess) return fail(e, "hipMalloc(Q)");
    // Band-boundary buffers: two columns of n doubles (two kinds each for the
    // fused bound sweep); progress flags: one per band of one row per lane
    if ((e = hipMalloc(&s->slots, sizeof(double) * 4 * n)) != hipSuccess) return fail(e, "hipMalloc(slots)");
    // (the windowed decision sweep keeps a 64-bit word per band there).
    if ((e = hipMalloc(&s->progress, sizeof(int) * 2 * (nbands_for(m) + 2))) != hipSuccess)
        return fail(e, "hipMalloc(progress)");
    if ((e = hipMalloc(&s->queue, sizeof(int))) != hipSuccess) return fail(e, "hipMalloc(queue)");
    if ((e = hipMalloc(&s->result, 4 * sizeof(double))) != hipSuccess) return fail(e, "hipMalloc(result)");
    if ((e = hipMalloc(&s->evals, 2 * sizeof(unsigned long long))) != hipSuccess) return fail(e, "hipMalloc(evals)");
    if ((e = hipMalloc(&s->abort, sizeof(int))) != hipSuccess) return fail(e, "hipMalloc(abort)");
    if ((e = hipMemcpyAsync(s->dP, P, sizeof(double) * m * dims, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
        return fail(e, "hipMemcpy(P)");
    if ((e = hipMemcpyAsync(s->dQ, Q, sizeof(double) * n * dims, hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
        return fail(e, "hipMemcpy(Q)");
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    *out = s;
    return RFD_OK;
}

rfd_status rfd_session_bounds(rfd_session *s, double *dfd, double *vefd, int run_vefd, int fp32) {
    if (!s || !dfd || !vefd) return RFD_ERR_ARG;
    rfd_ctx *ctx = s->ctx;
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    double F = 0;
    if (ctx->fused && run_vefd && !fp32 && !ctx->prune && s->m >= 2 && s->n >= 2) {
        rfd_status st = run_bounds(s, *vefd, dfd, &F, vefd);
        if (st == RFD_OK) s->dfd_sq = F;
        return st;
    }
    bool ran = false;
    rfd_status st = RFD_OK;
    // DFD in FP64 and VE-FD in FP32 by one fused sweep, where the
    // configuration has that form.
    if (ctx->fused && run_vefd && fp32 && !ctx->prune && s->m >= 2 && s->n >= 2) {
        st = run_bounds(s, *vefd, dfd, &F, vefd, rfd::kModeV32, &ran);
        if (st != RFD_OK || ran) {
            if (st == RFD_OK) s->dfd_sq = F;
            return st;
        }
    }
    // DFD alone: the fused sweep's DFD-only form where it has one, else the
    // band sweep.
    if (ctx->fused && s->m >= 2 && s->n >= 2) st = run_bounds(s, 0.0, dfd, &F, nullptr, rfd::kModeDfd, &ran);
    if (st != RFD_OK) return st;
    if (!ran) st = run_sweep<rfd::OP_DFD, double>(s, s->dP, s->dQ, nullptr, 0.0, RFD_K_DFD, dfd, INFINITY, &F);
    if (st != RFD_OK) return st;
    s->dfd_sq = F;
    if (run_vefd && s->m >= 2 && s->n >= 2) {
        const double init = *vefd;
        // VE-FD <= DFD: cells whose incoming bottleneck is at least DFD^2 (with a
        // margin far above rounding) cannot change the minimum and are skipped.
        // If the result ever reaches the threshold the sweep is rerun unpruned.
        const bool prune = ctx->prune;
        const double thr64 = prune ? F * (1.0 + 0x1p-40) : INFINITY;
        const double thr32 = prune ? (double)(float)(F * (1.0 + 0x1p-18)) : INFINITY;
        for (int attempt = 0; attempt < 2; ++attempt) {
            double sq = 0;
            if (fp32) {
                if ((st = ensure_f32(s)) != RFD_OK) return st;
                const double thr = attempt ? INFINITY : thr32;
                st = run_sweep<rfd::OP_VEFD, float>(s, s->fP, s->fQ, nullptr, init, RFD_K_VEFD32, vefd,
                                                    thr, &sq);
                if (st != RFD_OK) return st;
                if (sq < thr) break;
            } else {
                const double thr = attempt ? INFINITY : thr64;
                st = run_sweep<rfd::OP_VEFD, double>(s, s->dP, s->dQ, nullptr, init, RFD_K_VEFD64, vefd,
                                                     thr, &sq);
                if (st != RFD_OK) return st;
                if (sq < thr) break;
            }
        }
    }
    return RFD_OK;
}

rfd_status rfd_session_decide(rfd_session *s, double radius, int fp32, int *reachable,
                              double *terminal_ub) {
    if (!s || !reachable) return RFD_ERR_ARG;
    rfd_ctx *ctx = s->ctx;
    if (s->m < 2 || s->n < 2) {
        set_err(ctx, "the decision sweep needs at least one segment per curve");
        return RFD_ERR_ARG;
    }
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    double ub = 0;
    rfd_status st;
    if (fp32) {
        if ((st = ensure_f32(s)) != RFD_OK) return st;
        if ((st = ensure_qseg<float>(s, &s->qseg32, s->fQ)) != RFD_OK) return st;
        st = run_sweep<rfd::OP_DEC, float>(s, s->fP, s->fQ, s->qseg32, radius, RFD_K_DEC32, &ub);
    } else {
        if ((st = ensure_qseg<double>(s, &s->qseg64, s->dQ)) != RFD_OK) return st;
        bool ran = false;
        if (ctx->dec_cfg != 0 && ctx->sync == rfd::SYNC_SCOPED && (st = run_dec(s, radius, &ub, &ran)) != RFD_OK)
            return st;
        if (!ran) st = run_sweep<rfd::OP_DEC, double>(s, s->dP, s->dQ, s->qseg64, radius, RFD_K_DEC64, &ub);
    }
    if (st != RFD_OK) return st;
    *reachable = (ub == 1.0) ? 1 : 0;  // .cu:1592
    if (terminal_ub) *terminal_ub = ub;
    return RFD_OK;
}


rfd_status rfd_session_refined_ub(rfd_session *s, int32_t levels, double *ub) {
    if (!s || !ub || levels < 0 || levels > 8) return RFD_ERR_ARG;
    RFD_HIP(s->ctx, hipSetDevice(s->ctx->device));
    if (levels == 0 || s->m < 2 || s->n < 2) {
        double F = 0;
        return run_sweep<rfd::OP_DFD, double>(s, s->dP, s->dQ, nullptr, 0.0, RFD_K_DFD, ub, INFINITY, &F);
    }
    return refined_ub(s, levels, ub);
}

rfd_status rfd_session_set_decide_form(rfd_session *s, int32_t form, int32_t *effective) {
    if (!s || form < RFD_DECIDE_FUSED || form > RFD_DECIDE_REFERENCE) return RFD_ERR_ARG;
    RFD_HIP(s->ctx, hipSetDevice(s->ctx->device));
    s->dec_form = rfd_host::resolve_form(s, form);
    if (effective) *effective = s->dec_form;
    return RFD_OK;
}

static rfd_status session_distance_impl(rfd_session *s, const rfd_opts *opts, rfd_result *res) {
    if (!s || !res) return RFD_ERR_ARG;
    rfd_opts o;
    if (opts) o = *opts; else rfd_default_opts(&o);
    rfd_ctx *ctx = s->ctx;
    std::memset(res, 0, sizeof(*res));
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    // Wall time of the call: HIP events on the stream, or, where nothing may be
    // created (noalloc), the host clock around the synchronous call.
    hipEvent_t t0 = nullptr, t1 = nullptr;
    const auto h0 = std::chrono::steady_clock::now();
    if (!s->noalloc) {
        RFD_HIP(ctx, hipEventCreate(&t0));
        RFD_HIP(ctx, hipEventCreate(&t1));
        RFD_HIP(ctx, hipEventRecord(t0, ctx->stream));
    }
    if (o.verbose) std::printf("%lld, %lld\n", (long long)s->m, (long long)s->n);  // frechet.cpp:21

    double lb = s->endpoint_lb;
    double ub = 0;
    res->endpoint_lb = lb;
    if (o.verbose) std::printf("Naive LB: %.14f\n", lb);  // frechet.cpp:31
    const bool degenerate = s->m < 2 || s->n < 2;
    const bool run_vefd = !degenerate && (!o.vefd_gate || s->m * s->n > 10000);
    const bool mixed = o.precision == RFD_PREC_MIXED;
    double vefd = lb;
    rfd_status st = rfd_session_bounds(s, &ub, &vefd, run_vefd ? 1 : 0, mixed ? 1 : 0);
    if (st != RFD_OK) return st;
    res->dfd = ub;
    if (degenerate) {
        // One point against a curve: the DFD is the distance.
        lb = ub;
    } else {
        lb = vefd;
    }
    res->vefd = lb;
    if (o.verbose) std::printf("LB: %.14f\nUB: %.14f\n", lb, ub);  // frechet.cpp:34
    const double eps = o.eps;
    if (o.refine_ub > 0 && !degenerate && keep_bisecting(lb, ub, o, eps)) {
        // The tighter upper bound (opt-in): DFD of the midpoint-refined curves,
        // at most DFD and, up to the midpoints' rounding, at least FD.
        double rub = ub;
        if ((st = refined_ub(s, o.refine_ub, &rub)) != RFD_OK) return st;
        ub = std::max(lb, std::min(ub, rub));
        if (o.verbose) std::printf("Refined UB: %.14f\n", ub);
    }
    int64_t iter = 0, iter32 = 0, cert = 0;
    if (mixed && !degenerate) {
        // FP32 VE-FD is an estimate, not a certified bound: clamp to [endpoint, DFD].
        lb = std::min(std::max(lb, s->endpoint_lb), ub);
        // FP32 bisection while the bracket is wide.
        while (keep_bisecting(lb, ub, o, eps) && (ub - lb) > o.mixed_switch * ub) {
            double split = (ub + lb) / 2.0;
            int r = 0;
            if ((st = rfd_session_decide(s, split, 1, &r, nullptr)) != RFD_OK) return st;
            ++iter;
            ++iter32;
            if (r) ub = split; else lb = split;
        }
        // The FP32 lower bound rarely meets DFD even where the FP64 one would
        // (the FP64 bracket then closes at once).  One FP64 decision just
        // below DFD settles the distance when it lies at DFD, the common end
        // of open brackets; otherwise it is a certified upper end.
        bool hi_certified = false;
        if (keep_bisecting(lb, ub, o, eps) && ub == res->dfd) {
            const double r = (o.stop_mode == RFD_STOP_ABS) ? ub - 0.5 * eps : ub * (1.0 - 0.5 * eps);
            if (r > lb && r > s->endpoint_lb) {
                int reach = 0;
                if ((st = rfd_session_decide(s, r, 0, &reach, nullptr)) != RFD_OK) return st;
                ++cert;
                if (reach) {
                    ub = r;
                    hi_certified = true;
                } else {
                    lb = r;  // the distance lies in (r, DFD]
                }
            }
        }
        // Certify the FP32 bracket with FP64 decisions at its ends: lb must be
        // unreachable and ub reachable.  Widen on failure.
        if (keep_bisecting(lb, ub, o, eps)) {
            double lo = lb, hi = ub;
            double w = std::max(hi - lo, 1e-12 * std::max(1.0, hi));
            for (int guard = 0; guard < 64; ++guard) {
                int r = 0;
                if (lo > s->endpoint_lb) {
                    if ((st = rfd_session_decide(s, lo, 0, &r, nullptr)) != RFD_OK) return st;
                    ++cert;
                    if (r) { hi = lo; lo = std::max(s->endpoint_lb, lo - w); w *= 2; continue; }
                }
                break;
            }
            for (int guard = 0; guard < 64; ++guard) {
                int r = 0;
                if (hi < res->dfd && !(hi_certified && hi == ub)) {
                    if ((st = rfd_session_decide(s, hi, 0, &r, nullptr)) != RFD_OK) return st;
                    ++cert;
                    if (!r) { lo = hi; hi = std::min(res->dfd, hi + w); w *= 2; continue; }
                }
                break;
            }
            lb = lo;
            ub = hi;
        }
    }
    while (keep_bisecting(lb, ub, o, eps)) {  // frechet.cpp:41-60
        double split = (ub + lb) / 2.0;
        int r = 0;
        if ((st = rfd_session_decide(s, split, 0, &r, nullptr)) != RFD_OK) return st;
        ++iter;
        if (r) ub = split; else lb = split;
    }
    double value = (ub + lb) / 2.0;  // frechet.cpp:63
    float ms = 0;
    if (!s->noalloc) {
        RFD_HIP(ctx, hipEventRecord(t1, ctx->stream));
        RFD_HIP(ctx, hipEventSynchronize(t1));
        (void)hipEventElapsedTime(&ms, t0, t1);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
    } else {
        RFD_HIP(ctx, hipStreamSynchronize(ctx->stream));
        ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    }
    res->distance = value;
    res->lb = lb;
    res->ub = ub;
    res->err_bound = rfd_host::error_bound(lb, ub, value);
    res->iterations = iter;
    res->iterations_fp32 = iter32;
    res->certify_steps = cert;
    res->ms_total = ms;
    if (o.verbose) {
        std::printf("Iteration: %lld\n", (long long)iter);  // frechet.cpp:65-66
        std::printf("Result: %.7f\n", value);
        std::printf("Total Time: %f ms\n", (double)ms);
    }
    return RFD_OK;
}

rfd_status rfd_session_distance(rfd_session *s, const rfd_opts *opts, rfd_result *res) {
    if (!s || !res) return RFD_ERR_ARG;
    rfd_opts o;
    if (opts) o = *opts; else rfd_default_opts(&o);
    if (o.decide_form < RFD_DECIDE_FUSED || o.decide_form > RFD_DECIDE_REFERENCE) return RFD_ERR_ARG;
    RFD_HIP(s->ctx, hipSetDevice(s->ctx->device));
    // The call's decisions use opts->decide_form; the session's own form
    // (rfd_session_set_decide_form) is restored afterwards.
    const int saved = s->dec_form;
    s->dec_form = rfd_host::resolve_form(s, o.decide_form);
    const rfd_status st = session_distance_impl(s, &o, res);
    s->dec_form = saved;
    return st;
}

rfd_status rfd_distance(rfd_ctx *ctx, const double *P, int64_t m, const double *Q, int64_t n,
                        int dims, const rfd_opts *opts, rfd_result *res) {
    rfd_session *s = nullptr;
    rfd_status st = rfd_session_create(ctx, P, m, Q, n, dims, &s);
    if (st != RFD_OK) return st;
    st = rfd_session_distance(s, opts, res);
    rfd_session_destroy(s);
    return st;
}

rfd_status rfd_distance_noalloc(rfd_ctx *ctx, const double *dP, int64_t m, const double *dQ, int64_t n,
                                int dims, const rfd_opts *opts, rfd_result *res) {
    if (!ctx || !dP || !dQ || !res) return RFD_ERR_ARG;
    rfd_opts o;
    if (opts) o = *opts; else rfd_default_opts(&o);
    if (m < 1 || n < 1 || (dims != 2 && dims != 3)) {
        set_err(ctx, "need non-empty curves and dims in {2,3} (m=%lld n=%lld dims=%d)", (long long)m,
                (long long)n, dims);
        return RFD_ERR_ARG;
    }
    if (m > kNaMax || n > kNaMax) {
        set_err(ctx, "rfd_distance_noalloc: curves of %lld and %lld points exceed its fixed working array "
                "(at most %lld points per curve)", (long long)m, (long long)n, (long long)kNaMax);
        return RFD_ERR_ARG;
    }
    if (o.precision != RFD_PREC_FP64) {
        set_err(ctx, "rfd_distance_noalloc computes in FP64 only (precision must be RFD_PREC_FP64)");
        return RFD_ERR_ARG;
    }
    if (o.refine_ub != 0) {
        set_err(ctx, "rfd_distance_noalloc does not refine the upper bound (refine_ub must be 0)");
        return RFD_ERR_ARG;
    }
    if (o.decide_form < RFD_DECIDE_FUSED || o.decide_form > RFD_DECIDE_REFERENCE) return RFD_ERR_ARG;
    if (!ctx->scratch) {
        set_err(ctx, "the fixed working array has no device address");
        return RFD_ERR_STATE;
    }
    RFD_HIP(ctx, hipSetDevice(ctx->device));
    char *base = static_cast<char *>(ctx->scratch);
    int *abort = reinterpret_cast<int *>(base);
    int *queue = reinterpret_cast<int *>(base + 4);
    int *flag = reinterpret_cast<int *>(base + 8);
    double *result = reinterpret_cast<double *>(base + 64);
    unsigned long long *evals = reinterpret_cast<unsigned long long *>(base + 96);
    // Non-finite coordinates are refused (as by rfd_session_create), found by
    // a scan on the device; the endpoints come back for the endpoint bound.
    hipStream_t st = ctx->stream;
    hipLaunchKernelGGL(rfd::clear_flag, dim3(1), dim3(1), 0, st, flag);
    const long long cp = m * dims, cq = n * dims;
    hipLaunchKernelGGL(rfd::scan_finite, dim3((unsigned)std::min<long long>((cp + 255) / 256, 1024)), dim3(256), 0,
                       st, dP, cp, flag);
    hipLaunchKernelGGL(rfd::scan_finite, dim3((unsigned)std::min<long long>((cq + 255) / 256, 1024)), dim3(256), 0,
                       st, dQ, cq, flag);
    RFD_HIP(ctx, hipGetLastError());
    double ends[12] = {0};  // P first, P last, Q first, Q last
    int hflag = 0;
    RFD_HIP(ctx, hipMemcpyAsync(ends, dP, sizeof(double) * dims, hipMemcpyDeviceToHost, st));
    RFD_HIP(ctx, hipMemcpyAsync(ends + dims, dP + (m - 1) * dims, sizeof(double) * dims, hipMemcpyDeviceToHost, st));
    RFD_HIP(ctx, hipMemcpyAsync(ends + 2 * dims, dQ, sizeof(double) * dims, hipMemcpyDeviceToHost, st));
    RFD_HIP(ctx, hipMemcpyAsync(ends + 3 * dims, dQ + (n - 1) * dims, sizeof(double) * dims, hipMemcpyDeviceToHost, st));
    RFD_HIP(ctx, hipMemcpyAsync(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    RFD_HIP(ctx, hipStreamSynchronize(st));
    if (hflag) {
        set_err(ctx, "a curve holds a non-finite coordinate");
        return RFD_ERR_ARG;
    }
    // A session over the caller's curves whose buffers are the array's.
    rfd_session s;
    s.ctx = ctx;
    s.m = m;
    s.n = n;
    s.dims = dims;
    s.dP = const_cast<double *>(dP);
    s.dQ = const_cast<double *>(dQ);
    s.slots = base + kNaSlotsOff;
    s.progress = reinterpret_cast<int *>(base + kNaProgressOff);
    s.queue = queue;
    s.result = result;
    s.evals = evals;
    s.abort = abort;
    s.endpoint_lb = endpoint_bound(ends, 2, ends + 2 * dims, 2, dims);
    s.noalloc = true;
    if (n >= 2) {
        s.qseg64 = reinterpret_cast<double *>(base + kNaQsegOff);
        const unsigned g = (unsigned)((n - 1 + 255) / 256);
        if (dims == 2)
            hipLaunchKernelGGL((rfd::qseg_terms<2, double>), dim3(g), dim3(256), 0, st, dQ, (long long)(n - 1), s.qseg64);
        else
            hipLaunchKernelGGL((rfd::qseg_terms<3, double>), dim3(g), dim3(256), 0, st, dQ, (long long)(n - 1), s.qseg64);
        RFD_HIP(ctx, hipGetLastError());
    }
    s.dec_form = rfd_host::resolve_form(&s, o.decide_form);
    return session_distance_impl(&s, &o, res);
}

}  // extern "C"
