/*
 * rfd.h - C ABI of the MI355X (gfx950) continuous Frechet distance library.
 *
 * Plain C types only: pointers, sizes, scalars.  Curves are AoS FP64 arrays,
 * point i at [i*dims + k], dims in {2, 3}.  Every entry point returns an
 * rfd_status (0 = success); none throws.  One rfd_ctx per host thread; calls
 * on one context are serialised on that context's HIP stream.
 *
 * Reference interfaces each entry point replaces (RobotFD):
 *   src/frechet.cpp, include/frechet.hpp, src/compute_distance_wrapper.cpp,
 *   include/intersection_algorithm_in_parallel.hpp (class Cuda_intersection).
 */
#ifndef RFD_H
#define RFD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RFD_OK = 0,
    RFD_ERR_ARG = 1,        /* bad argument (null pointer, dims not 2/3, empty curve) */
    RFD_ERR_HIP = 2,        /* a HIP runtime call failed */
    RFD_ERR_NO_DEVICE = 3,  /* no GPU visible */
    RFD_ERR_STATE = 4,      /* inconsistent result from a sweep (lb > ub, missing result) */
    RFD_ERR_NOMEM = 5       /* device allocation failed */
} rfd_status;

/* Stop rule of the bisection (frechet.cpp:41). */
typedef enum {
    RFD_STOP_ABS = 0,  /* while (ub - lb > eps): the reference, eps = FLT_EPSILON */
    RFD_STOP_REL = 1   /* while (ub - lb > eps * ub) */
} rfd_stop;

/* Arithmetic of the sweeps. */
typedef enum {
    RFD_PREC_FP64 = 0,   /* every sweep in FP64 */
    RFD_PREC_MIXED = 1   /* FP32 VE-FD and FP32 decisions while the bracket is wide,
                            FP64 decisions for the end game (certified bracket) */
} rfd_precision;

/* Interval form of the decision sweep.  The reference has two decision
 * paths with the same reachability rules and slightly different interval
 * arithmetic: reachable<D> (.cu:960-1256), and the two-phase
 * interval_calculation<D> + reachability (.cu:674-958), which it runs when its
 * four interval arrays fit in free device memory (.cu:87-96) unless
 * forceOversize is set.  The two differ only in rounding near a critical
 * radius and at zero-length segments (NaN intervals in the two-phase form). */
typedef enum {
    RFD_DECIDE_FUSED = 0,      /* reachable<D>'s intervals (default) */
    RFD_DECIDE_TWO_PHASE = 1,  /* interval_calculation<D>'s intervals */
    RFD_DECIDE_REFERENCE = 2   /* as the reference picks for each pair: two-phase when
                                  4 * ((m-1)(n+m-4) + m) doubles plus CUs * 1024 bytes fit
                                  in free device memory, else fused */
} rfd_decide_form;

typedef struct {
    int32_t stop_mode;     /* rfd_stop */
    int32_t precision;     /* rfd_precision */
    double eps;            /* stop tolerance */
    int32_t vefd_gate;     /* 1: run VE-FD only when m*n > 10000 (.cu:1343); 0: always */
    int32_t verbose;       /* 1: print the reference's progress lines (frechet.cpp:21-71) */
    double mixed_switch;   /* MIXED: relative bracket width below which FP64 takes over */
    int32_t rows_per_lane; /* sweep tile height per lane: 0 = library choice, else 1..4 */
    int32_t use_graph;     /* batch: 1 = replay the captured hipGraph of all rounds */
    int32_t prune;         /* VE-FD cells above DFD^2 skipped: -1 = library default, 0, 1 */
    int32_t spec_depth;    /* batch: levels of the bisection tree one decision round may
                              evaluate when few pairs remain (0 = library default, 1 = plain) */
    int32_t decide_form;   /* rfd_decide_form of the decisions (default RFD_DECIDE_FUSED) */
    int32_t refine_ub;     /* single pair: 0 = off (default); k > 0 = when the bracket is open,
                              lower DFD to the DFD of both curves refined k times at segment
                              midpoints (rfd_session_refined_ub) before bisecting */
} rfd_opts;

typedef struct {
    double distance;       /* (lb + ub) / 2 at the end (frechet.cpp:63) */
    double endpoint_lb;    /* frechet.cpp:30 */
    double dfd;            /* upper bound from DFD<D> */
    double vefd;           /* lower bound from VEFD<D> (endpoint_lb when gated off) */
    double lb, ub;         /* final bracket */
    double err_bound;      /* max(ub - distance, distance - lb) / lb: bounds |distance - FD| / FD
                              for the distance FD, which lies in [lb, ub] */
    int64_t iterations;    /* bisection steps (all precisions) */
    int64_t iterations_fp32;
    int64_t certify_steps; /* FP64 decisions spent certifying an FP32 bracket */
    double ms_total;       /* wall time of the call on the device stream */
} rfd_result;

/* Per-kernel timing accumulated on a context's stream (HIP events). */
typedef struct {
    int64_t launches[12];  /* index: rfd_kernel_kind */
    double ms[12];
    double cells[12];      /* grid cells swept (the algorithmic count) */
    double evaluated[12];  /* cells whose weights / free intervals were computed;
                              the rest were skipped as unable to change the result */
    double skipped_blocks[12];  /* band sweeps: whole blocks of dead steps skipped at once
                                   (every lane's cells and inputs unreachable or pruned) */
} rfd_stats;

typedef enum {
    RFD_K_DFD = 0, RFD_K_VEFD64 = 1, RFD_K_DEC64 = 2, RFD_K_VEFD32 = 3, RFD_K_DEC32 = 4,
    RFD_K_BATCH = 5, RFD_K_MLP = 6, RFD_K_OTHER = 7,
    RFD_K_BOUNDS64 = 8,    /* fused DFD + VE-FD sweep (FP64); cells = m*n + (m-1)*(n-1) */
    RFD_K_BOUNDS_MIXED = 9 /* fused sweep, DFD in FP64 and VE-FD in FP32; cells as RFD_K_BOUNDS64 */
} rfd_kernel_kind;

typedef struct rfd_ctx rfd_ctx;
typedef struct rfd_session rfd_session;

/* Library context: device, stream, timing.  No counterpart in the reference,
 * which hard-codes device 0 and the default stream (.cu:32-37, 52). */
rfd_status rfd_create(rfd_ctx **ctx, int device);
rfd_status rfd_destroy(rfd_ctx *ctx);
const char *rfd_status_string(rfd_status s);
const char *rfd_last_error(rfd_ctx *ctx);
void rfd_default_opts(rfd_opts *opts);
rfd_status rfd_get_stats(rfd_ctx *ctx, rfd_stats *out);
rfd_status rfd_reset_stats(rfd_ctx *ctx);

/* Replaces Cuda_intersection::Cuda_intersection + data_move
 * (.cu:23-142, 144-160): flattens nothing (input is already AoS), allocates
 * the device buffers and copies both curves to the device. */
rfd_status rfd_session_create(rfd_ctx *ctx, const double *P, int64_t m, const double *Q,
                              int64_t n, int dims, rfd_session **out);
/* Replaces Cuda_intersection::free_memory (.cu:162-175). */
rfd_status rfd_session_destroy(rfd_session *s);

/* Replaces Cuda_intersection::Bound_call_gpu(double *dfd, double *vefd)
 * (.cu:1258-1395): *dfd <- DFD; *vefd is read as the endpoint bound and
 * overwritten by VE-FD when run_vefd != 0 (the caller applies the m*n > 10000
 * gate of .cu:1343).  fp32 != 0 evaluates VE-FD in FP32. */
rfd_status rfd_session_bounds(rfd_session *s, double *dfd, double *vefd, int run_vefd,
                              int fp32);

/* Replaces Cuda_intersection::intersection_interval_call_gpu(double radius)
 * (.cu:1398-1597): *reachable <- 1 if the free-space sweep at `radius` reaches
 * (m-1, n-1), i.e. the terminal rightUB == 1.0 (.cu:1592).  *terminal_ub (may
 * be NULL) receives that rightUB.  fp32 != 0 evaluates the sweep in FP32. */
rfd_status rfd_session_decide(rfd_session *s, double radius, int fp32, int *reachable,
                              double *terminal_ub);

/* The decision path Cuda_intersection's constructor fixes for the session
 * (.cu:87-96, 1433-1558): form is an rfd_decide_form; RFD_DECIDE_REFERENCE
 * applies the reference's free-memory rule now, as its constructor does.
 * *effective (may be NULL) receives the form in use (FUSED or TWO_PHASE).
 * rfd_session_decide uses it; rfd_session_distance uses opts->decide_form. */
rfd_status rfd_session_set_decide_form(rfd_session *s, int32_t form, int32_t *effective);

/* A tighter upper bound than DFD, an addition beside Bound_call_gpu
 * (.cu:1262-1326), with no counterpart in the reference: *ub <- DFD of both
 * curves with every segment split at its midpoint, `levels` times (0..8; the
 * point a + 0.5 (b - a) between consecutive points a, b).  A coupling of the
 * refined vertices is a matching of the curves and the distance of two points
 * moving linearly is convex, so FD <= *ub <= DFD, up to the rounding of the
 * midpoints.  Costs a DFD sweep of 4^levels times the cells. */
rfd_status rfd_session_refined_ub(rfd_session *s, int32_t levels, double *ub);

/* Replaces Frechet::Continuous::distance_cuda (frechet.cpp:18-75) on a
 * session whose curves are already resident on the device. */
rfd_status rfd_session_distance(rfd_session *s, const rfd_opts *opts, rfd_result *res);

/* Whole call: create a session, run distance_cuda, destroy the session.
 * compute_distance_parallel (compute_distance_wrapper.cpp:8-12) is
 * rfd_distance with stop_mode = RFD_STOP_ABS, eps = FLT_EPSILON. */
rfd_status rfd_distance(rfd_ctx *ctx, const double *P, int64_t m, const double *Q, int64_t n,
                        int dims, const rfd_opts *opts, rfd_result *res);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): rfd_distance on curves that are already in device memory
 * (dP, dQ: device pointers, AoS FP64), allocating nothing.  Every working
 * buffer lives in a fixed array the library holds on each GPU it runs on until
 * the process ends (RFD_NOALLOC_BYTES), which limits each curve to
 * RFD_NOALLOC_MAX_POINTS points (larger curves: RFD_ERR_ARG).  FP64 only
 * (opts->precision must be RFD_PREC_FP64); RFD_TLOG is ignored.  The result
 * equals rfd_distance's on the same curves and options.  All calls on one GPU
 * share the array: calls that overlap on one GPU (other host threads or
 * contexts) race and must be serialised by the caller. */
#define RFD_NOALLOC_MAX_POINTS (1 << 19)
#define RFD_NOALLOC_BYTES (24 * 1024 * 1024 + 65536)
rfd_status rfd_distance_noalloc(rfd_ctx *ctx, const double *dP, int64_t m, const double *dQ, int64_t n,
                                int dims, const rfd_opts *opts, rfd_result *res);

/* ---- Batches: many pairs, resident on the device, one launch per round ----
 * No counterpart in the reference, which computes one pair per call.  All
 * rounds (DFD, VE-FD, the bisection's decisions) of all pairs run as
 * persistent kernels on the context's stream; each pair keeps its own bracket
 * on the device, so no host round trip happens between rounds. */
typedef struct rfd_batch rfd_batch;

/* P[k], Q[k]: AoS curves of pair k with m[k], n[k] points of dims[k] in {2,3}. */
rfd_status rfd_batch_create(rfd_ctx *ctx, int64_t npairs, const double *const *P,
                            const int64_t *m, const double *const *Q, const int64_t *n,
                            const int32_t *dims, rfd_batch **out);
/* Distances of every pair (results[k], frechet.cpp semantics per pair).
 * latency_ms (may be NULL): per pair, time from the batch start to the
 * round that finished that pair. */
rfd_status rfd_batch_run(rfd_batch *b, const rfd_opts *opts, rfd_result *results,
                         double *latency_ms);
/* Times the batch once for each rows-per-lane choice (1..4), keeps the
 * fastest for later runs and returns it; ms (may be NULL) gets 4 timings. */
rfd_status rfd_batch_tune(rfd_batch *b, const rfd_opts *opts, int32_t *rows_per_lane, double *ms);
rfd_status rfd_batch_destroy(rfd_batch *b);

/* ---- Bracket-position network (learned bracket) ----
 * No counterpart in the reference.  A BF16 MLP on MFMA (386 features -> 4 x 256
 * ReLU -> sigmoid) predicts where the distance lies in [VE-FD, DFD] as a
 * fraction f.  A learned batch run certifies a bracket around the prediction
 * for each bisecting pair and falls back to bisection of whatever bracket the
 * certifying decisions leave when the prediction misses.  Policy: f < tau ->
 * [lb, lb + t] (one decision; t = half the stop tolerance, so a success ends
 * the pair), f > 1 - tau -> [ub - t, ub], otherwise [f -/+ delta] of the
 * bracket (two decisions) when delta > 0. */
typedef struct rfd_mlp rfd_mlp;

int64_t rfd_mlp_param_count(void);
rfd_status rfd_mlp_create(rfd_ctx *ctx, uint64_t seed, rfd_mlp **out);
rfd_status rfd_mlp_destroy(rfd_mlp *net);
/* FP32 parameters: W1[400x256] b1 W2[256x256] b2 W3 b3 W4 b4 w5[256] b5, row-major. */
rfd_status rfd_mlp_get_params(rfd_mlp *net, float *out);
rfd_status rfd_mlp_set_params(rfd_mlp *net, const float *in);
rfd_status rfd_mlp_save(rfd_mlp *net, const char *path);
rfd_status rfd_mlp_load(rfd_mlp *net, const char *path);
rfd_status rfd_mlp_set_policy(rfd_mlp *net, double tau, double delta);
rfd_status rfd_mlp_get_policy(rfd_mlp *net, double *tau, double *delta);
/* features: rows x 386 FP32; out: rows sigmoid outputs. */
rfd_status rfd_mlp_forward(rfd_mlp *net, const float *features, int64_t rows, float *out);
/* Binary cross-entropy against labels in [0,1]; grads: rfd_mlp_param_count() FP32. */
rfd_status rfd_mlp_gradients(rfd_mlp *net, const float *features, const float *labels,
                             int64_t rows, float *grads, float *loss);
/* Adam training on the device; loss_per_epoch (may be NULL) gets `epochs` values. */
rfd_status rfd_mlp_train(rfd_mlp *net, const float *features, const float *labels, int64_t rows,
                         int32_t epochs, int32_t batch, float lr, uint64_t seed,
                         float *loss_per_epoch);
/* Features of every pair from the bounds of the batch's last run (npairs x 386). */
rfd_status rfd_batch_features(rfd_batch *b, float *out);
/* rfd_batch_run with the network's predicted bracket certified after the
 * bounds; certified (may be NULL) gets 1 for pairs whose prediction held. */
rfd_status rfd_batch_run_learned(rfd_batch *b, rfd_mlp *net, const rfd_opts *opts,
                                 rfd_result *results, double *latency_ms, int32_t *certified);

#ifdef __cplusplus
}
#endif
#endif /* RFD_H */
