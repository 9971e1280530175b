/*
 * CPU restatement of RobotFD's continuous Frechet distance path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * library is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product path never does.
 *
 * Curves are AoS FP64 arrays, point i at [i*dims + k], dims in {2, 3}.
 * P has m points (rows of the free-space grid), Q has n points (columns).
 *
 * Every function cites the reference lines it restates
 * (src/intersection_algorithm_in_parallel.cu = ".cu").
 */
#ifndef FRECHET_ORACLE_H
#define FRECHET_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* frechet.cpp:30  sqrt(max(|p0-q0|^2, |p_{m-1}-q_{n-1}|^2)), host arithmetic
 * (Point::dist_sqr, geometry_basics.hpp:72-81, no FMA). */
double or_endpoint_lb(const double *P, int64_t m, const double *Q, int64_t n, int dims);

/* DFD<D> (.cu:458-641): discrete Frechet distance on squared distances. */
double or_dfd(const double *P, int64_t m, const double *Q, int64_t n, int dims);

/* A tighter upper bound than DFD, an addition beside Bound_call_gpu
 * (.cu:1262-1326): DFD of both curves with every segment split at its
 * midpoint, `levels` times (or_refine: point a + 0.5 (b - a) between
 * consecutive points a, b; returns the refined length, writes it to out when
 * out != NULL).  FD <= or_dfd_refined <= DFD. */
int64_t or_refine(const double *P, int64_t m, int dims, int levels, double *out);
double or_dfd_refined(const double *P, int64_t m, const double *Q, int64_t n, int dims, int levels);

/* VEFD<D> (.cu:200-432, README.md:77-126): vertex-edge Frechet bound.
 * init_lb is the endpoint bound; cell (0,0) enters with init_lb*init_lb. */
double or_vefd(const double *P, int64_t m, const double *Q, int64_t n, int dims,
               double init_lb);

/* reachable<D> (.cu:960-1256): terminal rightUB of the free-space sweep at
 * radius r (r2 = r*r, .cu:1440).  The decision is (return == 1.0), .cu:1592. */
double or_decide_ub(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                    double r);
int or_decide(const double *P, int64_t m, const double *Q, int64_t n, int dims, double r);

/* Same formulas evaluated in FP32 (points rounded to float, r rounded to float);
 * in the decision, the square root of the discriminant and the two interval
 * ends are computed in FP64 from the FP32 terms and rounded to FP32 once. */
double or_vefd_f32(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                   double init_lb);
double or_decide_ub_f32(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                        double r);

/* Free interval of one cell edge, the per-edge part of reachable<D>
 * (.cu:1173-1230).  which = 0: right edge of cell (i,j) (vertex p_{i+1} on
 * segment q_j q_{j+1}); which = 1: top edge (vertex q_{j+1} on p_i p_{i+1}).
 * out[0] = LB, out[1] = UB ((-1,-2) when empty). */
void or_free_interval(const double *P, const double *Q, int dims, int64_t i, int64_t j,
                      double r, int which, double *out);

/* The two-phase decision path, interval_calculation<D> + reachability
 * (.cu:674-958), which the reference runs when its four interval arrays fit
 * in device memory (.cu:87-96, 1450-1558): the same reachability rules on
 * intervals whose C term is (sum mi^2) - r2 and whose clamps keep a NaN end
 * (zero-length segments).  Terminal rightUB, and the interval of one edge. */
double or_decide_ub_2ph(const double *P, int64_t m, const double *Q, int64_t n, int dims, double r);
void or_free_interval_2ph(const double *P, const double *Q, int dims, int64_t i, int64_t j,
                          double r, int which, double *out);

/* Vertex-edge weights of cell (i,j) (.cu:374-416, README.md:79-83):
 * out[0] = w_top = dist^2(q_{j+1}, p_i p_{i+1}), out[1] = w_right = dist^2(p_{i+1}, q_j q_{j+1}). */
void or_ve_weights(const double *P, const double *Q, int dims, int64_t i, int64_t j,
                   double *out);

/* distance_cuda (frechet.cpp:18-75) with the reference's bound order and
 * split rule.  stop_mode 0: while (ub - lb > eps)  (reference, eps = FLT_EPSILON);
 * stop_mode 1: while (ub - lb > eps * ub)  (relative).
 * vefd_gate: 1 runs VE-FD only when m*n > 10000 (.cu:1343, evaluated in 64 bits).
 * out[0] distance, out[1] endpoint lb, out[2] DFD, out[3] VE-FD (or lb if
 * skipped), out[4] final lb, out[5] final ub.  splits/decisions (may be NULL)
 * receive up to max_trace entries of the bisection trajectory.
 * Returns the iteration count. */
int64_t or_distance(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                    int stop_mode, double eps, int vefd_gate, double *out,
                    double *splits, int32_t *decisions, int64_t max_trace);
/* or_distance with the decision form: 0 = reachable<D> (fused), 1 = two-phase. */
int64_t or_distance_form(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                         int stop_mode, double eps, int vefd_gate, int form, double *out,
                         double *splits, int32_t *decisions, int64_t max_trace);

/* Rows [r0, r1) of one sweep, for bounded CPU timing samples.  kind 0 = VE-FD,
 * 1 = decision at r, 2 = DFD.  Returns a checksum of the last row's outputs. */
double or_sweep_rows(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                     int kind, double r, int64_t r0, int64_t r1);

/* The same bound sweeps (kind 0 = VE-FD, 2 = DFD) over rows [r0, r1) on
 * `threads` CPU threads: contiguous row blocks, one per thread, pipelined in
 * column chunks (block t consumes block t-1's last-row outputs chunk by
 * chunk), for the multi-core CPU timing.  Bit-identical to or_sweep_rows. */
double or_sweep_rows_mt(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                        int kind, double r, int64_t r0, int64_t r1, int threads);

/* The terminal value of one whole sweep on `threads` CPU threads, with the
 * row-block pipeline of or_sweep_rows_mt: kind 0 = VE-FD (r = the endpoint
 * bound, as or_vefd), 1 = decision at radius r (terminal rightUB, as
 * or_decide_ub; cells whose inputs are both unreachable skip their interval
 * arithmetic, whose outputs there are (-1, -2, -1) either way), 2 = DFD (as
 * or_dfd).  Bit-identical to the single-thread functions; used to compute the
 * reference values of long pairs. */
double or_sweep_mt(const double *P, int64_t m, const double *Q, int64_t n, int dims, int kind, double r,
                   int threads);

#ifdef __cplusplus
}
#endif
#endif
