// Band-wavefront sweep kernels for gfx950 (CDNA4).
//
// One wave (64 lanes) owns a band of 64 consecutive rows of the free-space
// grid; lane l owns row 64*band + l.  At step s, lane l evaluates cell
// (row, s - l): an anti-diagonal that slides one column per step.  The value a
// row hands to the row above (DFD: min(top, bot); VE-FD: R; decision: rightLB)
// moves one lane up per step through a cross-lane shift, so a cell's two
// inputs are its own previous cell (register) and lane l-1's previous output.
//
// Bands are taken from an in-order queue (one atomic per band).  Band b reads
// row 64b-1's outputs from a column buffer that band b-1 fills in chunks of
// kHand columns; band b-1 publishes the count of finished columns with an agent-scope
// release store, band b polls it with acquire loads.  Because the queue hands
// out bands in order, every band a wave waits on is already held by a running
// wave, so the pipeline cannot deadlock whatever the grid size.  Two column
// buffers suffice: band b+2 overwrites band b's buffer column c only after
// band b+1 has consumed column c (it needs band b+1's output at c first).
//
// Per-cell arithmetic follows the reference kernels with every FMA explicit
// (the file is built with -ffp-contract=off):
//   DFD      DFD<D>        .cu:458-641
//   VE-FD    VEFD<D>       .cu:200-432, README.md:77-126
//   decision reachable<D>  .cu:960-1256
#pragma once

#include <hip/hip_runtime.h>
#include <cfloat>
#include <cstdint>

namespace rfd {

enum SweepOp { OP_DFD = 0, OP_VEFD = 1, OP_DEC = 2 };

// Columns per band hand-off (a power of two, at most 64): a band trails its
// predecessor by about 64 + kHand steps, and publishes (a memory wait and a
// flag store) every kHand columns.  Measured at 200k x 200k and on a batch of
// 4096 pairs: for decisions and VE-FD 16 beats 64 (shorter pipelines) and 8
// (the publishes dominate).  DFD, whose cells are cheap and never skipped,
// keeps 64 (133 ms at 200k x 200k against 145 ms at 16).
template <int OP> constexpr int kHand = (OP == OP_DFD) ? 64 : 16;

template <typename T> struct Num;
template <> struct Num<double> {
    __device__ static inline double fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
    __device__ static inline double min(double a, double b) { return __builtin_fmin(a, b); }
    __device__ static inline double max(double a, double b) { return __builtin_fmax(a, b); }
    __device__ static inline double sqrt(double a) { return __builtin_sqrt(a); }
    __device__ static inline double big() { return (double)FLT_MAX; }
};
template <> struct Num<float> {
    __device__ static inline float fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
    __device__ static inline float min(float a, float b) { return __builtin_fminf(a, b); }
    __device__ static inline float max(float a, float b) { return __builtin_fmaxf(a, b); }
    __device__ static inline float sqrt(float a) { return __builtin_sqrtf(a); }
    __device__ static inline float big() { return FLT_MAX; }
};

// Whole-wave lane read with a wave-uniform lane index.
__device__ inline double lane_read(double v, int lane) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(unsigned)(b & 0xffffffffull), lane);
    int hi = __builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}
__device__ inline float lane_read(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// Value of lane l-1, by DPP wave_shr:1 (one VALU move per dword, no LDS
// round trip).  Lane 0 has no source lane and keeps `v`; the caller overrides it.
__device__ inline int dpp_shr1(int v) { return __builtin_amdgcn_update_dpp(v, v, 0x138, 0xf, 0xf, false); }
__device__ inline double shift_up(double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    int lo = dpp_shr1((int)(unsigned)(b & 0xffffffffull));
    int hi = dpp_shr1((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}
__device__ inline float shift_up(float v) { return __int_as_float(dpp_shr1(__float_as_int(v))); }

// ---- per-cell formulas (identical op order to the CPU oracle) ----

template <typename T, int D>
__device__ inline T dist2(const T *a, const T *b) {
    using N = Num<T>;
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        T t = a[k] - b[k];
        acc = N::fma(t, t, acc);
    }
    return acc;
}

template <typename T, int D>
__device__ inline T seg_len2(const T *a, const T *b) {
    using N = Num<T>;
    T acc = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        T d = b[k] - a[k];
        acc = N::fma(d, d, acc);
    }
    return acc;
}

// Squared distance of point x to segment [a, b] (.cu:375-394).
template <typename T, int D>
__device__ inline T point_seg2(const T *x, const T *a, const T *b) {
    using N = Num<T>;
    T ab[D];
    T ab2 = T(0), apab = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        ab[k] = b[k] - a[k];
        T ap = x[k] - a[k];
        ab2 = N::fma(ab[k], ab[k], ab2);
        apab = N::fma(ap, ab[k], apab);
    }
    // t = min(1, max(0, apab / ab2)) is 0 for apab <= 0 (or NaN) and 1 for
    // apab >= ab2 whatever the quotient rounds to; the divide runs only for
    // projections strictly inside the segment (a branch: most waves skip it).
    T t = (apab > T(0)) ? T(1) : T(0);
    if (apab > T(0) && apab < ab2) {
        asm volatile("");
        t = apab / ab2;
        t = N::min(T(1), N::max(T(0), t));
    }
    T w = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        T q = N::fma(t, ab[k], a[k]);
        T d = x[k] - q;
        w = N::fma(d, d, w);
    }
    return w;
}

// Decision interval forms (SweepArgs::dec_form).
enum DecForm { DEC_FUSED = 0, DEC_TWO_PHASE = 1 };

// Free interval [lb, ub] of segment [s0, s1] around vertex c at radius^2 r2,
// with A4 = 4|s1 - s0|^2 and inv2A = 0.5 / |s1 - s0|^2 (the per-segment
// terms); (-1, -2) when the segment misses the disc or the interval leaves
// [0, 1].  Two forms, as the reference's two decision paths compute it:
//  fused (TWO_PHASE = false), reachable<D> (.cu:1173-1200): the C term is a
//    sum started at -r2, and the ends are clamped with max / min;
//  two-phase (true), interval_calculation<D> (.cu:722-782): the C term is
//    (sum of mi^2) - r2 (.cu:733-742), and the clamps are comparisons
//    (.cu:751-752), so a NaN end (0 * inf at a zero-length segment) stays NaN.
// In three phases (iv_disc, iv_root, iv_ends) so that a caller can run the
// phases of all its intervals side by side, and without branches: the square
// root is taken of disc where disc >= 0 (sign of a zero included) and of 0
// elsewhere, and where disc < 0 selects give (-1, -2), the values a branch
// around the root gives.
template <typename T, int D, bool TWO_PHASE>
__device__ inline void iv_disc(const T *c, const T *s0, const T *s1, T r2, T A4, T &B, T &disc) {
    using N = Num<T>;
    T b = T(0), C = TWO_PHASE ? T(0) : -r2;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        T di = s1[k] - s0[k];
        T mi = s0[k] - c[k];
        b = N::fma(mi, di, b);
        C = N::fma(mi, mi, C);
    }
    B = b * T(2);
    if (TWO_PHASE) C = C - r2;
    disc = N::fma(B, B, -(A4 * C));
}

// The square root and the interval ends are FP64 in both precisions: in FP32
// (the precision form's early decisions) they are computed in FP64 from the
// FP32 B, disc and inv2A, and the ends rounded to FP32 once.
template <typename T> struct RootT { using type = T; };
template <> struct RootT<float> { using type = double; };

template <typename T>
__device__ inline typename RootT<T>::type iv_root(T disc) {
    using R = typename RootT<T>::type;
    return Num<R>::sqrt(disc >= T(0) ? (R)disc : R(0));
}

// dead_lb: the lower end an empty interval gets (-1, the reference's; the
// windowed decision sweep's chain takes +inf).
template <typename T, bool TWO_PHASE>
__device__ inline void iv_ends(T B, T disc, typename RootT<T>::type s, T inv2A, T &lb, T &ub,
                               const T dead_lb = T(-1)) {
    using N = Num<T>;
    using R = typename RootT<T>::type;
    const T ta = (T)((-(R)B - s) * (R)inv2A);
    const T tb = (T)((-(R)B + s) * (R)inv2A);
    // Bitwise tests: with && the compiler sinks the root and the ends into a
    // branch on disc >= 0, which keeps the cells' chains from interleaving.
    const bool valid = (disc >= T(0)) & !((tb < T(0)) | (ta > T(1)));
    if (TWO_PHASE) {
        lb = valid ? ((ta < T(0)) ? T(0) : ta) : dead_lb;
        ub = valid ? ((tb > T(1)) ? T(1) : tb) : T(-2);
    } else {
        lb = valid ? N::max(ta, T(0)) : dead_lb;
        ub = valid ? N::min(tb, T(1)) : T(-2);
    }
}

// Bounded hand-off waits.  Every wait of a sweep counts its polls; at `limit`
// polls it raises the launch's abort word and gives up.  A wave that gave up
// waits no more in its band: it runs its remaining steps on whatever values
// are in its buffers (every index stays within its ring or column), so every
// wave ends and the kernel exits; the host reads the word after the launch and
// reports the sweep as failed.  The first kQuickPolls polls of a wait touch no
// memory and add only a count and a compare to the poll: an access to the word
// would wait for the wave's own outstanding loads and stores, and heavier
// polls take issue slots from the working waves that share a SIMD with many
// waiting ones (a batch round's long chains of bands).  Past them a wait naps
// longer between polls and gives up as soon as it sees the word raised (read
// every 64th poll), so once one wait gives up, the others follow quickly.
struct Watch {
    int *abort;       // 0 until a wait gives up
    unsigned limit;   // polls one wait may take (at least 1)
};

constexpr unsigned kQuickPolls = 1024;

// One failed poll of a wait: true when the wait must give up.
__device__ inline bool watch_expired(const Watch &w, unsigned &polls) {
    if (!w.abort) return false;  // no word to raise: the launch waits without a bound
    ++polls;
    if (polls >= w.limit) {
        __hip_atomic_store(w.abort, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // Nothing left in flight on this rare path: where it rejoins the
        // common one, the compiler's counter waits stay those of the common
        // path (an outstanding access here would make later waits stricter).
        __builtin_amdgcn_s_waitcnt(0);
        return true;
    }
    if (polls < kQuickPolls) return false;
    __builtin_amdgcn_s_sleep(16);  // a long wait: poll seldom
    if ((polls & 63) != 0) return false;
    const bool raised = __hip_atomic_load(w.abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    __builtin_amdgcn_s_waitcnt(0);
    return raised;
}

struct SweepArgs {
    const void *P;       // AoS points of the row curve (T)
    const void *Q;       // AoS points of the column curve (T)
    const void *qseg;    // decision: per column j, {4*|q_{j+1}-q_j|^2, 0.5/|q_{j+1}-q_j|^2} (T)
    void *slots;         // 2 * cols values of T: band-boundary columns
    int *progress;       // per band: columns published
    int *queue;          // next band to hand out
    double *result;      // terminal value
    long long rows, cols;
    int nbands;
    double param;        // VE-FD: endpoint bound; decision: radius
    double thr;          // VE-FD: prune threshold on incoming values (+inf: none)
    unsigned long long *evals;  // [0] cells whose weights / free intervals were evaluated,
                                //  [1] whole dead blocks skipped
    unsigned long long *tlog;   // optional, 4 per band: wave start, first input ready, end, hw id
    int dec_form;        // decision: DEC_FUSED or DEC_TWO_PHASE (selects the kernel variant)
    Watch watch;         // bound on the hand-off waits
};

// "Dead" values: those that can no longer change the sweep's result.
//  VE-FD: an incoming bottleneck >= thr, where thr is slightly above DFD^2 >= VE-FD^2;
//    a path through such a cell costs more than the optimum, so the cell's weights
//    are not needed and it passes on its incoming value (itself dead).
//  decision: a negative lower end (unreachable); such a cell's outputs are
//    (-1, -2, -1), exactly what the full formula gives.
//  DFD: nothing is dead.
template <int OP, typename T>
__device__ inline bool is_dead(T v, T thr) {
    if (OP == OP_VEFD) return v >= thr;
    if (OP == OP_DEC) return v < T(0);
    return false;
}

// Band hand-off protocol.
//  SYNC_ACQREL: agent-scope release store / acquire load of the progress flag;
//    boundary values are plain loads/stores.  The compiler lowers these to an L2
//    write-back (buffer_wbl2) before the flag store and an L2 invalidate
//    (buffer_inv) after each flag load that succeeds or fails.
//  SYNC_SCOPED: every access to the flag and to the boundary columns is an
//    agent-scope atomic (sc1: coherent at device scope, no whole-cache
//    maintenance).  The producer waits for its column stores to complete
//    (s_waitcnt vmcnt(0)) before it stores the flag; the consumer issues its
//    column loads only after the flag load has returned the needed count.
enum SyncMode { SYNC_ACQREL = 0, SYNC_SCOPED = 1 };

template <int SYNC>
__device__ inline int flag_load(const int *p) {
    if (SYNC == SYNC_ACQREL) return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int SYNC, typename T>
__device__ inline T column_load(const T *p) {
    if (SYNC == SYNC_ACQREL) return *p;
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int SYNC, typename T>
__device__ inline void column_store(T *p, T v) {
    if (SYNC == SYNC_ACQREL) *p = v;
    else __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <int SYNC>
__device__ inline void flag_store(int *p, int v) {
    if (SYNC == SYNC_ACQREL) {
        __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        __builtin_amdgcn_s_waitcnt(0);  // column stores of every lane have completed
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// One pair's sweep as the band code sees it.
struct SweepJob {
    const void *P, *Q, *qseg;  // T arrays (see SweepArgs)
    void *slots;               // 2 * cols values of T
    int *progress;             // per band of this pair
    double *result;            // [0] terminal value, [1] its square (DFD, VE-FD)
    long long rows, cols;
    int nbands;
    double param, thr;
    Watch watch;
};

template <int OP> struct Staged { static constexpr int value = (OP == OP_DEC) ? 2 : 1; };

// Rows per lane.  Lane l of band b owns the R consecutive rows
// 64*R*b + R*l + t, t = 0..R-1.  At step s it evaluates column s - l for all
// R rows, top to bottom: row t's output feeds row t+1 in a register, and only
// the last row's output crosses to lane l+1 (one shift per step for R cells).
// The R weight evaluations of a step are independent, which gives each wave
// instruction-level parallelism; the min/max chain is the only serial part.
//
// Q points (and, for the decision, the per-column terms) reach the lanes
// through a ring of four 64-point chunks in LDS, structure-of-arrays so that
// consecutive lanes read consecutive addresses.  Chunk k+2 is loaded into
// registers at step 64k and written to LDS at step 64(k+1), so global-memory
// latency stays off the per-step path.  At step s lane l reads points s-l and
// s-l+1: chunks k-1, k, k+1 are live, the fourth slot is free.
// VAR selects the operation's variant: VE-FD skips cells whose incoming value
// is at least thr (exact pruning); the decision uses the two-phase interval
// form (iv_disc / iv_ends with TWO_PHASE).
// PIPE selects the decision step's form: true (single pair: a wave or two per
// SIMD, latency-bound) computes every row's intervals without branches one
// step ahead; false (batch rounds: several waves per SIMD, throughput-bound)
// computes a row's intervals only where the row is live somewhere in the wave.
template <int OP, int D, typename T, int SYNC, int R, bool VAR = false, bool PIPE = true>
__device__ inline void run_band(const SweepJob &a, const int band, T (&qring)[4][D][64],
                                T (&sring)[4][Staged<OP>::value][64], unsigned long long &evals,
                                unsigned long long *tlog, unsigned long long &skips) {
    using N = Num<T>;
    constexpr int NS = (OP == OP_DEC) ? 2 : 0;  // per-column terms staged
    constexpr int NP = (OP == OP_DFD) ? R : R + 1;  // P points a lane holds
    const int lane = (int)threadIdx.x;
    const T *P = (const T *)a.P;
    const T *Q = (const T *)a.Q;
    const T *qseg = (const T *)a.qseg;
    T *slots = (T *)a.slots;
    const long long rows = a.rows, cols = a.cols;
    // Points the sweep reads: DFD reads q_j (j < cols = n), the others q_j, q_{j+1}.
    const long long npts = (OP == OP_DFD) ? cols : cols + 1;
    const long long prow_max = (OP == OP_DFD) ? rows - 1 : rows;  // last P index read
    const T big = N::big();
    const T par = (T)a.param;
    const T r2 = par * par;
    const T thr = (T)a.thr;
    const T dead_out = (OP == OP_DEC) ? T(-1) : big;  // value a dead cell passes on

    if (tlog && lane == 0) tlog[4 * band] = wall_clock64();

    const long long i0 = (long long)band * 64 * R + (long long)R * lane;  // first row
    T p[NP][D];
#pragma unroll
    for (int t = 0; t < NP; ++t) {
        long long pi = i0 + t;
        pi = pi < prow_max ? pi : prow_max;
#pragma unroll
        for (int k = 0; k < D; ++k) p[t][k] = P[pi * D + k];
    }
    T A4a[R], inv2a[R], bot[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
        A4a[t] = T(0);
        inv2a[t] = T(0);
        if (OP == OP_DEC) {
            T Aa = seg_len2<T, D>(p[t], p[t + (OP == OP_DFD ? 0 : 1)]);
            A4a[t] = T(4) * Aa;
            inv2a[t] = T(0.5) / Aa;
        }
        const long long it = i0 + t;
        if (OP == OP_DFD) bot[t] = big;
        else if (OP == OP_VEFD) bot[t] = (it == 0) ? par * par : big;
        else bot[t] = (it == 0) ? T(0) : T(-1);
    }
    T out = (OP == OP_DEC) ? T(-1) : big;

    // Q staging: chunk 0 straight into LDS, chunk 1 into registers.
    T pre[D], pres[NS > 0 ? NS : 1];
    {
        const long long p0 = lane < npts ? lane : npts - 1;
#pragma unroll
        for (int k = 0; k < D; ++k) qring[0][k][lane] = Q[p0 * D + k];
        if (NS > 0) {
            const long long s0 = lane < cols ? lane : cols - 1;
            sring[0][0][lane] = qseg[2 * s0];
            sring[0][NS - 1][lane] = qseg[2 * s0 + 1];
        }
        long long p1 = 64 + lane;
        p1 = p1 < npts ? p1 : npts - 1;
#pragma unroll
        for (int k = 0; k < D; ++k) pre[k] = Q[p1 * D + k];
        if (NS > 0) {
            const long long s1 = p1 < cols ? p1 : cols - 1;
            pres[0] = qseg[2 * s1];
            pres[NS - 1] = qseg[2 * s1 + 1];
        }
    }
    const bool has_pred = band > 0;
    const bool has_succ = band + 1 < a.nbands;
    const T *sin = slots + (long long)((band + 1) & 1) * cols;
    T *sout = slots + (long long)(band & 1) * cols;
    T bin = T(0), bout = T(0);
    const long long nsteps = cols + 63;
    // q_j of the lane's current column; every lane starts at column 0.
    T qcar[D];
#pragma unroll
    for (int k = 0; k < D; ++k) qcar[k] = qring[0][k][0];
    bool reload = false;
    bool gave_up = false;  // a wait exceeded its bound: wait no more (see Watch)
    // Decision: the free intervals of the current step's cells, once computed
    // (have_iv), R rows each: right edge [rl, ru], top edge [tl, tu].
    T rl[R], ru[R], tl[R], tu[R];
    bool have_iv = false;
    // Decision: the terminal cell's rightUB, held by the lane that owns row
    // rows - 1 and stored after the band's last step.
    T term = T(-3);
    // The 2R free intervals of the lane's cells in column jc (rows i0 + t),
    // phase by phase so that their square-root chains sit side by side for
    // the scheduler.  Columns outside the grid read ring slots that hold other
    // columns' points; those intervals are never used.
    auto dec_ivs = [&](const long long jc, T (&l1)[R], T (&u1)[R], T (&l2)[R], T (&u2)[R])
                       __attribute__((always_inline)) {
      if constexpr (OP == OP_DEC) {
        const int ra = (int)((jc >> 6) & 3), oa = (int)(jc & 63);
        const long long jb = jc + 1;
        const int rb = (int)((jb >> 6) & 3), ob = (int)(jb & 63);
        T qa[D], qb[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            qa[k] = qring[ra][k][oa];
            qb[k] = qring[rb][k][ob];
        }
        const T A4b = sring[ra][0][oa];
        const T inv2b = sring[ra][NS - 1][oa];
        T Bv[2 * R], Dv[2 * R];
        typename RootT<T>::type Sv[2 * R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            iv_disc<T, D, VAR>(p[t + 1], qa, qb, r2, A4b, Bv[2 * t], Dv[2 * t]);
            iv_disc<T, D, VAR>(qb, p[t], p[t + 1], r2, A4a[t], Bv[2 * t + 1], Dv[2 * t + 1]);
        }
#pragma unroll
        for (int i = 0; i < 2 * R; ++i) Sv[i] = iv_root(Dv[i]);
#pragma unroll
        for (int t = 0; t < R; ++t) {
            iv_ends<T, VAR>(Bv[2 * t], Dv[2 * t], Sv[2 * t], inv2b, l1[t], u1[t]);
            iv_ends<T, VAR>(Bv[2 * t + 1], Dv[2 * t + 1], Sv[2 * t + 1], inv2a[t], l2[t], u2[t]);
        }
      }
    };

    for (long long s = 0; s < nsteps; ++s) {
        const long long j = s - lane;
        if ((s & (kHand<OP> - 1)) == 0) {
            const long long kc = s >> 6;
            if ((s & 63) == 0) {
                // Stage chunk k+1 (in registers) into LDS.
                const int slot = (int)((kc + 1) & 3);
#pragma unroll
                for (int k = 0; k < D; ++k) qring[slot][k][lane] = pre[k];
                if (NS > 0) {
                    sring[slot][0][lane] = pres[0];
                    sring[slot][NS - 1][lane] = pres[NS - 1];
                }
            }
            // Boundary values of columns s .. s+kHand-1, into lanes (s mod 64)
            // .. (s mod 64) + kHand - 1 of `bin` (lane l holds column
            // 64 floor(s / 64) + l), kHand columns at a time so that a band
            // trails its predecessor by about 64 + kHand steps, not 64 + 64.
            // Every lane loads (the other
            // lanes' values are consumed or not yet needed, and are reloaded
            // before use), clamped, so the load count is fixed and the wait for
            // `bin` does not cover the prefetch issued after it.
            const bool in_slice = (lane & (64 - kHand<OP>)) == (int)(s & (64 - kHand<OP>));
            if (has_pred && s < cols) {
                const long long need = (s + kHand<OP> < cols) ? s + kHand<OP> : cols;
                unsigned polls = 0;
                // Every lane reads the same flag: taking lane 0's value makes
                // the wait a scalar loop (no exec-mask bookkeeping per poll).
                while (!gave_up && __builtin_amdgcn_readfirstlane(flag_load<SYNC>(&a.progress[band - 1])) < (int)need) {
                    if (watch_expired(a.watch, polls)) gave_up = true;
                    else __builtin_amdgcn_s_sleep(1);
                }
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                const long long c = ((s & ~63LL) + lane < cols) ? (s & ~63LL) + lane : cols - 1;
                bin = column_load<SYNC>(sin + c);
                if (tlog && lane == 0 && s == 0) tlog[4 * band + 1] = wall_clock64();
            }
            // Dead block: every lane's state and in-flight value is dead and so
            // is the boundary input for columns s..s+kHand-1.  Then every cell
            // of steps s..s+kHand-1 is dead; skip them, publishing dead outputs
            // for columns s-63..s-64+kHand (lane 63's outputs of those steps).
            bool skip = false;
            if (OP != OP_DFD && s >= 64 && s + 128 < cols) {
                bool mine = is_dead<OP>(out, thr);
#pragma unroll
                for (int t = 0; t < R; ++t) mine = mine && is_dead<OP>(bot[t], thr);
                const T edge = (has_pred && in_slice) ? bin : dead_out;
                mine = mine && is_dead<OP>(edge, thr);
                skip = __builtin_amdgcn_read_exec() == __ballot(mine);
            }
            if ((s & 63) == 0) {
                // Fetch chunk k+2 into registers; it is written to LDS 64 steps later.
                long long pn = (kc + 2) * 64 + lane;
                pn = pn < npts ? pn : npts - 1;
#pragma unroll
                for (int k = 0; k < D; ++k) pre[k] = Q[pn * D + k];
                if (NS > 0) {
                    const long long ps = pn < cols ? pn : cols - 1;
                    pres[0] = qseg[2 * ps];
                    pres[NS - 1] = qseg[2 * ps + 1];
                }
            }
            if (skip) {
                ++skips;  // whole dead blocks skipped (wave-uniform)
                if (has_succ) {
                    // Columns s-64 (lane s mod 64 of `bout`, output at step
                    // s-1) .. s-65+kHand form a full chunk.
                    if (in_slice && lane != (int)(s & 63)) bout = dead_out;
                    if (in_slice) column_store<SYNC>(sout + (s & ~63LL) - 64 + lane, bout);
                    __syncthreads();
                    if (lane == 0) flag_store<SYNC>(&a.progress[band], (int)(s - 64 + kHand<OP>));
                    bout = dead_out;  // column s-64+kHand
                }
                reload = true;
                s += kHand<OP> - 1;
                continue;
            }
        }
        T bval;
        if (has_pred) bval = lane_read(bin, (int)(s & 63));
        else if (OP == OP_DFD) bval = (s == 0) ? T(0) : big;
        else if (OP == OP_VEFD) bval = big;
        else bval = T(-1);
        T left = shift_up(out);
        if (lane == 0) left = bval;

        if (OP == OP_DEC) {
            // Decision cells (.cu:1173-1244).  A lane can hold a live cell
            // this step only if its input from below or one of its rows' own
            // values is not unreachable (< 0); when no lane can, every cell of
            // the step gives (-1, -2, -1) and the step costs nothing more.
            // Otherwise the rows' free intervals are propagated up the rows;
            // the propagation rules give a dead cell's (-1, -2, -1) from any
            // intervals, so no lane needs a branch of its own.  The intervals
            // depend on the points only: a live step computes the next step's
            // (software pipelining), so that only the propagation, a few
            // selects per row, sits on the step-to-step path.  A step whose
            // intervals were not computed the step before (the first live step
            // after dead ones) computes its own first.
            const bool inr = j >= 0 && j < cols && i0 < rows;
          if constexpr (!PIPE) {
            // Throughput form (batch rounds, several waves per SIMD): row by
            // row, a row's free intervals computed only where one of its
            // lanes is live (the branch skips rows dead across the wave).
            if (inr) {
                const int ra = (int)((j >> 6) & 3), oa = (int)(j & 63);
                const long long jb = j + 1;
                const int rb = (int)((jb >> 6) & 3), ob = (int)(jb & 63);
                T qa[D], qb[D];
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    qa[k] = qring[ra][k][oa];
                    qb[k] = qring[rb][k][ob];
                }
                const T A4b = sring[ra][0][oa];
                const T inv2b = sring[ra][NS - 1][oa];
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const long long it = i0 + t;
                    if (R > 1 && it >= rows) break;
                    const T leftLB = left, botLB = bot[t];
                    T rLB = T(-1), rUB = T(-2), tLB = T(-1);
                    // A NaN input (two-phase form) is not unreachable: the
                    // test is on < 0, not on >= 0.
                    if (!(leftLB < T(0) && botLB < T(0))) {
                        ++evals;
                        T B1, D1, B2, D2, tUB = T(-2);
                        iv_disc<T, D, VAR>(p[t + 1], qa, qb, r2, A4b, B1, D1);
                        iv_disc<T, D, VAR>(qb, p[t], p[t + 1], r2, A4a[t], B2, D2);
                        // The root only where the segment meets the disc (a
                        // branch: waves whose lanes all miss skip it).
                        if (D1 >= T(0)) iv_ends<T, VAR>(B1, D1, iv_root(D1), inv2b, rLB, rUB);
                        if (D2 >= T(0)) iv_ends<T, VAR>(B2, D2, iv_root(D2), inv2a[t], tLB, tUB);
                        if (botLB < T(0)) {
                            const bool inv = leftLB > rUB || leftLB < T(0);
                            const T t2 = N::max(rLB, leftLB);
                            rLB = inv ? T(-1) : t2;
                            rUB = inv ? T(-2) : rUB;
                        }
                        if (leftLB < T(0)) {
                            const bool inv = botLB > tUB || botLB < T(0);
                            const T t2 = N::max(tLB, botLB);
                            tLB = inv ? T(-1) : t2;
                        }
                    }
                    bot[t] = tLB;
                    left = rLB;
                    if (it == rows - 1 && j == cols - 1) term = rUB;
                }
                out = left;
            }
          } else {
            bool maybe = inr && !(left < T(0));
#pragma unroll
            for (int t = 0; t < R; ++t) maybe = maybe || (inr && !(bot[t] < T(0)));
            if (__ballot(maybe) != 0) {
                if (!have_iv || reload) dec_ivs(j, rl, ru, tl, tu);
                T nrl[R], nru[R], ntl[R], ntu[R];
                dec_ivs(j + 1, nrl, nru, ntl, ntu);
                // Every lane propagates (no branch, so the next step's
                // intervals above and this chain interleave); lanes outside
                // the grid keep their state.  Rows past the last (padding of
                // the last band) are propagated too, but keep their state and
                // hand -1 upward: they stay dead, so whole dead blocks of the
                // last band can still be skipped.
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    const long long it = i0 + t;
                    const T leftLB = left, botLB = bot[t];
                    // Cells whose inputs are both unreachable are not
                    // counted; a NaN input (two-phase form) is not
                    // unreachable, so the test is on < 0, not on >= 0.
                    evals += (inr && it < rows && !((leftLB < T(0)) & (botLB < T(0)))) ? 1 : 0;
                    // (.cu:1231-1244) as selects: where the input along the
                    // row is unreachable, the interval on the right edge
                    // starts at the input from below, if that lies in it;
                    // where the input from below is unreachable, the top
                    // edge's starts at the input along the row.
                    const bool bdead = botLB < T(0), ldead = leftLB < T(0);
                    const bool rinv = (leftLB > ru[t]) | ldead;
                    const bool tinv = (botLB > tu[t]) | bdead;
                    const T rmax = N::max(rl[t], leftLB);
                    const T tmax = N::max(tl[t], botLB);
                    const T rLB = bdead ? (rinv ? T(-1) : rmax) : rl[t];
                    const T rUB = (bdead & rinv) ? T(-2) : ru[t];
                    const T tLB = ldead ? (tinv ? T(-1) : tmax) : tl[t];
                    bot[t] = (inr && it < rows) ? tLB : bot[t];
                    left = (it < rows) ? rLB : T(-1);
                    term = (inr && it == rows - 1 && j == cols - 1) ? rUB : term;
                }
                out = inr ? left : out;
#pragma unroll
                for (int t = 0; t < R; ++t) {
                    rl[t] = nrl[t];
                    ru[t] = nru[t];
                    tl[t] = ntl[t];
                    tu[t] = ntu[t];
                }
                have_iv = true;
            } else {
                have_iv = false;
                // Every row of the lane stays at -1 (already so); the terminal
                // cell, if this is it, has rightUB -2.
                out = inr ? T(-1) : out;
                term = (inr && i0 <= rows - 1 && rows - 1 < i0 + R && j == cols - 1) ? T(-2) : term;
            }
          }
            reload = false;
        }
        if (OP != OP_DEC && j >= 0 && j < cols && i0 < rows) {
            // q_{j+1} from the LDS ring; q_j is last step's q_{j+1}, carried in
            // registers (re-read from the ring after a skipped chunk).
            const int ra = (int)((j >> 6) & 3), oa = (int)(j & 63);
            const long long jb = j + 1;
            const int rb = (int)((jb >> 6) & 3), ob = (int)(jb & 63);
            T qa[D], qb[D];
            if (OP == OP_DFD || reload) {
#pragma unroll
                for (int k = 0; k < D; ++k) qcar[k] = qring[ra][k][oa];
            }
#pragma unroll
            for (int k = 0; k < D; ++k) {
                qa[k] = qcar[k];
                qb[k] = (OP == OP_DFD) ? qa[k] : qring[rb][k][ob];
            }
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const long long it = i0 + t;
                if (R > 1 && it >= rows) break;
                const bool last = (it == rows - 1) && (j == cols - 1);
                if (OP == OP_DFD) {
                    T d = dist2<T, D>(p[t], qa);
                    T in = N::min(left, bot[t]);
                    T top = N::max(d, in);
                    left = N::min(top, bot[t]);
                    bot[t] = top;
                    if (last) {
                        a.result[0] = (double)N::sqrt(top);
                        a.result[1] = (double)top;
                    }
                } else if (OP == OP_VEFD) {
                    T in = N::min(left, bot[t]);
                    T Tv = in, Rv = in;
                    if (!VAR || in < thr) {
                        T wr = point_seg2<T, D>(p[t + 1], qa, qb);
                        T wt = point_seg2<T, D>(qb, p[t], p[t + 1]);
                        Tv = N::max(wt, in);
                        Rv = N::max(wr, in);
                        ++evals;
                    }
                    bot[t] = Tv;
                    left = Rv;
                    if (last) {
                        a.result[0] = (double)N::sqrt(N::min(Tv, Rv));
                        a.result[1] = (double)N::min(Tv, Rv);
                    }
                }
            }
            out = left;
#pragma unroll
            for (int k = 0; k < D; ++k) qcar[k] = qb[k];
        }
        reload = false;

        if (has_succ && s >= 63) {
            const long long col = s - 63;
            const T v = lane_read(out, 63);
            if (lane == (int)(col & 63)) bout = v;
            if ((col & (kHand<OP> - 1)) == kHand<OP> - 1 || col == cols - 1) {
                // Columns kHand floor(col / kHand) .. col, from their lanes.
                if ((lane & (64 - kHand<OP>)) == (int)(col & (64 - kHand<OP>)) && lane <= (int)(col & 63))
                    column_store<SYNC>(sout + (col & ~63LL) + lane, bout);
                __syncthreads();
                if (lane == 0) flag_store<SYNC>(&a.progress[band], (int)(col + 1));
            }
        }
    }
    if (OP == OP_DFD) evals += (unsigned long long)((i0 < rows) ? ((rows - i0 < R) ? rows - i0 : R) : 0) * cols;
    if (OP == OP_DEC && i0 <= rows - 1 && rows - 1 < i0 + R) *a.result = (double)term;
    if (tlog && lane == 0) {
        tlog[4 * band + 2] = wall_clock64();
        tlog[4 * band + 3] = ((unsigned long long)__builtin_amdgcn_s_getreg((20 << 0) | (0 << 6) | (31 << 11)) << 32) |
                              (unsigned long long)__builtin_amdgcn_s_getreg((4 << 0) | (0 << 6) | (31 << 11));
    }
}

template <int OP, int D, typename T, int SYNC, int R, bool VAR = false>
// A single-pair sweep launches at most about two waves per SIMD, so the
// compiler may spend registers on overlapping a step's independent chains
// rather than on fitting more waves.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 2))) void band_sweep(SweepArgs a) {
    __shared__ T qring[4][D][64];
    __shared__ T sring[4][Staged<OP>::value][64];
    const int lane = (int)threadIdx.x;
    SweepJob job{a.P, a.Q, a.qseg, a.slots, a.progress, a.result, a.rows, a.cols, a.nbands,
                 a.param, a.thr, a.watch};
    unsigned long long evals = 0, skips = 0;
    for (;;) {
        int band = 0;
        if (lane == 0) band = atomicAdd(a.queue, 1);
        band = __builtin_amdgcn_readfirstlane(band);
        if (band >= a.nbands) break;
        run_band<OP, D, T, SYNC, R, VAR>(job, band, qring, sring, evals, a.tlog, skips);
    }
    if (a.evals) {
        atomicAdd(a.evals, evals);
        if (lane == 0) atomicAdd(a.evals + 1, skips);
    }
}

// Per-column terms of the decision sweep (.cu:1129-1144): 4*Ab and 0.5/Ab.
template <int D, typename T>
__global__ void qseg_terms(const T *Q, long long cols, T *qseg) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cols) return;
    T Ab = seg_len2<T, D>(Q + j * D, Q + (j + 1) * D);
    qseg[2 * j] = T(4) * Ab;
    qseg[2 * j + 1] = T(0.5) / Ab;
}

}  // namespace rfd
