// Fused bound sweep: DFD (upper bound) and VE-FD (lower bound) of one pair in
// a single pass over the m x n point grid, with the cell weights computed off
// the serial min/max chain.
//
// Why: in band_sweep one wave both computes a cell's weights (two point-to-
// segment distances with a divide each for VE-FD) and runs the min/max
// recursion, so a band advances at one wave's latency-bound pace, bands wait
// for each other in a chain, and at 200k x 200k a third of the run is the
// pipeline filling and draining.  Here a workgroup owns a band and splits the
// work: producer waves compute the weights of the band's next steps into an
// LDS ring (they depend on the points only, never on the recursion), and one
// chain wave runs the recursion, a few min/max per cell.  A band then advances
// at the rate its CU can compute weights, and the chain's own latency is off
// the critical path, so the fill is short and every SIMD has work.
//
// Grid and cells.  DFD cell (i, j) couples p_i and q_j (DFD<D>, .cu:458-641).
// VE-FD cell (i', j') couples segments p_i'p_i'+1 and q_j'q_j'+1 (VEFD<D>,
// .cu:200-432); it is evaluated at DFD cell (i' + 1, j' + 1), whose two
// recursion inputs (the cells below and to the left) are exactly VE-FD's.  Row
// 0 and column 0 carry VE-FD's boundary: weights there are set so that the
// uniform min/max formula reproduces the boundary values (row 0 passes +big
// upward; column 0 leaves row 1 holding r^2 and every other row +big).  Rows
// past m - 1 (padding of the last band) get weights that pass VE-FD values
// through unchanged.
//
// Layout.  Band b holds rows 64Rb .. 64R(b+1)-1; lane l of the chain wave owns
// rows 64Rb + Rl + t, t < R, and at step s evaluates column s - l for all of
// them (the anti-diagonal of band_sweep).  Producer wave w computes, for every
// lane and step, the weights of a block of consecutive rows t into slot s mod
// KS of the ring: three doubles per cell (DFD distance^2, VE-FD right and top
// weights), lane-contiguous.  Producers read q_{s-l} from a ring of Q columns
// in LDS (each wave stores every 64-column chunk itself, one chunk ahead, from
// a coalesced load issued a chunk before that), and carry q_{j-1} and the
// differences p_i - q_{j-1} from the step before.
//
// Hand-offs.  Inside the workgroup: producer w publishes the count of steps it
// has written (an LDS word stored after its weights: a wave's LDS operations
// complete in order); the chain publishes the count of steps whose ring slot
// it has read (once its last loads from the slot are issued) and the count of
// steps it has finished; a producer writes step s only once the chain has read
// step s - KS.
// Between bands: as band_sweep (scoped atomics, in-order band queue), with two
// values per column (DFD, VE-FD), but carried by a third role, the courier
// wave, through two LDS rings, so that the chain never waits on global memory.
#pragma once

#include <type_traits>

#include "sweep_kernels.hpp"

namespace rfd {

struct BoundsArgs {
    const double *P, *Q;   // AoS points (m and n)
    double *slots;         // [2 buffers][2 kinds: DFD, VE-FD][cols]
    int *progress;         // per band: columns published
    int *queue;            // next band to hand out
    double *result;        // [0] DFD, [1] DFD^2, [2] VE-FD, [3] VE-FD^2
    long long rows, cols;  // m, n
    int nbands;
    double r2;             // endpoint bound squared: VE-FD's entry value
    unsigned long long *tlog;  // optional, 4 per band (as band_sweep)
    Watch watch;           // bound on the hand-off waits
};

// Workgroup hand-off words in LDS.  Relaxed atomics (not volatile) so that,
// once inlined, they lower to ds_read/ds_write: a generic (flat) access would
// also wait on this wave's outstanding global-memory operations.
__device__ inline int lds_get(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline void lds_put(int *p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Columns of Q staged in LDS for the producers: 4C chunks of 64 columns,
// component-major (component k of column j at k * qring_cols(C) + (j mod
// qring_cols(C))).
constexpr int qring_cols(int C) { return 256 * C; }
// Components per Q-ring column: q (3), and with five or more producers also
// the column's segment terms q_j - q_{j-1} (3), |q_j - q_{j-1}|^2 and the
// weight test's limit (far_end_lim), staged once per chunk instead of
// recomputed by every producer wave every step.
// Sweep modes: DFD and VE-FD in FP64 (three weights per cell: DFD distance^2,
// VE-FD right and top weights); DFD alone (one weight per cell, q alone in the
// ring); DFD in FP64 and VE-FD in FP32 (three weights per cell, the VE-FD ones
// FP32 values held in doubles, so the chain's min / max, which are exact, give
// the FP32 recursion's values; the ring holds q and, packed two to a double,
// q, its segment terms and test limit rounded to FP32).
constexpr int kModeF64 = 3, kModeDfd = 1, kModeV32 = 2;
constexpr int mode_weights(int mode) { return mode == kModeDfd ? 1 : 3; }
constexpr bool mode_ve(int mode) { return mode != kModeDfd; }
constexpr int qring_comps(int NPW, int mode = kModeF64) {
    return NPW < 5 ? 3 : mode == kModeF64 ? 8 : mode == kModeV32 ? 7 : 3;
}

// LDS plan of one workgroup.  The weight ring holds KS steps of 512 W R C bytes (W weights per cell)
// (C columns per step), as many as fit (at most 16) beside the Q ring, the
// boundary rings and the hand-off words: within 160 KB for the configurations
// with five or more producers (one workgroup per CU), within a third of that
// otherwise (three workgroups per CU).  The chain's step loop runs in trips
// of KS steps, one per ring slot; the boundary rings hold KB positions, a
// multiple of C KS (so that a trip's positions never wrap), at least 128
// (64 with few producers, whose steps are short).
template <int R, int NPW, int C = 1, int MODE = kModeF64> struct BoundsRing {
    static constexpr int kb(int ks) {
        const int min_cols = (NPW < 5) ? 64 : 128;
        return ((min_cols + C * ks - 1) / (C * ks)) * (C * ks);
    }
    static constexpr int bytes(int ks) {
        return ks * 512 * mode_weights(MODE) * R * C + qring_comps(NPW, MODE) * qring_cols(C) * 8 + 2 * 2 * kb(ks) * 8 +
               2 * (64 + C * ks) * 8 + 2 * 64 * 4 + 16 * 4 + 64;
    }
    static constexpr int cap = (NPW < 5) ? 163840 / 3 : 163840;
    static constexpr int pick() {
        int k = 16;
        while (k > 2 && bytes(k) > cap) --k;
        return k;
    }
    static constexpr int KS = pick();
    static constexpr int KB = kb(KS);
    static_assert(bytes(KS) <= cap, "ring needs two slots");
};

__device__ inline double dpp_f64(double old, double src, int ctrl_shl) {
    unsigned long long o = (unsigned long long)__double_as_longlong(old);
    unsigned long long v = (unsigned long long)__double_as_longlong(src);
    int lo, hi;
    if (ctrl_shl) {  // wave_shl:1, lane l takes lane l+1; lane 63 keeps `old`
        lo = __builtin_amdgcn_update_dpp((int)(unsigned)(o & 0xffffffffull), (int)(unsigned)(v & 0xffffffffull),
                                         0x130, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(v >> 32), 0x130, 0xf, 0xf, false);
    } else {         // wave_shr:1, lane l takes lane l-1; lane 0 keeps `old`
        lo = __builtin_amdgcn_update_dpp((int)(unsigned)(o & 0xffffffffull), (int)(unsigned)(v & 0xffffffffull),
                                         0x138, 0xf, 0xf, false);
        hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(v >> 32), 0x138, 0xf, 0xf, false);
    }
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) |
                                            (unsigned long long)(unsigned)lo));
}

// Two FP32 values in one double's bits (low, high), for the Q ring's FP32
// entries.
__device__ inline double pack_f2(float lo, float hi) {
    return __longlong_as_double((long long)(((unsigned long long)__float_as_uint(hi) << 32) |
                                            (unsigned long long)__float_as_uint(lo)));
}
__device__ inline float unpack_lo(double v) { return __uint_as_float((unsigned)(unsigned long long)__double_as_longlong(v)); }
__device__ inline float unpack_hi(double v) {
    return __uint_as_float((unsigned)((unsigned long long)__double_as_longlong(v) >> 32));
}

// |x - fma(t, b, a)|^2: a point-segment weight at clamped parameter t, the
// operation order of point_seg2 (.cu:375-394) with the segment terms hoisted;
// T is double, or float for the FP32 VE-FD.
__device__ inline double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ inline float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
template <int D, typename T>
__device__ inline T seg_w_at(const T *x, const T *a, const T *b, const T t) {
    T w = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const T e = x[k] - fma_t(t, b[k], a[k]);
        w = fma_t(e, e, w);
    }
    return w;
}

// A VE-FD weight is the squared distance of a point x to a segment [a, a + b]
// (b = e - a rounded, e the far end, l = |b|^2, ap = x - a), with clamped
// parameter t = min(1, max(0, (ap . b) / l)): 0 when ap . b <= 0 and 1 when
// ap . b >= l, whatever the quotient rounds to.  The producer already holds
// |x - a|^2 and |x - e|^2 as DFD values (of the cell and its neighbours, each
// an FMA chain in the weight's own order), so:
//  - at t = 0 the projection fma(0, b, a) is a (a zero's sign aside, which no
//    square sees), and the weight is |x - a|^2, bit for bit;
//  - at t = 1 it is a + b rounded, which is e exactly where that sum rounds
//    back to it; `lim` is l for such segments and +inf for the others, so
//    ap . b >= lim means the weight is |x - e|^2, bit for bit;
//  - otherwise (a projection strictly inside the segment, or t = 1 where
//    a + b misses e) the weight is computed, by seg_w_at at seg_t_slow's t,
//    with the correctly rounded divide for the interior.
// The last case sits behind one branch per wave step (every row's two
// weights); on the 200k-point random walks a wave of one row takes it in
// about one step in six, and every column's and row's far end is exact.
__device__ inline double seg_t_slow(const double p, const double l) {
    return (p < l) ? __builtin_fmin(1.0, __builtin_fmax(0.0, p / l)) : 1.0;
}
__device__ inline float seg_t_slow(const float p, const float l) {
    return (p < l) ? __builtin_fminf(1.0f, __builtin_fmaxf(0.0f, p / l)) : 1.0f;
}

// The weight test's limit for a segment from a to a + b (b = e - a rounded,
// l = |b|^2): l where a + b rounds back to the far end e in every coordinate,
// else +inf (see seg_t_slow).
template <int D, typename T>
__device__ inline T far_end_lim(const T *a, const T *b, const T *e, const T l) {
    bool ex = true;
#pragma unroll
    for (int k = 0; k < D; ++k) ex = ex && (a[k] + b[k] == e[k]);
    return ex ? l : T(__builtin_huge_val());
}

// min / max of doubles that are never NaN.  Where the translation unit is
// compiled without NaN semantics (rfd_bounds.hip) these are the compiler's
// fmin / fmax, which then lower to bare v_min_f64 / v_max_f64 with known
// latency.  Elsewhere they are inline asm: __builtin_fmin would first quiet
// each operand the compiler cannot prove canonical (LDS loads, DPP results,
// loop-carried values), one extra instruction per operand.
#if defined(RFD_BOUNDS_NO_NANS)
static __device__ inline double dmin(double a, double b) { return __builtin_fmin(a, b); }
static __device__ inline double dmax(double a, double b) { return __builtin_fmax(a, b); }
#else
static __device__ inline double dmin(double a, double b) {
    double r;
    asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
static __device__ inline double dmax(double a, double b) {
    double r;
    asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
    return r;
}
#endif

// The lane index as a value the compiler cannot hoist out of the band loop:
// each role derives its addresses from it inside its own band, instead of at
// kernel entry, where the set-up of every role would stay live (and spill)
// across the others' loops.
__device__ inline int role_lane() {
    int lane = (int)(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    return lane;
}

// Row groups of the chain's software pipeline: NG groups of at most GMAX
// rows, group g holding rows start(g) .. start(g + 1) - 1.  Three groups at
// one column per step; groups of at most three rows at two, or where the
// registers are tight (G = 3: four waves per SIMD, 128 VGPRs).
template <int R, int C, int G = 0> struct ChainPlan {
    static constexpr int GT = (G > 0) ? G : (C == 1) ? (R + 2) / 3 : 3;
    static constexpr int NG = (R + GT - 1) / GT;
    static constexpr int GMAX = (R + NG - 1) / NG;
    static constexpr int start(int g) { return (R * g) / NG; }
    static constexpr int size(int g) { return (R * (g + 1)) / NG - (R * g) / NG; }
};

// Weights of group g for the C columns of one step, issued as one batch of LDS
// loads.  A slot holds [column][row][d, wr, wt][lane].
template <int R, int C, int G, int g, int MODE>
__device__ inline void chain_load(const double *w, double (&wv)[C][ChainPlan<R, C, G>::GMAX][mode_weights(MODE)]) {
    using PL = ChainPlan<R, C, G>;
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int i = 0; i < PL::size(g); ++i)
#pragma unroll
            for (int k = 0; k < mode_weights(MODE); ++k)
                wv[c][i][k] = w[(c * R + PL::start(g) + i) * mode_weights(MODE) * 64 + k * 64];
}

// Orders a group's results before any later memory access: the next loads
// into the group's registers must not be hoisted above the compute that
// frees them.
template <int R, int C, int G, int g, int MODE>
__device__ inline void chain_done(const double (&dbot)[R], const double (&vbot)[R], const double (&dl)[C],
                                  const double (&vl)[C]) {
    using PL = ChainPlan<R, C, G>;
#pragma unroll
    for (int c = 0; c < C; ++c) {
        if constexpr (mode_ve(MODE)) asm volatile("" ::"v"(dl[c]), "v"(vl[c]) : "memory");
        else asm volatile("" ::"v"(dl[c]) : "memory");
    }
#pragma unroll
    for (int i = 0; i < PL::size(g); ++i) {
        if constexpr (mode_ve(MODE)) asm volatile("" ::"v"(dbot[PL::start(g) + i]), "v"(vbot[PL::start(g) + i]) : "memory");
        else asm volatile("" ::"v"(dbot[PL::start(g) + i]) : "memory");
    }
}

// Group g's rows of one chain lane for the C columns of one step, column by
// column (a row's state after column c is its input at column c + 1).  Per
// row, with dl / vl the values from the row below and dbot / vbot the row's
// own values from the previous column:
//   DFD   (.cu:458-641)  top = max(d, min(dl, dbot)),  dl' = min(top, dbot)
//   VE-FD (.cu:200-432)  in = min(vl, vbot),  T = max(wt, in),  vl' = R = max(wr, in)
// (VE-FD only where the cell carries its weights).
template <int R, int C, int G, int g, int MODE>
__device__ inline void chain_rows(const double (&wv)[C][ChainPlan<R, C, G>::GMAX][mode_weights(MODE)], double (&dbot)[R],
                                  double (&vbot)[R], double (&dl)[C], double (&vl)[C]) {
    using PL = ChainPlan<R, C, G>;
    // One wait for the whole group (its loads were issued a group earlier),
    // not one counter wait per load.
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int i = 0; i < PL::size(g); ++i)
#pragma unroll
            for (int k = 0; k < mode_weights(MODE); ++k) asm volatile("" ::"v"(wv[c][i][k]));
#pragma unroll
    for (int c = 0; c < C; ++c)
#pragma unroll
        for (int i = 0; i < PL::size(g); ++i) {
            const int t = PL::start(g) + i;
            const double d = wv[c][i][0];
            // dl' = min(top, dbot) = min(max(d, dl), dbot) (min distributes
            // over max, and these doubles are never NaN): two operations on
            // the serial path instead of three, and no copy of the old dbot.
            const double dn = dmin(dmax(d, dl[c]), dbot[t]);
            dbot[t] = dmax(d, dmin(dl[c], dbot[t]));
            dl[c] = dn;
            if constexpr (mode_ve(MODE)) {
                const double wr = wv[c][i][1], wt = wv[c][i][2];
                const double vin = dmin(vl[c], vbot[t]);
                vbot[t] = dmax(wt, vin);
                vl[c] = dmax(wr, vin);
            }
        }
}

// Calls f(integral_constant<int, i>) for i = I .. N - 1.
template <int I, int N, typename F>
__device__ inline void static_for(F &&f) {
    if constexpr (I < N) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, N>(f);
    }
}

// LDS of one workgroup of the fused sweep.  The boundary rings hold, per
// position, the DFD and VE-FD values of one column: incoming column c at
// position c mod KB, outgoing column c (lane 63's output at step c / C + 63)
// at position (c + 63 C) mod KB, so that a step's positions follow its index
// and a trip's never wrap.  Lanes 0-62 of the chain store their outputs to
// slots of their own (odummy), so that the chain's store needs no lane mask.
template <int R, int NPW, int C = 1, int MODE = kModeF64> struct BoundsShared {
    static constexpr int KS = BoundsRing<R, NPW, C, MODE>::KS;
    static constexpr int KB = BoundsRing<R, NPW, C, MODE>::KB;
    // The small arrays first, so that their addresses fit a DS instruction's
    // 16-bit offset.
    alignas(16) int sprod[16];  // steps each producer has written
    int cons[64];               // steps the chain has finished (one copy per lane; read [0])
    int wfree[64];              // steps whose ring slot the chain has read (one copy per lane; read [0])
    int in_avail;               // incoming columns in the ring
    int out_done;               // outgoing columns the courier has published
    int item;
    alignas(16) double inring[2 * KB];
    alignas(16) double outring[2 * KB];
    alignas(16) double odummy[2 * (64 + C * KS)];
    alignas(16) double qring[qring_comps(NPW, MODE) * qring_cols(C)];
    alignas(16) double wring[KS * C * R * mode_weights(MODE) * 64];
};

// Steps of band `band`.  Lane l evaluates columns C (s - l) .. C (s - l) + C - 1
// at step s, so a band takes n / C + 63 steps.  The band holding row m - 1
// (the last band, which has no successor) stops at the step where that row's
// lane reaches column n - 1.
template <int R, int C>
__device__ inline int band_steps(const BoundsArgs &a, const int band) {
    const int cols = (int)a.cols;
    const long long toff = a.rows - 1 - (long long)band * 64 * R;
    return (toff >= 0 && toff < 64LL * R) ? cols / C + (int)(toff / R) : cols / C + 63;
}

template <int R, int NPW, int C, int MODE>
__device__ inline void bounds_chain(const BoundsArgs &a, const int band, BoundsShared<R, NPW, C, MODE> &sh) {
    constexpr int GP = (NPW + 2 > 12) ? 3 : 0;  // four waves per SIMD: 128 VGPRs
    using PL = ChainPlan<R, C, GP>;
    constexpr int NG = PL::NG;
    constexpr int KS = BoundsShared<R, NPW, C, MODE>::KS;
    constexpr int KB = BoundsShared<R, NPW, C, MODE>::KB;
    constexpr int SLOT = C * R * mode_weights(MODE) * 64;  // doubles per ring slot
    constexpr int NONE = 0x3fffffff;      // no limit
    static_assert(NG >= 2, "the pipeline needs two row groups");
    const int lane = role_lane();
    const double big = (double)FLT_MAX;
    // Steps and columns in 32 bits (the launch requires n < 2^31 - 64 and, at
    // two columns per step, n even), so that the loop's bookkeeping stays on
    // the scalar unit.
    const int cols = (int)a.cols;
    const int nsteps = band_steps<R, C>(a, band);
    const bool has_succ = band + 1 < a.nbands;
    // Terminal cell (m-1, n-1): when this band holds row m - 1, it is row tl
    // of lane tlane, which reaches column n - 1 (the last column of a step) at
    // the band's last step.
    const long long toff = a.rows - 1 - (long long)band * 64 * R;
    const bool has_term = toff >= 0 && toff < 64LL * R;
    const int tlane = has_term ? (int)(toff / R) : -1;
    const int tl = has_term ? (int)(toff - (long long)tlane * R) : 0;
    double dbot[R], vbot[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
        dbot[t] = big;
        vbot[t] = -big;  // column 0's weights then set the row's VE-FD entry value
    }
    // Row 0 of the pair: its VE-FD values are +big at every column (the
    // weights' boundary rule).  Entering it at +big, with +big from row -1,
    // gives exactly that whatever its weights, so the producers apply no rule
    // to it.
    if (band == 0 && lane == 0) vbot[0] = big;
    double dout[C], vout[C];
#pragma unroll
    for (int c = 0; c < C; ++c) dout[c] = vout[c] = big;
    if (a.tlog && lane == 0) a.tlog[4 * band] = wall_clock64();
    // The chain is the band's serial path and every later band waits on it:
    // its SIMD issues for it first, producers fill the remaining cycles.
    __builtin_amdgcn_s_setprio(3);

    // Hand-offs.  Step s reads its weights (loaded a step earlier) and loads
    // step s + 1's: the producers must have written min(s + 2, nsteps) steps.
    // It loads step s + 1's input from below: the courier must have filled
    // min(C (s + 2), n) incoming columns.  It stores lane 63's columns
    // C (s - 63) .. into the outgoing ring, over the columns KB earlier: the
    // courier must have published C (s - 62) - KB.  Steps below `lim` meet all
    // three by the counts last read, so a step checks one scalar; at `lim` the
    // chain reads all the counts (waiting while one falls short) and moves it.
    int avail = 0, in_known = 0, out_known = 0;
    // The producers' counts as plain loads, which the compiler merges into
    // wide LDS reads (an atomic load per word costs the chain 1% here, where
    // it reaches its limit every few steps).  Each count is one 32-bit word
    // stored by one producer, and LDS dword accesses are single-copy atomic,
    // so each element reads the old count or the new one; the clobber makes
    // every poll read the words again.
    auto read_prod = [&]() __attribute__((always_inline)) {
        asm volatile("" ::: "memory");
        int mn = sh.sprod[0];
#pragma unroll
        for (int k = 1; k < NPW; ++k) mn = sh.sprod[k] < mn ? sh.sprod[k] : mn;
        return __builtin_amdgcn_readfirstlane(mn);
    };
    auto wait_for = [&](const int s) __attribute__((always_inline)) {
        const int np = (s + 2 < nsteps) ? s + 2 : nsteps;
        const int ni = (C * (s + 2) < cols) ? C * (s + 2) : cols;
        const int no = C * (s - 62) - KB;
        unsigned polls = 0;
        for (;;) {
            avail = read_prod();
            in_known = __builtin_amdgcn_readfirstlane(lds_get(&sh.in_avail));
            out_known = __builtin_amdgcn_readfirstlane(lds_get(&sh.out_done));
            if (avail >= np && in_known >= ni && (!has_succ || out_known >= no)) break;
            if (watch_expired(a.watch, polls)) {
                // Given up (Watch): counts that never hold the chain again.
                avail = nsteps;
                in_known = cols;
                out_known = NONE;
                break;
            }
            // A longer nap: a chain that polls fast takes issue slots and LDS
            // reads from the producers it waits for.
            __builtin_amdgcn_s_sleep(8);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    auto limit = [&]() __attribute__((always_inline)) {
        const int lp = (avail >= nsteps) ? NONE : avail - 1;
        const int li = (in_known >= cols) ? NONE : in_known / C - 1;
        const int lo = has_succ ? (out_known + KB) / C + 63 : NONE;
        const int m = lp < li ? lp : li;
        return __builtin_amdgcn_readfirstlane(m < lo ? m : lo);
    };

    // Software pipeline over the row groups: a group's LDS loads are issued
    // while the group before it computes (group 0's, with the input from
    // below, while the last group of the step before computes), so no LDS
    // round trip sits between two steps' min/max chains, and two groups'
    // weights occupy registers at a time.  Loads are issued on every path
    // (the last step reads a slot and ring positions nobody uses), so that
    // the compiler's LDS counter waits stay exact.
    double w[NG][C][PL::GMAX][mode_weights(MODE)];
    double bD[C], bV[C];  // the next step's input from below (lane 0)
    auto load_in = [&](const double *p) __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < C; ++c) {
            bD[c] = p[2 * c];
            if constexpr (mode_ve(MODE)) bV[c] = p[2 * c + 1];
        }
    };
    wait_for(-1);
    if (a.tlog && lane == 0 && band > 0) a.tlog[4 * band + 1] = wall_clock64();
    int lim = limit();
    // Ring slots through NBASE base registers, NB slots each, so that every
    // offset fits a DS instruction's 16-bit field (the compiler folds no
    // larger one).  The offsets are hidden from it, or it would fold them back
    // into one base.
    constexpr int SLOTB = SLOT * 8;
    constexpr int NB = 65536 / SLOTB;
    constexpr int NBASE = (KS + NB - 1) / NB;
    const double *wb[NBASE];
#pragma unroll
    for (int i = 0; i < NBASE; ++i) {
        int off = i * NB * SLOT;
        asm volatile("" : "+v"(off));
        wb[i] = sh.wring + lane + off;
    }
    // Slot k: a compile-time k (integral_constant) through its base register,
    // a run-time one (the last steps) directly.
    auto slot_ptr = [&](auto k) __attribute__((always_inline)) {
        if constexpr (std::is_integral_v<decltype(k)>) {
            return (const double *)(sh.wring + lane + k * SLOT);
        } else {
            constexpr int kk = decltype(k)::value;
            return wb[kk / NB] + (kk % NB) * SLOT;
        }
    };
    chain_load<R, C, GP, 0, MODE>(slot_ptr(std::integral_constant<int, 0>{}), w[0]);
    load_in(sh.inring);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    // Steps finished, in a vector register: the store of the count then needs
    // no copy from a scalar one.
    int vdone = 0;
    asm volatile("" : "+v"(vdone));

    // Step s0 + k from ring slot k; the next step's weights come from slot
    // kn and its input from below from in_next; `ow` is this step's output
    // slot; lrel = lim - s0.  Every lane runs every step: before its first
    // column the producers' weights keep its rows at their initial values,
    // and after its last column it computes values nobody reads.
    auto step = [&](const int s0, auto k, auto kn, const double *in_next, double *ow, int &lrel)
                    __attribute__((always_inline)) {
        const double *wcur = slot_ptr(k);
        const double *wnext = slot_ptr(kn);
        // Input from below: lane l-1's outputs of the step before (same
        // columns), and for lane 0 the row under the band.
        double dl[C], vl[C];
#pragma unroll
        for (int c = 0; c < C; ++c) {
            dl[c] = dpp_f64(bD[c], dout[c], 0);
            if constexpr (mode_ve(MODE)) vl[c] = dpp_f64(bV[c], vout[c], 0);
        }
        static_for<0, NG>([&](auto G) __attribute__((always_inline)) {
            constexpr int g = decltype(G)::value;
            if constexpr (g + 1 < NG) {
                chain_load<R, C, GP, g + 1, MODE>(wcur, w[g + 1]);
                if constexpr (g + 2 == NG) {
                    // The step's last loads from its slot are issued, and a
                    // wave's LDS operations complete in order: producers may
                    // overwrite the slot once they see this count.
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    lds_put(&sh.wfree[lane], vdone + 1);
                }
            } else {
                if (__builtin_expect(k >= lrel, 0)) {
                    wait_for(s0 + k);
                    lim = limit();
                    lrel = lim - s0;
                }
                chain_load<R, C, GP, 0, MODE>(wnext, w[0]);
                load_in(in_next);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            chain_rows<R, C, GP, g, MODE>(w[g], dbot, vbot, dl, vl);
            chain_done<R, C, GP, g, MODE>(dbot, vbot, dl, vl);
        });
#pragma unroll
        for (int c = 0; c < C; ++c) {
            dout[c] = dl[c];
            ow[2 * c] = dout[c];
            if constexpr (mode_ve(MODE)) {
                vout[c] = vl[c];
                ow[2 * c + 1] = vout[c];
            }
        }
        // The outputs are in the ring and step s0 + k's weights and input in
        // registers: the step is finished.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        vdone += 1;
        lds_put(&sh.cons[lane], vdone);
    };

    // Trips of KS steps, one per ring slot, with every LDS offset constant;
    // bi = C s0 mod KB.
    const int nfull = nsteps - nsteps % KS;
    int s0 = 0, bi = 0;
    for (; s0 < nfull; s0 += KS) {
        const int bin = (bi + C * KS == KB) ? 0 : bi + C * KS;
        double *ow = (lane == 63) ? sh.outring + 2 * bi : sh.odummy + 2 * lane;
        int lrel = lim - s0;
        static_for<0, KS>([&](auto K) __attribute__((always_inline)) {
            constexpr int k = decltype(K)::value;
            step(s0, K, std::integral_constant<int, (k + 1) % KS>{},
                 sh.inring + 2 * (k + 1 < KS ? bi + C * (k + 1) : bin), ow + 2 * C * k, lrel);
        });
        bi = bin;
    }
    {
        // The last nsteps mod KS (< KS) steps, from slots 0 .. KS - 2.
        double *ow = (lane == 63) ? sh.outring + 2 * bi : sh.odummy + 2 * lane;
        int lrel = lim - s0;
        for (int k = 0; s0 + k < nsteps; ++k)
            step(s0, k, k + 1, sh.inring + 2 * (bi + C * (k + 1)), ow + 2 * C * k, lrel);  // k + 1 < KS
    }
    __builtin_amdgcn_s_setprio(0);
    if (has_term && lane == tlane) {
        // Terminal cell: its DFD value is the row's top; its VE-FD value is
        // min(T, R), where R left row tl and, rows above tl being padding that
        // passes VE-FD values through, is the lane's output.
        double topD = dbot[0], topV = vbot[0];
#pragma unroll
        for (int t = 1; t < R; ++t) {
            topD = (tl == t) ? dbot[t] : topD;
            topV = (tl == t) ? vbot[t] : topV;
        }
        a.result[0] = __builtin_sqrt(topD);
        a.result[1] = topD;
        if constexpr (mode_ve(MODE)) {
            const double resV = dmin(topV, vout[C - 1]);
            // The FP32 VE-FD's square root is FP32's (resV is an FP32 value).
            a.result[2] = (MODE == kModeV32) ? (double)__builtin_sqrtf((float)resV) : __builtin_sqrt(resV);
            a.result[3] = resV;
        }
    }
    if (a.tlog && lane == 0) {
        a.tlog[4 * band + 2] = wall_clock64();
        a.tlog[4 * band + 3] = ((unsigned long long)__builtin_amdgcn_s_getreg((20 << 0) | (0 << 6) | (31 << 11)) << 32) |
                               (unsigned long long)__builtin_amdgcn_s_getreg((4 << 0) | (0 << 6) | (31 << 11));
    }
}

// The band's global-memory hand-offs, kept off the chain: the courier copies
// the predecessor's published boundary columns into the incoming ring (the
// first band's row -1 it writes itself: DFD 0 at column 0 and +big after it,
// VE-FD +big) and publishes the chain's outgoing columns (scoped stores,
// s_waitcnt, flag), up to 64 per round.  Slots are [2 buffers][DFD, VE-FD]
// [cols].
// Every flag poll and column access is a device-scope access that L2 does not
// keep, so each is memory traffic.  Bands of 13 rows per lane, whose steps are
// long, therefore publish whole pieces of PUB = 16 columns (the band's last
// columns excepted), copy them in such pieces, and poll the predecessor only
// once every column it has published is in the ring and the chain has fewer
// than 32 incoming columns left: at 200k x 200k the sweep's time is unchanged
// and its memory traffic falls from 5.9 to 2.5 GB.  Shorter steps publish
// every column as soon as it is ready, since there the batches lengthen the
// pipeline's fill (5% at 7 rows per lane on a 100k x 120k pair).
template <int R, int NPW, int C, int MODE>
__device__ inline void bounds_courier(const BoundsArgs &a, const int band, BoundsShared<R, NPW, C, MODE> &sh) {
    constexpr int KB = BoundsShared<R, NPW, C, MODE>::KB;
    constexpr bool BATCH = R >= 13;
    constexpr int PUB = BATCH ? 16 : 1;
    const int lane = role_lane();
    const int cols = (int)a.cols;
    const double big = (double)FLT_MAX;
    const double *inD = a.slots + (long long)((band + 1) & 1) * 2 * cols;
    const double *inV = inD + cols;
    double *outD = a.slots + (long long)(band & 1) * 2 * cols;
    double *outV = outD + cols;
    const bool first = band == 0;
    int in_next = 0;
    int known = first ? cols : 0;  // the predecessor's published count, as last read
    int out_next = band + 1 < a.nbands ? 0 : cols;
    unsigned polls = 0;
    while (in_next < cols || out_next < cols) {
        bool moved = false;
        // Steps the chain has finished: their incoming columns are consumed
        // and their outputs are in the ring.
        const int done = __builtin_amdgcn_readfirstlane(lds_get(&sh.cons[0]));
        if (in_next < cols) {
            const int room = C * done + KB - in_next;
            if (room > 0) {
                if (known <= in_next && (!BATCH || in_next - C * done < 32))
                    known = __builtin_amdgcn_readfirstlane(flag_load<SYNC_SCOPED>(&a.progress[band - 1]));
                int n = known - in_next;
                n = n < room ? n : room;
                n = n < 64 ? n : 64;
                if (BATCH && in_next + n < known) n &= ~(PUB - 1);  // whole 128-byte pieces
                if (n > 0) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane < n) {
                        const int c = in_next + lane;
                        double vD, vV = big;
                        if (first) {
                            vD = (c == 0) ? 0.0 : big;
                        } else {
                            vD = column_load<SYNC_SCOPED>(inD + c);
                            if constexpr (mode_ve(MODE)) vV = column_load<SYNC_SCOPED>(inV + c);
                        }
                        const int p = c % KB;
                        sh.inring[2 * p] = vD;
                        if constexpr (mode_ve(MODE)) sh.inring[2 * p + 1] = vV;
                    }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the ring writes are done
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane == 0) lds_put(&sh.in_avail, in_next + n);
                    in_next += n;
                    moved = true;
                }
            }
        }
        if (out_next < cols) {
            // Lane 63 writes columns C (s - 63) .. at step s.
            int n = C * (done - 63) - out_next;
            n = n < 64 ? n : 64;
            // Whole 16-column (128-byte) pieces, the band's last columns
            // excepted: a store that covers part of a line makes the memory
            // side read the line to merge it.
            if (BATCH && out_next + n < cols) n &= ~(PUB - 1);
            if (n >= PUB || (n > 0 && out_next + n == cols)) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (lane < n) {
                    const int c = out_next + lane;
                    const int p = (c + 63 * C) % KB;
                    column_store<SYNC_SCOPED>(outD + c, sh.outring[2 * p]);
                    if constexpr (mode_ve(MODE)) column_store<SYNC_SCOPED>(outV + c, sh.outring[2 * p + 1]);
                }
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) flag_store<SYNC_SCOPED>(&a.progress[band], out_next + n);
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (lane == 0) lds_put(&sh.out_done, out_next + n);
                out_next += n;
                moved = true;
            }
        }
        if (moved) {
            polls = 0;
        } else {
            // Given up (Watch): the chain and producers run their steps
            // without this band's columns; the courier's part ends here.
            if (watch_expired(a.watch, polls)) break;
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// Rows of producer wave pw within a lane: a block of consecutive rows, RP of
// them for the first R - (RP - 1) NPW waves and RP - 1 for the others, RP =
// ceil(R / NPW).
template <int R, int NPW> struct ProducerRows {
    static constexpr int RP = (R + NPW - 1) / NPW;
    static constexpr int NBIG = R - (RP - 1) * NPW;  // waves with RP rows
    static constexpr int first(int pw) { return pw < NBIG ? pw * RP : NBIG * RP + (pw - NBIG) * (RP - 1); }
};

// Producer wave pw: rows t = first(pw) .. first(pw) + NR - 1 of every lane.
// `cons`: the chain's count of steps whose ring slot it has read.  Points are
// DS doubles apart, of which the first D are read (the batch stores every
// curve 3D: its 2D pairs run D = 2 over DS = 3, whose extra coordinate is 0).
template <int D, int R, int NPW, int KS, int C, int NR, int DS = D, int MODE = kModeF64>
__device__ inline void bounds_producer(const BoundsArgs &a, const int band, const int pw, double *wring,
                                       double *qring, int *prod, int *cons) {
    constexpr int QR = qring_cols(C);
    constexpr int RP = NR;
    const int tfirst = ProducerRows<R, NPW>::first(pw);
    const int lane = role_lane();
    const long long rows = a.rows;
    const int cols = (int)a.cols;
    const double big = (double)FLT_MAX;
    const double *P = a.P, *Q = a.Q;
    const long long i0 = (long long)band * 64 * R + (long long)R * lane;
    // Row terms: p_i, p_{i-1}, p_i - p_{i-1}, its squared length and the top
    // weight test's limit.  A row after the wave's first takes p_{i-1} from
    // the row before it (both indices clamp to m - 1 alike in the padding).
    double pi[RP][D], pm[RP][D], ab[RP][D], ab2[RP], lim2[RP];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        const long long i = i0 + tfirst + k;
        const long long ic = i < rows ? i : rows - 1;
        long long im = i - 1 < rows ? i - 1 : rows - 1;
        im = im > 0 ? im : 0;
        double acc = 0.0;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            pi[k][c] = P[ic * DS + c];
            pm[k][c] = (k == 0) ? P[im * DS + c] : pi[k - 1][c];
            ab[k][c] = pi[k][c] - pm[k][c];
            acc = __builtin_fma(ab[k][c], ab[k][c], acc);
        }
        ab2[k] = acc;
        lim2[k] = far_end_lim<D>(pm[k], ab[k], pi[k], acc);
    }
    // The same terms of the points rounded to FP32, for the FP32 VE-FD.
    float pif[RP][D], pmf[RP][D], abf[RP][D], ab2f[RP], lim2f[RP];
    if constexpr (MODE == kModeV32) {
#pragma unroll
        for (int k = 0; k < RP; ++k) {
            float acc = 0.0f;
#pragma unroll
            for (int c = 0; c < D; ++c) {
                pif[k][c] = (float)pi[k][c];
                pmf[k][c] = (k == 0) ? (float)pm[k][c] : pif[k - 1][c];
                abf[k][c] = pif[k][c] - pmf[k][c];
                acc = __builtin_fmaf(abf[k][c], abf[k][c], acc);
            }
            ab2f[k] = acc;
            lim2f[k] = far_end_lim<D>(pmf[k], abf[k], pif[k], acc);
        }
    }
    // Q ring.  Lane l reads columns C(s - l) .. C(s - l) + C - 1 at step s.
    // Every producer wave stores each chunk itself (the same values): the
    // chunk after the one holding column Cs, whenever Cs is a multiple of 64,
    // from registers loaded one chunk earlier.  Producers stay within KS steps
    // of each other (the chain waits for the slowest, and each runs at most KS
    // steps ahead of the chain), so no wave still reads the chunk 4C chunks
    // back, whose slots the stored chunk takes, and none reads a chunk before
    // its own store of it (a wave's LDS operations complete in order).
    constexpr bool QT = qring_comps(NPW, MODE) == 8;
    constexpr bool QT32 = qring_comps(NPW, MODE) == 7;
    static_assert(MODE != kModeV32 || QT32, "the FP32 VE-FD needs staged column terms (five or more producers)");
    // Stores column c's entries: q_c, and with the segment terms q_c -
    // q_{c-1}, its squared length and the weight test's limit (the operations
    // of the column terms below, so the same values), q_{c-1} being `qm`; for
    // the FP32 VE-FD the same terms of q rounded to FP32, packed two FP32
    // values to a double (q_c, the segment terms, length, limit).
    auto put_column = [&](const int idx, const double(&qc)[D], const double(&qm)[D]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < D; ++k) qring[k * QR + idx] = qc[k];
        if constexpr (QT32) {
            float F[8] = {0, 0, 0, 0, 0, 0, 0, 0}, qcf[D], qmf[D], cabf[D], c2 = 0.0f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                qcf[k] = (float)qc[k];
                qmf[k] = (float)qm[k];
                cabf[k] = qcf[k] - qmf[k];
                c2 = __builtin_fmaf(cabf[k], cabf[k], c2);
                F[k] = qcf[k];
                F[D + k] = cabf[k];
            }
            F[2 * D] = c2;
            F[2 * D + 1] = far_end_lim<D>(qmf, cabf, qcf, c2);
#pragma unroll
            for (int u = 0; u < 4; ++u) qring[(3 + u) * QR + idx] = pack_f2(F[2 * u], F[2 * u + 1]);
        }
        if constexpr (QT) {
            double c2 = 0.0, cab[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                cab[k] = qc[k] - qm[k];
                c2 = __builtin_fma(cab[k], cab[k], c2);
                qring[(3 + k) * QR + idx] = cab[k];
            }
            qring[6 * QR + idx] = c2;
            qring[7 * QR + idx] = far_end_lim<D>(qm, cab, qc, c2);
        }
    };
    double qn[D];
    {
        const int c0 = lane < cols ? lane : cols - 1;
        const int c1 = 64 + lane < cols ? 64 + lane : cols - 1;
        double q0[D], qm[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            q0[k] = Q[(long long)c0 * DS + k];
            qn[k] = Q[(long long)c1 * DS + k];
        }
        // Column 0 has no q_{-1}; its right weights are boundary values.
#pragma unroll
        for (int k = 0; k < D; ++k) qm[k] = dpp_f64(q0[k], q0[k], 0);
        put_column(lane, q0, qm);
    }
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    const double r2 = a.r2;
    const int nsteps = band_steps<R, C>(a, band);
    int cons_seen = 0;
    int slot = 0;
    // Carried from the column before: q_{j-1}, and per row p_i - q_{j-1},
    // which is the right weight's p_i - a term (the same operands and
    // operation as the DFD difference one column earlier, so the same value),
    // and that DFD value |p_i - q_{j-1}|^2, the right weight at t = 0.
    // Values at j <= 0 feed only weights that the boundary rules below
    // overwrite.
    // (For the FP32 VE-FD the same differences and values in FP32.)
    struct Carry {
        double q[D], x[RP][D], d[RP];
        float xf[RP][D], df[RP];
    };
    Carry cA, cB;
#pragma unroll
    for (int c = 0; c < D; ++c) cA.q[c] = 0.0;
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        cA.d[k] = 0.0;
        cA.df[k] = 0.0f;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            cA.x[k][c] = 0.0;
            cA.xf[k][c] = 0.0f;
        }
    }

    // q of the wave's next column, loaded a column ahead (a wave's columns
    // follow each other: j, j + 1, ...).  At j < 0 it reads ring slots not yet
    // holding Q; those values feed only weights that the boundary rules
    // overwrite.
    double qp[D], cp[D], c2p = 0.0, lp = 0.0, fp[4];
    auto prefetch = [&](const int qi) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < D; ++k) qp[k] = qring[k * QR + qi];
        if constexpr (QT32) {
#pragma unroll
            for (int u = 0; u < 4; ++u) fp[u] = qring[(3 + u) * QR + qi];
        }
        if constexpr (QT) {
#pragma unroll
            for (int k = 0; k < D; ++k) cp[k] = qring[(3 + k) * QR + qi];
            c2p = qring[6 * QR + qi];
            lp = qring[7 * QR + qi];
        }
    };
    prefetch((C * (0 - lane)) & (QR - 1));

    // Weights of column j for this wave's rows into w (one column of a slot).
    auto column = [&](const int j, const Carry &cp_, Carry &cn_, double *w, const bool edge)
                      __attribute__((always_inline)) {
        const double(&qa)[D] = cp_.q;
        double(&qb)[D] = cn_.q;
        const double(&xp)[RP][D] = cp_.x;
        const double(&dp)[RP] = cp_.d;
        double(&xn)[RP][D] = cn_.x;
        double(&dn)[RP] = cn_.d;
        if constexpr (MODE == kModeDfd) {
            // The DFD-only sweep: each cell's distance^2, and +big before a
            // lane's first column (the row's initial value).
#pragma unroll
            for (int k = 0; k < D; ++k) qb[k] = qp[k];
            prefetch((j + 1) & (QR - 1));
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                double d = 0.0;
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    const double x = pi[k][c] - qb[c];
                    d = __builtin_fma(x, x, d);
                }
                w[(tfirst + k) * 64] = (edge && j < 0) ? big : d;
            }
            return;
        }
        if constexpr (MODE == kModeV32) {
            // DFD in FP64; VE-FD weights in FP32 from the points rounded to
            // FP32, by the same rules, in FP32's operations.
#pragma unroll
            for (int k = 0; k < D; ++k) qb[k] = qp[k];
            float F[8];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                F[2 * u] = unpack_lo(fp[u]);
                F[2 * u + 1] = unpack_hi(fp[u]);
            }
            float qf[D], cabf[D];
#pragma unroll
            for (int c = 0; c < D; ++c) {
                qf[c] = F[c];
                cabf[c] = F[D + c];
            }
            const float c2f = F[2 * D], climf = F[2 * D + 1];
            prefetch((j + 1) & (QR - 1));
            float a1f[RP], a2f[RP], wrf[RP], wtf[RP];
            bool s1[RP], s2[RP], slow = false;
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                double d = 0.0;
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    const double x = pi[k][c] - qb[c];
                    d = __builtin_fma(x, x, d);
                }
                dn[k] = d;
                float e = 0.0f;
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    cn_.xf[k][c] = pif[k][c] - qf[c];
                    e = __builtin_fmaf(cn_.xf[k][c], cn_.xf[k][c], e);
                }
                float dtf[D];
#pragma unroll
                for (int c = 0; c < D; ++c) dtf[c] = (k == 0) ? qf[c] - pmf[k][c] : -cn_.xf[k - 1][c];
                float nt = 0.0f;
                if (k == 0) {
#pragma unroll
                    for (int c = 0; c < D; ++c) nt = __builtin_fmaf(dtf[c], dtf[c], nt);
                } else {
                    nt = cn_.df[k - 1];
                }
                float x1 = 0.0f, x2 = 0.0f;
#pragma unroll
                for (int c = 0; c < D; ++c) x1 = __builtin_fmaf(cp_.xf[k][c], cabf[c], x1);
#pragma unroll
                for (int c = 0; c < D; ++c) x2 = __builtin_fmaf(dtf[c], abf[k][c], x2);
                a1f[k] = x1;
                a2f[k] = x2;
                wrf[k] = x1 > 0.0f ? e : cp_.df[k];
                wtf[k] = x2 > 0.0f ? e : nt;
                s1[k] = x1 > 0.0f && x1 < climf;
                s2[k] = x2 > 0.0f && x2 < lim2f[k];
                slow = slow || s1[k] || s2[k];
                cn_.df[k] = e;
            }
            if (slow) {
                asm volatile("");  // one branch for the wave's rows, not selects
                float qmf[D];  // q_{j-1} rounded to FP32
#pragma unroll
                for (int c = 0; c < D; ++c) qmf[c] = (float)qa[c];
#pragma unroll
                for (int k = 0; k < RP; ++k) {
                    if (s1[k]) {
                        asm volatile("");
                        wrf[k] = seg_w_at<D>(pif[k], qmf, cabf, seg_t_slow(a1f[k], c2f));
                    }
                    if (s2[k]) {
                        asm volatile("");
                        wtf[k] = seg_w_at<D>(qf, pmf[k], abf[k], seg_t_slow(a2f[k], ab2f[k]));
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                const int t = tfirst + k;
                w[(t * 3 + 0) * 64] = dn[k];
                w[(t * 3 + 1) * 64] = (double)wrf[k];
                w[(t * 3 + 2) * 64] = (double)wtf[k];
            }
        } else {
            // Column terms of segment [q_{j-1}, q_j].
            double cab[D], cab2 = 0.0, clim;
#pragma unroll
            for (int k = 0; k < D; ++k) qb[k] = qp[k];
            if constexpr (QT) {
#pragma unroll
                for (int k = 0; k < D; ++k) cab[k] = cp[k];
                cab2 = c2p;
                clim = lp;
            } else {
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    cab[k] = qb[k] - qa[k];
                    cab2 = __builtin_fma(cab[k], cab[k], cab2);
                }
                clim = far_end_lim<D>(qa, cab, qb, cab2);
            }
            prefetch((j + 1) & (QR - 1));
            // Per row: the projections' dot products, the weights at a clamped
            // parameter (seg_t_slow's rule) and whether a weight needs computing.
            double p1[RP], p2[RP], dt[RP][D], wr[RP], wt[RP];
            bool s1[RP], s2[RP], slow = false;
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                double d = 0.0;
#pragma unroll
                for (int c = 0; c < D; ++c) {
                    xn[k][c] = pi[k][c] - qb[c];
                    d = __builtin_fma(xn[k][c], xn[k][c], d);
                }
                // The top weight's q_j - p_{i-1}: for a row after the wave's first,
                // -(p_{i-1} - q_j), the row before's difference (round to nearest
                // is symmetric; a zero's sign cannot change the projection's sum,
                // which starts at +0).
#pragma unroll
                for (int c = 0; c < D; ++c) dt[k][c] = (k == 0) ? qb[c] - pm[k][c] : -xn[k - 1][c];
                // |q_j - p_{i-1}|^2, the top weight at t = 0: for a row after the
                // wave's first, the row before's DFD value (its terms are dt's
                // negations, whose squares are the same).
                double nt = 0.0;
                if (k == 0) {
#pragma unroll
                    for (int c = 0; c < D; ++c) nt = __builtin_fma(dt[k][c], dt[k][c], nt);
                } else {
                    nt = dn[k - 1];
                }
                // Right weight: p_i to [q_{j-1}, q_j] (ap = p_i - q_{j-1}); top
                // weight: q_j to [p_{i-1}, p_i] (ap = q_j - p_{i-1}).
                double a1 = 0.0, a2 = 0.0;
#pragma unroll
                for (int c = 0; c < D; ++c) a1 = __builtin_fma(xp[k][c], cab[c], a1);
#pragma unroll
                for (int c = 0; c < D; ++c) a2 = __builtin_fma(dt[k][c], ab[k][c], a2);
                p1[k] = a1;
                p2[k] = a2;
                wr[k] = a1 > 0.0 ? d : dp[k];
                wt[k] = a2 > 0.0 ? d : nt;
                s1[k] = a1 > 0.0 && a1 < clim;
                s2[k] = a2 > 0.0 && a2 < lim2[k];
                slow = slow || s1[k] || s2[k];
                dn[k] = d;
            }
            if (slow) {
                asm volatile("");  // one branch for the wave's rows, not selects
#pragma unroll
                for (int k = 0; k < RP; ++k) {
                    if (s1[k]) {
                        asm volatile("");
                        wr[k] = seg_w_at<D>(pi[k], qa, cab, seg_t_slow(p1[k], cab2));
                    }
                    if (s2[k]) {
                        asm volatile("");
                        wt[k] = seg_w_at<D>(qb, pm[k], ab[k], seg_t_slow(p2[k], ab2[k]));
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                const int t = tfirst + k;
                w[(t * 3 + 0) * 64] = dn[k];
                w[(t * 3 + 1) * 64] = wr[k];
                w[(t * 3 + 2) * 64] = wt[k];
            }
        }
        if (edge) {
            // Boundary and padding cells overwrite their VE-FD weights.  A real
            // branch (the asm keeps the compiler from turning it into selects
            // on every step): it is taken while the band fills and at the last
            // band's closing steps.
            asm volatile("");
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                const int t = tfirst + k;
                const long long i = i0 + t;
                const bool pad = i >= rows;  // padding rows of the last band pass VE-FD values up unchanged (at its closing steps)
                if (j < 0) {
                    // Before the lane's first column: weights that leave the
                    // chain's row state at its initial values (DFD +big,
                    // VE-FD -big), so the chain needs no per-lane branch.
                    w[(t * 3 + 0) * 64] = big;
                    w[(t * 3 + 1) * 64] = -big;
                    w[(t * 3 + 2) * 64] = -big;
                } else if (j == 0 || i == 0 || pad) {
                    w[(t * 3 + 1) * 64] = pad ? -big : big;
                    w[(t * 3 + 2) * 64] = (!pad && j == 0 && i == 1) ? r2 : big;
                }
            }
        }
    };
    // `ahead`: steps past s whose ring slots must be free too (the step pair
    // of the one-column loop checks and publishes once).
    auto step_begin = [&](const int s, const int ahead) __attribute__((always_inline)) {
        if (((C * s) & 63) == 0) {
            const int c = C * s + 64 + lane;  // the chunk after the one holding column Cs
            int cn = c + 64;                  // and the chunk after that, stored 64 columns later
            cn = cn < cols ? cn : cols - 1;
            // q_{c-1}: lane l-1's q, and for lane 0 the chunk before's last
            // column, already in the ring.
            double qm[D];
            if constexpr (QT || QT32) {
#pragma unroll
                for (int k = 0; k < D; ++k) qm[k] = dpp_f64(qring[k * QR + ((C * s + 63) & (QR - 1))], qn[k], 0);
            }
            put_column(c & (QR - 1), qn, qm);
#pragma unroll
            for (int k = 0; k < D; ++k) qn[k] = Q[(long long)cn * DS + k];
        }
        if (s + ahead >= cons_seen + KS) {
            unsigned polls = 0;
            for (;;) {
                cons_seen = __builtin_amdgcn_readfirstlane(lds_get(cons));
                if (s + ahead < cons_seen + KS) break;
                if (watch_expired(a.watch, polls)) {
                    cons_seen = 0x7fff0000;  // given up (Watch): never wait for the chain again
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
    };
    auto step_end = [&](const int s) __attribute__((always_inline)) {
        // A wave's LDS operations complete in order, so a reader that sees
        // the count sees this step's weights.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) lds_put(&prod[pw], s + 1);
        slot = (slot + 1 == KS) ? 0 : slot + 1;
    };
    // Steps that apply the boundary rules: the first 64 (lanes before or at
    // their first column) and the last steps of the last band (padding rows).
    // Row 0's rule needs no weights after column 0 (the chain enters the row
    // at +big).  A padding row only has to pass
    // the terminal row's VE-FD value R up through the terminal lane at the
    // band's last step: its right weight -big there, and its own value +big
    // from the step before, which its top weight +big one step earlier gives.
    // Nothing else reads padding rows (the terminal DFD value is its row's,
    // lanes above the terminal lane feed only each other, and the last band
    // has no successor), so the last band runs its other steps without the
    // rules.  Run with them on every step, the first band set the pace of the
    // bands above it and the last band fell behind the band below: 109.5
    // against 100.5 ms at 200k x 200k.
    const int tail = band + 1 == a.nbands ? nsteps - 2 : nsteps;
    if constexpr (C == 2) {
        // The carried values alternate between the A and B registers from
        // one column to the next instead of being copied.
        for (int s = 0; s < nsteps; ++s) {
            step_begin(s, 0);
            double *w = wring + slot * (C * R * mode_weights(MODE) * 64) + lane;
            const bool edge = s < 64 || s >= tail;
            const int j = C * (s - lane);
            column(j, cA, cB, w, edge);
            column(j + 1, cB, cA, w + R * mode_weights(MODE) * 64, edge);
            step_end(s);
        }
    } else {
        // Two steps per trip, the carried values alternating between the A
        // and B registers instead of being copied.  The pair waits for ring
        // room and publishes once: the chain then sees the producers two
        // steps at a time, which needs a ring of three steps or more (a
        // chain at step c needs c + 2 steps written; producers short of
        // that are at most c + 1 < c + KS - 1, so they are not waiting).
        // Chunk stores fall on multiples of 64, the pair's first step.
        // Configurations with a deep ring (12 steps or more) and five or more
        // producers take four steps per wait and count instead: free of
        // deadlock from a ring of five (the same argument with c + 1 <
        // c + KS - 3), and worth it only where the ring still decouples the
        // roles after the coarser counts: 57.5 -> 54.8 ms at 100k x 120k (7
        // rows per lane, a ring of 13), where rings of 7 took 147 ms (13 rows
        // per lane, as a ring of 4 does with pairs) and 16% longer (4 rows).
        static_assert(KS >= 3, "step pairs need a ring of three steps");
        constexpr bool QUADS = KS >= 12 && NPW >= 5;
        // Steps with the boundary rules (above) run a copy of the loop that
        // checks them, the others one without the check.
        auto one = [&](auto edge_c, const int s, const Carry &cp_, Carry &cn_, const bool first, const bool last,
                       const int ahead) __attribute__((always_inline)) {
            if (first) step_begin(s, last ? 0 : ahead);
            double *w = wring + slot * (R * mode_weights(MODE) * 64) + lane;
            column(s - lane, cp_, cn_, w, decltype(edge_c)::value);
            if (last) {
                step_end(s);
            } else {
                slot = (slot + 1 == KS) ? 0 : slot + 1;
            }
        };
        int s = 0;
        auto steps_to = [&](auto e, const int end) __attribute__((always_inline)) {
            if constexpr (QUADS) {
                for (; s + 3 < end; s += 4) {
                    one(e, s, cA, cB, true, false, 3);
                    one(e, s + 1, cB, cA, false, false, 0);
                    one(e, s + 2, cA, cB, false, false, 0);
                    one(e, s + 3, cB, cA, false, true, 0);
                }
            }
            for (; s + 1 < end; s += 2) {
                one(e, s, cA, cB, true, false, 1);
                one(e, s + 1, cB, cA, false, true, 0);
            }
            if (s < end) {
                one(e, s, cA, cB, true, true, 0);
                ++s;
            }
        };
        // Chunk stores are checked at a trip's first step, so every trip
        // starts at a multiple of its length: 64 is one, and the last band's
        // closing steps start at a multiple of four.
        const int edge_end = nsteps < 64 ? nsteps : 64;
        int tail_start = tail < nsteps ? (tail & ~3) : nsteps;
        tail_start = tail_start > edge_end ? tail_start : edge_end;
        steps_to(std::true_type{}, edge_end);
        steps_to(std::false_type{}, tail_start);
        steps_to(std::true_type{}, nsteps);
    }
}

// One band of one pair, by every wave of the workgroup.  The caller has
// cleared the sync words and ends with a barrier before reusing the LDS.
// C = 2 (two columns per step) requires n even.  DS: the points' stride.  MODE:
// the sweep mode (kModeF64, kModeDfd, kModeV32).
template <int D, int R, int NPW, int C = 1, int DS = D, int MODE = kModeF64>
__device__ inline void bounds_band(const BoundsArgs &a, const int band, BoundsShared<R, NPW, C, MODE> &sh) {
    constexpr int KS = BoundsRing<R, NPW, C, MODE>::KS;
    // The wave index as a value the compiler cannot hoist out of the band
    // loop, so that no role's set-up is computed at kernel entry and held
    // (spilled) across the other roles' loops.
    int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    asm volatile("" : "+s"(wave));
    // Waves of a workgroup are dealt to the CU's four SIMDs in turn, so wave 4
    // shares a SIMD with wave 0: that is the courier, which mostly sleeps, and
    // the chain's SIMD keeps its issue slots for the chain.
    constexpr int cw = (NPW + 2 > 4) ? 4 : NPW + 1;
    constexpr int RP = (R + NPW - 1) / NPW;
    if (wave == 0) {
        bounds_chain<R, NPW, C, MODE>(a, band, sh);
    } else if (wave == cw) {
        bounds_courier<R, NPW, C, MODE>(a, band, sh);
    } else {
        const int pw = wave - 1 - (wave > cw ? 1 : 0);
        if (RP * NPW == R || pw < R - (RP - 1) * NPW) {
            bounds_producer<D, R, NPW, KS, C, RP, DS, MODE>(a, band, pw, sh.wring, sh.qring, sh.sprod, &sh.wfree[0]);
        } else if constexpr (RP > 1) {
            bounds_producer<D, R, NPW, KS, C, RP - 1, DS, MODE>(a, band, pw, sh.wring, sh.qring, sh.sprod,
                                                                &sh.wfree[0]);
        }
    }
}

template <int R, int NPW, int C, int MODE>
__device__ inline void bounds_clear(BoundsShared<R, NPW, C, MODE> &sh) {
    for (int w = 0; w < NPW; ++w) sh.sprod[w] = 0;
    sh.cons[0] = 0;
    sh.wfree[0] = 0;
    sh.in_avail = 0;
    sh.out_done = 0;
}

// MODE: kModeF64 and kModeV32 give DFD and VE-FD (result[0..3]), kModeDfd DFD
// alone (result[0..1]).
template <int D, int R, int NPW, int C, int MODE = kModeF64>
__global__ __launch_bounds__(64 * (NPW + 2)) void bounds_sweep(BoundsArgs a) {
    static_assert(NPW >= 1 && NPW <= R, "one to R producer waves");
    static_assert(C == 1 || C == 2, "one or two columns per step");
    static_assert(MODE == kModeF64 || MODE == kModeDfd || MODE == kModeV32, "a sweep mode");
    __shared__ BoundsShared<R, NPW, C, MODE> sh;
    for (;;) {
        if (threadIdx.x == 0) {
            sh.item = atomicAdd(a.queue, 1);
            bounds_clear(sh);
        }
        __syncthreads();
        const int band = sh.item;
        if (band >= a.nbands) break;
        bounds_band<D, R, NPW, C, D, MODE>(a, band, sh);
        __syncthreads();
    }
}

}  // namespace rfd
