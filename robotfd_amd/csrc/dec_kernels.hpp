// Windowed decision sweep: the free-space decision at radius r (reachable<D>,
// .cu:960-1256) on the (m-1) x (n-1) segment grid of one pair, with the free
// intervals computed off the serial propagation and every band confined to the
// columns where it can hold a reachable cell.
//
// Why.  In band_sweep one wave computes a cell's two free intervals (two
// square roots) and applies the propagation rules, so a step of a band of 64 R
// rows costs about 0.6-0.7 us, and a sweep's critical path is about (columns +
// 80 x bands) such steps, however few cells are reachable: at 200k x 200k near
// the critical radius 19% of the cells are reachable, yet the sweep takes about
// 200 ms.  Here a workgroup owns a band: producer waves compute the intervals
// (they depend on the points and r only) into an LDS ring, one chain wave
// applies the rules (a few selects per cell), and a courier wave carries the
// hand-offs between bands, so a step costs about what the chain needs.
//
// Windows.  Band b's cells at columns below the first reachable cell of band
// b-1's top row are unreachable (they are fed from the left by rows that start
// unreachable and from below by unreachable cells), and once every row of the
// band is unreachable and so are all later inputs from below, so is every
// later cell.  So a band starts at the first column its predecessor can make
// reachable and stops as soon as it and its inputs are dead; a band that never
// holds a reachable cell makes every band above it dead, and the terminal
// cell's rightUB is -2 (what the full sweep gives an unreachable cell).
//
// Band hand-off.  Band b publishes one 64-bit word: the low half counts the
// columns of its top row that are final (published, or dead without being
// stored), the high half is `lo`, the first column whose value it stored and
// may be reachable (kLoUnknown until it stores one).  Columns below
// min(lo, count) are unreachable and are never read.  A band's last word has
// count = cols + 1 + h, h being one past its last reachable output column:
// the columns from h on are unreachable.  Values are stored with device-scope atomics and the word after
// an s_waitcnt, as in band_sweep.
//
// Layout in a band (as band_sweep): lane l of the chain owns rows
// 64 R b + R l + t, t < R, and at step s evaluates column s - l.  Producer w
// computes every R rows of every lane at steps s = s0 + w (mod NP) into ring
// slot s mod KS: four values per cell, the right edge's free interval (vertex
// p_{i+1} on segment q_j q_{j+1}) and the top edge's (vertex q_{j+1} on p_i
// p_{i+1}), in the operation order of free_iv (.cu:1173-1200).
#pragma once

#include "bounds_kernels.hpp"
#include "dec_types.hpp"

namespace rfd {

// LDS of one workgroup.  A ring slot holds [row t][rl, ru, tl, tu][lane].
template <int D, typename T, int R, int NP> struct DecShared {
    static constexpr int SLOT = R * 4 * 64;  // values per ring slot
    static constexpr int KS0 = 65536 / (SLOT * (int)sizeof(T));
    static constexpr int KS = KS0 < 16 ? KS0 : 16;  // at most 64 KB of ring
    static constexpr int KB = 128;           // boundary ring positions
    static constexpr int QC = D + 2;         // Q-ring entries per column: q, A4b, inv2b
    int prod[16];   // first step each producer has not written
    int done[64];   // steps the chain has finished (absolute; one copy per lane, read [0])
    int start;      // first step of the band; -1: the band is dead; -2: not known yet
    int stop;       // step at which the chain stopped (kDecNone: it has not)
    int in_avail;   // incoming columns in the ring (absolute)
    int in_hi;      // incoming columns from in_hi on are unreachable (kDecNone: not known)
    int out_done;   // outgoing columns published
    int item;
    alignas(16) T inring[KB];
    alignas(16) T outring[KB];
    alignas(16) T odummy[64];
    alignas(16) T qring[QC * 256];
    alignas(16) T wring[KS * SLOT];
    static_assert(KS >= 4, "ring needs four slots");
};

// Steps of band `band`: lane l evaluates column s - l at step s; the band
// holding row rows - 1 stops where that row's lane reaches column cols - 1.
template <int R>
__device__ inline int dec_steps(const DecArgs &a, const int band, int *tlane_out) {
    const int cols = (int)a.cols;
    const long long toff = a.rows - 1 - (long long)band * 64 * R;
    const bool has_term = toff >= 0 && toff < 64LL * R;
    const int tlane = has_term ? (int)(toff / R) : -1;
    if (tlane_out) *tlane_out = tlane;
    return has_term ? cols + tlane : cols + 63;
}

// v, or (keep false) v with the high word of 4.0: a value in [4, 8), above
// every interval end.  One select of one word.
__device__ inline double dead_unless(const bool keep, const double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned hi = keep ? (unsigned)(b >> 32) : 0x40100000u;
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | (b & 0xffffffffull)));
}

template <typename SH>
__device__ inline int dec_wait_start(const DecArgs &a, SH &sh) {
    unsigned polls = 0;
    for (;;) {
        const int s0 = __builtin_amdgcn_readfirstlane(lds_get(&sh.start));
        if (s0 != -2) return s0;
        if (watch_expired(a.watch, polls)) return -1;
        __builtin_amdgcn_s_sleep(2);
    }
}

template <int D, typename T, int R, int NP, bool VAR>
__device__ inline void dec_chain(const DecArgs &a, const int band, DecShared<D, T, R, NP> &sh,
                                 unsigned long long &evals, unsigned long long &skips) {
    using N = Num<T>;
    using SH = DecShared<D, T, R, NP>;
    constexpr int KS = SH::KS, KB = SH::KB, SLOT = SH::SLOT;
    const int lane = (int)(threadIdx.x & 63);
    const long long rows = a.rows;
    const int cols = (int)a.cols;
    const long long i0 = (long long)band * 64 * R + (long long)R * lane;
    int tlane = -1;
    const int nsteps = dec_steps<R>(a, band, &tlane);
    const bool has_succ = band + 1 < a.nbands;
    const int s0 = dec_wait_start(a, sh);
    if (s0 < 0) {
        // Dead band: every cell is unreachable, the terminal one too.
        if (tlane >= 0 && lane == 0) a.result[0] = -2.0;
        skips += (unsigned long long)nsteps / 16;
        return;
    }
    lds_put(&sh.done[lane], s0);
    if (a.tlog && lane == 0) a.tlog[12 * band + 1] = wall_clock64();
    __builtin_amdgcn_s_setprio(3);

    int avail = 0, in_known = 0, out_known = 0;
    auto read_prod = [&]() __attribute__((always_inline)) {
        asm volatile("" ::: "memory");
        int mn = sh.prod[0];
#pragma unroll
        for (int k = 1; k < NP; ++k) mn = sh.prod[k] < mn ? sh.prod[k] : mn;
        return __builtin_amdgcn_readfirstlane(mn);
    };
    // Step s (about to load step s + 1's intervals and input): the producers
    // must have written min(s + 2, nsteps) steps, the courier must have filled
    // min(s + 2, cols) incoming columns, and the output position s mod KB
    // must be free (column s - 63 - KB published, where it is one).
    bool gave_up = false;
    // With a timestamp log: wall-clock ticks spent waiting while the
    // producers were short, and while the input from below was.
    unsigned long long w_prod = 0, w_in = 0, w_out = 0;
    auto wait_for = [&](const int s) __attribute__((always_inline)) {
        const int np = (s + 2 < nsteps) ? s + 2 : nsteps;
        const int ni = (s + 2 < cols) ? s + 2 : cols;
        const int no = s - 62 - KB;
        unsigned polls = 0;
        unsigned long long t0 = 0;
        for (;;) {
            avail = read_prod();
            in_known = __builtin_amdgcn_readfirstlane(lds_get(&sh.in_avail));
            out_known = __builtin_amdgcn_readfirstlane(lds_get(&sh.out_done));
            const bool ok = avail >= np && in_known >= ni && (!has_succ || out_known >= no);
            if (a.tlog) {
                const unsigned long long t = wall_clock64();
                const unsigned long long dt = polls > 0 ? t - t0 : 0;
                w_prod += (avail < np) ? dt : 0;
                w_in += (avail >= np && in_known < ni) ? dt : 0;
                w_out += (avail >= np && in_known >= ni) ? dt : 0;
                t0 = t;
            }
            if (ok) break;
            if (gave_up || watch_expired(a.watch, polls)) {
                gave_up = true;
                avail = nsteps;
                in_known = cols;
                out_known = kDecNone;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
    };
    auto limit = [&]() __attribute__((always_inline)) {
        const int lp = (avail >= nsteps) ? kDecNone : avail - 1;
        const int li = (in_known >= cols) ? kDecNone : in_known - 1;
        const int lo = has_succ ? out_known + KB + 62 : kDecNone;
        const int m = lp < li ? lp : li;
        return __builtin_amdgcn_readfirstlane(m < lo ? m : lo);
    };

    // Unreachable values.  Fused form (FAST): any value above 1 (the interval
    // ends lie in [0, 1]), so that a row's rules take max, one compare, a
    // select of one word and min, with no branch: the values are never NaN
    // there, and only comparisons and the terminal interval end reach the
    // result, so the encoding changes no bit of it.  Two-phase form: -1, with
    // the NaN-aware rules of band_sweep.  The band hand-off columns hold -1 in
    // both.
    constexpr bool FAST = !VAR;
    const T INF = __builtin_inf();
    auto live_v = [&](const T v) __attribute__((always_inline)) { return FAST ? (v <= T(1)) : !(v < T(0)); };
    T bot[R];
#pragma unroll
    for (int t = 0; t < R; ++t) bot[t] = (i0 + t == 0) ? T(0) : (FAST ? INF : T(-1));
    T out = FAST ? INF : T(-1);
    T term = T(-2);
    bool real[R];  // rows of the grid (the others pad the last band)
#pragma unroll
    for (int t = 0; t < R; ++t) real[t] = i0 + t < rows;
    // 2^100 and -2^100 in registers (hidden from the compiler, which would
    // otherwise rebuild the addend of every FMA from literals).
    T big = T(0x1p100), nbig = T(-0x1p100);
    asm volatile("" : "+v"(big), "+v"(nbig));

    // Step s0's intervals and input from below.
    T c0[R][4], c1[R][4];
    T b0 = T(-1), b1 = T(-1);
    {
        // Step s0 needs s0 + 1 steps written and column s0 in the ring.
        wait_for(s0 - 1);
        const T *w = sh.wring + (s0 % KS) * SLOT + lane;
#pragma unroll
        for (int t = 0; t < R; ++t)
#pragma unroll
            for (int c = 0; c < 4; ++c) c0[t][c] = w[(t * 4 + c) * 64];
        b0 = sh.inring[s0 % KB];
    }
    int lim = limit();
    int slot = s0 % KS, pos = s0 % KB;
    int s = s0;
    int stop_at = kDecNone;
    // Lane 63 stores its output to the outgoing ring, the others to slots of
    // their own, so the store needs no lane mask.
    T *const obase = (lane == 63) ? sh.outring : sh.odummy + lane;
    const int ostep = (lane == 63) ? 1 : 0;

    // Step s from the intervals in iv and the input from below bval (lane 0).
    auto step = [&](const T (&iv)[R][4], const T bval) __attribute__((always_inline)) {
        T x = shift_up(out);
        if constexpr (FAST) {
            if (lane == 0) x = bval;
            // Per row, with x the value from below and y the row's own from
            // the column before (.cu:1231-1244): reachable from the left (y),
            // the whole right interval [rl, ru]; else from below where
            // x <= ru, from max(rl, x) on.  The top edge likewise with the
            // roles swapped.  fma(v, BIG, -BIG) is <= 0 for a reachable v and
            // huge for an unreachable one, so max(rl, it) is rl or unreachable;
            // dead_unless keeps a value or makes it unreachable ([4, 8)).
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const T rl = iv[t][0], ru = iv[t][1], tl = iv[t][2], tu = iv[t][3];
                const T y = bot[t];
                const T Cr = dmax(rl, __builtin_fma(y, big, nbig));
                const T Ct = dmax(tl, __builtin_fma(x, big, nbig));
                const T g1 = dead_unless(x <= ru, dmax(rl, x));
                const T g2 = dead_unless(y <= tu, dmax(tl, y));
                x = dmin(Cr, g1);
                bot[t] = dmin(Ct, g2);
            }
            out = x;
        } else {
            T left = x;
            if (lane == 0) left = bval;
            const long long j = (long long)s - lane;
            const bool inr = j >= 0 && j < cols && i0 < rows;
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const long long it = i0 + t;
                const T leftLB = left, botLB = bot[t];
                // The propagation rules (.cu:1231-1244) as selects, as band_sweep.
                const bool bdead = botLB < T(0), ldead = leftLB < T(0);
                const bool rinv = (leftLB > iv[t][1]) | ldead;
                const bool tinv = (botLB > iv[t][3]) | bdead;
                const T rmax = N::max(iv[t][0], leftLB);
                const T tmax = N::max(iv[t][2], botLB);
                const T rLB = bdead ? (rinv ? T(-1) : rmax) : iv[t][0];
                const T rUB = (bdead & rinv) ? T(-2) : iv[t][1];
                const T tLB = ldead ? (tinv ? T(-1) : tmax) : iv[t][2];
                bot[t] = (inr && real[t]) ? tLB : bot[t];
                left = real[t] ? rLB : T(-1);
                term = (inr && it == rows - 1 && j == cols - 1) ? rUB : term;
            }
            out = inr ? left : out;
        }
    };

    // One step from intervals ca and input ba, prefetching the next step's
    // into cb and bb; false once the band is done.  The step first stores the
    // step before's output and count: the LDS operations then run stores,
    // next step's loads, and the compute waits only for loads a step old.
    // Two copies alternate the buffers, so no registers are copied between
    // steps.
    int ppos = (pos + KB - 1) % KB;  // the step before's output position (at s0: one nobody reads)
    auto one = [&](const T (&ca)[R][4], const T ba, T (&cb)[R][4], T &bb) __attribute__((always_inline)) {
        obase[ppos * ostep] = out;
        lds_put(&sh.done[lane], s);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (s >= nsteps) return false;
        // Every 64 steps: stop when every row and output of the band is
        // unreachable and so is every later input from below.
        if ((s & 63) == 0 && s != s0) {
            bool live = live_v(out);
#pragma unroll
            for (int t = 0; t < R; ++t) live = live || live_v(bot[t]);
            const int hi = __builtin_amdgcn_readfirstlane(lds_get(&sh.in_hi));
            if (__ballot(live) == 0 && s >= hi) {
                stop_at = s;
                return false;
            }
        }
        if (s >= lim) {
            wait_for(s);
            lim = limit();
        }
        const int slotn = (slot + 1 == KS) ? 0 : slot + 1;
        const int posn = (pos + 1 == KB) ? 0 : pos + 1;
        {
            const T *w = sh.wring + slotn * SLOT + lane;
#pragma unroll
            for (int t = 0; t < R; ++t)
#pragma unroll
                for (int c = 0; c < 4; ++c) cb[t][c] = w[(t * 4 + c) * 64];
            bb = sh.inring[posn];
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        step(ca, ba);
        ++s;
        ppos = pos;
        slot = slotn;
        pos = posn;
        return true;
    };
    for (;;) {
        if (!one(c0, b0, c1, b1)) break;
        if (!one(c1, b1, c0, b0)) break;
    }
    __builtin_amdgcn_s_setprio(0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (stop_at != kDecNone) {
        lds_put(&sh.stop, stop_at);
        skips += (unsigned long long)(nsteps - stop_at) / 16;
        term = T(-2);  // the terminal cell, if this band holds it, is unreachable
    } else if (FAST && tlane >= 0 && lane == tlane) {
        // The terminal cell (row rows - 1, column cols - 1) was the lane's
        // last step; the padding rows above it pass its right lower end on
        // unchanged.  Reachable: its right interval's upper end, else -2.
        const int tl = (int)(rows - 1 - i0);
        const T ru = sh.wring[((nsteps - 1) % KS) * SLOT + (tl * 4 + 1) * 64 + lane];
        term = (out <= T(1)) ? ru : T(-2);
    }
    {
        // Cells whose intervals were computed and propagated: steps s0 .. s - 1
        // of the lane's rows inside the grid.
        const long long lo = (s0 > lane) ? s0 : lane;
        const long long hi = (s < (long long)cols + lane) ? s : (long long)cols + lane;
        int nreal = 0;
#pragma unroll
        for (int t = 0; t < R; ++t) nreal += real[t] ? 1 : 0;
        evals += (unsigned long long)((hi > lo) ? (hi - lo) * nreal : 0);
    }
    skips += (unsigned long long)s0 / 16;
    if (tlane >= 0 && lane == tlane) a.result[0] = (double)term;
    if (a.tlog && lane == 0) {
        a.tlog[12 * band + 5] = (unsigned long long)s;
        a.tlog[12 * band + 6] = w_prod;
        a.tlog[12 * band + 7] = w_in;
        a.tlog[12 * band + 8] = w_out;
    }
}

// The band's global-memory hand-offs: decides where the band starts, fills
// the incoming ring from the predecessor's column (unreachable values below
// its lo and from its stop on, without loading them), and publishes the
// chain's outputs with the band's own lo and, if the chain stops early, its
// stop column.
template <int D, typename T, int R, int NP, bool VAR>
__device__ inline void dec_courier(const DecArgs &a, const int band, DecShared<D, T, R, NP> &sh) {
    using SH = DecShared<D, T, R, NP>;
    constexpr int KB = SH::KB;
    const int lane = (int)(threadIdx.x & 63);
    const int cols = (int)a.cols;
    const T *sin = (const T *)a.slots + (long long)((band + 1) & 1) * cols;
    T *sout = (T *)a.slots + (long long)(band & 1) * cols;
    const bool has_succ = band + 1 < a.nbands;
    unsigned long long *myflag = a.flags + band;
    unsigned polls = 0;
    // The rings hold the chain's encoding of an unreachable value (+inf in
    // the fused form, -1 in the two-phase form); the band-boundary columns
    // hold -1 in both.
    const T dead = VAR ? T(-1) : (T)__builtin_inf();

    // The predecessor's state: lo, final count, stop column.
    int lo_in = 0, hi_in = 0, pub_in = cols;
    bool fin_in = true;
    int s0 = 0;
    if (band > 0) {
        for (;;) {
            // One 64-bit load: the count and lo of the same store.
            const unsigned long long f = flag64_load(a.flags + band - 1);
            const unsigned cnt = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(f & 0xffffffffull));
            const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(f >> 32));
            if (lo != kLoUnknown || cnt >= (unsigned)cols) {
                lo_in = (int)lo;
                fin_in = cnt >= (unsigned)cols;
                hi_in = fin_in ? (cnt > (unsigned)cols ? (int)(cnt - (unsigned)cols - 1) : cols) : cols;
                pub_in = fin_in ? cols : (int)cnt;
                break;
            }
            if (watch_expired(a.watch, polls)) {
                lo_in = (int)kLoUnknown;
                fin_in = true;
                hi_in = 0;
                pub_in = cols;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (lo_in == (int)kLoUnknown) {
            // The predecessor never held a reachable cell: neither does this band.
            lds_put(&sh.start, -1);
            if (lane == 0) flag64_store(myflag, flag64((unsigned)cols + 1u, kLoUnknown));
            return;
        }
        s0 = lo_in & ~63;
    } else {
        // Row -1 does not exist: every input from below is unreachable.
        lo_in = 0;
        hi_in = 0;
        fin_in = true;
        pub_in = cols;
    }
    if (a.tlog && lane == 0) {
        a.tlog[12 * band] = wall_clock64();
        a.tlog[12 * band + 4] = (unsigned long long)s0;
    }
    if (fin_in) lds_put(&sh.in_hi, hi_in);
    lds_put(&sh.in_avail, s0);
    lds_put(&sh.out_done, s0);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    lds_put(&sh.start, s0);
    if (has_succ && s0 > 0 && lane == 0) flag64_store(myflag, flag64((unsigned)s0, kLoUnknown));

    unsigned lo_out = kLoUnknown;
    int last_live = -1;            // last reachable output column published
    unsigned n_pub = 0, n_in = 0;  // publishes and input batches (timestamp log)
    int in_next = s0;
    int out_next = has_succ ? s0 : cols;
    polls = 0;
    for (;;) {
        const int done = __builtin_amdgcn_readfirstlane(lds_get(&sh.done[0]));
        const int stop = __builtin_amdgcn_readfirstlane(lds_get(&sh.stop));
        bool moved = false;
        if (stop == kDecNone && in_next < cols) {
            const int room = done + KB - in_next;
            if (room > 0) {
                if (!fin_in && in_next >= pub_in) {
                    const unsigned long long f = flag64_load(a.flags + band - 1);
                    const unsigned cnt = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(f & 0xffffffffull));
                    if (cnt >= (unsigned)cols) {
                        fin_in = true;
                        hi_in = cnt > (unsigned)cols ? (int)(cnt - (unsigned)cols - 1) : cols;
                        pub_in = cols;
                    } else {
                        pub_in = (int)cnt;
                    }
                }
                int n = pub_in - in_next;
                n = n < room ? n : room;
                n = n < 64 ? n : 64;
                if (n > 0) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane < n) {
                        const int c = in_next + lane;
                        T v = dead;
                        if (c >= lo_in && c < hi_in) {
                            v = column_load<SYNC_SCOPED>(sin + c);
                            if (!VAR && v < T(0)) v = dead;
                        }
                        sh.inring[c % KB] = v;
                    }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the ring writes are done
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (fin_in) lds_put(&sh.in_hi, hi_in);
                    lds_put(&sh.in_avail, in_next + n);
                    in_next += n;
                    moved = true;
                    ++n_in;
                }
            }
        }
        if (out_next < cols) {
            // Lane 63 stored column c at step c + 63; a chain that stopped at
            // step h stored columns below h - 63 and the rest are unreachable.
            int ready = done - 63;
            const int hi_out = (stop != kDecNone) ? ((stop - 63 > s0) ? stop - 63 : s0) : cols;
            ready = ready < hi_out ? ready : hi_out;
            int n = ready - out_next;
            n = n < 64 ? n : 64;
            // The final word says where the band's reachable outputs end: one
            // past the last reachable column published, so that the band
            // above can stop there, not where this band's chain happened to
            // notice it was dead.
            if (n > 0) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                T v = T(-1);
                if (lane < n) {
                    const int c = out_next + lane;
                    v = sh.outring[(c + 63) % KB];
                    if (!VAR && v > T(1)) v = T(-1);
                    column_store<SYNC_SCOPED>(sout + c, v);
                }
                const unsigned long long liv = __ballot(lane < n && !(v < T(0)));
                if (liv != 0) {
                    if (lo_out == kLoUnknown) lo_out = (unsigned)(out_next + (int)__builtin_ctzll(liv));
                    last_live = out_next + 63 - (int)__builtin_clzll(liv);
                }
                __builtin_amdgcn_wave_barrier();
                const int pub = out_next + n;
                const bool last = pub >= hi_out;
                const unsigned cnt = last ? (unsigned)cols + 1u + (unsigned)(last_live + 1) : (unsigned)pub;
                if (lane == 0) flag64_store(myflag, flag64(cnt, lo_out));
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                lds_put(&sh.out_done, pub);
                out_next = last ? cols : pub;
                moved = true;
                ++n_pub;
            } else if (stop != kDecNone && out_next >= hi_out) {
                if (lane == 0) flag64_store(myflag, flag64((unsigned)cols + 1u + (unsigned)(last_live + 1), lo_out));
                out_next = cols;
                moved = true;
            }
        }
        // Done once the chain needs no more input (it stopped, or every
        // column is in the ring) and every output column is final.
        if ((stop != kDecNone || in_next >= cols) && out_next >= cols) {
            if (a.tlog && lane == 0) {
                a.tlog[12 * band + 9] = n_pub;
                a.tlog[12 * band + 10] = n_in;
            }
            break;
        }
        if (moved) {
            polls = 0;
        } else {
            if (watch_expired(a.watch, polls)) break;
            __builtin_amdgcn_s_sleep(2);
        }
    }
}

// Producer w: every R rows of every lane at steps s0 + w, s0 + w + NP, ...
template <int D, typename T, int R, int NP, bool VAR>
__device__ inline void dec_producer(const DecArgs &a, const int band, const int w, DecShared<D, T, R, NP> &sh) {
    using SH = DecShared<D, T, R, NP>;
    constexpr int KS = SH::KS, SLOT = SH::SLOT, QC = SH::QC;
    const int lane = (int)(threadIdx.x & 63);
    const long long rows = a.rows;
    const int cols = (int)a.cols;
    const T *P = (const T *)a.P, *Q = (const T *)a.Q, *qseg = (const T *)a.qseg;
    const long long i0 = (long long)band * 64 * R + (long long)R * lane;
    const T r2 = (T)a.r * (T)a.r;
    // The lane's rows: points p_i .. p_{i+R} (clamped to the last point) and
    // the segment terms of each row.
    T p[R + 1][D];
#pragma unroll
    for (int t = 0; t <= R; ++t) {
        long long pi = i0 + t;
        pi = pi < rows ? pi : rows;
#pragma unroll
        for (int k = 0; k < D; ++k) p[t][k] = P[pi * D + k];
    }
    T A4a[R], inv2a[R];
#pragma unroll
    for (int t = 0; t < R; ++t) {
        const T Aa = seg_len2<T, D>(p[t], p[t + 1]);
        A4a[t] = T(4) * Aa;
        inv2a[t] = T(0.5) / Aa;
    }
    const int s0 = dec_wait_start(a, sh);
    if (s0 < 0) return;
    const int nsteps = dec_steps<R>(a, band, nullptr);
    if (lane == 0) lds_put(&sh.prod[w], s0 + w);
    const T INF = __builtin_inf();
    const bool pad_band = (long long)(band + 1) * 64 * R > rows;  // the last band, with padding rows

    // Q ring: 256 columns, column c at c mod 256, QC entries (q, A4b, inv2b).
    // At steps 64k .. 64k + 63 lanes read columns 64k - 63 .. 64k + 64
    // (chunks k - 1 .. k + 1).  Entering block k a producer stores chunk k + 1
    // from registers and loads chunk k + 2; every producer stores every chunk
    // itself (the same values), and the producers stay within KS + NP < 64
    // steps of each other, so no wave still reads the chunk a store replaces.
    const int npts = cols + 1;
    T pre[QC];
    auto load_chunk = [&](const int k, T (&v)[QC]) __attribute__((always_inline)) {
        int c = 64 * k + lane;
        const int cq = c < 0 ? 0 : (c < npts ? c : npts - 1);
        const int cs = c < 0 ? 0 : (c < cols ? c : cols - 1);
#pragma unroll
        for (int kk = 0; kk < D; ++kk) v[kk] = Q[(long long)cq * D + kk];
        v[D] = qseg[2LL * cs];
        v[D + 1] = qseg[2LL * cs + 1];
    };
    auto put_chunk = [&](const int k, const T (&v)[QC]) __attribute__((always_inline)) {
        const int idx = (64 * k + lane) & 255;
#pragma unroll
        for (int kk = 0; kk < QC; ++kk) sh.qring[kk * 256 + idx] = v[kk];
    };
    int blk = s0 >> 6;
    {
        T v[QC];
        load_chunk(blk - 1, v);
        put_chunk(blk - 1, v);
        load_chunk(blk, v);
        put_chunk(blk, v);
        load_chunk(blk + 1, v);
        put_chunk(blk + 1, v);
        load_chunk(blk + 2, pre);
    }
    int done_seen = s0;
    for (int s = s0 + w; s < nsteps; s += NP) {
        while ((s >> 6) > blk) {
            ++blk;
            put_chunk(blk + 1, pre);
            load_chunk(blk + 2, pre);
        }
        if (s >= done_seen + KS) {
            unsigned polls = 0;
            for (;;) {
                done_seen = __builtin_amdgcn_readfirstlane(lds_get(&sh.done[0]));
                if (s < done_seen + KS) break;
                if (__builtin_amdgcn_readfirstlane(lds_get(&sh.stop)) != kDecNone) return;
                if (watch_expired(a.watch, polls)) {
                    done_seen = 0x3fff0000;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        // Columns outside the grid read ring entries of other columns; the
        // chain never uses those intervals.
        const int jc = s - lane;
        const int ia = jc & 255, ib = (jc + 1) & 255;
        T qa[D], qb[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            qa[k] = sh.qring[k * 256 + ia];
            qb[k] = sh.qring[k * 256 + ib];
        }
        const T A4b = sh.qring[D * 256 + ia];
        const T inv2b = sh.qring[(D + 1) * 256 + ia];
        T Bv[2 * R], Dv[2 * R];
        typename RootT<T>::type Sv[2 * R];
#pragma unroll
        for (int t = 0; t < R; ++t) {
            iv_disc<T, D, VAR>(p[t + 1], qa, qb, r2, A4b, Bv[2 * t], Dv[2 * t]);
            iv_disc<T, D, VAR>(qb, p[t], p[t + 1], r2, A4a[t], Bv[2 * t + 1], Dv[2 * t + 1]);
        }
#pragma unroll
        for (int i = 0; i < 2 * R; ++i) Sv[i] = iv_root(Dv[i]);
        T *wp = sh.wring + (s % KS) * SLOT + lane;
#pragma unroll
        for (int t = 0; t < R; ++t) {
            T rl, ru, tl, tu;
            // The chain's fused-form encoding: an empty interval's lower end
            // is +inf.
            const T dead_lb = VAR ? T(-1) : INF;
            iv_ends<T, VAR>(Bv[2 * t], Dv[2 * t], Sv[2 * t], inv2b, rl, ru, dead_lb);
            iv_ends<T, VAR>(Bv[2 * t + 1], Dv[2 * t + 1], Sv[2 * t + 1], inv2a[t], tl, tu, dead_lb);
            if constexpr (!VAR) {
                // A padding row of the last band passes the value from below
                // on unchanged and stays unreachable itself.
                if (pad_band && i0 + t >= rows) {
                    rl = T(0);
                    ru = T(2);
                    tl = INF;
                    tu = T(-2);
                }
            }
            wp[(t * 4 + 0) * 64] = rl;
            wp[(t * 4 + 1) * 64] = ru;
            wp[(t * 4 + 2) * 64] = tl;
            wp[(t * 4 + 3) * 64] = tu;
        }
        // A wave's LDS operations complete in order: a reader that sees the
        // count sees the intervals.
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) lds_put(&sh.prod[w], s + NP);
    }
}

// One band by every wave of the workgroup: wave 0 the chain, the courier in
// the fifth slot (the chain's SIMD: waves are dealt to the CU's SIMDs in turn,
// and the courier mostly sleeps) when there is one, the others producers.
// The caller has cleared the hand-off words and ends with a barrier.
template <int D, typename T, int R, int NP, bool VAR>
__device__ inline void dec_roles(const DecArgs &a, const int band, DecShared<D, T, R, NP> &sh,
                                 unsigned long long &evals, unsigned long long &skips) {
    const int wave = (int)(threadIdx.x >> 6);
    constexpr int cw = (NP + 2 > 4) ? 4 : NP + 1;
    if (wave == 0) {
        dec_chain<D, T, R, NP, VAR>(a, band, sh, evals, skips);
    } else if (wave == cw) {
        dec_courier<D, T, R, NP, VAR>(a, band, sh);
    } else {
        const int w = wave - 1 - (wave > cw ? 1 : 0);
        dec_producer<D, T, R, NP, VAR>(a, band, w, sh);
    }
}

template <int D, typename T, int R, int NP>
__device__ inline void dec_clear(DecShared<D, T, R, NP> &sh) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < 16; ++k) sh.prod[k] = 0;
        sh.start = -2;
        sh.stop = kDecNone;
        sh.in_avail = 0;
        sh.in_hi = kDecNone;
        sh.out_done = 0;
    }
    if (threadIdx.x < 64) sh.done[threadIdx.x] = 0;
}

template <int D, typename T, int R, int NP, bool VAR>
__global__ __launch_bounds__(64 * (NP + 2)) void dec_sweep(DecArgs a) {
    __shared__ DecShared<D, T, R, NP> sh;
    const int wave = (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    unsigned long long evals = 0, skips = 0;
    for (;;) {
        if (threadIdx.x == 0) sh.item = atomicAdd(a.queue, 1);
        dec_clear(sh);
        __syncthreads();
        const int band = sh.item;
        if (band >= a.nbands) break;
        dec_roles<D, T, R, NP, VAR>(a, band, sh, evals, skips);
        if (a.tlog && threadIdx.x == 0) {
            a.tlog[12 * band + 2] = wall_clock64();
            a.tlog[12 * band + 3] = ((unsigned long long)__builtin_amdgcn_s_getreg((20 << 0) | (0 << 6) | (31 << 11)) << 32) |
                                   (unsigned long long)__builtin_amdgcn_s_getreg((4 << 0) | (0 << 6) | (31 << 11));
        }
        __syncthreads();
    }
    if (wave == 0 && a.evals) {
        atomicAdd(a.evals, evals);
        if (lane == 0) atomicAdd(a.evals + 1, skips);
    }
}

}  // namespace rfd
