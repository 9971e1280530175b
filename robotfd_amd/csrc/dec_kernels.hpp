// Fused decision sweep: the free-space decision at radius r (reachable<D>,
// .cu:960-1256) on the (m-1) x (n-1) segment grid, split across the waves of a
// workgroup as bounds_sweep splits the bound sweeps.
//
//   Producers compute each cell's two free intervals, which depend on the
//   points and r only: the interval of q-segment j around vertex p_{i+1}
//   (rLB0, rUB0) and of p-segment i around vertex q_{j+1} (tLB0, tUB0), with
//   the operation order of free_iv (.cu:1173-1200), into an LDS ring (4
//   doubles per cell).
//   The chain applies the propagation rules (.cu:1202-1230) with the values
//   from the row below (lL) and the previous column (bL):
//     both negative              -> (-1, -2, -1)
//     bL >= 0                    -> right edge (rLB0, rUB0)
//     bL < 0, lL >= 0            -> right edge (max(rLB0, lL), rUB0), or
//                                   (-1, -2) when lL > rUB0
//     lL >= 0                    -> top lower end tLB0
//     lL < 0, bL >= 0            -> max(tLB0, bL), or -1 when bL > tUB0
//   evaluated without branches.  Only the terminal cell's rUB is kept, and it
//   equals rUB0 when the propagated right lower end is >= 0, else -2 (a cell
//   whose right interval is empty has rLB0 = -1 and rUB0 = -2).  Padding rows
//   of the last band pass the right lower end through unchanged, so the
//   terminal lane reads it after the sweep.
//   The courier carries the band hand-offs (one value per column).
#pragma once

#include "bounds_kernels.hpp"

namespace rfd {

// Columns of the band-boundary rings in LDS (incoming and outgoing).
constexpr int kDecRing = 128;

// The band's global-memory hand-offs, kept off the chain: the courier copies
// the predecessor's published boundary columns into the incoming ring and
// publishes the chain's outgoing columns (scoped stores, s_waitcnt, flag), as
// many columns per round as are ready, up to 64.  NK values per column (one
// for the decision); slots are [buffer][NK][cols].
template <int NK = 1, int C = 1>
__device__ inline void dec_courier(const BoundsArgs &a, const int band, double *inring,
                                      const double *outring, const BoundsSync &sy) {
    const int lane = (int)(threadIdx.x & 63);
    const long long cols = a.cols;
    const double *inD = a.slots + (long long)((band + 1) & 1) * NK * cols;
    const double *inV = inD + (NK - 1) * cols;
    double *outD = a.slots + (long long)(band & 1) * NK * cols;
    double *outV = outD + (NK - 1) * cols;
    long long in_next = band > 0 ? 0 : cols;
    long long out_next = band + 1 < a.nbands ? 0 : cols;
    while (in_next < cols || out_next < cols) {
        bool moved = false;
        if (in_next < cols) {
            // Ring slots free: the chain has consumed the columns of the
            // steps below `cons` (C per step).
            const long long room =
                (long long)C * __builtin_amdgcn_readfirstlane(lds_get(sy.cons)) + kDecRing - in_next;
            if (room > 0) {
                const long long f = __builtin_amdgcn_readfirstlane(flag_load<SYNC_SCOPED>(&a.progress[band - 1]));
                long long n = f - in_next;
                n = n < room ? n : room;
                n = n < 64 ? n : 64;
                if (n > 0) {
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane < n) {
                        const long long c = in_next + lane;
                        const double vD = column_load<SYNC_SCOPED>(inD + c);
                        inring[c % kDecRing] = vD;
                        if (NK > 1) {
                            const double vV = column_load<SYNC_SCOPED>(inV + c);
                            inring[kDecRing + c % kDecRing] = vV;
                        }
                    }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): the ring writes are done
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    if (lane == 0) lds_put(sy.in_avail, (int)(in_next + n));
                    in_next += n;
                    moved = true;
                }
            }
        }
        if (out_next < cols) {
            const long long oa = __builtin_amdgcn_readfirstlane(lds_get(sy.out_avail));
            long long n = oa - out_next;
            n = n < 64 ? n : 64;
            if (n > 0) {
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (lane < n) {
                    const long long c = out_next + lane;
                    column_store<SYNC_SCOPED>(outD + c, outring[c % kDecRing]);
                    if (NK > 1) column_store<SYNC_SCOPED>(outV + c, outring[kDecRing + c % kDecRing]);
                }
                __builtin_amdgcn_wave_barrier();
                if (lane == 0) flag_store<SYNC_SCOPED>(&a.progress[band], (int)(out_next + n));
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                if (lane == 0) lds_put(sy.out_done, (int)(out_next + n));
                out_next += n;
                moved = true;
            }
        }
        if (!moved) __builtin_amdgcn_s_sleep(2);
    }
}

// Ring depth in steps, 2048 R bytes each (as BoundsRing).
template <int R> struct DecRing {
    static constexpr int budget = (R <= 8) ? 40960 : 122880;
    static constexpr int KS = (budget / (2048 * R) < 16) ? budget / (2048 * R) : 16;
    static_assert(KS >= 2, "ring needs two slots");
};

template <int R>
__device__ inline void dec_cells(const double *w, double (&bot)[R], double &dl) {
    constexpr int G = (R <= 8) ? R : (R + 1) / 2;
#pragma unroll
    for (int g0 = 0; g0 < R; g0 += G) {
        double wv[G][4];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (g0 + g < R) wv[g][c] = w[((g0 + g) * 4 + c) * 64];
#pragma unroll
        for (int g = 0; g < G; ++g)
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (g0 + g < R) asm volatile("" ::"v"(wv[g][c]));
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int t = g0 + g;
            if (t >= R) break;
            const double rLB0 = wv[g][0], rUB0 = wv[g][1], tLB0 = wv[g][2], tUB0 = wv[g][3];
            const double lL = dl, bL = bot[t];
            const bool lp = lL >= 0.0, bp = bL >= 0.0;
            const bool rok = bp || (lp && lL <= rUB0);
            const bool tok = lp || (bp && bL <= tUB0);
            const double rl = bp ? rLB0 : dmax(rLB0, lL);
            const double tl = lp ? tLB0 : dmax(tLB0, bL);
            dl = rok ? rl : -1.0;
            bot[t] = tok ? tl : -1.0;
        }
    }
}

template <int D, int R, int KS>
__device__ inline void dec_chain(const BoundsArgs &a, const int band, const double *wring, const double *inring,
                                 double *outring, const BoundsSync &sy, const int npw) {
    const int lane = (int)(threadIdx.x & 63);
    const long long rows = a.rows, cols = a.cols;  // segments: m - 1, n - 1
    const long long i0 = (long long)band * 64 * R + (long long)R * lane;
    double bot[R];
#pragma unroll
    for (int t = 0; t < R; ++t) bot[t] = (i0 + t == 0) ? 0.0 : -1.0;  // .cu:1026: row 0 enters reachable
    double dout = -1.0;
    const bool has_pred = band > 0;
    const bool has_succ = band + 1 < a.nbands;
    const long long nsteps = cols + 63;
    const long long tl = rows - 1 - i0;
    long long avail = 0, in_known = 0, done_known = 0;
    int slot = 0;
    __builtin_amdgcn_s_setprio(3);
    for (long long s = 0; s < nsteps; ++s) {
        const long long j = s - lane;
        if (has_pred && s < cols && s >= in_known) {
            for (;;) {
                in_known = __builtin_amdgcn_readfirstlane(lds_get(sy.in_avail));
                if (s < in_known) break;
                __builtin_amdgcn_s_sleep(1);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        if (s >= avail) {
            for (;;) {
                int mn = lds_get(&sy.prod[0]);
                for (int w = 1; w < npw; ++w) {
                    const int v = lds_get(&sy.prod[w]);
                    mn = v < mn ? v : mn;
                }
                avail = __builtin_amdgcn_readfirstlane(mn);
                if (s < avail) break;
                __builtin_amdgcn_s_sleep(1);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        const double *w = wring + (long long)slot * (R * 4 * 64) + lane;
        double dl = shift_up(dout);
        const double bD = has_pred ? inring[(int)((s < cols ? s : cols - 1) % kDecRing)] : -1.0;
        if (lane == 0) dl = bD;
        if (j >= 0 && j < cols) {
            dec_cells<R>(w, bot, dl);
            dout = dl;
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) lds_put(sy.cons, (int)(s + 1));
        slot = (slot + 1 == KS) ? 0 : slot + 1;
        if (has_succ && s >= 63) {
            const long long col = s - 63;
            if (col >= done_known + kDecRing) {
                for (;;) {
                    done_known = __builtin_amdgcn_readfirstlane(lds_get(sy.out_done));
                    if (col < done_known + kDecRing) break;
                    __builtin_amdgcn_s_sleep(1);
                }
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
            }
            if (lane == 63) {
                outring[(int)(col % kDecRing)] = dout;
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                lds_put(sy.out_avail, (int)(col + 1));
            }
        }
    }
    __builtin_amdgcn_s_setprio(0);
    if (tl >= 0 && tl < R) {
        // rUB0 of the terminal cell (vertex p_{m-1}, segment q_{n-2} q_{n-1}),
        // formed as the producers form it.
        const double *pv = a.P + (rows)*D;
        const double *qa = a.Q + (cols - 1) * D;
        const double *qb = a.Q + cols * D;
        const double Ab = seg_len2<double, D>(qa, qb);
        double lb, ub;
        free_iv<double, D>(pv, qa, qb, a.r2, 4.0 * Ab, 0.5 / Ab, lb, ub);
        a.result[0] = (dout >= 0.0) ? ub : -2.0;
    }
}

template <int D, int R, int NPW, int KS>
__device__ inline void dec_producer(const BoundsArgs &a, const int band, const int pw, double *wring,
                                    int *prod, int *cons) {
    constexpr int RP = (R + NPW - 1) / NPW;
    const int nmine = (R - pw + NPW - 1) / NPW;
    const int lane = (int)(threadIdx.x & 63);
    const long long rows = a.rows, cols = a.cols;  // segments
    const double big = (double)FLT_MAX;
    const double *P = a.P, *Q = a.Q;
    const double r2 = a.r2;
    const long long i0 = (long long)band * 64 * R + (long long)R * lane;
    // Row terms: p_i, p_{i+1}, 4|p_{i+1} - p_i|^2 and 0.5 / |p_{i+1} - p_i|^2.
    double pi[RP][D], pn[RP][D], A4a[RP], inv2a[RP];
#pragma unroll
    for (int k = 0; k < RP; ++k) {
        const long long i = i0 + pw + (long long)k * NPW;
        const long long ic = i < rows ? i : rows - 1;
#pragma unroll
        for (int c = 0; c < D; ++c) {
            pi[k][c] = P[ic * D + c];
            pn[k][c] = P[(ic + 1) * D + c];
        }
        const double Aa = seg_len2<double, D>(pi[k], pn[k]);
        A4a[k] = 4.0 * Aa;
        inv2a[k] = 0.5 / Aa;
    }
    // Q queue (as bounds_producer), offset by one point: at step s lane 0
    // takes q_{s+1}, and a lane's previous point is q_j.
    double qq[D], qn[D], qa[D], qb[D];
    {
        const long long npts = cols + 1;
        long long c0 = 1 + lane;
        c0 = c0 < npts ? c0 : npts - 1;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            qq[k] = Q[c0 * D + k];
            qn[k] = qq[k];
            qb[k] = Q[k];  // q_0: lane 0's point before step 0
        }
    }
    const long long nsteps = cols + 63;
    long long cons_seen = 0;
    int slot = 0;
    for (long long s = 0; s < nsteps; ++s) {
        if ((s & 63) == 0) {
            if (s > 0) {
#pragma unroll
                for (int k = 0; k < D; ++k) qq[k] = qn[k];
            }
            long long c = s + 65 + lane;
            c = c < cols + 1 ? c : cols;
#pragma unroll
            for (int k = 0; k < D; ++k) qn[k] = Q[c * D + k];
        } else {
#pragma unroll
            for (int k = 0; k < D; ++k) qq[k] = dpp_f64(qq[k], qq[k], 1);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            qa[k] = qb[k];
            qb[k] = dpp_f64(qq[k], qb[k], 0);  // q_{j+1}, j = s - l
        }
        // Column terms of segment [q_j, q_{j+1}] (qseg_terms, .cu:1129-1144).
        const double Ab = seg_len2<double, D>(qa, qb);
        const double A4b = 4.0 * Ab, inv2b = 0.5 / Ab;
        if (s >= cons_seen + KS) {
            for (;;) {
                cons_seen = __builtin_amdgcn_readfirstlane(lds_get(cons));
                if (s < cons_seen + KS) break;
                __builtin_amdgcn_s_sleep(1);
            }
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
        }
        double *w = wring + (long long)slot * (R * 4 * 64) + lane;
#pragma unroll
        for (int k = 0; k < RP; ++k) {
            if (RP * NPW != R && k == RP - 1 && k >= nmine) break;
            const int t = pw + k * NPW;
            double rl, ru, tlb, tub;
            free_iv<double, D>(pn[k], qa, qb, r2, A4b, inv2b, rl, ru);
            free_iv<double, D>(qb, pi[k], pn[k], r2, A4a[k], inv2a[k], tlb, tub);
            w[(t * 4 + 0) * 64] = rl;
            w[(t * 4 + 1) * 64] = ru;
            w[(t * 4 + 2) * 64] = tlb;
            w[(t * 4 + 3) * 64] = tub;
        }
        if (band + 1 == a.nbands) {
            // Padding rows of the last band pass the right lower end through.
            asm volatile("");
#pragma unroll
            for (int k = 0; k < RP; ++k) {
                if (RP * NPW != R && k == RP - 1 && k >= nmine) break;
                const int t = pw + k * NPW;
                if (i0 + t >= rows) {
                    w[(t * 4 + 0) * 64] = -1.0;
                    w[(t * 4 + 1) * 64] = big;
                    w[(t * 4 + 2) * 64] = -1.0;
                    w[(t * 4 + 3) * 64] = -2.0;
                }
            }
        }
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this step's intervals are in LDS
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        if (lane == 0) lds_put(&prod[pw], (int)(s + 1));
        slot = (slot + 1 == KS) ? 0 : slot + 1;
    }
}

template <int R, int NPW> struct DecShared {
    static constexpr int KS = DecRing<R>::KS;
    double wring[KS * R * 4 * 64];
    double inring[kDecRing];
    double outring[kDecRing];
    int sprod[NPW];
    int swords[4];  // cons, in_avail, out_avail, out_done
    int item;
};

template <int R, int NPW>
__device__ inline void dec_clear(DecShared<R, NPW> &sh) {
    for (int w = 0; w < NPW; ++w) sh.sprod[w] = 0;
    for (int k = 0; k < 4; ++k) sh.swords[k] = 0;
}

// One band of one decision, by every wave of the workgroup (roles and SIMD
// placement as bounds_band).
template <int D, int R, int NPW>
__device__ inline void dec_band(const BoundsArgs &a, const int band, DecShared<R, NPW> &sh) {
    constexpr int KS = DecRing<R>::KS;
    const int wave = (int)(threadIdx.x >> 6);
    const BoundsSync sy{sh.sprod, &sh.swords[0], &sh.swords[1], &sh.swords[2], &sh.swords[3]};
    constexpr int cw = (NPW + 2 > 4) ? 4 : NPW + 1;
    if (wave == 0)
        dec_chain<D, R, KS>(a, band, sh.wring, sh.inring, sh.outring, sy, NPW);
    else if (wave == cw)
        dec_courier<1>(a, band, sh.inring, sh.outring, sy);
    else
        dec_producer<D, R, NPW, KS>(a, band, wave - 1 - (wave > cw ? 1 : 0), sh.wring, sh.sprod, sy.cons);
}

// a.rows, a.cols: segments (m - 1, n - 1); a.r2: radius squared; a.slots:
// 2 * cols values; a.result[0]: the terminal rightUB (1.0: reachable).
template <int D, int R, int NPW>
__global__ __launch_bounds__(64 * (NPW + 2)) void dec_sweep(BoundsArgs a) {
    static_assert(NPW >= 1 && NPW <= R, "one to R producer waves");
    __shared__ DecShared<R, NPW> sh;
    for (;;) {
        if (threadIdx.x == 0) {
            sh.item = atomicAdd(a.queue, 1);
            dec_clear(sh);
        }
        __syncthreads();
        const int band = sh.item;
        if (band >= a.nbands) break;
        dec_band<D, R, NPW>(a, band, sh);
        __syncthreads();
    }
}

}  // namespace rfd
