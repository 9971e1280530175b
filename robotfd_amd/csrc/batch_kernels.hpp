// Batched sweeps: many pairs of uneven length in one persistent launch per
// round, with each pair's bisection state kept on the device.
//
// A round is: plan (one workgroup: lists the bands of every pair still active,
// longest pairs first, and clears their progress flags) -> sweep (persistent
// waves take (pair, band) items from the list in order and run the same band
// code as the single-pair path) -> update (one thread per pair: consumes the
// sweep's terminal value, moves the bracket, decides whether the pair stays
// active).  All of it is enqueued on one stream and captured as a hipGraph.
//
// Curves are stored 3D: a 2D pair gets z = 0, which leaves every distance,
// weight and free interval bit-identical to the 2D evaluation (each extra term
// is fma(0, 0, acc) = acc).
#pragma once

#include "batch_types.hpp"
#include "dec_types.hpp"
#include "sweep_kernels.hpp"

namespace rfd {
__global__ void batch_reset(BatchArgs a) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k == 0) {
        *a.t0 = wall_clock64();
        *a.watch.abort = 0;
    }
    if (k >= a.npairs) return;
    BatchState &st = a.state[k];
    const BatchPair &p = a.pairs[k];
    st.lb = p.endpoint_lb;
    st.ub = 0;
    st.split = 0;
    st.dfd = 0;
    st.vefd = p.endpoint_lb;
    st.F = 0;
    st.distance = 0;
    for (int q = 0; q < kSpec; ++q) st.result[q][0] = st.result[q][1] = -3.0;
    st.thr = INFINITY;
    st.iterations = 0;
    st.t_done = 0;
    st.active = 0;
    st.need_redo = 0;
    st.cert_mask = 0;
    st.certified = 0;
    if (k == 0) *a.evals = 0;
}

// Radii of the speculative tree of pair state st, depth d; returns how many
// nodes need a sweep (nodes whose interval still fails the stop rule are
// skipped, and so are their descendants).  node_ok receives a mask.
__device__ inline unsigned spec_tree(BatchState &st, int d, int stop_mode, double eps) {
    double lo[kSpec], hi[kSpec];
    unsigned ok = 0;
    lo[0] = st.lb;
    hi[0] = st.ub;
    const int nodes = (1 << d) - 1;
    for (int i = 0; i < nodes; ++i) {
        if (i > 0) {
            const int par = (i - 1) / 2;
            if (!(ok & (1u << par))) continue;
            const double sp = st.radius[par];
            if (i == 2 * par + 1) { lo[i] = lo[par]; hi[i] = sp; }   // par reachable: ub = split
            else { lo[i] = sp; hi[i] = hi[par]; }                   // not reachable: lb = split
        }
        if (!batch_keep(lo[i], hi[i], stop_mode, eps)) continue;
        st.radius[i] = (hi[i] + lo[i]) / 2.0;  // frechet.cpp:42
        ok |= 1u << i;
    }
    return ok;
}

// One workgroup of 1024 threads: exclusive scan of per-pair band counts in
// `order`, then every thread writes its pairs' items and clears their flags.
// An item is (pair * kSpec + node, band).
__global__ __launch_bounds__(1024) void batch_plan(BatchArgs a, int phase) {
    __shared__ long long part[1024];
    __shared__ int depth_s;
    const int tid = (int)threadIdx.x;
    const int per = (a.npairs + 1023) / 1024;
    const int k0 = tid * per;
    // Decision rounds on the windowed sweep keep 64-bit band words instead
    // of the progress flags.
    const bool win = (phase == PH_DEC || phase == PH_CERT) && a.dec_win > 0;
    const int R = (phase == PH_BOUNDS) ? a.bounds_rows : win ? a.dec_win / 10 : a.rows_per_lane;
    long long mine = 0;
    for (int q = 0; q < per; ++q) {
        const int idx = k0 + q;
        if (idx >= a.npairs) break;
        const int k = a.order[idx];
        const BatchPair &p = a.pairs[k];
        if (batch_in_round(p, a.state[k], phase, a.gate)) {
            const long long rows = batch_rows(p, phase);
            mine += (rows + 64LL * R - 1) / (64LL * R);
        }
    }
    part[tid] = mine;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    if (tid == 0) {
        int d = 1;
        if (phase == PH_DEC) {
            const long long bands = part[1023];
            for (int c = a.max_depth; c > 1; --c)
                if (bands * ((1 << c) - 1) <= 2LL * a.max_waves) { d = c; break; }
        }
        depth_s = d;
        a.count[2] = d;
    }
    __syncthreads();
    const int d = depth_s;
    // Second pass: per-pair node counts (d > 1 changes them), then a second scan.
    long long mine2 = 0;
    for (int q = 0; q < per; ++q) {
        const int idx = k0 + q;
        if (idx >= a.npairs) break;
        const int k = a.order[idx];
        const BatchPair &p = a.pairs[k];
        BatchState &st = a.state[k];
        if (!batch_in_round(p, st, phase, a.gate)) continue;
        const long long rows = batch_rows(p, phase);
        const long long nb = (rows + 64LL * R - 1) / (64LL * R);
        unsigned ok = 1;
        if (phase == PH_DEC) ok = spec_tree(st, d, a.stop_mode, a.eps);
        if (phase == PH_CERT) ok = st.cert_mask;
        mine2 += nb * __builtin_popcount(ok);
    }
    __syncthreads();
    part[tid] = mine2;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        long long v = tid >= off ? part[tid - off] : 0;
        __syncthreads();
        part[tid] += v;
        __syncthreads();
    }
    long long pos = part[tid] - mine2;
    for (int q = 0; q < per; ++q) {
        const int idx = k0 + q;
        if (idx >= a.npairs) break;
        const int k = a.order[idx];
        const BatchPair &p = a.pairs[k];
        BatchState &st = a.state[k];
        if (!batch_in_round(p, st, phase, a.gate)) continue;
        const long long rows = batch_rows(p, phase);
        const int nb = (int)((rows + 64LL * R - 1) / (64LL * R));
        unsigned ok = 1;
        if (phase == PH_DEC) {
            // spec_tree is deterministic in (lb, ub): recompute the mask.
            ok = spec_tree(st, d, a.stop_mode, a.eps);
        }
        if (phase == PH_CERT) ok = st.cert_mask;
        for (int node = 0; node < kSpec; ++node) {
            if (!(ok & (1u << node))) continue;
            for (int b = 0; b < nb; ++b) {
                a.items[pos + b] = make_int2(k * kSpec + node, b);
                if (win) a.flags64[p.bandoff + node * p.bandstride + b] = flag64(0u, kLoUnknown);
                else a.progress[p.bandoff + node * p.bandstride + b] = 0;
            }
            pos += nb;
            st.result[node][0] = -3.0;
            st.result[node][1] = -3.0;
            if (phase == PH_BOUNDS) {  // the fused round writes nodes 0 and 1
                st.result[1][0] = -3.0;
                st.result[1][1] = -3.0;
            }
        }
        if (phase == PH_VEFD || phase == PH_VEFD_REDO) {
            const double F = st.F;
            st.thr = (phase == PH_VEFD && a.prune) ? F * (1.0 + 0x1p-40) : INFINITY;
        }
    }
    if (tid == 1023) {
        a.count[0] = (int)part[1023];
        a.count[1] = 0;
    }
}

template <int OP, int R, bool PRUNE = false>
__global__ __launch_bounds__(64) void batch_sweep(BatchArgs a) {
    __shared__ double qring[4][3][64];
    __shared__ double sring[4][Staged<OP>::value][64];
    const int lane = (int)threadIdx.x;
    unsigned long long evals = 0, skips = 0;
    const int total = __hip_atomic_load(&a.count[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
        int idx = 0;
        if (lane == 0) idx = atomicAdd(&a.count[1], 1);
        idx = __builtin_amdgcn_readfirstlane(idx);
        if (idx >= total) break;
        const int2 it = a.items[idx];
        const int kn = __builtin_amdgcn_readfirstlane(it.x);
        const int k = kn / kSpec, node = kn % kSpec;
        const int band = __builtin_amdgcn_readfirstlane(it.y);
        const BatchPair p = a.pairs[k];
        BatchState &st = a.state[k];
        const long long rows = (OP == OP_DFD) ? p.m : p.m - 1;
        const long long cols = (OP == OP_DFD) ? p.n : p.n - 1;
        SweepJob job;
        job.P = a.pts + p.poff;
        job.Q = a.pts + p.qoff;
        job.qseg = a.qseg + p.qsoff;
        job.slots = a.slots + p.sloff + (long long)node * 2 * p.n;
        job.progress = a.progress + p.bandoff + (long long)node * p.bandstride;
        job.result = st.result[node];
        job.rows = rows;
        job.cols = cols;
        job.nbands = (int)((rows + 64LL * R - 1) / (64LL * R));
        job.param = (OP == OP_DEC) ? st.radius[node] : (OP == OP_VEFD ? p.endpoint_lb : 0.0);
        job.thr = (OP == OP_VEFD) ? st.thr : INFINITY;
        job.watch = a.watch;
        // A decision item runs the interval form of its pair (wave-uniform).
        if (OP == OP_DEC && p.dec_form == DEC_TWO_PHASE)
            run_band<OP, 3, double, SYNC_SCOPED, R, true, false>(job, band, qring, sring, evals, nullptr, skips);
        else
            run_band<OP, 3, double, SYNC_SCOPED, R, PRUNE, false>(job, band, qring, sring, evals, nullptr, skips);
    }
    if (a.evals) atomicAdd(a.evals, evals);
}

// Decision per-column terms of every pair (4*Ab, 0.5/Ab), once per batch.
__global__ void batch_qseg(BatchArgs a) {
    const int k = (int)blockIdx.y;
    if (k >= a.npairs) return;
    const BatchPair p = a.pairs[k];
    const double *Q = a.pts + p.qoff;
    double *qs = const_cast<double *>(a.qseg) + p.qsoff;
    for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < p.n - 1;
         j += (long long)gridDim.x * blockDim.x) {
        double Ab = seg_len2<double, 3>(Q + j * 3, Q + (j + 1) * 3);
        qs[2 * j] = 4.0 * Ab;
        qs[2 * j + 1] = 0.5 / Ab;
    }
}

__global__ void batch_update(BatchArgs a, int phase) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= a.npairs) return;
    const BatchPair &p = a.pairs[k];
    BatchState &st = a.state[k];
    if (!batch_in_round(p, st, phase, a.gate)) return;
    if (phase == PH_BOUNDS) {
        st.dfd = st.result[0][0];
        st.F = st.result[0][1];
        st.ub = st.dfd;
        st.lb = p.endpoint_lb;
        st.need_redo = 0;
        if (p.m < 2 || p.n < 2) {
            // One point against a curve: the DFD is the distance.
            st.lb = st.ub;
            st.distance = st.dfd;
            st.active = 0;
            st.t_done = wall_clock64();
            return;
        }
        if (batch_runs_vefd(p, a.gate)) {
            st.vefd = st.result[1][0];
            st.lb = st.vefd;
        }
    } else if (phase == PH_DFD) {
        st.dfd = st.result[0][0];
        st.F = st.result[0][1];
        st.ub = st.dfd;
        st.lb = p.endpoint_lb;
        if (p.m < 2 || p.n < 2) {
            // One point against a curve: the DFD is the distance.
            st.lb = st.ub;
            st.distance = st.dfd;
            st.active = 0;
            st.t_done = wall_clock64();
            return;
        }
        if (!batch_runs_vefd(p, a.gate)) {
            st.active = batch_keep(st.lb, st.ub, a.stop_mode, a.eps) ? 1 : 0;
            st.split = 0.5 * (st.ub + st.lb);
            st.distance = (st.ub + st.lb) / 2.0;
            if (!st.active) st.t_done = wall_clock64();
        }
        return;
    }
    if (phase == PH_BOUNDS) {
        // the bracket is set above; the common tail below applies the stop rule
    } else if (phase == PH_VEFD || phase == PH_VEFD_REDO) {
        if (phase == PH_VEFD && st.result[0][1] >= st.thr) {  // pruning reached: rerun unpruned
            st.need_redo = 1;
            return;
        }
        st.need_redo = 0;
        st.vefd = st.result[0][0];
        st.lb = st.vefd;
    } else if (phase == PH_CERT) {
        // Decisions at the predicted bracket [lo, hi]: certified when lo is
        // unreachable and hi reachable; otherwise the bracket still shrinks to
        // the side the decisions point to and bisection continues from there.
        const bool has_lo = st.cert_mask & 1u, has_hi = st.cert_mask & 2u;
        const bool rlo = has_lo && st.result[0][0] == 1.0;
        const bool rhi = !has_hi || st.result[1][0] == 1.0;
        st.iterations += (has_lo ? 1 : 0) + (has_hi ? 1 : 0);
        if (rlo) {
            st.ub = st.radius[0];
        } else if (!rhi) {
            st.lb = st.radius[1];
        } else {
            if (has_lo) st.lb = st.radius[0];
            if (has_hi) st.ub = st.radius[1];
            st.certified = 1;
        }
        st.cert_mask = 0;
    } else {
        // Walk the speculative tree with its decisions (sequential bisection order).
        const int d = a.count[2];
        int node = 0;
        for (int level = 0; level < d; ++level) {
            if (!batch_keep(st.lb, st.ub, a.stop_mode, a.eps)) break;
            const double split = (st.ub + st.lb) / 2.0;  // frechet.cpp:42 (== radius[node])
            const bool reach = st.result[node][0] == 1.0;  // .cu:1592
            if (reach) st.ub = split; else st.lb = split;
            st.iterations += 1;
            node = reach ? 2 * node + 1 : 2 * node + 2;
        }
    }
    st.active = batch_keep(st.lb, st.ub, a.stop_mode, a.eps) ? 1 : 0;
    st.split = (st.ub + st.lb) / 2.0;     // frechet.cpp:42
    st.distance = (st.ub + st.lb) / 2.0;  // frechet.cpp:63
    if (!st.active) st.t_done = wall_clock64();
}

// Features of pair k, one workgroup of 128 threads: both curves resampled to
// 64 points by index (linear interpolation), centred on the mean of the 128
// samples, scaled by 1/DFD; then VE-FD/DFD and endpoint/DFD.  FP32 row of 386
// (for training records) and BF16 row of 400 (the network's input).
__global__ __launch_bounds__(128) void batch_features(BatchArgs a, float *fout, uint16_t *xout) {
    __shared__ double red[3][128];
    const int k = (int)blockIdx.x;
    const int t = (int)threadIdx.x;
    if (k >= a.npairs) return;
    const BatchPair p = a.pairs[k];
    const BatchState &st = a.state[k];
    const double *C = a.pts + (t < 64 ? p.poff : p.qoff);
    const long long len = t < 64 ? p.m : p.n;
    const int s = t & 63;
    double x[3];
    if (len > 1) {
        const double pos = (double)s * (double)(len - 1) / 63.0;
        long long i0 = (long long)pos;
        if (i0 > len - 2) i0 = len - 2;
        const double f = pos - (double)i0;
        for (int c = 0; c < 3; ++c) x[c] = C[3 * i0 + c] + f * (C[3 * (i0 + 1) + c] - C[3 * i0 + c]);
    } else {
        for (int c = 0; c < 3; ++c) x[c] = C[c];
    }
    for (int c = 0; c < 3; ++c) red[c][t] = x[c];
    __syncthreads();
    for (int off = 64; off > 0; off >>= 1) {
        if (t < off)
            for (int c = 0; c < 3; ++c) red[c][t] += red[c][t + off];
        __syncthreads();
    }
    const double scale = st.dfd > 0 ? st.dfd : 1.0;
    float *fr = fout + (long long)k * MLP_FEAT_ROW;
    uint16_t *xr = xout + (long long)k * MLP_IN_ROW;
    for (int c = 0; c < 3; ++c) {
        const double mean = red[c][0] / 128.0;
        const float v = (float)((x[c] - mean) / scale);
        const int col = (t < 64 ? 0 : 192) + s * 3 + c;
        fr[col] = v;
        xr[col] = f2bf_dev(v);
    }
    if (t == 0) {
        const float lbf = (float)(st.vefd / scale), epf = (float)(p.endpoint_lb / scale);
        fr[384] = lbf;
        fr[385] = epf;
        xr[384] = f2bf_dev(lbf);
        xr[385] = f2bf_dev(epf);
        for (int c = MLP_FEAT_ROW; c < MLP_IN_ROW; ++c) xr[c] = 0;
    }
}

// Learned run: the bracket to certify for every pair still bisecting, from
// the predicted fraction f of [lb, ub] where the distance lies.
//  f < tau:      [lb, lb + t]  one decision at the top end; t meets the stop rule
//  f > 1 - tau:  [ub - t, ub]  one decision at the bottom end
//  otherwise:    [lb + (f - delta) w, lb + (f + delta) w] when delta > 0
// where t is half the stop tolerance at that end.  Node 0 holds the lower
// radius, node 1 the upper; a bracket end equal to lb or ub needs no decision.
__global__ void batch_learned_prepare(BatchArgs a, const float *pred, double tau, double delta) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= a.npairs) return;
    BatchState &st = a.state[k];
    st.cert_mask = 0;
    if (!st.active) return;
    const double f = (double)pred[k];
    const double w = st.ub - st.lb;
    unsigned mask = 0;
    if (f < tau) {
        const double t = a.stop_mode == 0 ? 0.5 * a.eps : 0.5 * a.eps * st.lb;
        st.radius[1] = st.lb + t;
        if (st.radius[1] > st.lb && st.radius[1] < st.ub) mask = 2u;
    } else if (f > 1.0 - tau) {
        const double t = a.stop_mode == 0 ? 0.5 * a.eps : 0.5 * a.eps * st.ub;
        st.radius[0] = st.ub - t;
        if (st.radius[0] > st.lb && st.radius[0] < st.ub) mask = 1u;
    } else if (delta > 0.0) {
        st.radius[0] = st.lb + (f - delta) * w;
        st.radius[1] = st.lb + (f + delta) * w;
        if (st.radius[0] > st.lb && st.radius[0] < st.ub) mask |= 1u;
        if (st.radius[1] > st.lb && st.radius[1] < st.ub) mask |= 2u;
        if ((mask & 3u) == 3u && !(st.radius[0] < st.radius[1])) mask = 2u;
    }
    st.cert_mask = mask;
}

}  // namespace rfd
