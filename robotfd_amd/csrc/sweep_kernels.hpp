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
// row 64b-1's outputs from a column buffer that band b-1 fills in chunks of 64
// columns; band b-1 publishes the count of finished columns with an agent-scope
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

// Value of lane l-1 (lane 0 receives its own value; the caller overrides it).
__device__ inline double shift_up(double v) { return __shfl_up(v, 1, 64); }
__device__ inline float shift_up(float v) { return __shfl_up(v, 1, 64); }

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
    T t = apab / ab2;
    t = N::min(T(1), N::max(T(0), t));
    T w = T(0);
#pragma unroll
    for (int k = 0; k < D; ++k) {
        T q = N::fma(t, ab[k], a[k]);
        T d = x[k] - q;
        w = N::fma(d, d, w);
    }
    return w;
}

// Free interval of segment [s0, s1] around vertex c (.cu:1173-1200).
template <typename T, int D>
__device__ inline void free_iv(const T *c, const T *s0, const T *s1, T r2, T A4, T inv2A,
                               T &lb, T &ub) {
    using N = Num<T>;
    T B = T(0), C = -r2;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        T di = s1[k] - s0[k];
        T mi = s0[k] - c[k];
        B = N::fma(mi, di, B);
        C = N::fma(mi, mi, C);
    }
    B = B * T(2);
    T disc = N::fma(B, B, -(A4 * C));
    lb = T(-1);
    ub = T(-2);
    if (disc >= T(0)) {
        T s = N::sqrt(disc);
        T ta = (-B - s) * inv2A;
        T tb = (-B + s) * inv2A;
        bool invalid = tb < T(0) || ta > T(1);
        lb = invalid ? T(-1) : N::max(ta, T(0));
        ub = invalid ? T(-2) : N::min(tb, T(1));
    }
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
};

// Q points (and, for the decision, the per-column terms) reach the lanes
// through a ring of four 64-point chunks in LDS, structure-of-arrays so that
// consecutive lanes read consecutive addresses.  Chunk k+2 is loaded into
// registers at step 64k and written to LDS at step 64(k+1), so global-memory
// latency stays off the per-step path.  At step s lane l reads point s-l+1
// (s-l for DFD): chunks k-1, k, k+1 are live, the fourth slot is free.
template <int OP, int D, typename T>
__global__ __launch_bounds__(64) void band_sweep(SweepArgs a) {
    using N = Num<T>;
    constexpr int NS = (OP == OP_DEC) ? 2 : 0;  // per-column terms staged
    __shared__ T qring[4][D][64];
    __shared__ T sring[4][NS > 0 ? NS : 1][64];
    const int lane = (int)threadIdx.x;
    const T *P = (const T *)a.P;
    const T *Q = (const T *)a.Q;
    const T *qseg = (const T *)a.qseg;
    T *slots = (T *)a.slots;
    const long long rows = a.rows, cols = a.cols;
    // Points the sweep reads: DFD reads q_j (j < cols = n), the others q_j, q_{j+1}.
    const long long npts = (OP == OP_DFD) ? cols : cols + 1;
    const T big = N::big();
    const T par = (T)a.param;
    const T r2 = par * par;

    for (;;) {
        int band = 0;
        if (lane == 0) band = atomicAdd(a.queue, 1);
        band = __builtin_amdgcn_readfirstlane(band);
        if (band >= a.nbands) break;

        const long long i = (long long)band * 64 + lane;
        const bool row_ok = i < rows;
        const long long ii = row_ok ? i : rows - 1;
        T pa[D], pb[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            pa[k] = P[ii * D + k];
            pb[k] = (OP == OP_DFD) ? T(0) : P[(ii + 1) * D + k];
        }
        T A4a = T(0), inv2a = T(0);
        if (OP == OP_DEC) {
            T Aa = seg_len2<T, D>(pa, pb);
            A4a = T(4) * Aa;
            inv2a = T(0.5) / Aa;
        }
        T bot;
        if (OP == OP_DFD) bot = big;
        else if (OP == OP_VEFD) bot = (i == 0) ? par * par : big;
        else bot = (i == 0) ? T(0) : T(-1);
        T out = (OP == OP_DEC) ? T(-1) : big;

        // Q staging: chunk 0 straight into LDS, chunk 1 into registers.
        T pre[D], pres[NS > 0 ? NS : 1];
        {
            const long long p0 = lane;
#pragma unroll
            for (int k = 0; k < D; ++k) qring[0][k][lane] = (p0 < npts) ? Q[p0 * D + k] : T(0);
            if (NS > 0) {
                sring[0][0][lane] = (p0 < cols) ? qseg[2 * p0] : T(0);
                sring[0][NS - 1][lane] = (p0 < cols) ? qseg[2 * p0 + 1] : T(0);
            }
            const long long p1 = 64 + lane;
#pragma unroll
            for (int k = 0; k < D; ++k) pre[k] = (p1 < npts) ? Q[p1 * D + k] : T(0);
            if (NS > 0) {
                pres[0] = (p1 < cols) ? qseg[2 * p1] : T(0);
                pres[NS - 1] = (p1 < cols) ? qseg[2 * p1 + 1] : T(0);
            }
        }
        // q_j of the lane's current column (VE-FD / decision); every lane starts at j = 0.
        T qa[D];
#pragma unroll
        for (int k = 0; k < D; ++k) qa[k] = Q[k];

        const bool has_pred = band > 0;
        const bool has_succ = band + 1 < a.nbands;
        const T *sin = slots + (long long)((band + 1) & 1) * cols;
        T *sout = slots + (long long)(band & 1) * cols;
        T bin = T(0), bout = T(0);
        const long long nsteps = cols + 63;

        for (long long s = 0; s < nsteps; ++s) {
            const long long j = s - lane;
            if ((s & 63) == 0) {
                // Stage chunk k+1 (in registers) and fetch chunk k+2.
                const long long kc = s >> 6;
                const int slot = (int)((kc + 1) & 3);
#pragma unroll
                for (int k = 0; k < D; ++k) qring[slot][k][lane] = pre[k];
                if (NS > 0) {
                    sring[slot][0][lane] = pres[0];
                    sring[slot][NS - 1][lane] = pres[NS - 1];
                }
                const long long pn = (kc + 2) * 64 + lane;
#pragma unroll
                for (int k = 0; k < D; ++k) pre[k] = (pn < npts) ? Q[pn * D + k] : T(0);
                if (NS > 0) {
                    pres[0] = (pn < cols) ? qseg[2 * pn] : T(0);
                    pres[NS - 1] = (pn < cols) ? qseg[2 * pn + 1] : T(0);
                }
                if (has_pred && s < cols) {
                    const long long need = (s + 64 < cols) ? s + 64 : cols;
                    while (__hip_atomic_load(&a.progress[band - 1], __ATOMIC_ACQUIRE,
                                             __HIP_MEMORY_SCOPE_AGENT) < (int)need)
                        __builtin_amdgcn_s_sleep(1);
                    const long long c = s + lane;
                    bin = (c < cols) ? sin[c] : T(0);
                }
            }
            T bval;
            if (has_pred) bval = lane_read(bin, (int)(s & 63));
            else if (OP == OP_DFD) bval = (s == 0) ? T(0) : big;
            else if (OP == OP_VEFD) bval = big;
            else bval = T(-1);
            T left = shift_up(out);
            if (lane == 0) left = bval;

            if (row_ok && j >= 0 && j < cols) {
                const bool last = (i == rows - 1) && (j == cols - 1);
                // LDS index of the point this lane reads this step.
                const long long pt = (OP == OP_DFD) ? j : j + 1;
                const int rs = (int)((pt >> 6) & 3), ro = (int)(pt & 63);
                T qb[D];
#pragma unroll
                for (int k = 0; k < D; ++k) qb[k] = qring[rs][k][ro];
                if (OP == OP_DFD) {
                    T d = dist2<T, D>(pa, qb);
                    T in = N::min(left, bot);
                    T top = N::max(d, in);
                    out = N::min(top, bot);
                    bot = top;
                    if (last) *a.result = (double)N::sqrt(top);
                } else if (OP == OP_VEFD) {
                    T in = N::min(left, bot);
                    T wr = point_seg2<T, D>(pb, qa, qb);
                    T wt = point_seg2<T, D>(qb, pa, pb);
                    T Tv = N::max(wt, in);
                    T Rv = N::max(wr, in);
                    bot = Tv;
                    out = Rv;
                    if (last) *a.result = (double)N::sqrt(N::min(Tv, Rv));
                } else {
                    const T leftLB = left, botLB = bot;
                    T rLB = T(-1), rUB = T(-2), tLB = T(-1);
                    if (leftLB >= T(0) || botLB >= T(0)) {
                        const int ss = (int)((j >> 6) & 3), so = (int)(j & 63);
                        const T A4b = sring[ss][0][so], inv2b = sring[ss][NS - 1][so];
                        T tUB;
                        free_iv<T, D>(pb, qa, qb, r2, A4b, inv2b, rLB, rUB);
                        free_iv<T, D>(qb, pa, pb, r2, A4a, inv2a, tLB, tUB);
                        if (botLB < T(0)) {
                            bool inv = leftLB > rUB || leftLB < T(0);
                            T t2 = N::max(rLB, leftLB);
                            rLB = inv ? T(-1) : t2;
                            rUB = inv ? T(-2) : rUB;
                        }
                        if (leftLB < T(0)) {
                            bool inv = botLB > tUB || botLB < T(0);
                            T t2 = N::max(tLB, botLB);
                            tLB = inv ? T(-1) : t2;
                        }
                    }
                    bot = tLB;
                    out = rLB;
                    if (last) *a.result = (double)rUB;
                }
                if (OP != OP_DFD) {
#pragma unroll
                    for (int k = 0; k < D; ++k) qa[k] = qb[k];
                }
            }

            if (has_succ && s >= 63) {
                const long long col = s - 63;
                const T v = lane_read(out, 63);
                if (lane == (int)(col & 63)) bout = v;
                if ((col & 63) == 63 || col == cols - 1) {
                    const long long base = col & ~63LL;
                    if (base + lane <= col) sout[base + lane] = bout;
                    __syncthreads();
                    if (lane == 0)
                        __hip_atomic_store(&a.progress[band], (int)(col + 1), __ATOMIC_RELEASE,
                                           __HIP_MEMORY_SCOPE_AGENT);
                }
            }
        }
    }
}

__global__ void sweep_init(int *progress, int nbands, int *queue, double *result, double sentinel) {
    const int tid = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int stride = (int)(gridDim.x * blockDim.x);
    for (int k = tid; k < nbands; k += stride) progress[k] = 0;
    if (tid == 0) {
        *queue = 0;
        *result = sentinel;
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

__global__ void to_f32(const double *src, float *dst, long long count) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < count) dst[k] = (float)src[k];
}

}  // namespace rfd
