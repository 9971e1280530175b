// CPU restatement of RobotFD's continuous Frechet distance path.
//
// TEST INFRASTRUCTURE ONLY: the checker for the GPU library.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// Built with -ffp-contract=off: every fused multiply-add below is written as an
// explicit std::fma, and every other product/sum is rounded on its own, so the
// GPU kernels (built the same way) can reproduce each value bit for bit.
//
// Canonical arithmetic.  The reference kernels are plain CUDA C compiled by
// nvcc with its default --fmad=true, which fuses "acc += a*b" and
// "t*x + y" into FMA.  This restatement makes those fusions explicit:
//   * every "acc += a*b" accumulation (.cu:382-384, 392, 1180-1181, 1210-1211,
//     619) is acc = fma(a, b, acc);
//   * the projection "t*(b-a) + a" (.cu:390, 412) is fma(t, b - a, a);
//   * the discriminant "B*B - 4.0*A*C" (.cu:1185, 1215) is
//     fma(B, B, -((4.0*A)*C)).
// Parity of these choices with nvcc's own contraction cannot be checked here
// (no CUDA toolchain); the GPU library follows this file exactly.
#include "frechet_oracle.h"

#include <cfloat>
#include <cmath>
#include <cstdlib>
#include <vector>
#include <algorithm>
#include <atomic>
#include <thread>

namespace {

template <typename T> struct Lim;
template <> struct Lim<double> { static double big() { return (double)FLT_MAX; } };
template <> struct Lim<float> { static float big() { return FLT_MAX; } };

template <typename T>
inline T fmaT(T a, T b, T c) { return std::fma(a, b, c); }

template <typename T>
inline T sqrtT(T a) { return std::sqrt(a); }

// Squared Euclidean distance as the DFD kernel forms it (.cu:615-619).
template <typename T>
inline T dist2(const T *a, const T *b, int D) {
    T acc = T(0);
    for (int k = 0; k < D; ++k) {
        T t = a[k] - b[k];
        acc = fmaT(t, t, acc);
    }
    return acc;
}

// Squared distance from point x to segment [a, b] (.cu:375-394 right weight;
// the top weight uses the same steps with the roles of the curves swapped,
// README.md:79).
template <typename T>
inline T point_seg2(const T *x, const T *a, const T *b, int D) {
    T ab2 = T(0), apab = T(0);
    T ab[3];
    for (int k = 0; k < D; ++k) {
        ab[k] = b[k] - a[k];
        T ap = x[k] - a[k];
        ab2 = fmaT(ab[k], ab[k], ab2);
        apab = fmaT(ap, ab[k], apab);
    }
    T t = apab / ab2;
    t = std::fmin(T(1), std::fmax(T(0), t));
    T w = T(0);
    for (int k = 0; k < D; ++k) {
        T q = fmaT(t, ab[k], a[k]);
        T d = x[k] - q;
        w = fmaT(d, d, w);
    }
    return w;
}

// The interval ends (-B -/+ sqrt(disc)) / (2A), with B, disc and inv2A in T.
// In FP64 every operation is FP64.  In FP32 (the precision form's early
// decisions) the square root and the two ends are computed in FP64 from the
// FP32 B, disc and inv2A, and the ends rounded to FP32 once.
inline void iv_ends(double B, double disc, double inv2A, double &ta, double &tb) {
    const double s = std::sqrt(disc);
    ta = (-B - s) * inv2A;
    tb = (-B + s) * inv2A;
}
inline void iv_ends(float B, float disc, float inv2A, float &ta, float &tb) {
    const double s = std::sqrt((double)disc);
    ta = (float)((-(double)B - s) * (double)inv2A);
    tb = (float)((-(double)B + s) * (double)inv2A);
}

// Free interval of an edge: the part of segment [s0, s1] within sqrt(r2) of
// vertex c (.cu:1173-1230).  A4 = 4*|s1-s0|^2 and inv2A = 0.5/|s1-s0|^2 are
// per-segment terms.  mi = s0 - c, di = s1 - s0 (.cu:1178-1179, 1208-1209).
template <typename T>
inline void free_iv(const T *c, const T *s0, const T *s1, int D, T r2, T A4, T inv2A,
                    T &lb, T &ub) {
    T B = T(0), C = -r2;
    for (int k = 0; k < D; ++k) {
        T di = s1[k] - s0[k];
        T mi = s0[k] - c[k];
        B = fmaT(mi, di, B);
        C = fmaT(mi, mi, C);
    }
    B = B * T(2);
    T disc = fmaT(B, B, -(A4 * C));
    lb = T(-1);
    ub = T(-2);
    if (disc >= T(0)) {
        T ta, tb;
        iv_ends(B, disc, inv2A, ta, tb);
        bool invalid = tb < T(0) || ta > T(1);
        lb = invalid ? T(-1) : std::fmax(ta, T(0));
        ub = invalid ? T(-2) : std::fmin(tb, T(1));
    }
}

// The same interval as interval_calculation<D> forms it (.cu:722-782), the
// decision's two-phase path for inputs whose interval arrays fit in device
// memory (.cu:1450-1558).  It differs from free_iv in two places: the C term is
// (sum of mi^2) - r2, the sum started at 0 (.cu:733-742, 765-770), where
// reachable<D> starts it at -r2; and the clamps are comparisons (.cu:751-752,
// 778-779), so a NaN end (0 * inf at a zero-length segment) stays NaN, where
// reachable<D>'s max / min take it to 0 / 1.
template <typename T>
inline void free_iv_2ph(const T *c, const T *s0, const T *s1, int D, T r2, T A4, T inv2A,
                        T &lb, T &ub) {
    T B = T(0), mm = T(0);
    for (int k = 0; k < D; ++k) {
        T di = s1[k] - s0[k];
        T mi = s0[k] - c[k];
        B = fmaT(mi, di, B);
        mm = fmaT(mi, mi, mm);
    }
    B = B * T(2);
    const T C = mm - r2;
    T disc = fmaT(B, B, -(A4 * C));
    lb = T(-1);
    ub = T(-2);
    if (disc >= T(0)) {
        T ta, tb;
        iv_ends(B, disc, inv2A, ta, tb);
        if (!(tb < T(0) || ta > T(1))) {
            if (ta < T(0)) ta = T(0);
            if (tb > T(1)) tb = T(1);
            lb = ta;
            ub = tb;
        }
    }
}

// Squared length of a segment as the decision kernel forms it
// (.cu:1001-1005 for P, .cu:1138-1141 for Q).
template <typename T>
inline T seg_len2(const T *a, const T *b, int D) {
    T acc = T(0);
    for (int k = 0; k < D; ++k) {
        T d = b[k] - a[k];
        acc = fmaT(d, d, acc);
    }
    return acc;
}

template <typename T>
std::vector<T> to_t(const double *X, int64_t cnt) {
    std::vector<T> v((size_t)cnt);
    for (int64_t i = 0; i < cnt; ++i) v[(size_t)i] = (T)X[i];
    return v;
}

// DFD<D> (.cu:458-641) row by row.  colv[j] carries right = min(top, bot) of
// row i-1 into row i as "left" (.cu:624-626); (0,0) enters with left = 0,
// row 0 with left = FLT_MAX, column 0 with bot = FLT_MAX (.cu:620-622).
template <typename T>
T dfd_impl(const T *P, int64_t m, const T *Q, int64_t n, int D) {
    const T big = Lim<T>::big();
    std::vector<T> colv((size_t)n, big);
    T top = T(0);
    for (int64_t i = 0; i < m; ++i) {
        T bot = big;
        const T *pi = P + i * D;
        for (int64_t j = 0; j < n; ++j) {
            T left = (i == 0) ? ((j == 0) ? T(0) : big) : colv[(size_t)j];
            T d = dist2(pi, Q + j * D, D);
            T in = std::fmin(left, bot);
            top = std::fmax(d, in);
            colv[(size_t)j] = std::fmin(top, bot);
            bot = top;
        }
    }
    return sqrtT(top);
}

// VEFD<D> (.cu:200-432).  Cell (i,j) pairs segment p_i p_{i+1} with q_j q_{j+1}.
// T = max(w_top, min(bot, left)), R = max(w_right, ...) (.cu:373, 418-419);
// row 0 starts with bot = init^2 (.cu:258-259), other rows with FLT_MAX;
// result sqrt(min(T, R)) of cell (m-2, n-2) (.cu:426-429).
template <typename T>
T vefd_rows(const T *P, int64_t /*m*/, const T *Q, int64_t n, int D, T init,
            int64_t r0, int64_t r1, std::vector<T> &colv, T *lastTR) {
    const T big = Lim<T>::big();
    const int64_t cols = n - 1;
    T Tv = T(0), Rv = T(0);
    for (int64_t i = r0; i < r1; ++i) {
        T bot = (i == 0) ? init * init : big;
        const T *pa = P + i * D, *pb = P + (i + 1) * D;
        for (int64_t j = 0; j < cols; ++j) {
            T left = (i == 0) ? big : colv[(size_t)j];
            const T *qa = Q + j * D, *qb = Q + (j + 1) * D;
            T in = std::fmin(left, bot);
            T wr = point_seg2(pb, qa, qb, D);
            T wt = point_seg2(qb, pa, pb, D);
            Tv = std::fmax(wt, in);
            Rv = std::fmax(wr, in);
            bot = Tv;
            colv[(size_t)j] = Rv;
        }
    }
    if (lastTR) { lastTR[0] = Tv; lastTR[1] = Rv; }
    return std::fmin(Tv, Rv);
}

template <typename T>
T vefd_impl(const T *P, int64_t m, const T *Q, int64_t n, int D, T init) {
    std::vector<T> colv((size_t)(n - 1), Lim<T>::big());
    T v = vefd_rows(P, m, Q, n, D, init, 0, m - 1, colv, (T *)nullptr);
    return sqrtT(v);
}

// reachable<D> (.cu:960-1256) row by row.  Only lower ends travel between
// cells: rightLB to the next row (as leftLB), topLB along the row (as botLB)
// (.cu:1231-1249).  Row 0 starts with botLB = 0, others with -1 (.cu:1026-1027);
// the left boundary is -1 (.cu:656).  Returns rightUB of cell (m-2, n-2).
// TWO: free intervals of the two-phase path (free_iv_2ph); the reachability
// rules are the same in both paths (.cu:932-947 and .cu:1231-1244).
template <typename T, bool TWO = false>
T decide_rows(const T *P, int64_t m, const T *Q, int64_t n, int D, T r,
              int64_t r0, int64_t r1, std::vector<T> &colv,
              const std::vector<T> &A4b, const std::vector<T> &inv2b) {
    (void)m;
    const int64_t cols = n - 1;
    const T r2 = r * r;
    T rightUB = T(-2);
    for (int64_t i = r0; i < r1; ++i) {
        const T *pa = P + i * D, *pb = P + (i + 1) * D;
        T Aa = seg_len2(pa, pb, D);
        T A4a = T(4) * Aa, inv2a = T(0.5) / Aa;
        T botLB = (i == 0) ? T(0) : T(-1);
        for (int64_t j = 0; j < cols; ++j) {
            T leftLB = (i == 0) ? T(-1) : colv[(size_t)j];
            const T *qa = Q + j * D, *qb = Q + (j + 1) * D;
            T rLB, rUB, tLB, tUB;
            if (TWO) {
                free_iv_2ph(pb, qa, qb, D, r2, A4b[(size_t)j], inv2b[(size_t)j], rLB, rUB);
                free_iv_2ph(qb, pa, pb, D, r2, A4a, inv2a, tLB, tUB);
            } else {
                free_iv(pb, qa, qb, D, r2, A4b[(size_t)j], inv2b[(size_t)j], rLB, rUB);
                free_iv(qb, pa, pb, D, r2, A4a, inv2a, tLB, tUB);
            }
            if (botLB < T(0)) {
                bool inv = leftLB > rUB || leftLB < T(0);
                T t2 = std::fmax(rLB, leftLB);
                rLB = inv ? T(-1) : t2;
                rUB = inv ? T(-2) : rUB;
            }
            if (leftLB < T(0)) {
                bool inv = botLB > tUB || botLB < T(0);
                T t2 = std::fmax(tLB, botLB);
                tLB = inv ? T(-1) : t2;
                tUB = inv ? T(-2) : tUB;
            }
            botLB = tLB;
            colv[(size_t)j] = rLB;
            rightUB = rUB;
        }
    }
    return rightUB;
}

template <typename T>
void q_seg_terms(const T *Q, int64_t n, int D, std::vector<T> &A4b, std::vector<T> &inv2b) {
    A4b.resize((size_t)(n - 1));
    inv2b.resize((size_t)(n - 1));
    for (int64_t j = 0; j + 1 < n; ++j) {
        // .cu:1138-1141 forms q_j - q_{j+1}; its square equals that of
        // q_{j+1} - q_j bit for bit.
        T Ab = seg_len2(Q + j * D, Q + (j + 1) * D, D);
        A4b[(size_t)j] = T(4) * Ab;
        inv2b[(size_t)j] = T(0.5) / Ab;
    }
}

template <typename T, bool TWO = false>
T decide_impl(const T *P, int64_t m, const T *Q, int64_t n, int D, T r) {
    std::vector<T> A4b, inv2b;
    q_seg_terms(Q, n, D, A4b, inv2b);
    std::vector<T> colv((size_t)(n - 1), T(-1));
    return decide_rows<T, TWO>(P, m, Q, n, D, r, 0, m - 1, colv, A4b, inv2b);
}

inline bool continue_bisect(double lb, double ub, int stop_mode, double eps) {
    return stop_mode == 0 ? (ub - lb > eps) : (ub - lb > eps * ub);
}

}  // namespace

extern "C" {

double or_endpoint_lb(const double *P, int64_t m, const double *Q, int64_t n, int dims) {
    // Point::dist_sqr (geometry_basics.hpp:72-81): host code, no FMA.
    auto d2 = [&](const double *a, const double *b) {
        double result = 0.0;
        for (int k = 0; k < dims; ++k) {
            double ds = a[k] - b[k];
            double sq = ds * ds;
            result = result + sq;
        }
        return result;
    };
    double a = d2(P, Q);
    double b = d2(P + (m - 1) * dims, Q + (n - 1) * dims);
    return std::sqrt(std::max(a, b));
}

// Midpoint refinement, `levels` times: between consecutive points a, b the
// point a + 0.5 (b - a) (0.5 (b - a) is exact, so one rounding).  out holds
// (m - 1) 2^levels + 1 points.
int64_t or_refine(const double *P, int64_t m, int dims, int levels, double *out) {
    std::vector<double> cur(P, P + m * dims), nxt;
    int64_t len = m;
    for (int l = 0; l < levels && len >= 2; ++l) {
        nxt.assign((size_t)((2 * len - 1) * dims), 0.0);
        for (int64_t i = 0; i + 1 < len; ++i)
            for (int k = 0; k < dims; ++k) {
                const double a = cur[(size_t)(i * dims + k)], b = cur[(size_t)((i + 1) * dims + k)];
                nxt[(size_t)(2 * i * dims + k)] = a;
                nxt[(size_t)((2 * i + 1) * dims + k)] = a + 0.5 * (b - a);
            }
        for (int k = 0; k < dims; ++k) nxt[(size_t)((2 * len - 2) * dims + k)] = cur[(size_t)((len - 1) * dims + k)];
        cur.swap(nxt);
        len = 2 * len - 1;
    }
    if (out) std::copy(cur.begin(), cur.end(), out);
    return len;
}

// Upper bound of the continuous distance at most DFD: DFD of both curves
// refined `levels` times.  A coupling of the refined vertex sequences is a
// matching of the curves, and the distance of two points moving linearly is
// convex, so each refined pair is at most the larger of the pairs it lies
// between: FD <= this <= DFD (up to the rounding of the midpoints).
double or_dfd_refined(const double *P, int64_t m, const double *Q, int64_t n, int dims, int levels) {
    const int64_t m2 = levels > 0 && m >= 2 ? ((m - 1) << levels) + 1 : m;
    const int64_t n2 = levels > 0 && n >= 2 ? ((n - 1) << levels) + 1 : n;
    std::vector<double> Pr((size_t)(m2 * dims)), Qr((size_t)(n2 * dims));
    or_refine(P, m, dims, levels, Pr.data());
    or_refine(Q, n, dims, levels, Qr.data());
    return dfd_impl<double>(Pr.data(), m2, Qr.data(), n2, dims);
}

double or_dfd(const double *P, int64_t m, const double *Q, int64_t n, int dims) {
    return dfd_impl<double>(P, m, Q, n, dims);
}

double or_vefd(const double *P, int64_t m, const double *Q, int64_t n, int dims,
               double init_lb) {
    return vefd_impl<double>(P, m, Q, n, dims, init_lb);
}

double or_vefd_f32(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                   double init_lb) {
    auto Pf = to_t<float>(P, m * dims);
    auto Qf = to_t<float>(Q, n * dims);
    return (double)vefd_impl<float>(Pf.data(), m, Qf.data(), n, dims, (float)init_lb);
}

double or_decide_ub(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                    double r) {
    return decide_impl<double>(P, m, Q, n, dims, r);
}

int or_decide(const double *P, int64_t m, const double *Q, int64_t n, int dims, double r) {
    return or_decide_ub(P, m, Q, n, dims, r) == 1.0 ? 1 : 0;
}

double or_decide_ub_f32(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                        double r) {
    auto Pf = to_t<float>(P, m * dims);
    auto Qf = to_t<float>(Q, n * dims);
    return (double)decide_impl<float>(Pf.data(), m, Qf.data(), n, dims, (float)r);
}

void or_free_interval(const double *P, const double *Q, int dims, int64_t i, int64_t j,
                      double r, int which, double *out) {
    double lb, ub;
    if (which == 0) {
        const double *qa = Q + j * dims, *qb = Q + (j + 1) * dims;
        double Ab = seg_len2(qa, qb, dims);
        free_iv(P + (i + 1) * dims, qa, qb, dims, r * r, 4.0 * Ab, 0.5 / Ab, lb, ub);
    } else {
        const double *pa = P + i * dims, *pb = P + (i + 1) * dims;
        double Aa = seg_len2(pa, pb, dims);
        free_iv(Q + (j + 1) * dims, pa, pb, dims, r * r, 4.0 * Aa, 0.5 / Aa, lb, ub);
    }
    out[0] = lb;
    out[1] = ub;
}

void or_free_interval_2ph(const double *P, const double *Q, int dims, int64_t i, int64_t j,
                          double r, int which, double *out) {
    double lb, ub;
    if (which == 0) {
        const double *qa = Q + j * dims, *qb = Q + (j + 1) * dims;
        double Ab = seg_len2(qa, qb, dims);
        free_iv_2ph(P + (i + 1) * dims, qa, qb, dims, r * r, 4.0 * Ab, 0.5 / Ab, lb, ub);
    } else {
        const double *pa = P + i * dims, *pb = P + (i + 1) * dims;
        double Aa = seg_len2(pa, pb, dims);
        free_iv_2ph(Q + (j + 1) * dims, pa, pb, dims, r * r, 4.0 * Aa, 0.5 / Aa, lb, ub);
    }
    out[0] = lb;
    out[1] = ub;
}

double or_decide_ub_2ph(const double *P, int64_t m, const double *Q, int64_t n, int dims, double r) {
    return decide_impl<double, true>(P, m, Q, n, dims, r);
}

void or_ve_weights(const double *P, const double *Q, int dims, int64_t i, int64_t j,
                   double *out) {
    const double *pa = P + i * dims, *pb = P + (i + 1) * dims;
    const double *qa = Q + j * dims, *qb = Q + (j + 1) * dims;
    out[0] = point_seg2(qb, pa, pb, dims);
    out[1] = point_seg2(pb, qa, qb, dims);
}

int64_t or_distance(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                    int stop_mode, double eps, int vefd_gate, double *out,
                    double *splits, int32_t *decisions, int64_t max_trace) {
    return or_distance_form(P, m, Q, n, dims, stop_mode, eps, vefd_gate, 0, out, splits, decisions,
                            max_trace);
}

int64_t or_distance_form(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                         int stop_mode, double eps, int vefd_gate, int form, double *out,
                         double *splits, int32_t *decisions, int64_t max_trace) {
    double lb0 = or_endpoint_lb(P, m, Q, n, dims);
    double ub = or_dfd(P, m, Q, n, dims);
    double lb = lb0;
    out[1] = lb0;
    out[2] = ub;
    if (m < 2 || n < 2) {
        // A single point against a curve: the distance is the DFD exactly.
        out[0] = ub;
        out[3] = lb0;
        out[4] = ub;
        out[5] = ub;
        return 0;
    }
    if (!vefd_gate || m * n > 10000) lb = or_vefd(P, m, Q, n, dims, lb0);
    out[3] = lb;
    int64_t iter = 0;
    std::vector<double> A4b, inv2b;
    q_seg_terms(Q, n, dims, A4b, inv2b);
    std::vector<double> colv((size_t)(n - 1));
    while (continue_bisect(lb, ub, stop_mode, eps)) {
        double split = (ub + lb) / 2.0;
        std::fill(colv.begin(), colv.end(), -1.0);
        double rub = form == 1
                         ? decide_rows<double, true>(P, m, Q, n, dims, split, 0, m - 1, colv, A4b, inv2b)
                         : decide_rows<double>(P, m, Q, n, dims, split, 0, m - 1, colv, A4b, inv2b);
        int dec = rub == 1.0;
        if (splits && iter < max_trace) splits[iter] = split;
        if (decisions && iter < max_trace) decisions[iter] = dec;
        ++iter;
        if (dec) ub = split; else lb = split;
    }
    out[0] = (ub + lb) / 2.0;
    out[4] = lb;
    out[5] = ub;
    return iter;
}

double or_sweep_rows(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                     int kind, double r, int64_t r0, int64_t r1) {
    if (kind == 0) {
        std::vector<double> colv((size_t)(n - 1), (double)FLT_MAX);
        double tr[2];
        vefd_rows<double>(P, m, Q, n, dims, r, r0, r1, colv, tr);
        double s = 0;
        for (double v : colv) s += v;
        return s;
    }
    if (kind == 1) {
        std::vector<double> A4b, inv2b;
        q_seg_terms(Q, n, dims, A4b, inv2b);
        std::vector<double> colv((size_t)(n - 1), 0.0);
        decide_rows<double>(P, m, Q, n, dims, r, r0, r1, colv, A4b, inv2b);
        double s = 0;
        for (double v : colv) s += v;
        return s;
    }
    return dfd_impl<double>(P + r0 * dims, r1 - r0, Q, n, dims);
}

// Row-block pipeline of the bound sweeps for the multi-core timing: thread t
// owns rows [r0 + t*B, ...) and walks the columns in chunks; before chunk c it
// waits until thread t-1 has published its last row's outputs for chunk c.
// Per cell the arithmetic is that of vefd_rows / dfd_impl, in the same order.
namespace {

struct alignas(128) Progress { std::atomic<int64_t> v{0}; };

struct BlockJob {
    const double *P, *Q;
    int D, kind;
    double r, big;
    int64_t a, b, cols, chunk;
    const double *edge_in;     // previous block's last-row outputs, or null for the first block
    const Progress *ready_in;  // columns of edge_in already published
    double *edge_out;
    Progress *ready_out;
    double *last;              // DFD: top of the block's last cell; VE-FD: its min(T, R);
                               // decision: its rightUB
    const double *A4b = nullptr, *inv2b = nullptr;  // decision: per-column terms
};

// One row of one column chunk.  `left` is read from lin (the row above), the
// row's outputs go to lout (lin == lout is allowed: each column is read
// before it is written).
inline double row_chunk(const BlockJob &J, int64_t i, int64_t c0, int64_t c1, double bt,
                        const double *lin, double *lout, double *lastv = nullptr) {
    const int D = J.D;
    const double *pa = J.P + i * D, *pb = pa + D;
    if (J.kind == 1) {
        // decide_rows (reachable<D>, .cu:1173-1244).  A cell whose two inputs
        // are unreachable (< 0) gives (-1, -2, -1) whatever its intervals (the
        // rules below yield exactly that), so its intervals are not computed.
        const double r2 = J.r * J.r;
        const double Aa = seg_len2(pa, pb, D);
        const double A4a = 4.0 * Aa, inv2a = 0.5 / Aa;
        double rUBl = -2.0;
        for (int64_t j = c0; j < c1; ++j) {
            const double leftLB = lin[j - c0], botLB = bt;
            double rLB = -1.0, rUB = -2.0, tLB = -1.0, tUB = -2.0;
            if (!(leftLB < 0.0 && botLB < 0.0)) {
                const double *qa = J.Q + j * D, *qb = qa + D;
                free_iv(pb, qa, qb, D, r2, J.A4b[j], J.inv2b[j], rLB, rUB);
                free_iv(qb, pa, pb, D, r2, A4a, inv2a, tLB, tUB);
                if (botLB < 0.0) {
                    const bool inv = leftLB > rUB || leftLB < 0.0;
                    const double t2 = std::fmax(rLB, leftLB);
                    rLB = inv ? -1.0 : t2;
                    rUB = inv ? -2.0 : rUB;
                }
                if (leftLB < 0.0) {
                    const bool inv = botLB > tUB || botLB < 0.0;
                    const double t2 = std::fmax(tLB, botLB);
                    tLB = inv ? -1.0 : t2;
                }
            }
            bt = tLB;
            lout[j - c0] = rLB;
            rUBl = rUB;
        }
        if (lastv) *lastv = rUBl;
        return bt;
    }
    if (J.kind == 0) {
        for (int64_t j = c0; j < c1; ++j) {
            const double *qa = J.Q + j * D, *qb = qa + D;
            const double in = std::fmin(lin[j - c0], bt);
            const double wr = point_seg2(pb, qa, qb, D);
            const double wt = point_seg2(qb, pa, pb, D);
            bt = std::fmax(wt, in);
            lout[j - c0] = std::fmax(wr, in);
        }
        if (lastv && c1 > c0) *lastv = std::fmin(bt, lout[c1 - 1 - c0]);
    } else {
        for (int64_t j = c0; j < c1; ++j) {
            const double d = dist2(pa, J.Q + j * D, D);
            const double in = std::fmin(lin[j - c0], bt);
            const double top = std::fmax(d, in);
            lout[j - c0] = std::fmin(top, bt);
            bt = top;
        }
    }
    return bt;
}

void run_block(BlockJob J) {
    if (J.a >= J.b) {
        J.ready_out->v.store(J.cols, std::memory_order_release);
        return;
    }
    const int64_t rows = J.b - J.a;
    std::vector<double> bot((size_t)rows);
    for (int64_t i = J.a; i < J.b; ++i) {
        double v = J.big;                          // DFD, VE-FD rows after the first
        if (J.kind == 0 && i == 0) v = J.r * J.r;  // VE-FD row 0 (.cu:258-259)
        if (J.kind == 1) v = (i == 0) ? 0.0 : -1.0;  // decision (.cu:1026-1027)
        bot[(size_t)(i - J.a)] = v;
    }
    std::vector<double> buf((size_t)J.chunk), first((size_t)J.chunk);
    double lastv = 0.0;
    for (int64_t c0 = 0; c0 < J.cols; c0 += J.chunk) {
        const int64_t c1 = std::min(J.cols, c0 + J.chunk);
        if (J.edge_in) {
            while (J.ready_in->v.load(std::memory_order_acquire) < c1) std::this_thread::yield();
            std::copy(J.edge_in + c0, J.edge_in + c1, first.begin());
        } else {
            // vefd_rows: the first row enters with left = big; dfd_impl on
            // P + r0: (r0, 0) enters with 0, the rest of the row with big;
            // decide_rows: the left boundary is -1 (.cu:656).
            std::fill(first.begin(), first.begin() + (c1 - c0), J.kind == 1 ? -1.0 : J.big);
            if (J.kind == 2 && c0 == 0) first[0] = 0.0;
        }
        double *lv = (c1 == J.cols && rows == 1) ? &lastv : nullptr;
        bot[0] = row_chunk(J, J.a, c0, c1, bot[0], first.data(), buf.data(), lv);
        for (int64_t i = J.a + 1; i < J.b; ++i) {
            lv = (c1 == J.cols && i + 1 == J.b) ? &lastv : nullptr;
            bot[(size_t)(i - J.a)] = row_chunk(J, i, c0, c1, bot[(size_t)(i - J.a)], buf.data(), buf.data(), lv);
        }
        std::copy(buf.begin(), buf.begin() + (c1 - c0), J.edge_out + c0);
        J.ready_out->v.store(c1, std::memory_order_release);
    }
    *J.last = (J.kind == 2) ? bot[(size_t)rows - 1] : lastv;
}

}  // namespace

double or_sweep_rows_mt(const double *P, int64_t m, const double *Q, int64_t n, int dims,
                        int kind, double r, int64_t r0, int64_t r1, int threads) {
    if ((kind != 0 && kind != 2) || r1 <= r0) return or_sweep_rows(P, m, Q, n, dims, kind, r, r0, r1);
    const int64_t cols = (kind == 0) ? n - 1 : n;
    const int64_t rows = r1 - r0;
    if (threads < 1) threads = 1;
    if (threads > rows) threads = (int)std::max<int64_t>(rows, 1);
    const int64_t B = (rows + threads - 1) / threads;
    threads = (int)((rows + B - 1) / B);  // every block non-empty
    std::vector<std::vector<double>> edge((size_t)threads, std::vector<double>((size_t)cols));
    std::vector<Progress> ready((size_t)threads);
    std::vector<double> last((size_t)threads * 16, 0.0);  // one cache line apart
    std::vector<BlockJob> jobs((size_t)threads);
    for (int t = 0; t < threads; ++t) {
        BlockJob &J = jobs[(size_t)t];
        J.P = P; J.Q = Q; J.D = dims; J.kind = kind; J.r = r; J.big = (double)FLT_MAX;
        J.a = r0 + t * B; J.b = std::min(r1, J.a + B); J.cols = cols; J.chunk = 256;
        J.edge_in = t > 0 ? edge[(size_t)t - 1].data() : nullptr;
        J.ready_in = t > 0 ? &ready[(size_t)t - 1] : nullptr;
        J.edge_out = edge[(size_t)t].data();
        J.ready_out = &ready[(size_t)t];
        J.last = &last[(size_t)t * 16];
    }
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(run_block, jobs[(size_t)t]);
    run_block(jobs[0]);
    for (auto &th : pool) th.join();
    if (kind == 2) return std::sqrt(last[(size_t)(threads - 1) * 16]);
    double s = 0;
    for (double v : edge[(size_t)threads - 1]) s += v;
    return s;
}

double or_sweep_mt(const double *P, int64_t m, const double *Q, int64_t n, int dims, int kind, double r,
                   int threads) {
    const int64_t rows = (kind == 2) ? m : m - 1;
    const int64_t cols = (kind == 2) ? n : n - 1;
    if (rows < 1 || cols < 1 || kind < 0 || kind > 2) return std::nan("");
    if (threads < 1) threads = 1;
    if (threads > rows) threads = (int)rows;
    const int64_t B = (rows + threads - 1) / threads;
    threads = (int)((rows + B - 1) / B);
    std::vector<double> A4b, inv2b;
    if (kind == 1) q_seg_terms(Q, n, dims, A4b, inv2b);
    std::vector<std::vector<double>> edge((size_t)threads, std::vector<double>((size_t)cols));
    std::vector<Progress> ready((size_t)threads);
    std::vector<double> last((size_t)threads * 16, 0.0);
    std::vector<BlockJob> jobs((size_t)threads);
    for (int t = 0; t < threads; ++t) {
        BlockJob &J = jobs[(size_t)t];
        J.P = P; J.Q = Q; J.D = dims; J.kind = kind; J.r = r; J.big = (double)FLT_MAX;
        J.a = t * B; J.b = std::min(rows, J.a + B); J.cols = cols; J.chunk = 256;
        J.edge_in = t > 0 ? edge[(size_t)t - 1].data() : nullptr;
        J.ready_in = t > 0 ? &ready[(size_t)t - 1] : nullptr;
        J.edge_out = edge[(size_t)t].data();
        J.ready_out = &ready[(size_t)t];
        J.last = &last[(size_t)t * 16];
        if (kind == 1) { J.A4b = A4b.data(); J.inv2b = inv2b.data(); }
    }
    std::vector<std::thread> pool;
    for (int t = 1; t < threads; ++t) pool.emplace_back(run_block, jobs[(size_t)t]);
    run_block(jobs[0]);
    for (auto &th : pool) th.join();
    const double v = last[(size_t)(threads - 1) * 16];
    return kind == 1 ? v : std::sqrt(v);
}

}  // extern "C"
