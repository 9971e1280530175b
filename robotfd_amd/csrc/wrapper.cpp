// C++ drop-in API over the C ABI: compute_distance_parallel
// (compute_distance_wrapper.cpp:8-12) and Frechet::Continuous::distance_cuda
// (frechet.cpp:18-75).  Curves are flattened to AoS FP64 here and handed to
// rfd_distance; one lazily created context per host thread on device 0.
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <limits>
#include <vector>

#include "rfd.h"
#include "wrapper.hpp"

namespace {

struct ThreadCtx {
    rfd_ctx *ctx = nullptr;
    rfd_status status = RFD_OK;
    ThreadCtx() { status = rfd_create(&ctx, 0); }
    ~ThreadCtx() {
        if (ctx) rfd_destroy(ctx);
    }
};

rfd_ctx *thread_ctx() {
    thread_local ThreadCtx t;
    if (t.status != RFD_OK) {
        std::fprintf(stderr, "robotfd: no GPU context (%s)\n", rfd_status_string(t.status));
        return nullptr;
    }
    return t.ctx;
}

std::vector<double> flatten(const Curve &c, int dims) {
    std::vector<double> out(c.size() * (std::size_t)dims);
    for (std::size_t i = 0; i < c.size(); ++i)
        for (int k = 0; k < dims; ++k) out[i * dims + k] = c[i].coordinates[k];
    return out;
}

}  // namespace

namespace Frechet {
namespace Continuous {

double distance_cuda(const Curve &curve1, const Curve &curve2, const double eps,
                     bool forceOversize) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (curve1.empty() || curve2.empty()) {
        std::fprintf(stderr, "robotfd: empty curve\n");
        return nan;
    }
    // The dimension comes from the first point (.cu:25).
    const int dims = (int)curve1[0].coordinates.size();
    if ((int)curve2[0].coordinates.size() != dims) {
        std::fprintf(stderr, "robotfd: curves of different dimension\n");
        return nan;
    }
    rfd_ctx *ctx = thread_ctx();
    if (!ctx) return nan;
    std::vector<double> P = flatten(curve1, dims), Q = flatten(curve2, dims);
    rfd_opts o;
    rfd_default_opts(&o);
    o.stop_mode = RFD_STOP_ABS;
    o.eps = eps;
    o.verbose = 1;
    // The reference's decision path: reachable<D> when forceOversize is set
    // or its interval arrays do not fit, else the two-phase path (.cu:87-96).
    o.decide_form = forceOversize ? RFD_DECIDE_FUSED : RFD_DECIDE_REFERENCE;
    rfd_result r;
    rfd_status st = rfd_distance(ctx, P.data(), (int64_t)curve1.size(), Q.data(),
                                 (int64_t)curve2.size(), dims, &o, &r);
    if (st != RFD_OK) {
        std::fprintf(stderr, "robotfd: %s: %s\n", rfd_status_string(st), rfd_last_error(ctx));
        return nan;
    }
    return r.distance;
}

}  // namespace Continuous
}  // namespace Frechet

double compute_distance_parallel(const Curve &curve1, const Curve &curve2, bool forceOversize) {
    return Frechet::Continuous::distance_cuda(curve1, curve2, (double)FLT_EPSILON, forceOversize);
}
