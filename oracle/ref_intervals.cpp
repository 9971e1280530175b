// C entry point over the reference's own free-interval routine
// (include/geometry_basics.hpp:161-298, IntersectionAlgorithm::intersection_interval),
// compiled from the reference header where it lies.  Test infrastructure only:
// used to cross-check the oracle's free intervals (containment, tolerance 2.5e-4).
#include "geometry_basics.hpp"

extern "C" void ref_intersection_interval(const double *center, const double *seg0,
                                          const double *seg1, int dims, double radius,
                                          double *out) {
    Point c(std::vector<double>(center, center + dims));
    Point a(std::vector<double>(seg0, seg0 + dims));
    Point b(std::vector<double>(seg1, seg1 + dims));
    Interval iv = IntersectionAlgorithm::intersection_interval(c, radius, a, b);
    out[0] = iv.begin;
    out[1] = iv.end;
}

// Point::dist_sqr (geometry_basics.hpp:72-81), the endpoint-bound formula of
// frechet.cpp:30.
extern "C" double ref_dist_sqr(const double *a, const double *b, int dims) {
    Point pa(std::vector<double>(a, a + dims));
    Point pb(std::vector<double>(b, b + dims));
    return pa.dist_sqr(pb);
}
