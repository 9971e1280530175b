// A C++ caller of the drop-in, written as the reference's users write it
// (README.md:38-51): it builds Curve objects point by point and calls
// compute_distance_parallel (include/wrapper.hpp:4) with and without
// forceOversize, and Frechet::Continuous::distance_cuda (include/frechet.hpp:14)
// with its own eps.  Input: a binary file of int64 m, n, dims followed by the
// m + n points (FP64, point-major).  Output: one line "RESULT a b c" with the
// three distances as hexadecimal floats (exact), after the library's own
// progress lines.
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <fstream>
#include <vector>

#include "curve.hpp"
#include "frechet.hpp"
#include "wrapper.hpp"

int main(int argc, char **argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s <points.bin> <eps>\n", argv[0]);
        return 1;
    }
    std::ifstream f(argv[1], std::ios::binary);
    int64_t hdr[3];
    if (!f.read(reinterpret_cast<char *>(hdr), sizeof(hdr))) return 1;
    const int64_t m = hdr[0], n = hdr[1], dims = hdr[2];
    std::vector<double> xs((size_t)((m + n) * dims));
    if (!f.read(reinterpret_cast<char *>(xs.data()), (std::streamsize)(xs.size() * sizeof(double)))) return 1;
    Curve c1((unsigned)dims), c2((unsigned)dims);
    for (int64_t i = 0; i < m + n; ++i) {
        Point p(std::vector<double>(xs.begin() + i * dims, xs.begin() + (i + 1) * dims));
        if (i < m) c1.push_back(p);
        else c2.push_back(p);
    }
    const double eps = std::strtod(argv[2], nullptr);
    const double a = compute_distance_parallel(c1, c2);
    const double b = compute_distance_parallel(c1, c2, true);
    const double c = Frechet::Continuous::distance_cuda(c1, c2, eps, true);
    std::printf("RESULT %a %a %a\n", a, b, c);
    return (a == a && b == b && c == c) ? 0 : 2;
}
