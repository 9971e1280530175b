// The CSV loader (include/txtReader.hpp) and the Curve helpers the reference
// declares (curve.hpp:25-30, 49), without a GPU: loads a file with
// txtLoader2 / txtLoader3, prints every point as hexadecimal floats, then
// prints the curve with operator<< and recentres it on its first point.
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "curve.hpp"
#include "txtReader.hpp"

int main(int argc, char **argv) {
    if (argc < 3) return 1;
    const int dims = std::atoi(argv[2]);
    Points pts = dims == 2 ? txtLoader2(argv[1]) : txtLoader3(argv[1]);
    std::printf("POINTS %zu\n", pts.size());
    for (const Point &p : pts) {
        for (double c : p.coordinates) std::printf(" %a", c);
        std::printf("\n");
    }
    Curve c(pts, (unsigned)dims);
    std::cout << "CURVE\n" << c << std::flush;
    if (c.size() >= 2) {
        std::printf("RECENTER\n");
        c.recenter(c.front());
        std::printf("FIRST %a %a\n", c.front()[0], c.front()[1]);
    }
    return 0;
}
