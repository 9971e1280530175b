// fd <curve1.txt> <curve2.txt> <dim> [1]: the reference CLI (src/main.cpp:13-44)
// on the gfx950 library.  The shorter curve goes first (main.cpp:34,39).
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>

#include "txtReader.hpp"
#include "wrapper.hpp"

int main(int argc, char **argv) {
    if (argc < 4) {
        std::cerr << "Usage:\n  " << argv[0] << " <curve1.txt> <curve2.txt> [dim]\n";
        return 1;
    }
    const std::string path1 = argv[1], path2 = argv[2];
    const int dim = (int)std::strtoll(argv[3], nullptr, 10);
    bool forceOversize = argc >= 5 && std::string(argv[4]) == "1";
    std::printf("%d\n", forceOversize ? 1 : 0);
    if (dim != 2 && dim != 3) {
        std::cerr << "dim must be 2 or 3\n";
        return 1;
    }
    Curve c1(txtLoad(path1, dim), dim), c2(txtLoad(path2, dim), dim);
    if (c1.empty() || c2.empty()) return 1;
    double d = (c1.size() > c2.size()) ? compute_distance_parallel(c2, c1, forceOversize)
                                       : compute_distance_parallel(c1, c2, forceOversize);
    return d == d ? 0 : 2;
}
