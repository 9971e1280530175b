// CSV trajectory loader with the reference's behaviour (include/txtReader.hpp:6-88):
// one point per line, comma-separated coordinates; a point equal to the
// previous one on its first `dims` coordinates is dropped; lines before
// START are skipped and at most LIMIT points are read.
#pragma once

#include <fstream>
#include <iostream>
#include <sstream>
#include <string>

#include "geometry_basics.hpp"

inline Points txtLoad(const std::string &path, int dims, int START = 0, int LIMIT = 10000000) {
    Points ret;
    std::ifstream fin(path);
    if (!fin.is_open()) {
        std::cerr << "Failed to open file: " << path << std::endl;
        return ret;
    }
    std::string line;
    int count = 0;
    while (std::getline(fin, line)) {
        if (count >= LIMIT) break;
        std::stringstream ss(line);
        std::string token;
        Point pt;
        while (std::getline(ss, token, ',')) {
            try {
                pt.coordinates.push_back(std::stod(token));
            } catch (const std::exception &) {
                std::cerr << "Warning: invalid double in line: " << line << std::endl;
            }
        }
        if (pt.coordinates.empty()) continue;
        ++count;
        if (count < START) continue;
        bool same = !ret.empty();
        for (int k = 0; same && k < dims && k < (int)pt.coordinates.size(); ++k)
            same = ret.back().coordinates[k] == pt.coordinates[k];
        if (!same) ret.push_back(std::move(pt));
    }
    return ret;
}

inline Points txtLoader2(const std::string &path, int START = 0, int LIMIT = 10000000) {
    return txtLoad(path, 2, START, LIMIT);
}
inline Points txtLoader3(const std::string &path, int START = 0, int LIMIT = 10000000) {
    return txtLoad(path, 3, START, LIMIT);
}
