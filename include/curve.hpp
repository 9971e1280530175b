// Curve: source-compatible with the reference's include/curve.hpp:9-40.
#pragma once

#include <cstddef>
#include <cstdio>
#include <iostream>
#include <vector>

#include "geometry_basics.hpp"

class Curve {
public:
    typedef unsigned long index_type;

    explicit Curve(unsigned int dimensions) : number_dimensions(dimensions) {}
    Curve(const Points &pts, unsigned int dimensions) : points(pts), number_dimensions(dimensions) {
        if (points.empty()) std::cerr << "warning: constructed empty curve" << std::endl;
    }

    std::size_t size() const { return points.size(); }
    bool empty() const { return points.empty(); }
    std::size_t dimensions() const { return number_dimensions; }
    const Point &operator[](std::size_t i) const { return points[i]; }
    Point front() const { return points.front(); }
    Point back() const { return points.back(); }
    void push_back(const Point &p) { points.push_back(p); }
    // Translates every point by -point, printing the second point before and
    // after and the offset, as the reference does (curve.hpp:25-30); needs at
    // least two points.
    void recenter(const Point &point) {
        std::printf("%f, %f\n", points[1][0], points[1][1]);
        std::printf("%f, %f\n", point[0], point[1]);
        for (auto &p : points) p -= point;
        std::printf("%f, %f\n", points[1][0], points[1][1]);
    }

    Points::iterator begin() { return points.begin(); }
    Points::iterator end() { return points.end(); }
    Points::const_iterator begin() const { return points.cbegin(); }
    Points::const_iterator end() const { return points.cend(); }

private:
    Points points;
    unsigned int number_dimensions;
};

class Curves : public std::vector<Curve> {
public:
    Curve get(std::size_t i) const { return this->operator[](i); }
};

// The reference declares this printer (curve.hpp:49) without defining it: one
// point per line, coordinates separated by commas (the loader's format).
inline std::ostream &operator<<(std::ostream &out, const Curve &curve) {
    for (const Point &p : curve) out << p << '\n';
    return out;
}
