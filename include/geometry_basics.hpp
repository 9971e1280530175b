// Point / Points: source-compatible with the reference's
// include/geometry_basics.hpp:11-99 (the members user code calls).
#pragma once

#include <cmath>
#include <cstddef>
#include <vector>

struct Point {
    std::vector<double> coordinates;

    Point() = default;
    explicit Point(const std::vector<double> &in) : coordinates(in) {}

    Point &operator-=(const Point &p) {
        for (std::size_t i = 0; i < coordinates.size(); ++i) coordinates[i] -= p.coordinates[i];
        return *this;
    }
    Point operator-(const Point &p) const { Point r = *this; r -= p; return r; }
    Point &operator+=(const Point &p) {
        for (std::size_t i = 0; i < coordinates.size(); ++i) coordinates[i] += p.coordinates[i];
        return *this;
    }
    Point operator+(const Point &p) const { Point r = *this; r += p; return r; }
    Point operator*(double s) const {
        Point r = *this;
        for (double &c : r.coordinates) c *= s;
        return r;
    }
    Point &operator/=(double s) {
        for (double &c : coordinates) c /= s;
        return *this;
    }
    // Sum of squared coordinate differences, each square rounded before the sum.
    double dist_sqr(const Point &p) const {
        double result = 0;
        for (std::size_t i = 0; i < coordinates.size(); ++i) {
            double d = coordinates[i] - p.coordinates[i];
            double sq = d * d;
            result = result + sq;
        }
        return result;
    }
    double dist(const Point &p) const { return std::sqrt(dist_sqr(p)); }
    std::vector<double> &getCoordinates() { return coordinates; }
    const std::vector<double> &getCoordinates() const { return coordinates; }
    double operator[](std::size_t i) const { return coordinates[i]; }
    int size() const { return (int)coordinates.size(); }
};

// The reference declares this printer (geometry_basics.hpp:103) without
// defining it: the coordinates separated by commas.
#include <ostream>
inline std::ostream &operator<<(std::ostream &out, const Point &p) {
    for (std::size_t k = 0; k < p.coordinates.size(); ++k) out << (k ? "," : "") << p.coordinates[k];
    return out;
}

typedef std::vector<Point> Points;
typedef std::vector<Points> multiPoints;
