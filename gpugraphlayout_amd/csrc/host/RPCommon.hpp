// Common helpers of the host layer. Mirrors the public helpers of the reference's
// src/RPCommon.hpp:43-96 that callers use (get_random, Coordinate, Real2DVector,
// is_file_exists, basename); no CUDA error macro, since every engine call returns a status.
#ifndef RPCommon_hpp
#define RPCommon_hpp
#include <string>

bool is_file_exists(std::string filepath);
std::string basename(std::string filepath);

namespace RPGraph
{
    // lowerbound + (upperbound-lowerbound)*float(random())/float(RAND_MAX), in float
    // (src/RPCommon.cpp:50-53); seeded by srandom().
    float get_random(float lowerbound, float upperbound);

    struct Real2DVector
    {
        float x, y;
        Real2DVector(float x, float y) : x(x), y(y) {}
        float magnitude() const;
        Real2DVector operator*(float b) const { return Real2DVector(x * b, y * b); }
        Real2DVector operator+(Real2DVector b) const { return Real2DVector(x + b.x, y + b.y); }
        Real2DVector operator-(Real2DVector b) const { return Real2DVector(x - b.x, y - b.y); }
        void operator+=(Real2DVector b) { x += b.x; y += b.y; }
    };

    struct Coordinate
    {
        float x, y;
        Coordinate(float x, float y) : x(x), y(y) {}
        bool operator==(Coordinate b) const { return x == b.x && y == b.y; }
        Coordinate operator+(Real2DVector b) const { return Coordinate(x + b.x, y + b.y); }
        Coordinate operator-(Coordinate b) const { return Coordinate(x - b.x, y - b.y); }
    };
}

#endif
