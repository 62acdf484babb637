// Minimal 3-vector with the reference's public fields (Solaris/Vector.h).
#ifndef VECTOR_H_
#define VECTOR_H_
#include <cmath>
class Vector {
public:
    double x, y, z;
    Vector() : x(0.0), y(0.0), z(0.0) {}
    Vector(double a, double b, double c) : x(a), y(b), z(c) {}
    double Length() { return std::sqrt(x * x + y * y + z * z); }
    double LengthSQR() { return x * x + y * y + z * z; }
};
#endif
