#pragma once
// Three-vector arithmetic with a fixed evaluation order (component-wise, left to right, no contraction),
// used wherever results must be bit-identical across the host setup, the oracle and the device kernels.
#include <cmath>

#ifdef __HIPCC__
#define MHDG_HD __host__ __device__ inline
#else
#define MHDG_HD inline
#endif

namespace mhdg
{
struct V3
{
    double x = 0, y = 0, z = 0;
    MHDG_HD V3() = default;
    MHDG_HD V3(double a, double b, double c) : x(a), y(b), z(c) {}
    MHDG_HD V3 &operator+=(const V3 &o) { x += o.x; y += o.y; z += o.z; return *this; }
    MHDG_HD V3 &operator-=(const V3 &o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    MHDG_HD V3 &operator*=(double s) { x *= s; y *= s; z *= s; return *this; }
};
MHDG_HD V3 operator+(V3 a, const V3 &b) { return a += b; }
MHDG_HD V3 operator-(V3 a, const V3 &b) { return a -= b; }
MHDG_HD V3 operator*(V3 a, double s) { return a *= s; }
MHDG_HD V3 operator*(double s, V3 a) { return a *= s; }
MHDG_HD double dot(const V3 &a, const V3 &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
MHDG_HD V3 cross(const V3 &a, const V3 &b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
MHDG_HD double norm(const V3 &a) { return sqrt(dot(a, a)); }
MHDG_HD V3 ld3(const double *p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
MHDG_HD void st3(double *p, long i, const V3 &v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
} // namespace mhdg
