#pragma once
// Forward-mode (tangent-linear) arithmetic for the explicit-stage kernels: a value and one directional
// derivative.  The kernels are templates on their real type R; R = double is the production path, R = Dual
// propagates d(output)/d(seed) through every operation, with branches (limiters, upwinding, floors, the CFL
// minimum) taken on the values, so the tangent is that of the branch the value computation takes.
#include <hip/hip_runtime.h>

#include "../host/vec3.h"

namespace mhdg
{
// the double overloads stay visible next to the Dual ones declared below
using ::fabs;
using ::sqrt;

struct Dual
{
    double v, d; // trivially constructible (LDS arrays), uninitialised like a double
    __host__ __device__ Dual() = default;
    __host__ __device__ Dual(double a) : v(a), d(0) {}
    __host__ __device__ Dual(double a, double b) : v(a), d(b) {}
    __host__ __device__ Dual &operator+=(const Dual &o) { v += o.v; d += o.d; return *this; }
    __host__ __device__ Dual &operator-=(const Dual &o) { v -= o.v; d -= o.d; return *this; }
    __host__ __device__ Dual &operator*=(const Dual &o) { d = d * o.v + v * o.d; v *= o.v; return *this; }
    __host__ __device__ Dual &operator/=(const Dual &o) { const double q = v / o.v; d = (d - q * o.d) / o.v; v = q; return *this; }
};
__host__ __device__ inline Dual operator+(Dual a, const Dual &b) { return a += b; }
__host__ __device__ inline Dual operator-(Dual a, const Dual &b) { return a -= b; }
__host__ __device__ inline Dual operator*(Dual a, const Dual &b) { return a *= b; }
__host__ __device__ inline Dual operator/(Dual a, const Dual &b) { return a /= b; }
__host__ __device__ inline Dual operator-(const Dual &a) { return {-a.v, -a.d}; }
__host__ __device__ inline bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }
__host__ __device__ inline bool operator<=(const Dual &a, const Dual &b) { return a.v <= b.v; }
__host__ __device__ inline bool operator>=(const Dual &a, const Dual &b) { return a.v >= b.v; }
__host__ __device__ inline bool operator==(const Dual &a, const Dual &b) { return a.v == b.v; }
__host__ __device__ inline bool operator!=(const Dual &a, const Dual &b) { return a.v != b.v; }
__device__ inline Dual sqrt(const Dual &a)
{
    const double s = ::sqrt(a.v);
    return {s, s > 0 ? a.d / (2.0 * s) : 0.0};
}
__device__ inline Dual fabs(const Dual &a) { return a.v < 0 ? -a : a; }
// std::min / std::max operand semantics on the values
__device__ inline Dual mn(const Dual &a, const Dual &b) { return (b < a) ? b : a; }
__device__ inline Dual mx(const Dual &a, const Dual &b) { return (a < b) ? b : a; }

__host__ __device__ inline double val(double x) { return x; }
__host__ __device__ inline double val(const Dual &x) { return x.v; }

// A real of type R from a value and its tangent (the tangent is dropped for R = double)
template <class R> __host__ __device__ inline R seeded(double v, double d);
template <> __host__ __device__ inline double seeded<double>(double v, double) { return v; }
template <> __host__ __device__ inline Dual seeded<Dual>(double v, double d) { return {v, d}; }

// Three-vector of duals; V3T<double> is V3 itself, so the production arithmetic is unchanged
struct V3D
{
    Dual x, y, z;
    __host__ __device__ V3D() : x(0.0), y(0.0), z(0.0) {}
    __host__ __device__ V3D(const Dual &a, const Dual &b, const Dual &c) : x(a), y(b), z(c) {}
    __host__ __device__ V3D(const V3 &o) : x(o.x), y(o.y), z(o.z) {}
    __host__ __device__ V3D &operator+=(const V3D &o) { x += o.x; y += o.y; z += o.z; return *this; }
    __host__ __device__ V3D &operator-=(const V3D &o) { x -= o.x; y -= o.y; z -= o.z; return *this; }
    __host__ __device__ V3D &operator*=(const Dual &s) { x *= s; y *= s; z *= s; return *this; }
};
__host__ __device__ inline V3D operator+(V3D a, const V3D &b) { return a += b; }
__host__ __device__ inline V3D operator-(V3D a, const V3D &b) { return a -= b; }
__host__ __device__ inline V3D operator*(V3D a, const Dual &s) { return a *= s; }
__host__ __device__ inline V3D operator*(const Dual &s, V3D a) { return a *= s; }
__host__ __device__ inline Dual dot(const V3D &a, const V3D &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline V3D cross(const V3D &a, const V3D &b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline Dual norm(const V3D &a) { return sqrt(dot(a, a)); }

template <class R> struct V3Of { using type = V3; };
template <> struct V3Of<Dual> { using type = V3D; };
template <class R> using V3T = typename V3Of<R>::type;

__host__ __device__ inline V3D L3(const Dual *p, long i) { return {p[3 * i], p[3 * i + 1], p[3 * i + 2]}; }
__host__ __device__ inline void st3(Dual *p, long i, const V3D &v) { p[3 * i] = v.x; p[3 * i + 1] = v.y; p[3 * i + 2] = v.z; }
} // namespace mhdg
