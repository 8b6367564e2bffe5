// dual.hpp -- forward-mode derivative numbers for the tangent form of the step.
//
// A Dual carries a value and its derivative along one input direction (here: a parameter
// such as a filament's Young's modulus).  Every operation computes the value exactly as the
// plain double operation would (same operands, same order, no fused multiply-add in the
// translation units that use it), so a kernel instantiated on Dual reproduces the FP64
// kernel's values bit for bit and carries the exact derivative of that computation beside
// them.  Derivative rules: (a b)' = a' b + a b', (a / b)' = (a' - (a/b) b') / b.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace lumagpu {

struct DualF;

struct Dual {
    double v, d;
    __host__ __device__ constexpr Dual() : v(0.0), d(0.0) {}
    __host__ __device__ constexpr Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ constexpr Dual(double x, double dx) : v(x), d(dx) {}
    __host__ __device__ explicit constexpr Dual(const DualF &x);
    __host__ __device__ explicit constexpr operator double() const { return v; }
    __host__ __device__ Dual &operator+=(const Dual &b) {
        v += b.v;
        d += b.d;
        return *this;
    }
    __host__ __device__ Dual &operator-=(const Dual &b) {
        v -= b.v;
        d -= b.d;
        return *this;
    }
    __host__ __device__ Dual &operator*=(const Dual &b) {
        const double nd = d * b.v + v * b.d;
        v *= b.v;
        d = nd;
        return *this;
    }
    __host__ __device__ Dual &operator/=(const Dual &b) {
        const double q = v / b.v;
        d = (d - q * b.d) / b.v;
        v = q;
        return *this;
    }
};

__host__ __device__ inline Dual operator+(const Dual &a, const Dual &b) { return Dual(a.v + b.v, a.d + b.d); }
__host__ __device__ inline Dual operator-(const Dual &a, const Dual &b) { return Dual(a.v - b.v, a.d - b.d); }
__host__ __device__ inline Dual operator-(const Dual &a) { return Dual(-a.v, -a.d); }
__host__ __device__ inline Dual operator*(const Dual &a, const Dual &b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__host__ __device__ inline Dual operator/(const Dual &a, const Dual &b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
// comparisons look at values only (branches follow the primal computation)
__host__ __device__ inline bool operator<(const Dual &a, const Dual &b) { return a.v < b.v; }
__host__ __device__ inline bool operator>(const Dual &a, const Dual &b) { return a.v > b.v; }

// The FP32 counterpart for the shifted tangent form (FP32 storage and collision of values and
// derivatives).  fmac and mul_split mirror the fused operations of the FP32 collision.
struct DualF {
    float v, d;
    __host__ __device__ constexpr DualF() : v(0.f), d(0.f) {}
    __host__ __device__ constexpr DualF(float x) : v(x), d(0.f) {}
    __host__ __device__ constexpr DualF(int x) : v(float(x)), d(0.f) {}
    __host__ __device__ constexpr DualF(float x, float dx) : v(x), d(dx) {}
    __host__ __device__ explicit constexpr DualF(double x) : v(float(x)), d(0.f) {}
    __host__ __device__ explicit constexpr DualF(const Dual &x) : v(float(x.v)), d(float(x.d)) {}
    __host__ __device__ explicit constexpr operator double() const { return v; }
    __host__ __device__ DualF &operator+=(const DualF &b) {
        v += b.v;
        d += b.d;
        return *this;
    }
    __host__ __device__ DualF &operator-=(const DualF &b) {
        v -= b.v;
        d -= b.d;
        return *this;
    }
    __host__ __device__ DualF &operator*=(const DualF &b) {
        const float nd = d * b.v + v * b.d;
        v *= b.v;
        d = nd;
        return *this;
    }
    __host__ __device__ DualF &operator/=(const DualF &b) {
        const float q = v / b.v;
        d = (d - q * b.d) / b.v;
        v = q;
        return *this;
    }
};
__host__ __device__ constexpr Dual::Dual(const DualF &x) : v(x.v), d(x.d) {}

__host__ __device__ inline DualF operator+(const DualF &a, const DualF &b) { return DualF(a.v + b.v, a.d + b.d); }
__host__ __device__ inline DualF operator-(const DualF &a, const DualF &b) { return DualF(a.v - b.v, a.d - b.d); }
__host__ __device__ inline DualF operator-(const DualF &a) { return DualF(-a.v, -a.d); }
__host__ __device__ inline DualF operator*(const DualF &a, const DualF &b) {
    return DualF(a.v * b.v, a.d * b.v + a.v * b.d);
}
__host__ __device__ inline DualF operator/(const DualF &a, const DualF &b) {
    const float q = a.v / b.v;
    return DualF(q, (a.d - q * b.d) / b.v);
}

// a x + y with one rounding of the value (fmaf); the derivative a x' + y'
__device__ inline float fmac(float a, float x, float y) { return fmaf(a, x, y); }
__device__ inline DualF fmac(float a, const DualF &x, const DualF &y) {
    return DualF(fmaf(a, x.v, y.v), fmaf(a, x.d, y.d));
}

// a b + c with one rounding of the value (fmaf); the derivative a' b + a b' + c'
__device__ inline float fma3(float a, float b, float c) { return fmaf(a, b, c); }
__device__ inline DualF fma3(const DualF &a, const DualF &b, const DualF &c) {
    return DualF(fmaf(a.v, b.v, c.v), fmaf(a.d, b.v, fmaf(a.v, b.d, c.d)));
}

// reciprocal, correctly rounded to within one ulp (v_rcp_f32); the derivative -x'/x^2
__device__ inline float recip(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline DualF recip(const DualF &x) {
    const float r = __builtin_amdgcn_rcpf(x.v);
    return DualF(r, -x.d * r * r);
}

// value part of every number type
__host__ __device__ inline double value_of(double x) { return x; }
__host__ __device__ inline double value_of(const Dual &x) { return x.v; }
__host__ __device__ inline double value_of(const DualF &x) { return x.v; }
__host__ __device__ inline double deriv_of(double) { return 0.0; }
__host__ __device__ inline double deriv_of(const Dual &x) { return x.d; }
__host__ __device__ inline double deriv_of(const DualF &x) { return x.d; }
// both parts zero
__host__ __device__ inline bool is_zero(double x) { return x == 0.0; }
__host__ __device__ inline bool is_zero(const Dual &x) { return x.v == 0.0 && x.d == 0.0; }

// host math on Duals (the FEM beam's corotational geometry); found by argument-dependent
// lookup next to std:: for doubles
inline Dual sqrt(const Dual &a) {
    const double s = std::sqrt(a.v);
    return Dual(s, a.d / (2.0 * s));
}
inline Dual cos(const Dual &a) { return Dual(std::cos(a.v), -std::sin(a.v) * a.d); }
inline Dual sin(const Dual &a) { return Dual(std::sin(a.v), std::cos(a.v) * a.d); }
inline Dual atan2(const Dual &y, const Dual &x) {
    const double r2 = x.v * x.v + y.v * y.v;
    return Dual(std::atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / r2);
}
inline Dual pow(const Dual &a, int n) {
    // integer powers as the double code calls them (std::pow(x, 2))
    const double p = std::pow(a.v, n);
    return Dual(p, n * std::pow(a.v, n - 1) * a.d);
}
inline Dual fabs(const Dual &a) { return a.v < 0.0 ? -a : a; }

}  // namespace lumagpu
