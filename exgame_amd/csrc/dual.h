// Forward-mode dual number with ONE derivative direction: each lane of the
// Jacobian kernel carries the value and the derivative along one of the 24 axes,
// so a joint's full Jacobian is 24 lanes instead of one 25-wide carrier.
//
// The arithmetic rules are those of the reference's Tangent carrier
// (crates/clifford/src/tangent.rs:145-310): a product's derivative is
// b.d·a.v + a.d·b.v; a quotient is a·recip(b) with recip(b).d = b.d·(−v²); sqrt'
// is 1/(√a+√a). The value part is therefore computed exactly as the 24-axis
// carrier computes it, whichever axis a lane differentiates along.
#pragma once
#include <hip/hip_runtime.h>

namespace exg {

template <class T>
struct Dual {
    T v, d;
    __host__ __device__ Dual() : v(T(0)), d(T(0)) {}
    __host__ __device__ explicit Dual(T value) : v(value), d(T(0)) {}
    __host__ __device__ Dual(T value, T deriv) : v(value), d(deriv) {}
};

template <class T> __host__ __device__ inline Dual<T> operator+(const Dual<T>& a, const Dual<T>& b) { return {a.v + b.v, a.d + b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(const Dual<T>& a, const Dual<T>& b) { return {a.v - b.v, a.d - b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(const Dual<T>& a) { return {-a.v, -a.d}; }
template <class T> __host__ __device__ inline Dual<T> operator*(const Dual<T>& a, const Dual<T>& b) { return {a.v * b.v, b.d * a.v + a.d * b.v}; }
template <class T> __host__ __device__ inline Dual<T> recip(const Dual<T>& b) {
    T r = T(1) / b.v;
    return {r, b.d * (-(r * r))};
}
template <class T> __host__ __device__ inline Dual<T> operator/(const Dual<T>& a, const Dual<T>& b) { return a * recip(b); }

// Mixed dual × constant forms (a constant has no tangent): the same values and
// derivatives as promoting the constant to Dual(c, 0), without the products by zero.
template <class T> __host__ __device__ inline Dual<T> operator*(const Dual<T>& a, T b) { return {a.v * b, a.d * b}; }
template <class T> __host__ __device__ inline Dual<T> operator*(T a, const Dual<T>& b) { return {a * b.v, b.d * a}; }
template <class T> __host__ __device__ inline Dual<T> operator+(const Dual<T>& a, T b) { return {a.v + b, a.d}; }
template <class T> __host__ __device__ inline Dual<T> operator+(T a, const Dual<T>& b) { return {a + b.v, b.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(const Dual<T>& a, T b) { return {a.v - b, a.d}; }
template <class T> __host__ __device__ inline Dual<T> operator-(T a, const Dual<T>& b) { return {a - b.v, -b.d}; }

template <class T> __host__ __device__ inline T value_of(const Dual<T>& x) { return x.v; }
__host__ __device__ inline double value_of(double x) { return x; }
__host__ __device__ inline float value_of(float x) { return x; }

__host__ __device__ inline double s_sqrt(double x) { return sqrt(x); }
__host__ __device__ inline float s_sqrt(float x) { return sqrtf(x); }
// sin and cos of the half-angles of one step's rotation (|x| ≤ 1/4 in practice): the
// Maclaurin series through x¹⁵ (sin) and x¹⁶ (cos), whose first omitted terms are
// below 2⁻⁸⁰ of the result there, evaluated by Horner's rule; the library functions
// (with their range reduction) beyond 1/4.
__host__ __device__ inline double s_sin(double x) {
    if (!(fabs(x) <= 0.25)) return sin(x);
    const double z = x * x;
    double p = -1.0 / 1307674368000.0;          // −1/15!
    p = fma(p, z, 1.0 / 6227020800.0);          // 1/13!
    p = fma(p, z, -1.0 / 39916800.0);           // −1/11!
    p = fma(p, z, 1.0 / 362880.0);              // 1/9!
    p = fma(p, z, -1.0 / 5040.0);               // −1/7!
    p = fma(p, z, 1.0 / 120.0);                 // 1/5!
    p = fma(p, z, -1.0 / 6.0);                  // −1/3!
    return z == 0.0 ? x : fma(x * z, p, x);     // x itself (and its sign) when x² underflows
}
__host__ __device__ inline double s_cos(double x) {
    if (!(fabs(x) <= 0.25)) return cos(x);
    const double z = x * x;
    double p = 1.0 / 20922789888000.0;          // 1/16!
    p = fma(p, z, -1.0 / 87178291200.0);        // −1/14!
    p = fma(p, z, 1.0 / 479001600.0);           // 1/12!
    p = fma(p, z, -1.0 / 3628800.0);            // −1/10!
    p = fma(p, z, 1.0 / 40320.0);               // 1/8!
    p = fma(p, z, -1.0 / 720.0);                // −1/6!
    p = fma(p, z, 1.0 / 24.0);                  // 1/4!
    p = fma(p, z, -0.5);
    return fma(z, p, 1.0);
}
__host__ __device__ inline float s_sin(float x) { return sinf(x); }
__host__ __device__ inline float s_cos(float x) { return cosf(x); }

template <class T> __host__ __device__ inline Dual<T> s_sqrt(const Dual<T>& a) {
    T v = s_sqrt(a.v);
    return {v, a.d * (T(1) / (v + v))};
}
template <class T> __host__ __device__ inline Dual<T> s_sin(const Dual<T>& a) { return {s_sin(a.v), a.d * s_cos(a.v)}; }
template <class T> __host__ __device__ inline Dual<T> s_cos(const Dual<T>& a) { return {s_cos(a.v), a.d * (-s_sin(a.v))}; }

// Real type underneath a scalar (the seam selector compares the value).
template <class S> struct RealOf { using type = S; };
template <class T> struct RealOf<Dual<T>> { using type = T; };

}  // namespace exg
