// ORACLE — test infrastructure only. Forward-mode dual number Tangent<N> over f64,
// restating crates/clifford/src/tangent.rs:14-310 (the carrier of the reference's
// 24-axis Jacobian, crates/clifford/src/pga3/ad.rs:119-150):
//   mul   value a·b, grad b.grad·a + a.grad·b            tangent.rs:145-156
//   div   a · recip(b), recip grad = grad·(−v²)          tangent.rs:157-174, 215-226
//   sqrt  grad · 1/(√a + √a)                             tangent.rs:257-264
//   sin/cos chain rule                                    tangent.rs:265-278
// standard_part (the seam selector) is the value.
#pragma once
#include <cmath>
#include "pga3.hpp"

namespace oracle {

template <int N>
struct Tangent {
    double v;
    double g[N];
    Tangent() : v(0.0) { for (auto& x : g) x = 0.0; }
    explicit Tangent(double value) : v(value) { for (auto& x : g) x = 0.0; }
    static Tangent seed(double value, int axis) { Tangent t(value); t.g[axis] = 1.0; return t; }
    Tangent operator+(const Tangent& o) const { Tangent r; r.v = v + o.v; for (int k = 0; k < N; ++k) r.g[k] = g[k] + o.g[k]; return r; }
    Tangent operator-(const Tangent& o) const { Tangent r; r.v = v - o.v; for (int k = 0; k < N; ++k) r.g[k] = g[k] - o.g[k]; return r; }
    Tangent operator-() const { Tangent r; r.v = -v; for (int k = 0; k < N; ++k) r.g[k] = -g[k]; return r; }
    Tangent operator*(const Tangent& o) const {
        Tangent r;
        r.v = v * o.v;
        for (int k = 0; k < N; ++k) r.g[k] = o.g[k] * v + g[k] * o.v;
        return r;
    }
    Tangent recip() const {
        Tangent r;
        r.v = 1.0 / v;
        double s = -(r.v * r.v);
        for (int k = 0; k < N; ++k) r.g[k] = g[k] * s;
        return r;
    }
    Tangent operator/(const Tangent& o) const { return (*this) * o.recip(); }
};

template <int N> inline double standard_part(const Tangent<N>& x) { return x.v; }

template <int N>
Tangent<N> sqrt_t(const Tangent<N>& x) {
    Tangent<N> r;
    r.v = std::sqrt(x.v);
    double d = 1.0 / (r.v + r.v);
    for (int k = 0; k < N; ++k) r.g[k] = x.g[k] * d;
    return r;
}
template <int N>
Tangent<N> sin_t(const Tangent<N>& x) {
    Tangent<N> r;
    double d = std::cos(x.v);
    r.v = std::sin(x.v);
    for (int k = 0; k < N; ++k) r.g[k] = x.g[k] * d;
    return r;
}
template <int N>
Tangent<N> cos_t(const Tangent<N>& x) {
    Tangent<N> r;
    double d = -std::sin(x.v);
    r.v = std::cos(x.v);
    for (int k = 0; k < N; ++k) r.g[k] = x.g[k] * d;
    return r;
}

// Hook the Study-form helpers of pga3.hpp onto the dual carrier.
template <> inline Tangent<24> sqrt_s(const Tangent<24>& x) { return sqrt_t(x); }
template <> inline Tangent<24> sin_s(const Tangent<24>& x) { return sin_t(x); }
template <> inline Tangent<24> cos_s(const Tangent<24>& x) { return cos_t(x); }
template <> inline Tangent<1> sqrt_s(const Tangent<1>& x) { return sqrt_t(x); }
template <> inline Tangent<1> sin_s(const Tangent<1>& x) { return sin_t(x); }
template <> inline Tangent<1> cos_s(const Tangent<1>& x) { return cos_t(x); }

}  // namespace oracle
