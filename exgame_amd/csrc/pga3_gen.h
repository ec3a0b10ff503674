// GENERATED by gen_pga3.py — do not edit by hand.
// Closed-form PGA3 operations (exact polynomial expansions of the reference's
// multivector expressions) for the device kernels, templated on the scalar.
#pragma once
#include "dual.h"

namespace exg {

template <class S>
__host__ __device__ inline void motor_mul(const S a[8], const S b[8], S c[8]) {
    c[0] = a[0]*b[0] - a[1]*b[1] - a[2]*b[2] - a[3]*b[3];
    c[1] = a[0]*b[1] + a[1]*b[0] + a[3]*b[2] - a[2]*b[3];
    c[2] = a[0]*b[2] + a[1]*b[3] + a[2]*b[0] - a[3]*b[1];
    c[3] = a[0]*b[3] + a[2]*b[1] + a[3]*b[0] - a[1]*b[2];
    c[4] = a[0]*b[4] + a[1]*b[5] + a[2]*b[6] + a[4]*b[0] - a[3]*b[7] - a[5]*b[1] - a[6]*b[2] - a[7]*b[3];
    c[5] = a[0]*b[5] + a[2]*b[7] + a[3]*b[6] + a[4]*b[1] + a[5]*b[0] + a[7]*b[2] - a[1]*b[4] - a[6]*b[3];
    c[6] = a[0]*b[6] + a[4]*b[2] + a[5]*b[3] + a[6]*b[0] - a[1]*b[7] - a[2]*b[4] - a[3]*b[5] - a[7]*b[1];
    c[7] = a[0]*b[7] + a[1]*b[6] + a[3]*b[4] + a[4]*b[3] + a[6]*b[1] + a[7]*b[0] - a[2]*b[5] - a[5]*b[2];
}

template <class S>
__host__ __device__ inline void motor_rev(const S a[8], S r[8]) {
    r[0] = a[0];
    r[1] = -a[1];
    r[2] = -a[2];
    r[3] = -a[3];
    r[4] = -a[4];
    r[5] = -a[5];
    r[6] = -a[6];
    r[7] = a[7];
}

template <class S>
__host__ __device__ inline void ad_matrix(const S m[8], S ad[6][6]) {
    ad[0][0] = m[0]*m[0] + m[3]*m[3] - m[1]*m[1] - m[2]*m[2];
    ad[1][0] = -S(2.0)*m[0]*m[1] - S(2.0)*m[2]*m[3];
    ad[2][0] = -S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    ad[3][0] = S(0.0);
    ad[4][0] = S(0.0);
    ad[5][0] = S(0.0);
    ad[0][1] = -S(2.0)*m[2]*m[3] + S(2.0)*m[0]*m[1];
    ad[1][1] = m[0]*m[0] + m[2]*m[2] - m[1]*m[1] - m[3]*m[3];
    ad[2][1] = -S(2.0)*m[0]*m[3] - S(2.0)*m[1]*m[2];
    ad[3][1] = S(0.0);
    ad[4][1] = S(0.0);
    ad[5][1] = S(0.0);
    ad[0][2] = S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    ad[1][2] = -S(2.0)*m[1]*m[2] + S(2.0)*m[0]*m[3];
    ad[2][2] = m[0]*m[0] + m[1]*m[1] - m[2]*m[2] - m[3]*m[3];
    ad[3][2] = S(0.0);
    ad[4][2] = S(0.0);
    ad[5][2] = S(0.0);
    ad[0][3] = -S(2.0)*m[0]*m[7] - S(2.0)*m[1]*m[6] + S(2.0)*m[2]*m[5] + S(2.0)*m[3]*m[4];
    ad[1][3] = -S(2.0)*m[0]*m[6] - S(2.0)*m[2]*m[4] + S(2.0)*m[1]*m[7] + S(2.0)*m[3]*m[5];
    ad[2][3] = S(2.0)*m[0]*m[5] + S(2.0)*m[1]*m[4] + S(2.0)*m[2]*m[7] + S(2.0)*m[3]*m[6];
    ad[3][3] = m[0]*m[0] + m[3]*m[3] - m[1]*m[1] - m[2]*m[2];
    ad[4][3] = -S(2.0)*m[0]*m[1] - S(2.0)*m[2]*m[3];
    ad[5][3] = -S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    ad[0][4] = -S(2.0)*m[1]*m[7] - S(2.0)*m[2]*m[4] + S(2.0)*m[0]*m[6] + S(2.0)*m[3]*m[5];
    ad[1][4] = -S(2.0)*m[0]*m[7] - S(2.0)*m[1]*m[6] - S(2.0)*m[2]*m[5] - S(2.0)*m[3]*m[4];
    ad[2][4] = -S(2.0)*m[0]*m[4] - S(2.0)*m[2]*m[6] + S(2.0)*m[1]*m[5] + S(2.0)*m[3]*m[7];
    ad[3][4] = -S(2.0)*m[2]*m[3] + S(2.0)*m[0]*m[1];
    ad[4][4] = m[0]*m[0] + m[2]*m[2] - m[1]*m[1] - m[3]*m[3];
    ad[5][4] = -S(2.0)*m[0]*m[3] - S(2.0)*m[1]*m[2];
    ad[0][5] = -S(2.0)*m[0]*m[5] - S(2.0)*m[2]*m[7] + S(2.0)*m[1]*m[4] + S(2.0)*m[3]*m[6];
    ad[1][5] = -S(2.0)*m[2]*m[6] - S(2.0)*m[3]*m[7] + S(2.0)*m[0]*m[4] + S(2.0)*m[1]*m[5];
    ad[2][5] = -S(2.0)*m[0]*m[7] - S(2.0)*m[3]*m[4] + S(2.0)*m[1]*m[6] + S(2.0)*m[2]*m[5];
    ad[3][5] = S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    ad[4][5] = -S(2.0)*m[1]*m[2] + S(2.0)*m[0]*m[3];
    ad[5][5] = m[0]*m[0] + m[1]*m[1] - m[2]*m[2] - m[3]*m[3];
}

template <class S>
__host__ __device__ inline void coad_matrix(const S m[8], S cad[6][6]) {
    cad[0][0] = m[0]*m[0] + m[3]*m[3] - m[1]*m[1] - m[2]*m[2];
    cad[1][0] = -S(2.0)*m[0]*m[1] - S(2.0)*m[2]*m[3];
    cad[2][0] = -S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    cad[3][0] = -S(2.0)*m[0]*m[7] - S(2.0)*m[1]*m[6] + S(2.0)*m[2]*m[5] + S(2.0)*m[3]*m[4];
    cad[4][0] = -S(2.0)*m[0]*m[6] - S(2.0)*m[2]*m[4] + S(2.0)*m[1]*m[7] + S(2.0)*m[3]*m[5];
    cad[5][0] = S(2.0)*m[0]*m[5] + S(2.0)*m[1]*m[4] + S(2.0)*m[2]*m[7] + S(2.0)*m[3]*m[6];
    cad[0][1] = -S(2.0)*m[2]*m[3] + S(2.0)*m[0]*m[1];
    cad[1][1] = m[0]*m[0] + m[2]*m[2] - m[1]*m[1] - m[3]*m[3];
    cad[2][1] = -S(2.0)*m[0]*m[3] - S(2.0)*m[1]*m[2];
    cad[3][1] = -S(2.0)*m[1]*m[7] - S(2.0)*m[2]*m[4] + S(2.0)*m[0]*m[6] + S(2.0)*m[3]*m[5];
    cad[4][1] = -S(2.0)*m[0]*m[7] - S(2.0)*m[1]*m[6] - S(2.0)*m[2]*m[5] - S(2.0)*m[3]*m[4];
    cad[5][1] = -S(2.0)*m[0]*m[4] - S(2.0)*m[2]*m[6] + S(2.0)*m[1]*m[5] + S(2.0)*m[3]*m[7];
    cad[0][2] = S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    cad[1][2] = -S(2.0)*m[1]*m[2] + S(2.0)*m[0]*m[3];
    cad[2][2] = m[0]*m[0] + m[1]*m[1] - m[2]*m[2] - m[3]*m[3];
    cad[3][2] = -S(2.0)*m[0]*m[5] - S(2.0)*m[2]*m[7] + S(2.0)*m[1]*m[4] + S(2.0)*m[3]*m[6];
    cad[4][2] = -S(2.0)*m[2]*m[6] - S(2.0)*m[3]*m[7] + S(2.0)*m[0]*m[4] + S(2.0)*m[1]*m[5];
    cad[5][2] = -S(2.0)*m[0]*m[7] - S(2.0)*m[3]*m[4] + S(2.0)*m[1]*m[6] + S(2.0)*m[2]*m[5];
    cad[0][3] = S(0.0);
    cad[1][3] = S(0.0);
    cad[2][3] = S(0.0);
    cad[3][3] = m[0]*m[0] + m[3]*m[3] - m[1]*m[1] - m[2]*m[2];
    cad[4][3] = -S(2.0)*m[0]*m[1] - S(2.0)*m[2]*m[3];
    cad[5][3] = -S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    cad[0][4] = S(0.0);
    cad[1][4] = S(0.0);
    cad[2][4] = S(0.0);
    cad[3][4] = -S(2.0)*m[2]*m[3] + S(2.0)*m[0]*m[1];
    cad[4][4] = m[0]*m[0] + m[2]*m[2] - m[1]*m[1] - m[3]*m[3];
    cad[5][4] = -S(2.0)*m[0]*m[3] - S(2.0)*m[1]*m[2];
    cad[0][5] = S(0.0);
    cad[1][5] = S(0.0);
    cad[2][5] = S(0.0);
    cad[3][5] = S(2.0)*m[0]*m[2] + S(2.0)*m[1]*m[3];
    cad[4][5] = -S(2.0)*m[1]*m[2] + S(2.0)*m[0]*m[3];
    cad[5][5] = m[0]*m[0] + m[1]*m[1] - m[2]*m[2] - m[3]*m[3];
}

// The point p may be of a plain type P while the motor carries tangents (S = Dual):
// its terms then multiply by a constant.
template <class S, class P = S>
__host__ __device__ inline void point_sandwich(const S m[8], const P p[3], S& w, S q[3]) {
    using R = typename RealOf<S>::type;
    w = m[0]*m[0] + m[1]*m[1] + m[2]*m[2] + m[3]*m[3];
    q[0] = p[0]*m[0]*m[0] + p[0]*m[3]*m[3] - p[0]*m[1]*m[1] - p[0]*m[2]*m[2] - R(2.0)*m[0]*m[4] - R(2.0)*m[1]*m[5] - R(2.0)*m[2]*m[6] - R(2.0)*m[3]*m[7] - R(2.0)*m[2]*m[3]*p[1] + R(2.0)*m[0]*m[1]*p[1] + R(2.0)*m[0]*m[2]*p[2] + R(2.0)*m[1]*m[3]*p[2];
    q[1] = p[1]*m[0]*m[0] + p[1]*m[2]*m[2] - p[1]*m[1]*m[1] - p[1]*m[3]*m[3] - R(2.0)*m[0]*m[5] - R(2.0)*m[3]*m[6] + R(2.0)*m[1]*m[4] + R(2.0)*m[2]*m[7] - R(2.0)*m[0]*m[1]*p[0] - R(2.0)*m[1]*m[2]*p[2] - R(2.0)*m[2]*m[3]*p[0] + R(2.0)*m[0]*m[3]*p[2];
    q[2] = p[2]*m[0]*m[0] + p[2]*m[1]*m[1] - p[2]*m[2]*m[2] - p[2]*m[3]*m[3] - R(2.0)*m[0]*m[6] - R(2.0)*m[1]*m[7] + R(2.0)*m[2]*m[4] + R(2.0)*m[3]*m[5] - R(2.0)*m[0]*m[2]*p[0] - R(2.0)*m[0]*m[3]*p[1] - R(2.0)*m[1]*m[2]*p[1] + R(2.0)*m[1]*m[3]*p[0];
}

template <class S>
__host__ __device__ inline void join_points(const S& wa, const S qa[3], const S& wb, const S qb[3], S l[6]) {
    l[0] = qb[2]*wa - qa[2]*wb;
    l[1] = qa[1]*wb - qb[1]*wa;
    l[2] = qb[0]*wa - qa[0]*wb;
    l[3] = qa[1]*qb[2] - qa[2]*qb[1];
    l[4] = qa[2]*qb[0] - qa[0]*qb[2];
    l[5] = qa[0]*qb[1] - qa[1]*qb[0];
}

template <class S>
__host__ __device__ inline S line_norm2(const S l[6]) {
    return l[0]*l[0] + l[1]*l[1] + l[2]*l[2];
}

template <class S>
__host__ __device__ inline void wrench_from_line(const S l[6], S w[6]) {
    w[0] = -l[2];
    w[1] = l[1];
    w[2] = -l[0];
    w[3] = -l[3];
    w[4] = -l[4];
    w[5] = -l[5];
}

template <class S>
__host__ __device__ inline void line_direction(const S l[6], S d[3]) {
    d[0] = l[2];
    d[1] = -l[1];
    d[2] = l[0];
}

template <class S>
__host__ __device__ inline S pairing(const S w[6], const S t[6]) {
    return t[0]*w[0] + t[1]*w[1] + t[2]*w[2] + t[3]*w[3] + t[4]*w[4] + t[5]*w[5];
}

template <class S>
__host__ __device__ inline void velocity_at(const S t[6], const S& pw, const S pq[3], S v[3]) {
    v[0] = (-S(2.0)*pq[1]*t[5] + S(2.0)*pq[2]*t[4] + S(2.0)*pw*t[0]) * S(0.5);
    v[1] = (-S(2.0)*pq[2]*t[3] + S(2.0)*pq[0]*t[5] + S(2.0)*pw*t[1]) * S(0.5);
    v[2] = (-S(2.0)*pq[0]*t[4] + S(2.0)*pq[1]*t[3] + S(2.0)*pw*t[2]) * S(0.5);
}

template <class S>
__host__ __device__ inline void coad_bracket(const S w[6], const S t[6], S o[6]) {
    o[0] = (-S(2.0)*t[5]*w[1] + S(2.0)*t[4]*w[2]) * S(0.5);
    o[1] = (-S(2.0)*t[3]*w[2] + S(2.0)*t[5]*w[0]) * S(0.5);
    o[2] = (-S(2.0)*t[4]*w[0] + S(2.0)*t[3]*w[1]) * S(0.5);
    o[3] = (-S(2.0)*t[2]*w[1] - S(2.0)*t[5]*w[4] + S(2.0)*t[1]*w[2] + S(2.0)*t[4]*w[5]) * S(0.5);
    o[4] = (-S(2.0)*t[0]*w[2] - S(2.0)*t[3]*w[5] + S(2.0)*t[2]*w[0] + S(2.0)*t[5]*w[3]) * S(0.5);
    o[5] = (-S(2.0)*t[1]*w[0] - S(2.0)*t[4]*w[3] + S(2.0)*t[0]*w[1] + S(2.0)*t[3]*w[4]) * S(0.5);
}

template <class S>
__host__ __device__ inline void exp_parts(const S t[6], S& u, S& p, S bv[8], S ibv[8]) {
    u = t[3]*t[3] + t[4]*t[4] + t[5]*t[5];
    p = S(2.0)*t[0]*t[3] + S(2.0)*t[1]*t[4] + S(2.0)*t[2]*t[5];
    bv[0] = S(0.0);
    ibv[0] = S(0.0);
    bv[1] = -t[5];
    ibv[1] = S(0.0);
    bv[2] = t[4];
    ibv[2] = S(0.0);
    bv[3] = -t[3];
    ibv[3] = S(0.0);
    bv[4] = -t[0];
    ibv[4] = t[3];
    bv[5] = -t[1];
    ibv[5] = t[4];
    bv[6] = -t[2];
    ibv[6] = t[5];
    bv[7] = S(0.0);
    ibv[7] = S(0.0);
}

template <class S>
__host__ __device__ inline void screw_to_pod(const S c[6], S pod[6]) {
    pod[0] = -c[5];
    pod[1] = c[4];
    pod[2] = -c[3];
    pod[3] = -c[0];
    pod[4] = -c[1];
    pod[5] = -c[2];
}

template <class S>
__host__ __device__ inline void screw_from_pod(const S q[6], S c[6]) {
    c[0] = -q[3];
    c[1] = -q[4];
    c[2] = -q[5];
    c[3] = -q[2];
    c[4] = q[1];
    c[5] = -q[0];
}

}  // namespace exg
