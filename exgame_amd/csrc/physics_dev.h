// Per-body and per-joint physics of the dynamics hot path, written once over the
// scalar type S (double, float, Dual<double>, Dual<float>) for the HIP kernels.
//
// Semantics follow the reference's scalar-generic Rust bodies:
//   Study-form seams        crates/peano/src/scalar.rs:67-135, thresholds support.rs:38-47
//   Motor::exp / compose    crates/clifford/src/pga3/motor.rs:149-221
//   screw transport         crates/clifford/src/pga3/screw.rs:86-101
//   axial spring-damper     crates/joints/src/axial_spring_damper/mod.rs:315-385 (+ force laws)
//   perpendicular damper    crates/joints/src/perpendicular_damper_warped.rs:107-162
//   torsional damper        crates/joints/src/torsional_damper_warped.rs:93-117
//   body_post               crates/functions/src/lib.rs:170-247
//   commit_span             crates/newton/src/integrator/implicit/newton.rs:708-753
// The geometric products come from pga3_gen.h (closed forms of the same
// multivector expressions).
#pragma once
#include <hip/hip_runtime.h>
#include "dual.h"
#include "pga3_gen.h"

namespace exg {

enum JointKind : int {
    SIMPLE_SPRING_DAMPER = 0,
    CRITICALLY_DAMPED_WARPED = 1,
    PERPENDICULAR_DAMPER_WARPED = 2,
    TORSIONAL_DAMPER_WARPED = 3,
};

// EffectiveZero: 1e-9 on f64, 1e-5 on f32 (support.rs:38-47)
__host__ __device__ inline bool effective_zero(double x) { return fabs(x) < 1e-9; }
__host__ __device__ inline bool effective_zero(float x) { return fabsf(x) < 1e-5f; }

template <class S>
__host__ __device__ inline S cst(double c) {
    using R = typename RealOf<S>::type;
    return S(R(c));
}
template <class S>
__host__ __device__ inline S ratio(int num, int den) {  // from_rational: num as R / den as R
    using R = typename RealOf<S>::type;
    return S(R(num) / R(den));
}

template <class S>
__host__ __device__ inline S sinc_sq(const S& u) {
    if (effective_zero(value_of(u))) {
        S c1 = ratio<S>(1, 6), c2 = ratio<S>(1, 120), c3 = ratio<S>(1, 5040);
        return cst<S>(1.0) - u * (c1 - u * (c2 - u * c3));
    }
    S l = s_sqrt(u);
    return s_sin(l) / l;
}
template <class S>
__host__ __device__ inline S cos_sq(const S& u) {
    if (effective_zero(value_of(u))) {
        S c1 = ratio<S>(1, 2), c2 = ratio<S>(1, 24), c3 = ratio<S>(1, 720);
        return cst<S>(1.0) - u * (c1 - u * (c2 - u * c3));
    }
    return s_cos(s_sqrt(u));
}
template <class S>
__host__ __device__ inline S dsinc_sq(const S& u) {
    if (effective_zero(value_of(u * ratio<S>(1, 100000)))) {
        S c0 = ratio<S>(1, 6), c1 = ratio<S>(1, 60), c2 = ratio<S>(1, 1680), c3 = ratio<S>(1, 90720);
        return -c0 + u * (c1 - u * (c2 - u * c3));
    }
    S l = s_sqrt(u);
    return (l * s_cos(l) - s_sin(l)) / (cst<S>(2.0) * l * l * l);
}

// Motor::exp of the bivector with screw coordinates t·k (motor.rs:194-221).
template <class S>
__host__ __device__ inline void motor_exp_scaled(const S t[6], const S& k, S m[8]) {
    S tt[6];
    for (int i = 0; i < 6; ++i) tt[i] = t[i] * k;
    S u, p, bv[8], ibv[8];
    exp_parts(tt, u, p, bv, ibv);
    S cu = cos_sq(u), su = sinc_sq(u), du = dsinc_sq(u);
    S pdu = p * du;
    m[0] = cu;
    for (int i = 1; i < 7; ++i) m[i] = bv[i] * su - ibv[i] * pdu;
    m[7] = p * ratio<S>(1, 2) * su;
}
// Twist::exp(dt) = Motor::exp(T·(dt·½)) (screw.rs:254-260)
template <class S>
__host__ __device__ inline void twist_exp(const S t[6], const S& dt, S m[8]) {
    motor_exp_scaled(t, dt * ratio<S>(1, 2), m);
}
// self.compose(rhs) = rhs·self (motor.rs:149-158)
template <class S>
__host__ __device__ inline void compose(const S self[8], const S rhs[8], S out[8]) { motor_mul(rhs, self, out); }

template <class S>
__host__ __device__ inline void mat6_vec(const S a[6][6], const S x[6], S y[6]) {
    for (int r = 0; r < 6; ++r) {
        S acc = a[r][0] * x[0];
        for (int c = 1; c < 6; ++c) acc = acc + a[r][c] * x[c];
        y[r] = acc;
    }
}
// Contra (twist) and Co (wrench) transport under a motor.
template <class S>
__host__ __device__ inline void transport_contra(const S t[6], const S m[8], S out[6]) {
    S ad[6][6];
    ad_matrix(m, ad);
    mat6_vec(ad, t, out);
}
template <class S>
__host__ __device__ inline void transport_co(const S w[6], const S m[8], S out[6]) {
    S cad[6][6];
    coad_matrix(m, cad);
    mat6_vec(cad, w, out);
}

// ───────────────────────── inertia (inertia.rs:11-47, 292-333) ─────────────────────────
template <class S>
__host__ __device__ inline void inertia_apply(const S& mass, const S t9[9], const S v[6], S p[6]) {
    for (int i = 0; i < 3; ++i) p[i] = v[i] * mass;
    for (int r = 0; r < 3; ++r) p[3 + r] = t9[3 * r] * v[3] + t9[3 * r + 1] * v[4] + t9[3 * r + 2] * v[5];
}
template <class S>
__host__ __device__ inline void inertia_apply_inverse(const S& mass, const S a[9], const S p[6], S v[6]) {
    S inv = cst<S>(1.0) / mass;
    for (int i = 0; i < 3; ++i) v[i] = p[i] * inv;
    // solve3 by cofactors (inertia.rs:23-47)
    S c00 = a[4] * a[8] - a[5] * a[7];
    S c01 = a[5] * a[6] - a[3] * a[8];
    S c02 = a[3] * a[7] - a[4] * a[6];
    S c10 = a[2] * a[7] - a[1] * a[8];
    S c11 = a[0] * a[8] - a[2] * a[6];
    S c12 = a[1] * a[6] - a[0] * a[7];
    S c20 = a[1] * a[5] - a[2] * a[4];
    S c21 = a[2] * a[3] - a[0] * a[5];
    S c22 = a[0] * a[4] - a[1] * a[3];
    S det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    const S* b = p + 3;
    v[3] = (c00 * b[0] + c10 * b[1] + c20 * b[2]) / det;
    v[4] = (c01 * b[0] + c11 * b[1] + c21 * b[2]) / det;
    v[5] = (c02 * b[0] + c12 * b[1] + c22 * b[2]) / det;
}

// World twist Ad_M(I⁻¹P) and world momentum Ad*_M P (body.rs:145-197). A kinematic body
// (mass 0) has the identity for I⁻¹: its momentum slot holds its body-frame velocity
// (crates/newton/src/inertia.rs:103-151).
template <class S>
__host__ __device__ inline void world_velocity(const S pose[8], const S& mass, const S t9[9], const S mom[6], S out[6]) {
    S vb[6];
    if (mass == cst<S>(0.0))
        for (int k = 0; k < 6; ++k) vb[k] = mom[k];
    else
        inertia_apply_inverse(mass, t9, mom, vb);
    transport_contra(vb, pose, out);
}
template <class S>
__host__ __device__ inline void world_momentum(const S pose[8], const S mom[6], S out[6]) {
    transport_co(mom, pose, out);
}
// Position of the body origin: (M·O·M̃).coords()
template <class S>
__host__ __device__ inline void body_position(const S pose[8], S pos[3]) {
    S z[3] = {cst<S>(0.0), cst<S>(0.0), cst<S>(0.0)};
    S w, q[3];
    point_sandwich(pose, z, w, q);
    S winv = cst<S>(1.0) / w;
    for (int i = 0; i < 3; ++i) pos[i] = q[i] * winv;
}

// ───────────────────────── joints ─────────────────────────
// A value of the pose type as a value of the velocity type: identity when they are
// the same type, a constant (zero tangent) when the pose part carries no tangent.
template <class SV, class SP>
__host__ __device__ inline SV promote(const SP& x) {
    return SV(x);
}
template <>
__host__ __device__ inline Dual<double> promote<Dual<double>, double>(const double& x) {
    return Dual<double>(x);
}
template <>
__host__ __device__ inline Dual<float> promote<Dual<float>, float>(const float& x) {
    return Dual<float>(x);
}

// The axial spring-damper's wrenches from its two world anchor points (homogeneous
// weight and coordinates): line, soft length, unit wrench, length rate, force law
// (axial_spring_damper/mod.rs:315-385; the laws at :47-61 and
// critically_damped_warped.rs:58-90).
template <class SP, class SV, class P>
__host__ __device__ inline void axial_from_points(int kind, const P* params, const SP& wa, const SP qa[3],
                                                  const SP& wb, const SP qb[3], const SV vrel[6], const P& dt,
                                                  const P& warp, SV wa_out[6], SV wb_out[6]) {
    using RP = typename RealOf<SP>::type;
    using RV = typename RealOf<SV>::type;
    SP l[6];
    join_points(wa, qa, wb, qb, l);
    SP eps = SP(RP(params[9]));
    SP dp = s_sqrt(line_norm2(l) + eps * eps);
    SP raw[6];
    wrench_from_line(l, raw);
    SP inv_d = cst<SP>(1.0) / dp;
    SV unit[6];
    for (int i = 0; i < 6; ++i) unit[i] = promote<SV>(raw[i] * inv_d);
    SV d = promote<SV>(dp);
    SV d_dot = pairing(unit, vrel);
    SV rest = SV(RV(params[6])), k = SV(RV(params[7])), c = SV(RV(params[8]));
    SV f;
    if (kind == SIMPLE_SPRING_DAMPER) {
        f = k * (d - rest) + c * d_dot;
    } else {
        SV wp = SV(RV(warp));
        SV c0 = k / (wp * wp);
        SV omega_0 = s_sqrt(c0);
        SV dts = SV(RV(dt));
        SV den = cst<SV>(1.0) + omega_0 * dts;
        f = (c0 * (d - rest) + (cst<SV>(1.0) + cst<SV>(1.0)) * omega_0 * d_dot + c0 * dts * d_dot) / (den * den);
    }
    SV nf = -f;
    for (int i = 0; i < 6; ++i) {
        wa_out[i] = unit[i] * nf;
        wb_out[i] = -wa_out[i];
    }
}

// The perpendicular damper's wrenches from its two world anchor points (the COMs):
// soft length, unit direction, the perpendicular part of the relative velocity at a
// (perpendicular_damper_warped.rs:107-162).
template <class SP, class SV, class P>
__host__ __device__ inline void perp_from_points(const P* params, const SP& wap, const SP qap[3], const SP& wb,
                                                 const SP qb[3], const SV vrel[6], const P& warp, SV wa_out[6],
                                                 SV wb_out[6]) {
    using RP = typename RealOf<SP>::type;
    using RV = typename RealOf<SV>::type;
    SP l[6];
    join_points(wap, qap, wb, qb, l);
    SP eps = SP(RP(params[1]));
    SP dp = s_sqrt(line_norm2(l) + eps * eps);
    SV wp = SV(RV(warp));
    SV b_eff = SV(RV(params[0])) / (wp * wp);
    SP dir[3];
    line_direction(l, dir);
    SP inv_d = cst<SP>(1.0) / dp;
    SV n[3] = {promote<SV>(dir[0] * inv_d), promote<SV>(dir[1] * inv_d), promote<SV>(dir[2] * inv_d)};
    SV wa = promote<SV>(wap), qa[3] = {promote<SV>(qap[0]), promote<SV>(qap[1]), promote<SV>(qap[2])};
    SV v_rel[3];
    velocity_at(vrel, wa, qa, v_rel);
    SV v_dot_n = v_rel[0] * n[0] + (v_rel[1] * n[1] + (v_rel[2] * n[2] + cst<SV>(0.0)));
    SV nb = -b_eff;
    SV winv = cst<SV>(1.0) / wa;
    SV tip[3];
    for (int i = 0; i < 3; ++i) {
        SV v_perp = v_rel[i] - n[i] * v_dot_n;
        SV f = v_perp * nb;
        tip[i] = qa[i] * winv + f;
    }
    SV l2[6];
    join_points(cst<SV>(1.0), tip, wa, qa, l2);
    wrench_from_line(l2, wa_out);
    for (int i = 0; i < 6; ++i) wb_out[i] = -wa_out[i];
}

// Wrenches [on a, on b] at world poses and world velocities (a pure world function).
// `params` in the canonical order of each joint (joints/src/lib.rs:273-299).
// The geometry (poses) is computed in SP and the velocity-dependent part in SV; with
// SP = SV this is one type throughout. A Jacobian column along a velocity axis needs
// tangents only in SV, so its geometry runs on plain values.
// `params`, `dt` and `warp` are doubles in the kernels (P = double) and symbolic
// inputs when the law is traced for code generation.
template <class SP, class SV, class P>
__host__ __device__ inline void joint_wrench2(int kind, const P* params, const SP pa_pose[8],
                                              const SP pb_pose[8], const SV va[6], const SV vb[6], const P& dt,
                                              const P& warp, SV wa_out[6], SV wb_out[6]) {
    using RP = typename RealOf<SP>::type;
    using RV = typename RealOf<SV>::type;
    SV vrel[6];
    for (int i = 0; i < 6; ++i) vrel[i] = va[i] - vb[i];
    if (kind == SIMPLE_SPRING_DAMPER || kind == CRITICALLY_DAMPED_WARPED) {
        RP la[3] = {RP(params[0]), RP(params[1]), RP(params[2])};
        RP lb[3] = {RP(params[3]), RP(params[4]), RP(params[5])};
        SP wa, qa[3], wb, qb[3];
        point_sandwich(pa_pose, la, wa, qa);
        point_sandwich(pb_pose, lb, wb, qb);
        axial_from_points<SP, SV>(kind, params, wa, qa, wb, qb, vrel, dt, warp, wa_out, wb_out);
    } else if (kind == PERPENDICULAR_DAMPER_WARPED) {
        const RP z[3] = {RP(0.0), RP(0.0), RP(0.0)};
        SP wap, qap[3], wb, qb[3];
        point_sandwich(pa_pose, z, wap, qap);
        point_sandwich(pb_pose, z, wb, qb);
        perp_from_points<SP, SV>(params, wap, qap, wb, qb, vrel, warp, wa_out, wb_out);
    } else {
        SV wp = SV(RV(warp));
        SV b_eff = SV(RV(params[0])) / (wp * wp);
        SV nb = -b_eff;
        for (int i = 0; i < 3; ++i) {
            wa_out[i] = cst<SV>(0.0);
            wa_out[3 + i] = vrel[3 + i] * nb;
        }
        for (int i = 0; i < 6; ++i) wb_out[i] = -wa_out[i];
    }
}

template <class S, class P>
__host__ __device__ inline void joint_wrench(int kind, const P* params, const S pa_pose[8], const S pb_pose[8],
                                             const S va[6], const S vb[6], const P& dt, const P& warp, S wa_out[6],
                                             S wb_out[6]) {
    joint_wrench2<S, S, P>(kind, params, pa_pose, pb_pose, va, vb, dt, warp, wa_out, wb_out);
}

// Derivative seeds of axis `axis` (0..23) for the unit retraction base∘exp(δ):
// exp(δ) = 1 + δ at δ = 0 with a unit derivative, so the seeded motor is base with
// derivative gp(B_k, base), B_k the screw basis of the pose axis (ad.rs:119-150).
template <class T>
__host__ __device__ inline void seed_pose(const T base[8], int dof, Dual<T> out[8]) {
    T t6[6];
    for (int k = 0; k < 6; ++k) t6[k] = k == dof ? T(1) : T(0);
    T u, p, bv[8], ibv[8];
    exp_parts(t6, u, p, bv, ibv);  // bv = the basis bivector in even-store layout
    T d[8];
    motor_mul(bv, base, d);
    for (int i = 0; i < 8; ++i) out[i] = Dual<T>(base[i], d[i]);
}

// One Jacobian column of a joint (forward-mode tangents along one axis) and, on
// request, the wrench values. Axes 0-5 / 6-11 perturb the pose of end a / b (unit
// left retraction), axes 12-17 / 18-23 its velocity. In the island kernel all
// pointers address LDS.
// Inlined: an out-of-line version, storing its results to LDS through flat
// pointers from inside the call, gave wrong solves in one kernel instantiation
// (those stores were not reliably visible after the caller's barrier).
// T = float is the mixed-precision Jacobian: inputs narrowed to f32, f32 tangents.
template <class T>
__host__ __device__ __forceinline__ void joint_column(int kind, const double* params, const double* pose_a,
                                          const double* pose_b, const double* vel_a, const double* vel_b, int ax,
                                          double dt, double warp, T* da, T* db, double* va, double* vb) {
    using D = Dual<T>;
    D vv[2][6], wa[6], wb[6];
    const double* poses[2] = {pose_a, pose_b};
    const double* vels[2] = {vel_a, vel_b};
    for (int e = 0; e < 2; ++e)
        for (int k = 0; k < 6; ++k) vv[e][k] = D((T)vels[e][k], (ax >= 12 && (ax - 12) == 6 * e + k) ? T(1) : T(0));
    const bool axial = kind == SIMPLE_SPRING_DAMPER || kind == CRITICALLY_DAMPED_WARPED;
    if (ax < 12 && kind != TORSIONAL_DAMPER_WARPED) {
        // pose axis: only the perturbed end's anchor carries a tangent; the other
        // end's anchor is a plain sandwich (ends selected per lane). Anchors: the
        // joint's attachment points (axial) or the COMs (perpendicular).
        const int e = ax / 6;
        T sb[8], ob[8], sa[3], oa[3];
        for (int k = 0; k < 8; ++k) {
            sb[k] = (T)(e ? pose_b[k] : pose_a[k]);
            ob[k] = (T)(e ? pose_a[k] : pose_b[k]);
        }
        for (int k = 0; k < 3; ++k) {
            sa[k] = axial ? (T)params[e ? 3 + k : k] : T(0);
            oa[k] = axial ? (T)params[e ? k : 3 + k] : T(0);
        }
        D ps[8];
        seed_pose(sb, ax % 6, ps);
        D ws, qs[3];
        point_sandwich(ps, sa, ws, qs);
        T wo, qo[3];
        point_sandwich(ob, oa, wo, qo);
        const D wod(wo);
        D pa_w = e ? wod : ws, pb_w = e ? ws : wod, pa_q[3], pb_q[3];
        for (int k = 0; k < 3; ++k) {
            const D qk(qo[k]);
            pa_q[k] = e ? qk : qs[k];
            pb_q[k] = e ? qs[k] : qk;
        }
        D vrel[6];
        for (int k = 0; k < 6; ++k) vrel[k] = vv[0][k] - vv[1][k];
        if (axial) axial_from_points<D, D>(kind, params, pa_w, pa_q, pb_w, pb_q, vrel, dt, warp, wa, wb);
        else perp_from_points<D, D>(params, pa_w, pa_q, pb_w, pb_q, vrel, warp, wa, wb);
    } else {  // velocity axis, or a torsional damper (no pose dependence): plain geometry
        T pz[2][8];
        for (int e = 0; e < 2; ++e)
            for (int k = 0; k < 8; ++k) pz[e][k] = (T)poses[e][k];
        joint_wrench2<T, D>(kind, params, pz[0], pz[1], vv[0], vv[1], dt, warp, wa, wb);
    }
    for (int k = 0; k < 6; ++k) {
        da[k] = wa[k].d;
        db[k] = wb[k].d;
    }
    if (va)
        for (int k = 0; k < 6; ++k) {
            va[k] = wa[k].v;
            vb[k] = wb[k].v;
        }
}

// ───────────────────────── body_post (functions/src/lib.rs:170-247) ─────────────────────────
// World mass block Ad*_M ∘ I ∘ Ad_{M⁻¹}, the residual −(𝕀V − P_n − half·W) and its
// self-scale (lin², ang²).
template <class S>
__host__ __device__ inline void body_post(const S pose_mid[8], const S vel[6], const S& mass, const S t9[9],
                                          const S snap[6], const S total[6], const S& half, const S& floor2,
                                          S block[6][6], S rhs[6], S scale2[2]) {
    S inv[8];
    motor_rev(pose_mid, inv);
    S adi[6][6], cad[6][6];
    ad_matrix(inv, adi);
    coad_matrix(pose_mid, cad);
    for (int col = 0; col < 6; ++col) {
        S x[6];
        for (int i = 0; i < 3; ++i) x[i] = adi[i][col] * mass;
        for (int r = 0; r < 3; ++r)
            x[3 + r] = t9[3 * r] * adi[3][col] + t9[3 * r + 1] * adi[4][col] + t9[3 * r + 2] * adi[5][col];
        for (int r = 0; r < 6; ++r) {
            S acc = cad[r][0] * x[0];
            for (int k = 1; k < 6; ++k) acc = acc + cad[r][k] * x[k];
            block[r][col] = acc;
        }
    }
    S p[6], hh[6];
    for (int r = 0; r < 6; ++r) {
        S acc = cst<S>(0.0);
        for (int c = 0; c < 6; ++c) acc = acc + block[r][c] * vel[c];
        p[r] = acc;
        hh[r] = total[r] * half;
        rhs[r] = -(p[r] - snap[r] - hh[r]);
    }
    for (int o = 0; o < 2; ++o) {
        int b = 3 * o;
        S sp_ = p[b] * p[b] + p[b + 1] * p[b + 1] + p[b + 2] * p[b + 2];
        S sn = snap[b] * snap[b] + snap[b + 1] * snap[b + 1] + snap[b + 2] * snap[b + 2];
        S sh = hh[b] * hh[b] + hh[b + 1] * hh[b + 1] + hh[b + 2] * hh[b + 2];
        scale2[o] = fmax(fmax(fmax(sp_, sn), sh), floor2);
    }
}

// commit_span for one body (newton.rs:708-753): pose for the full span by left
// composition, momentum by the midpoint rule (2·P_mid − P_n when converged, else
// the implicit-Euler reading P_mid), back to the body frame.
template <class S>
__host__ __device__ __forceinline__ void commit_body(S pose[8], S mom[6], const S& mass, const S t9[9], const S vs[6],
                                            const S snap[6], const S& span, const S& half, bool converged) {
    S e[8], pose_mid[8], inv[8];
    twist_exp(vs, half, e);
    compose(pose, e, pose_mid);
    motor_rev(pose_mid, inv);
    S v_body[6], pb[6], p_mid[6];
    transport_contra(vs, inv, v_body);
    inertia_apply(mass, t9, v_body, pb);
    transport_co(pb, pose_mid, p_mid);
    S p_new[6];
    for (int i = 0; i < 6; ++i) p_new[i] = converged ? p_mid[i] * cst<S>(2.0) - snap[i] : p_mid[i];
    S e2[8], m_new[8], inv2[8];
    twist_exp(vs, span, e2);
    compose(pose, e2, m_new);
    motor_rev(m_new, inv2);
    S mb[6];
    transport_co(p_new, inv2, mb);
    for (int i = 0; i < 8; ++i) pose[i] = m_new[i];
    for (int i = 0; i < 6; ++i) mom[i] = mb[i];
}

}  // namespace exg
