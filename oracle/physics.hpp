// ORACLE — test infrastructure only. CPU f64 restatement of the reference's
// per-body and per-joint physics: inertia, rigid-body conversions, the joint force
// elements, the solver-stage bodies of crates/functions, gravity's pair law and
// the explicit integrators.
#pragma once
#include <cmath>
#include <vector>
#include "pga3.hpp"
#include "tangent.hpp"

namespace oracle {

using M16 = Mv<double>;
using Block = std::array<std::array<double, 6>, 6>;

inline Block zero_block() { Block b; for (auto& r : b) r.fill(0.0); return b; }

// ───────────────────────── inertia (crates/newton/src/inertia.rs) ─────────────────────────
struct Inertia {
    double mass;
    double t[3][3];  // body-frame angular tensor (a diagonal one is passed as a diagonal matrix)
};

template <class S>
V3<S> matvec3(const double a[3][3], const V3<S>& v) {  // inertia.rs:11-21
    return V3<S>(S(a[0][0]) * v[0] + S(a[0][1]) * v[1] + S(a[0][2]) * v[2],
                 S(a[1][0]) * v[0] + S(a[1][1]) * v[1] + S(a[1][2]) * v[2],
                 S(a[2][0]) * v[0] + S(a[2][1]) * v[1] + S(a[2][2]) * v[2]);
}

inline V3<double> solve3(const double a[3][3], const V3<double>& b) {  // inertia.rs:23-47
    double c00 = a[1][1] * a[2][2] - a[1][2] * a[2][1];
    double c01 = a[1][2] * a[2][0] - a[1][0] * a[2][2];
    double c02 = a[1][0] * a[2][1] - a[1][1] * a[2][0];
    double c10 = a[0][2] * a[2][1] - a[0][1] * a[2][2];
    double c11 = a[0][0] * a[2][2] - a[0][2] * a[2][0];
    double c12 = a[0][1] * a[2][0] - a[0][0] * a[2][1];
    double c20 = a[0][1] * a[1][2] - a[0][2] * a[1][1];
    double c21 = a[0][2] * a[1][0] - a[0][0] * a[1][2];
    double c22 = a[0][0] * a[1][1] - a[0][1] * a[1][0];
    double det = a[0][0] * c00 + a[0][1] * c01 + a[0][2] * c02;
    return V3<double>((c00 * b[0] + c10 * b[1] + c20 * b[2]) / det,
                      (c01 * b[0] + c11 * b[1] + c21 * b[2]) / det,
                      (c02 * b[0] + c12 * b[1] + c22 * b[2]) / det);
}

// P = I·V in the body frame (inertia.rs:292-311, Full variant)
// A kinematic body (Inertia::Kinematic, inertia.rs:103-151) is one of mass 0 here: its
// momentum slot holds its velocity, apply and apply_inverse are the identity on the
// components (inertia.rs:135, 321), and the integrators leave it alone.
inline bool is_kinematic(const Inertia& in) { return in.mass == 0.0; }
inline M16 inertia_apply(const Inertia& in, const M16& twist) {
    V3<double> v = screw_linear(twist), w = screw_angular(twist);
    if (is_kinematic(in)) return screw_new(v, w);
    return screw_new(v.scale(in.mass), matvec3(in.t, w));
}
// V = I⁻¹·P (inertia.rs:313-333, Full variant)
inline M16 inertia_apply_inverse(const Inertia& in, const M16& mom) {
    V3<double> f = screw_linear(mom), t = screw_angular(mom);
    if (is_kinematic(in)) return screw_new(f, t);
    return screw_new(f.scale(1.0 / in.mass), solve3(in.t, t));
}

// ───────────────────────── rigid body (crates/newton/src/body.rs) ─────────────────────────
struct Body {
    M16 pose;       // motor, island frame
    M16 momentum;   // wrench, body frame
    Inertia inertia;
    double size;    // effective_size (m/4)^(1/3), body.rs:70-77
};

inline M16 body_velocity(const Body& b) { return inertia_apply_inverse(b.inertia, b.momentum); }          // body.rs:145-148
inline M16 world_velocity(const Body& b) { return transport_contra(body_velocity(b), b.pose); }           // body.rs:150-157
inline M16 world_momentum(const Body& b) { return transport_co(b.momentum, b.pose); }                    // body.rs:192-197
inline V3<double> body_position(const Body& b) { return point_coords(conjugate_mv(b.pose, make_point(V3<double>()))); }  // body.rs:167-173
inline M16 gyroscopic(const Body& b) { return -coad_bracket(b.momentum, body_velocity(b)); }              // body.rs:186-191
inline double effective_size(double mass) { return std::pow(mass / 4.0, 1.0 / 3.0); }                   // body.rs:70-77

// ───────────────────────── joints (crates/joints) ─────────────────────────
enum JointKind : int {
    SIMPLE_SPRING_DAMPER = 0,        // axial_spring_damper + simple_spring_damper.rs
    CRITICALLY_DAMPED_WARPED = 1,    // axial_spring_damper + critically_damped_warped.rs
    PERPENDICULAR_DAMPER_WARPED = 2, // perpendicular_damper_warped.rs
    TORSIONAL_DAMPER_WARPED = 3,     // torsional_damper_warped.rs
};

struct Joint {
    int kind;
    int a, b;            // body indices within the island (a is end 0, b end 1)
    double params[10];   // canonical order (joints/src/lib.rs:273-299)
};

inline int joint_n_params(int kind) {
    switch (kind) {
        case SIMPLE_SPRING_DAMPER: case CRITICALLY_DAMPED_WARPED: return 10;
        case PERPENDICULAR_DAMPER_WARPED: return 2;
        default: return 1;
    }
}

template <class S>
S powi2(const S& x) { return x * x; }  // powi(2): one squaring (tangent.rs:290-307; f64::powi)

// Both wrenches [on a, on b] of one joint at world poses/velocities (a pure world function).
template <class S>
void joint_wrench(const Joint& j, const Mv<S> poses[2], const Mv<S> vels[2], double dt, double warp,
                  Mv<S> out[2]) {
    const double* p = j.params;
    if (j.kind == SIMPLE_SPRING_DAMPER || j.kind == CRITICALLY_DAMPED_WARPED) {
        // axial_spring_damper/mod.rs:315-385
        V3<S> la{S(p[0]), S(p[1]), S(p[2])}, lb{S(p[3]), S(p[4]), S(p[5])};
        Mv<S> pa = conjugate_mv(poses[0], make_point(la));   // anchor_point (lib.rs:258-264)
        Mv<S> pb = conjugate_mv(poses[1], make_point(lb));
        Mv<S> line = join(pa, pb);
        S eps = S(p[9]);
        S d = sqrt_s(norm_squared(line) + eps * eps);        // soft_weight_norm (objects.rs:120-122)
        Mv<S> raw = wrench_from_line(line);
        Mv<S> unit = raw.scale(S(1.0) / d);
        S d_dot = pairing(unit, vels[0] - vels[1]);
        S rest = S(p[6]), k = S(p[7]), c = S(p[8]);
        S f;
        if (j.kind == SIMPLE_SPRING_DAMPER) {
            f = k * (d - rest) + c * d_dot;                   // mod.rs:47-61
        } else {
            // critically_damped_warped.rs:58-90: c0 = k/warp², Padé step
            S c0 = k / powi2(S(warp));
            S omega_0 = sqrt_s(c0);
            S omega_02 = c0;
            S dts = S(dt);
            f = (omega_02 * (d - rest) + (S(1.0) + S(1.0)) * omega_0 * d_dot + omega_02 * dts * d_dot) /
                powi2(S(1.0) + omega_0 * dts);
        }
        Mv<S> w = unit.scale(-f);
        out[0] = w;
        out[1] = -w;
    } else if (j.kind == PERPENDICULAR_DAMPER_WARPED) {
        // perpendicular_damper_warped.rs:107-162
        Mv<S> pa = conjugate_mv(poses[0], make_point(V3<S>()));
        Mv<S> pb = conjugate_mv(poses[1], make_point(V3<S>()));
        Mv<S> line = join(pa, pb);
        S d = sqrt_s(norm_squared(line) + S(p[1]) * S(p[1]));
        S b_eff = S(p[0]) / powi2(S(warp));
        V3<S> dir = line_direction(line);
        S inv_d = S(1.0) / d;
        V3<S> n = dir.scale(inv_d);
        V3<S> v_rel = velocity_at(vels[0] - vels[1], pa);
        S v_dot_n = v_rel.dot(n);
        V3<S> v_perp = v_rel - n.scale(v_dot_n);
        V3<S> f = v_perp.scale(-b_eff);
        V3<S> ca = point_coords(pa);
        Mv<S> tip = make_point(ca + f);
        Mv<S> wa = wrench_from_line(join(tip, pa));
        out[0] = wa;
        out[1] = -wa;
    } else {
        // torsional_damper_warped.rs:93-117
        V3<S> w_rel = screw_angular(vels[0] - vels[1]);
        S b_eff = S(p[0]) / powi2(S(warp));
        Mv<S> couple = screw_new(V3<S>(), w_rel.scale(-b_eff));
        out[0] = couple;
        out[1] = -couple;
    }
}

// Value and 24-axis Jacobian (clifford/src/pga3/ad.rs:119-150): pose i perturbed by
// the unit retraction base∘Motor::exp(δ) (δ axes 6i..6i+6), velocity additively
// (axes 12+6i..). jac[end][axis] is the wrench on `end` differentiated along `axis`.
struct JointJac {
    double value[2][6];
    double jac[2][24][6];
};

inline void joint_jacobian(const Joint& j, const M16 poses[2], const M16 vels[2], double dt, double warp,
                           JointJac& out) {
    using T = Tangent<24>;
    Mv<T> mp[2], mv[2];
    for (int i = 0; i < 2; ++i) {
        Mv<T> base, vel;
        for (int b = 0; b < DIM; ++b) { base.c[b] = T(poses[i].c[b]); vel.c[b] = T(vels[i].c[b]); }
        V3<T> sl(T::seed(0.0, 6 * i + 0), T::seed(0.0, 6 * i + 1), T::seed(0.0, 6 * i + 2));
        V3<T> sa(T::seed(0.0, 6 * i + 3), T::seed(0.0, 6 * i + 4), T::seed(0.0, 6 * i + 5));
        mp[i] = compose(base, motor_exp(screw_new(sl, sa)));
        V3<T> vl(T::seed(0.0, 12 + 6 * i + 0), T::seed(0.0, 12 + 6 * i + 1), T::seed(0.0, 12 + 6 * i + 2));
        V3<T> va(T::seed(0.0, 12 + 6 * i + 3), T::seed(0.0, 12 + 6 * i + 4), T::seed(0.0, 12 + 6 * i + 5));
        mv[i] = vel + screw_new(vl, va);
    }
    Mv<T> w[2];
    joint_wrench(j, mp, mv, dt, warp, w);
    for (int e = 0; e < 2; ++e) {
        V3<T> f = screw_linear(w[e]), t = screw_angular(w[e]);
        T comps[6] = {f[0], f[1], f[2], t[0], t[1], t[2]};
        for (int s = 0; s < 6; ++s) {
            out.value[e][s] = comps[s].v;
            for (int k = 0; k < 24; ++k) out.jac[e][k][s] = comps[s].g[k];
        }
    }
}

inline void joint_plain(const Joint& j, const M16 poses[2], const M16 vels[2], double dt, double warp,
                        double out[2][6]) {
    M16 w[2];
    joint_wrench(j, poses, vels, dt, warp, w);
    for (int e = 0; e < 2; ++e) screw_to6(w[e], out[e]);
}

// ───────────────────────── solver stage bodies (crates/functions/src/lib.rs) ─────────────────────────
// pre: (pose ∘ vel.exp(retraction), vel)  lib.rs:15-21
inline M16 fn_pre(const M16& pose, const M16& vel, double retraction) {
    return compose(pose, twist_exp(vel, retraction));
}

// body_post: world mass block, residual, self-scale (lib.rs:170-247)
struct PostOut { Block block; double rhs[6]; double scale[6]; };

inline PostOut fn_body_post(const M16& pose_mid, const M16& solve_vel, const Inertia& in, const double snap_mom[6],
                            const double total[6], double half, double floor2) {
    PostOut o;
    M16 inv = motor_inverse(pose_mid);
    for (int col = 0; col < 6; ++col) {
        M16 ej_body = transport_contra(twist_basis<double>(col), inv);
        V3<double> v = screw_linear(ej_body), w = screw_angular(ej_body);
        V3<double> l(in.t[0][0] * w[0] + in.t[0][1] * w[1] + in.t[0][2] * w[2],
                     in.t[1][0] * w[0] + in.t[1][1] * w[1] + in.t[1][2] * w[2],
                     in.t[2][0] * w[0] + in.t[2][1] * w[1] + in.t[2][2] * w[2]);
        M16 column = transport_co(screw_new(v.scale(in.mass), l), pose_mid);
        double vals[6];
        screw_to6(column, vals);
        for (int r = 0; r < 6; ++r) o.block[r][col] = vals[r];
    }
    double v[6];
    screw_to6(solve_vel, v);
    double p[6];
    for (int r = 0; r < 6; ++r) {
        double acc = 0.0;
        for (int c = 0; c < 6; ++c) acc += o.block[r][c] * v[c];
        p[r] = acc;
    }
    // hw = total·half through the wrench carrier, read back by force()/torque()
    double hh[6];
    {
        M16 hw = screw_from6(total).scale(half);
        screw_to6(hw, hh);
    }
    const double* nn = snap_mom;
    double res[6];
    for (int r = 0; r < 6; ++r) res[r] = -(p[r] - nn[r] - hh[r]);
    {
        M16 rw = screw_from6(res);
        screw_to6(rw, o.rhs);
    }
    auto sq = [](const double* x, int off) { return x[off] * x[off] + x[off + 1] * x[off + 1] + x[off + 2] * x[off + 2]; };
    double lin = std::fmax(std::fmax(std::fmax(sq(p, 0), sq(nn, 0)), sq(hh, 0)), floor2);
    double ang = std::fmax(std::fmax(std::fmax(sq(p, 3), sq(nn, 3)), sq(hh, 3)), floor2);
    for (int i = 0; i < 3; ++i) { o.scale[i] = lin; o.scale[3 + i] = ang; }
    return o;
}

// assemble_term (lib.rs:136-154): out[r][col] += pf·pose[col][r] + vf·vel[col][r]
inline void fn_assemble_term(Block& acc, const double pose[6][6], const double vel[6][6], double pf, double vf) {
    for (int col = 0; col < 6; ++col)
        for (int r = 0; r < 6; ++r) acc[r][col] = acc[r][col] + pf * pose[col][r] + vf * vel[col][r];
}

// gemm_term (lib.rs:63-89): acc + a·(sign·b + diag·2I)
inline void fn_gemm_term(Block& acc, const Block& a, const Block& b, double sign, double diag) {
    const double two = 1.0 + 1.0;
    Block right;
    for (int r = 0; r < 6; ++r) {
        for (int c = 0; c < 6; ++c) right[r][c] = sign * b[r][c];
        right[r][r] += diag * two;
    }
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double s = 0.0;
            for (int k = 0; k < 6; ++k) s += a[r][k] * right[k][c];
            acc[r][c] += s;
        }
}

// reduce_sq_term (lib.rs:104-115): acc + Σ ‖diag·I − b‖²
inline double fn_reduce_sq_term(double acc, const Block& b, double diag) {
    double out = acc;
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) {
            double e = (r == c) ? diag - b[r][c] : -b[r][c];
            out += e * e;
        }
    return out;
}

// matvec_term (lib.rs:35-55): acc + x·w
inline void fn_matvec_term(double acc[6], const Block& x, const double w[6]) {
    for (int r = 0; r < 6; ++r) {
        double s = 0.0;
        for (int c = 0; c < 6; ++c) s += x[r][c] * w[c];
        acc[r] += s;
    }
}

// ───────────────────────── gravity (crates/newton/src/gravity.rs) ─────────────────────────
struct Charge {
    double com[3];    // absolute position (S frame)
    double local[3];  // position relative to the owner's island origin (moment arm)
    double grav;      // G·m
    double size;      // effective size
};

inline Charge zero_charge() { Charge c{}; c.size = 1.0; return c; }  // gravity.rs:90-97

// pair_force: force on lo from hi before the 1/G factor; false when coincident
// (gravity.rs:110-140).
inline bool pair_force(const Charge& lo, const Charge& hi, double out[3]) {
    M16 plo = make_point(V3<double>(lo.com[0], lo.com[1], lo.com[2]));
    M16 phi = make_point(V3<double>(hi.com[0], hi.com[1], hi.com[2]));
    M16 line = join(plo, phi);
    double d = std::sqrt(norm_squared(line));
    if (is_effective_zero(d)) return false;
    double size_hi = hi.size, g_lo = lo.grav, g_hi = hi.grav;
    double g;
    if (d >= size_hi) {
        double g1 = g_lo / d, g2 = g_hi / d;
        double tmp = g1 * g2;
        g = tmp / d;
    } else {
        double g1 = g_lo / size_hi, g2 = g_hi / size_hi;
        double tmp = g1 * g2;
        g = tmp / size_hi;
    }
    V3<double> f = screw_linear(wrench_from_line(line).scale(-g));
    out[0] = f[0]; out[1] = f[1]; out[2] = f[2];
    return true;
}

// Row walk of body `me` in ascending slot order (gravity.rs:530-601): each pair's
// force is formed once on its lower slot and scaled by 1/G; the higher slot takes
// its negation. Moment about the owner's island origin: local × F.
inline void gravity_accumulate_for(int me, const std::vector<Charge>& front, double g_inv, double out[6]) {
    V3<double> acc_f, acc_t;
    M16 acc;  // accumulate as wrenches, in the reference's carrier
    const Charge& mine = front[me];
    for (int other = 0; other < (int)front.size(); ++other) {
        if (other == me) continue;
        int lo = me < other ? me : other, hi = me < other ? other : me;
        double raw[3];
        if (!pair_force(front[lo], front[hi], raw)) continue;
        double f[3] = {raw[0] * g_inv, raw[1] * g_inv, raw[2] * g_inv};
        if (me != lo) { f[0] = -f[0]; f[1] = -f[1]; f[2] = -f[2]; }
        // moment = local × f (peano cross: y·z' − z·y', …)
        const double* r = mine.local;
        double m[3] = {r[1] * f[2] - r[2] * f[1], r[2] * f[0] - r[0] * f[2], r[0] * f[1] - r[1] * f[0]};
        acc = acc + screw_new(V3<double>(f[0], f[1], f[2]), V3<double>(m[0], m[1], m[2]));
    }
    screw_to6(acc, out);
}

// ───────────────────────── explicit integrators (crates/newton/src/integrator/explicit) ─────────────────────────
// SymplecticEuler (symplectic_euler.rs:26-39); `wrench` in the body frame.
inline void symplectic_euler_step(Body& b, const M16& wrench, double dt) {
    if (is_kinematic(b.inertia)) return;   // symplectic_euler.rs:27-29
    M16 gyro = gyroscopic(b);
    b.momentum = b.momentum + (wrench + gyro).scale(dt);
    M16 v_new = body_velocity(b);
    b.pose = compose(twist_exp(v_new, dt), b.pose);
}
// ExplicitEuler (explicit_euler.rs:22-34)
inline void explicit_euler_step(Body& b, const M16& wrench, double dt) {
    if (is_kinematic(b.inertia)) return;   // explicit_euler.rs:22-24
    M16 v = body_velocity(b);
    M16 gyro = gyroscopic(b);
    b.pose = compose(twist_exp(v, dt), b.pose);
    b.momentum = b.momentum + (wrench + gyro).scale(dt);
}
// Motor::rotation_part (motor.rs:244-249): the scalar and Euclidean-bivector blades,
// normalized by √⟨M·M̃⟩₀ (motor.rs:231-242; left as is when that is effectively 0).
inline M16 rotation_part(const M16& m) {
    M16 r;
    r.c[0] = m.c[0];
    for (int b = 1; b < DIM; ++b)
        if (popc(b) == 2 && !(b & ZERO_MASK)) r.c[b] = m.c[b];
    double n2 = gp<EVEN, EVEN>(r, reverse(r)).c[0];
    if (is_effective_zero(n2)) return r;
    return r.scale(1.0 / std::sqrt(n2));
}
// LieEuler (lie_euler.rs:26-56): P += W·dt; δ = exp(dt·V(P)); P ← P transported by
// rot(δ)⁻¹; M ← δ∘M — the momentum update by conjugation with the pose increment.
inline void lie_euler_step(Body& b, const M16& wrench, double dt) {
    if (is_kinematic(b.inertia)) return;   // lie_euler.rs:32-34
    b.momentum = b.momentum + wrench.scale(dt);
    M16 delta = twist_exp(body_velocity(b), dt);
    M16 drot = rotation_part(delta);
    b.momentum = transport_co(b.momentum, motor_inverse(drot));
    b.pose = compose(delta, b.pose);
}

}  // namespace oracle
