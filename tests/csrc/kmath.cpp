// Test harness: the device physics headers (exgame_amd/csrc/physics_dev.h) compiled
// for the HOST, so every closed form the kernels use can be compared with the
// oracle on the CPU. Test infrastructure only; nothing in the product loads it.
#include <cstring>
#include "../../exgame_amd/csrc/physics_dev.h"

using namespace exg;

extern "C" {

void km_twist_exp(const double* t6, double dt, double* m8) { twist_exp(t6, dt, m8); }
void km_sincos(const double* x, int n, double* s, double* c) {
    for (int i = 0; i < n; ++i) {
        s[i] = s_sin(x[i]);
        c[i] = s_cos(x[i]);
    }
}
void km_compose(const double* a8, const double* b8, double* out8) { compose(a8, b8, out8); }
void km_point(const double* m8, const double* p3, double* out4) {
    double w, q[3];
    point_sandwich(m8, p3, w, q);
    out4[0] = w;
    for (int i = 0; i < 3; ++i) out4[1 + i] = q[i];
}
void km_transport_contra(const double* t6, const double* m8, double* out6) { transport_contra(t6, m8, out6); }
void km_transport_co(const double* w6, const double* m8, double* out6) { transport_co(w6, m8, out6); }
double km_pairing(const double* w6, const double* t6) { return pairing(w6, t6); }
void km_velocity_at(const double* t6, const double* p3, double* out3) {
    velocity_at(t6, 1.0, p3, out3);
}
void km_coad_bracket(const double* w6, const double* t6, double* out6) { coad_bracket(w6, t6, out6); }

void km_joint_plain(int kind, const double* params, const double* poses16, const double* vels12, double dt,
                    double warp, double* out12) {
    joint_wrench(kind, params, poses16, poses16 + 8, vels12, vels12 + 6, dt, warp, out12, out12 + 6);
}

// The island kernel's Jacobian lanes (joint_column), one per axis.
void km_joint_jacobian(int kind, const double* params, const double* poses16, const double* vels12, double dt,
                       double warp, double* value12, double* jac288) {
    for (int ax = 0; ax < 24; ++ax) {
        double da[6], db[6];
        joint_column<double>(kind, params, poses16, poses16 + 8, vels12, vels12 + 6, ax, dt, warp, da, db,
                             ax == 0 ? value12 : nullptr, ax == 0 ? value12 + 6 : nullptr);
        for (int k = 0; k < 6; ++k) {
            jac288[(0 * 24 + ax) * 6 + k] = da[k];
            jac288[(1 * 24 + ax) * 6 + k] = db[k];
        }
    }
}

// The mixed-precision Jacobian: FP32 forward-mode lanes on the same inputs narrowed
// to f32 (the island kernel's f32 Jacobian path), widened for comparison.
void km_joint_jacobian_f32(int kind, const double* params, const double* poses16, const double* vels12, double dt,
                           double warp, double* jac288) {
    for (int ax = 0; ax < 24; ++ax) {
        float da[6], db[6];
        joint_column<float>(kind, params, poses16, poses16 + 8, vels12, vels12 + 6, ax, dt, warp, da, db, nullptr,
                            nullptr);
        for (int k = 0; k < 6; ++k) {
            jac288[(0 * 24 + ax) * 6 + k] = da[k];
            jac288[(1 * 24 + ax) * 6 + k] = db[k];
        }
    }
}

void km_body_post(const double* pose8, const double* vel6, double mass, const double* t9, const double* snap6,
                  const double* total6, double half, double floor2, double* block36, double* rhs6, double* scale2) {
    double block[6][6];
    body_post(pose8, vel6, mass, t9, snap6, total6, half, floor2, block, rhs6, scale2);
    std::memcpy(block36, block, sizeof(block));
}

void km_world_velocity(const double* pose8, double mass, const double* t9, const double* mom6, double* out6) {
    world_velocity(pose8, mass, t9, mom6, out6);
}
void km_world_momentum(const double* pose8, const double* mom6, double* out6) { world_momentum(pose8, mom6, out6); }

void km_commit(double* pose8, double* mom6, double mass, const double* t9, const double* vs6, const double* snap6,
               double span, double half, int converged) {
    commit_body(pose8, mom6, mass, t9, vs6, snap6, span, half, converged != 0);
}

}  // extern "C"
