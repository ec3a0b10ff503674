// Explicit integrators through the reference's bridge, one thread per island:
// joints evaluated at the current pose and world velocity (retraction 0, value
// only), gathered onto the external wrench, pulled back to the body frame by the
// coadjoint transport, then SymplecticEuler, ExplicitEuler or LieEuler per body.
//   bridge::explicit_step_all            crates/newton/src/integrator/implicit/bridge.rs:19-63
//   SymplecticEuler::step                crates/newton/src/integrator/explicit/symplectic_euler.rs:26-39
//   ExplicitEuler::step                  crates/newton/src/integrator/explicit/explicit_euler.rs:22-34
//   LieEuler::step                       crates/newton/src/integrator/explicit/lie_euler.rs:26-56
//   Motor::rotation_part / normalized    crates/clifford/src/pga3/motor.rs:231-249
//   RigidBody::gyroscopic                crates/newton/src/body.rs:186-191
// followed by the same re-anchoring and charge publication as the implicit path.
#include <hip/hip_runtime.h>
#include "island.h"
#include "physics_dev.h"

namespace exg {

constexpr int EXPL_THREADS = 64;
constexpr int EXPL_MAXM = 8;
constexpr int EXPL_MAXJ = 16;

// The rotational part of a motor: its scalar and Euclidean-bivector components
// (1, e12, e13, e23), divided by √(s² + b12² + b13² + b23²) = √⟨M·M̃⟩₀ unless that
// is effectively zero.
__device__ void rotation_part(const double m[8], double r[8]) {
    for (int k = 0; k < 4; ++k) r[k] = m[k];
    for (int k = 4; k < 8; ++k) r[k] = 0.0;
    const double n2 = ((m[0] * m[0] + m[1] * m[1]) + m[2] * m[2]) + m[3] * m[3];
    if (effective_zero(n2)) return;
    const double inv = 1.0 / sqrt(n2);
    for (int k = 0; k < 4; ++k) r[k] = r[k] * inv;
}

__device__ void explicit_body(double pose[8], double mom[6], double mass, const double t9[9], const double w_body[6],
                              double dt, int integrator) {
    if (integrator == 3) {  // Lie: momentum by the wrench, then conjugated by the pose increment's rotation
        for (int k = 0; k < 6; ++k) mom[k] = mom[k] + w_body[k] * dt;
        double vn[6], delta[8], drot[8], dinv[8], pm[6], np[8];
        inertia_apply_inverse(mass, t9, mom, vn);
        twist_exp(vn, dt, delta);
        rotation_part(delta, drot);
        motor_rev(drot, dinv);
        transport_co(mom, dinv, pm);
        for (int k = 0; k < 6; ++k) mom[k] = pm[k];
        compose(delta, pose, np);
        for (int k = 0; k < 8; ++k) pose[k] = np[k];
        return;
    }
    double v[6], gyro[6], cb[6];
    inertia_apply_inverse(mass, t9, mom, v);
    coad_bracket(mom, v, cb);
    for (int k = 0; k < 6; ++k) gyro[k] = -cb[k];
    if (integrator == 1) {  // symplectic: momentum first, then the pose with the new velocity
        for (int k = 0; k < 6; ++k) mom[k] = mom[k] + (w_body[k] + gyro[k]) * dt;
        double vn[6], e[8], np[8];
        inertia_apply_inverse(mass, t9, mom, vn);
        twist_exp(vn, dt, e);
        compose(e, pose, np);  // v_new.exp(dt).compose(pose) = pose·exp
        for (int k = 0; k < 8; ++k) pose[k] = np[k];
    } else {  // explicit: the pose with the old velocity
        double e[8], np[8];
        twist_exp(v, dt, e);
        compose(e, pose, np);
        for (int k = 0; k < 8; ++k) pose[k] = np[k];
        for (int k = 0; k < 6; ++k) mom[k] = mom[k] + (w_body[k] + gyro[k]) * dt;
    }
}

__global__ __launch_bounds__(EXPL_THREADS) void explicit_step_kernel(IslandArgs a, int integrator) {
    const int isl = blockIdx.x * EXPL_THREADS + threadIdx.x;
    if (isl >= a.n_islands) return;
    const int b0 = a.island_start[isl], m = a.island_end[isl] - b0;
    const int j0 = a.island_jstart[isl], J = a.island_jend[isl] - j0;
    if (m > EXPL_MAXM || J > EXPL_MAXJ || m <= 0) return;  // the host routes islands by size; never taken
    double pose[EXPL_MAXM][8], mom[EXPL_MAXM][6], wv[EXPL_MAXM][6], total[EXPL_MAXM][6];
    for (int i = 0; i < m; ++i) {
        for (int k = 0; k < 8; ++k) pose[i][k] = a.pose[(size_t)(b0 + i) * 8 + k];
        for (int k = 0; k < 6; ++k) {
            mom[i][k] = a.mom[(size_t)(b0 + i) * 6 + k];
            total[i][k] = a.grav_tiles ? 0.0 : field_add(a, a.mass[b0 + i], k, a.ext[(size_t)(b0 + i) * 6 + k]);
        }
        world_velocity(pose[i], a.mass[b0 + i], a.tens + (size_t)(b0 + i) * 9, mom[i], wv[i]);
        if (a.grav_tiles) {
            // the gravity pass's partial sums added here in place of its reduce kernel:
            // the tile's chunks in ascending order and the same wrench arithmetic (the
            // same bits); sixteen chunks' loads in flight at a time
            const int gi = b0 + i;
            const GravTile& tl = a.grav_tiles[gi / GRAV_TILE];
            const double* p = a.grav_partial + (size_t)tl.e0 * GRAV_PARTIAL_PER_ENTRY + (gi - tl.tgt0);
            double f[3] = {0.0, 0.0, 0.0};
            for (int c0 = 0; c0 < tl.ne; c0 += 16) {
                double v[16][3];
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        v[u][k] = c0 + u < tl.ne ? p[(size_t)(c0 + u) * GRAV_PARTIAL_PER_ENTRY + k * GRAV_TILE] : 0.0;
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    if (c0 + u < tl.ne)
#pragma unroll
                        for (int k = 0; k < 3; ++k) f[k] += v[u][k];
            }
            gravity_wrench(a.grav_charges + (size_t)gi * 8, tl.g_inv, f, total[i]);
            for (int k = 0; k < 6; ++k) total[i][k] = field_add(a, a.mass[gi], k, total[i][k]);
        }
    }
    for (int q = 0; q < J; ++q) {
        int ia = a.ja[j0 + q] - b0, ib = a.jb[j0 + q] - b0;
        double wa[6], wb[6];
        joint_wrench(a.jkind[j0 + q], a.jparams + (size_t)(j0 + q) * 10, pose[ia], pose[ib], wv[ia], wv[ib], a.dt,
                     a.warp, wa, wb);
        for (int k = 0; k < 6; ++k) {
            total[ia][k] = total[ia][k] + wa[k];
            total[ib][k] = total[ib][k] + wb[k];
        }
    }
    for (int i = 0; i < m; ++i) {
        // a kinematic body is a prescribed reference: the integrators leave it
        // (symplectic_euler.rs:27-29, explicit_euler.rs:22-24, lie_euler.rs:32-34)
        if (a.mass[b0 + i] == 0.0) continue;
        double inv[8], wbody[6];
        motor_rev(pose[i], inv);
        transport_co(total[i], inv, wbody);
        explicit_body(pose[i], mom[i], a.mass[b0 + i], a.tens + (size_t)(b0 + i) * 9, wbody, a.dt, integrator);
    }
    // recompute_centroid + publish (island.rs:206-244, gravity.rs:485-496)
    double total_m = 0.0, wx[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < m; ++i) {
        double p[3];
        body_position(pose[i], p);
        total_m += a.mass[b0 + i];
        for (int k = 0; k < 3; ++k) wx[k] += p[k] * a.mass[b0 + i];
    }
    double org[3] = {a.origin[isl * 3], a.origin[isl * 3 + 1], a.origin[isl * 3 + 2]};
    if (!effective_zero(total_m)) {
        double inv = 1.0 / total_m;
        double sh[6] = {-(wx[0] * inv), -(wx[1] * inv), -(wx[2] * inv), 0.0, 0.0, 0.0};
        double shift[8];
        twist_exp(sh, 1.0, shift);
        for (int i = 0; i < m; ++i) {
            double np[8];
            compose(pose[i], shift, np);
            for (int k = 0; k < 8; ++k) pose[i][k] = np[k];
        }
        for (int k = 0; k < 3; ++k) {
            org[k] += wx[k] * inv;
            a.origin[isl * 3 + k] = org[k];
        }
    }
    for (int i = 0; i < m; ++i) {
        double p[3];
        body_position(pose[i], p);
        double* c = a.charges_out + (size_t)(b0 + i) * 8;
        for (int k = 0; k < 3; ++k) {
            c[k] = p[k] + org[k];
            c[3 + k] = p[k];
        }
        c[6] = a.G * a.mass[b0 + i];
        c[7] = a.size[b0 + i];
        for (int k = 0; k < 8; ++k) a.pose[(size_t)(b0 + i) * 8 + k] = pose[i][k];
        for (int k = 0; k < 6; ++k) a.mom[(size_t)(b0 + i) * 6 + k] = mom[i][k];
    }
    bool bad = false;
    for (int i = 0; i < m; ++i) {
        for (int k = 0; k < 8; ++k) bad |= !isfinite(pose[i][k]);
        for (int k = 0; k < 6; ++k) bad |= !isfinite(mom[i][k]);
    }
    if (bad) mark_fatal(a, isl);
}

void launch_explicit_step(const IslandArgs& a, int integrator, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    int blocks = (a.n_islands + EXPL_THREADS - 1) / EXPL_THREADS;
    launch_k(explicit_step_kernel, dim3(blocks), dim3(EXPL_THREADS), stream, start, stop, a, integrator);
}

}  // namespace exg
