// ORACLE — test infrastructure only. extern "C" surface of the CPU f64 restatement,
// loaded through ctypes by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg. Never linked into the product library.
//
// Screws cross this boundary as 6 doubles in constructor order (linear|force xyz,
// angular|torque xyz); motors as their 8 even-blade components in ascending blade
// order {1, e12, e13, e23, e01, e02, e03, e0123} (motor.rs:119-152).
#include <cstdint>
#include <cstring>
#include <memory>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "newton.hpp"
#include "rawmap.hpp"

using namespace oracle;

namespace {
M16 motor8(const double* c) { return motor_from8(c); }
void out8(const M16& m, double* c) { motor_to8(m, c); }
M16 screw6(const double* c) { return screw_from6(c); }
void out6(const M16& m, double* c) { screw_to6(m, c); }
}  // namespace

extern "C" {

int orc_version() { return 1; }

void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int orc_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

// ── scalar seams (peano/src/scalar.rs) ──
double orc_sinc_sq(double u) { return sinc_sq(u); }
double orc_cos_sq(double u) { return cos_sq(u); }
double orc_dsinc_sq(double u) { return dsinc_sq(u); }

// ── sign law introspection (for the convention tests) ──
void orc_blade_prod(int a, int b, int* sign, int* mask) { auto p = blade_prod(a, b); *sign = p.sign; *mask = p.mask; }
int orc_complement_sign(int a, int b) { return complement_sign(a, b); }
int orc_angular_embed_sign() { return angular_embed_sign(); }

// ── PGA primitives ──
void orc_motor_exp_twist(const double* twist6, double dt, double* motor8_out) { out8(twist_exp(screw6(twist6), dt), motor8_out); }
void orc_motor_exp_raw(const double* twist6, double* motor8_out) { out8(motor_exp(screw6(twist6)), motor8_out); }
void orc_motor_compose(const double* self8, const double* rhs8, double* out) { out8(compose(motor8(self8), motor8(rhs8)), out); }
void orc_motor_inverse(const double* m8, double* out) { out8(motor_inverse(motor8(m8)), out); }
void orc_motor_apply_point(const double* m8, const double* p3, double* out3) {
    out3[0] = 0;
    V3<double> c = point_coords(conjugate_mv(motor8(m8), make_point(V3<double>(p3[0], p3[1], p3[2]))));
    for (int i = 0; i < 3; ++i) out3[i] = c[i];
}
void orc_motor_point_raw(const double* m8, const double* p3, double* out16) {
    M16 r = conjugate_mv(motor8(m8), make_point(V3<double>(p3[0], p3[1], p3[2])));
    for (int i = 0; i < 16; ++i) out16[i] = r.c[i];
}
void orc_twist_transport(const double* t6, const double* m8, double* out6_) { out6(transport_contra(screw6(t6), motor8(m8)), out6_); }
void orc_wrench_transport(const double* w6, const double* m8, double* out6_) { out6(transport_co(screw6(w6), motor8(m8)), out6_); }
double orc_power(const double* w6, const double* t6) { return pairing(screw6(w6), screw6(t6)); }
// A screw's Pod store: its grade-2 blades in ascending order {e12, e13, e23, e01, e02,
// e03} as Twist::new / Wrench::new write them (screw.rs:240-290), and back.
void orc_screw_to_pod(const double* c6, double* pod6) {
    M16 m = screw6(c6);
    for (int k = 0; k < 6; ++k) pod6[k] = m.c[G2_BLADES[k]];
}
void orc_screw_from_pod(const double* pod6, double* c6) {
    M16 m;
    for (int k = 0; k < 6; ++k) m.c[G2_BLADES[k]] = pod6[k];
    out6(m, c6);
}
void orc_velocity_at(const double* t6, const double* p3, double* out3) {
    V3<double> v = velocity_at(screw6(t6), make_point(V3<double>(p3[0], p3[1], p3[2])));
    for (int i = 0; i < 3; ++i) out3[i] = v[i];
}
void orc_twist_bracket(const double* a6, const double* b6, double* out) { out6(twist_bracket(screw6(a6), screw6(b6)), out); }
void orc_coad_bracket(const double* w6, const double* t6, double* out) { out6(coad_bracket(screw6(w6), screw6(t6)), out); }
// join of two points: line direction, moment, and its weight norm
void orc_join_points(const double* p3, const double* q3, double* dir3, double* mom3, double* norm) {
    M16 l = join(make_point(V3<double>(p3[0], p3[1], p3[2])), make_point(V3<double>(q3[0], q3[1], q3[2])));
    V3<double> d = line_direction(l), mo = read_dual(l, radical_blade);
    for (int i = 0; i < 3; ++i) { dir3[i] = d[i]; mom3[i] = mo[i]; }
    *norm = std::sqrt(norm_squared(l));
}
// wrench of the line through p and q (Wrench::from_line)
void orc_wrench_from_points(const double* p3, const double* q3, double* out) {
    out6(wrench_from_line(join(make_point(V3<double>(p3[0], p3[1], p3[2])), make_point(V3<double>(q3[0], q3[1], q3[2])))), out);
}
// raw 16-blade vectors of Twist::new (for layout tests)
void orc_screw_raw(const double* s6, double* out16) { M16 m = screw6(s6); for (int i = 0; i < 16; ++i) out16[i] = m.c[i]; }

// ── joints (crates/joints) ──
static Joint mkjoint(int kind, const double* params) {
    Joint j{};
    j.kind = kind; j.a = 0; j.b = 1;
    for (int i = 0; i < joint_n_params(kind); ++i) j.params[i] = params[i];
    return j;
}
void orc_joint_plain(int kind, const double* params, const double* poses16, const double* vels12, double dt,
                     double warp, double* out12) {
    Joint j = mkjoint(kind, params);
    M16 p[2] = {motor8(poses16), motor8(poses16 + 8)};
    M16 v[2] = {screw6(vels12), screw6(vels12 + 6)};
    double o[2][6];
    joint_plain(j, p, v, dt, warp, o);
    std::memcpy(out12, o, sizeof(o));
}
// out_value12, out_jac: [end][axis 24][slot 6]
void orc_joint_jacobian(int kind, const double* params, const double* poses16, const double* vels12, double dt,
                        double warp, double* out_value12, double* out_jac288) {
    Joint j = mkjoint(kind, params);
    M16 p[2] = {motor8(poses16), motor8(poses16 + 8)};
    M16 v[2] = {screw6(vels12), screw6(vels12 + 6)};
    JointJac jj;
    joint_jacobian(j, p, v, dt, warp, jj);
    std::memcpy(out_value12, jj.value, sizeof(jj.value));
    std::memcpy(out_jac288, jj.jac, sizeof(jj.jac));
}

// ── solver stage bodies (crates/functions) ──
void orc_pre(const double* pose8, const double* vel6, double retraction, double* mid8) {
    out8(fn_pre(motor8(pose8), screw6(vel6), retraction), mid8);
}
void orc_body_post(const double* pose8, const double* vel6, double mass, const double* tensor9, const double* snap6,
                   const double* total6, double half, double floor2, double* block36, double* rhs6, double* scale6) {
    Inertia in{};
    in.mass = mass;
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) in.t[r][c] = tensor9[3 * r + c];
    PostOut o = fn_body_post(motor8(pose8), screw6(vel6), in, snap6, total6, half, floor2);
    for (int r = 0; r < 6; ++r) for (int c = 0; c < 6; ++c) block36[6 * r + c] = o.block[r][c];
    std::memcpy(rhs6, o.rhs, sizeof(o.rhs));
    std::memcpy(scale6, o.scale, sizeof(o.scale));
}

// ── gravity (crates/newton/src/gravity.rs) ──
// charges: n × 8 = com xyz, local xyz, grav, size
static std::vector<Charge> charges_from(int n, const double* c8) {
    std::vector<Charge> v(n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) { v[i].com[k] = c8[8 * i + k]; v[i].local[k] = c8[8 * i + 3 + k]; }
        v[i].grav = c8[8 * i + 6];
        v[i].size = c8[8 * i + 7];
    }
    return v;
}
int orc_pair_force(const double* lo8, const double* hi8, double* out3) {
    auto c = charges_from(1, lo8);
    auto d = charges_from(1, hi8);
    return pair_force(c[0], d[0], out3) ? 1 : 0;
}
void orc_gravity_accumulate(int n, const double* charges8, double G, double* out6n) {
    auto ch = charges_from(n, charges8);
    double g_inv = 1.0 / G;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < n; ++i) gravity_accumulate_for(i, ch, g_inv, out6n + 6 * i);
}

// ── the world: islands of bodies and joints plus N-body gravity ──
struct OrcWorld { World w; };

void* orc_world_new(int nb, const double* pose8, const double* mom6, const double* mass, const double* tensor9,
                    const double* size, int ni, const int* island_start, const double* origin3, int nj,
                    const int* jkind, const int* ja, const int* jb, const double* jparams10, double G, int gravity_on,
                    int integrator, double tol2, int ns_warm, int ns_seed) {
    auto* ow = new OrcWorld();
    World& w = ow->w;
    w.G = G;
    w.gravity = gravity_on != 0;
    w.integrator = integrator;
    w.prm.tol2 = tol2;
    w.prm.ns_warm = ns_warm;
    w.prm.ns_seed = ns_seed;
    std::vector<int> island_of(nb, -1);
    w.islands.resize(ni);
    for (int k = 0; k < ni; ++k) {
        Island& isl = w.islands[k];
        for (int c = 0; c < 3; ++c) isl.origin[c] = origin3[3 * k + c];
        for (int i = island_start[k]; i < island_start[k + 1]; ++i) {
            Body b;
            b.pose = motor8(pose8 + 8 * i);
            b.momentum = screw6(mom6 + 6 * i);
            b.inertia.mass = mass[i];
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) b.inertia.t[r][c] = tensor9[9 * i + 3 * r + c];
            b.size = size[i];
            isl.bodies.push_back(b);
            island_of[i] = k;
        }
    }
    for (int q = 0; q < nj; ++q) {
        int k = island_of[ja[q]];
        Island& isl = w.islands[k];
        Joint j{};
        j.kind = jkind[q];
        j.a = ja[q] - island_start[k];
        j.b = jb[q] - island_start[k];
        for (int p = 0; p < 10; ++p) j.params[p] = jparams10[10 * q + p];
        isl.joints.push_back(j);
    }
    w.init_front();
    return ow;
}
void orc_world_free(void* h) { delete static_cast<OrcWorld*>(h); }
void orc_world_step(void* h, int nsteps, double dt, double warp) {
    World& w = static_cast<OrcWorld*>(h)->w;
    for (int s = 0; s < nsteps; ++s) w.step(dt, warp);
}
// pose8/mom6 per body (island frame), origin3 per island, abs3 per body
void orc_world_get(void* h, double* pose8, double* mom6, double* origin3, double* abs3) {
    World& w = static_cast<OrcWorld*>(h)->w;
    int gi = 0;
    for (size_t k = 0; k < w.islands.size(); ++k) {
        const Island& isl = w.islands[k];
        for (int c = 0; c < 3; ++c) origin3[3 * k + c] = isl.origin[c];
        for (const Body& b : isl.bodies) {
            out8(b.pose, pose8 + 8 * gi);
            out6(b.momentum, mom6 + 6 * gi);
            V3<double> p = body_position(b);
            for (int c = 0; c < 3; ++c) abs3[3 * gi + c] = p[c] + isl.origin[c];
            ++gi;
        }
    }
}
// charges the next step will read: n × 8
void orc_world_front(void* h, double* out8n) {
    World& w = static_cast<OrcWorld*>(h)->w;
    for (size_t i = 0; i < w.front.size(); ++i) {
        const Charge& c = w.front[i];
        for (int k = 0; k < 3; ++k) { out8n[8 * i + k] = c.com[k]; out8n[8 * i + 3 + k] = c.local[k]; }
        out8n[8 * i + 6] = c.grav;
        out8n[8 * i + 7] = c.size;
    }
}
// per island: spans, newton iterations, ns iterations, reseeds, line-search rejects, unconverged commits
void orc_world_stats(void* h, int64_t* out6n) {
    World& w = static_cast<OrcWorld*>(h)->w;
    for (size_t k = 0; k < w.islands.size(); ++k) {
        const SolveStats& s = w.islands[k].stats;
        int64_t v[6] = {s.spans, s.newton_iterations, s.ns_iterations, s.reseeds, s.line_search_rejects, s.unconverged_commits};
        std::memcpy(out6n + 6 * k, v, sizeof(v));
    }
}

// ── Newton–Schulz on a given block matrix (newton.rs:209-327): transpose seed, then
// `iters` refinements R = A·X, X ← X·(2I − R) over dense block rows, then dv = X·rhs.
// a: m*m blocks of 36 (row-major blocks, row-major cells); rho_out[iters] = ‖I − A·X_k‖².
void orc_ns_solve(int m, const double* a, const double* rhs6m, int iters, double* x_out, double* dv_out,
                  double* rho_out) {
    Island isl;
    isl.bodies.resize(m);
    for (Body& b : isl.bodies) b.inertia.mass = 1.0;   // dynamic: one block row each
    cache_sync(isl);
    for (int b = 0; b < m * m; ++b)
        for (int r = 0; r < 6; ++r) for (int c = 0; c < 6; ++c) isl.A[b][r][c] = a[b * 36 + r * 6 + c];
    // dense left factor for A·X: every block column is a neighbour
    for (int i = 0; i < m; ++i) { isl.neighbours[i].clear(); for (int k = 0; k < m; ++k) isl.neighbours[i].push_back(k); }
    std::vector<std::array<double, 6>> ext(m);
    NewtonParams prm;
    NewtonCtx ctx(isl, prm, 0.0, 1.0, ext);
    ctx.seed_transpose();
    for (int k = 0; k < iters; ++k) rho_out[k] = ctx.refine_inverse();
    for (int i = 0; i < m; ++i) for (int s2 = 0; s2 < 6; ++s2) ctx.lanes[0].rhs[i][s2] = rhs6m[i * 6 + s2];
    ctx.block_matvec();
    for (int i = 0; i < m; ++i) for (int s2 = 0; s2 < 6; ++s2) dv_out[i * 6 + s2] = ctx.dv[i][s2];
    for (int b = 0; b < m * m; ++b)
        for (int r = 0; r < 6; ++r) for (int c = 0; c < 6; ++c) x_out[b * 36 + r * 6 + c] = isl.X[b][r][c];
}

// The spans of island k's last step, 4 ints each: depth, converged, committed,
// signal (NO_SIGNAL when none was read). Returns the number of spans (≤ max kept).
int orc_world_span_trace(void* h, int k, int32_t* out4, int max) {
    World& w = static_cast<OrcWorld*>(h)->w;
    const auto& tr = w.islands[k].trace;
    for (int i = 0; i < (int)tr.size() && i < max; ++i) {
        out4[4 * i] = tr[i].depth;
        out4[4 * i + 1] = tr[i].converged;
        out4[4 * i + 2] = tr[i].committed;
        out4[4 * i + 3] = tr[i].signal;
    }
    return (int)tr.size();
}

void orc_world_set_uniform(void* h, double ax, double ay, double az) {
    World& w = static_cast<OrcWorld*>(h)->w;
    w.uniform[0] = ax; w.uniform[1] = ay; w.uniform[2] = az;
}

// ── RawMap (crates/aristotle/src/world.rs) ──
void* orc_rawmap_new(uint64_t cap) { return new RawMap(cap); }
void orc_rawmap_free(void* h) { delete static_cast<RawMap*>(h); }
void orc_rawmap_insert(void* h, uint64_t value, uint32_t* index, uint32_t* generation) {
    MapKey k = static_cast<RawMap*>(h)->insert(value);
    *index = k.value;
    *generation = k.generation;
}
int orc_rawmap_remove(void* h, uint32_t index, uint32_t generation, uint64_t* value) {
    auto v = static_cast<RawMap*>(h)->remove({index, generation});
    if (!v) return 0;
    *value = *v;
    return 1;
}
int orc_rawmap_get(void* h, uint32_t index, uint32_t generation, uint64_t* value) {
    auto v = static_cast<RawMap*>(h)->get({index, generation});
    if (!v) return 0;
    *value = *v;
    return 1;
}
void orc_rawmap_info(void* h, uint64_t* capacity, uint64_t* storage_generation, uint64_t* len, uint32_t* first_free) {
    RawMap* m = static_cast<RawMap*>(h);
    *capacity = m->capacity;
    *storage_generation = m->storage_generation;
    *len = m->map.size();
    *first_free = m->first_free;
}

}  // extern "C"
