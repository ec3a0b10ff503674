// The joint (K1–K8) and Pre (K9) row kernels on the traced code: each row runs the
// device functions that exgame_amd/viete/gen.cpp emitted from the same templates the
// hand-written kernels instantiate (viete_gen.h), in the reference's kernel I/O
// layout (crates/newton/build.rs:13-48) — the route the reference's build takes
// for every connection kernel (crates/newton/build.rs:673-821). One thread per row;
// a Jacobian row evaluates all 24 axes at once. Selected per store with
// exg_store_set_codegen; rows, slots and results as rows.hip's kernels.
#include <hip/hip_runtime.h>
#include "row_slots.h"
#include "rows.h"
#include "viete_gen.h"

namespace exg {
namespace {

struct Into {
    double* v;
    __device__ __forceinline__ void operator()(int k, double x) const { v[k] = x; }
};

// output k of a _full function: component c = k / 25 (end c / 6, screw coordinate
// c % 6), slot k % 25 (0 the value, 1 + axis the partial), stored through the Pod map
template <class T>
struct JacOut {
    T* jac;
    T* pair;
    bool* bad;
    __device__ __forceinline__ void operator()(int k, double v) const {
        const int c = k / 25, slot = k - 25 * c, e = c / 6, coord = c - 6 * e;
        int q;
        double sign;
        pod_slot_of(coord, q, sign);
        if (slot == 0) pair[e * 6 + q] = (T)(sign * v);
        else jac[e * 144 + (slot - 1) * 6 + q] = (T)(sign * v);
        *bad = *bad || !isfinite(v);
    }
};

// in[30]: velocities of ends a and b (constructor coordinates), base poses of a and b,
// dt, warp; prm the joint's parameters
template <class T>
__device__ __forceinline__ void load_joint_inputs(const Slots<T>& s, const uint32_t* w, int first_param, int jk,
                                                  double in[30], double prm[10]) {
    ld_screw(s.twist + (size_t)w[0] * 6, in);
    ld_screw(s.twist + (size_t)w[1] * 6, in + 6);
    ld_n(s.motor + (size_t)w[2] * 8, in + 12, 8);
    ld_n(s.motor + (size_t)w[3] * 8, in + 20, 8);
    in[28] = ld(s.scalar + w[first_param - 2]);
    in[29] = ld(s.scalar + w[first_param - 1]);
    const int np = (jk == SIMPLE_SPRING_DAMPER || jk == CRITICALLY_DAMPED_WARPED) ? 10
                   : (jk == PERPENDICULAR_DAMPER_WARPED ? 2 : 1);
    for (int p = 0; p < 10; ++p) prm[p] = p < np ? ld(s.scalar + w[first_param + p]) : 0.0;
}

template <class T>
__global__ void k_joint_plain_traced(Slots<T> s, const uint32_t* slots, uint32_t n, int jk) {
    const uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double in[30], prm[10], out[12];
    load_joint_inputs(s, w, 7, jk, in, prm);
    switch (jk) {
        case SIMPLE_SPRING_DAMPER: viete_simple_spring_damper_plain(in, prm, Into{out}); break;
        case CRITICALLY_DAMPED_WARPED: viete_critically_damped_warped_plain(in, prm, Into{out}); break;
        case PERPENDICULAR_DAMPER_WARPED: viete_perpendicular_damper_warped_plain(in, prm, Into{out}); break;
        default: viete_torsional_damper_warped_plain(in, prm, Into{out}); break;
    }
    T* conn = s.pair + (size_t)w[4] * 12;
    st_screw(conn, out);
    st_screw(conn + 6, out + 6);
    mark(s, out, 12);
}

// one kernel per law (a law's 24 axes are one straight-line body of several
// thousand operations)
template <class T, int JK>
__global__ __launch_bounds__(RT) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_joint_jac_traced(
    Slots<T> s, const uint32_t* slots, uint32_t n) {
    const uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double in[30], prm[10];
    load_joint_inputs(s, w, 8, JK, in, prm);
    bool bad = false;
    const JacOut<T> out{s.jac + (size_t)w[4] * 288, s.pair + (size_t)w[5] * 12, &bad};
    if constexpr (JK == SIMPLE_SPRING_DAMPER) viete_simple_spring_damper_full(in, prm, out);
    else if constexpr (JK == CRITICALLY_DAMPED_WARPED) viete_critically_damped_warped_full(in, prm, out);
    else if constexpr (JK == PERPENDICULAR_DAMPER_WARPED) viete_perpendicular_damper_warped_full(in, prm, out);
    else viete_torsional_damper_warped_full(in, prm, out);
    if (bad) atomicAdd(s.fatal, 1u);
}

template <class T>
__global__ void k_pre_traced(Slots<T> s, const uint32_t* slots, uint32_t n) {
    const uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double in[15], out[14];
    ld_screw(s.twist + (size_t)w[0] * 6, in);
    ld_n(s.motor + (size_t)w[1] * 8, in + 6, 8);
    in[14] = ld(s.scalar + w[4]);
    viete_pre(in, Into{out});
    st_n(s.motor + (size_t)w[2] * 8, out, 8);
    st_screw(s.twist + (size_t)w[3] * 6, out + 8);
    mark(s, out, 8);
}

inline dim3 grid_of(uint32_t n) { return dim3((n + RT - 1) / RT); }

template <class T>
bool launch_traced_t(int kind, const Slots<T>& s, const uint32_t* slots, uint32_t n, hipStream_t st) {
    int jk = -1;
    bool jac = false;
    switch (kind) {
        case K_SSD_PLAIN: jk = SIMPLE_SPRING_DAMPER; break;
        case K_SSD_JAC: jk = SIMPLE_SPRING_DAMPER; jac = true; break;
        case K_CDW_PLAIN: jk = CRITICALLY_DAMPED_WARPED; break;
        case K_CDW_JAC: jk = CRITICALLY_DAMPED_WARPED; jac = true; break;
        case K_PERP_PLAIN: jk = PERPENDICULAR_DAMPER_WARPED; break;
        case K_PERP_JAC: jk = PERPENDICULAR_DAMPER_WARPED; jac = true; break;
        case K_TORS_PLAIN: jk = TORSIONAL_DAMPER_WARPED; break;
        case K_TORS_JAC: jk = TORSIONAL_DAMPER_WARPED; jac = true; break;
        case K_PRE:
            if (n) hipLaunchKernelGGL(k_pre_traced<T>, grid_of(n), dim3(RT), 0, st, s, slots, n);
            return true;
        default: return false;
    }
    if (n == 0) return true;
    if (!jac) {
        hipLaunchKernelGGL(k_joint_plain_traced<T>, grid_of(n), dim3(RT), 0, st, s, slots, n, jk);
        return true;
    }
    switch (jk) {
        case SIMPLE_SPRING_DAMPER:
            hipLaunchKernelGGL((k_joint_jac_traced<T, SIMPLE_SPRING_DAMPER>), grid_of(n), dim3(RT), 0, st, s, slots, n);
            break;
        case CRITICALLY_DAMPED_WARPED:
            hipLaunchKernelGGL((k_joint_jac_traced<T, CRITICALLY_DAMPED_WARPED>), grid_of(n), dim3(RT), 0, st, s, slots,
                               n);
            break;
        case PERPENDICULAR_DAMPER_WARPED:
            hipLaunchKernelGGL((k_joint_jac_traced<T, PERPENDICULAR_DAMPER_WARPED>), grid_of(n), dim3(RT), 0, st, s,
                               slots, n);
            break;
        default:
            hipLaunchKernelGGL((k_joint_jac_traced<T, TORSIONAL_DAMPER_WARPED>), grid_of(n), dim3(RT), 0, st, s, slots,
                               n);
            break;
    }
    return true;
}

}  // namespace

bool launch_rows_traced(int kind, const StorePtrs& p, const uint32_t* slots, uint32_t n, hipStream_t st) {
    if (p.scalar_bytes == 4) {
        Slots<float> s{(float*)p.scalar, (float*)p.motor, (float*)p.twist, (float*)p.wrench, (float*)p.pair,
                       (float*)p.jac, (float*)p.block, (float*)p.diag, (float*)p.full, p.row, p.fatal};
        return launch_traced_t(kind, s, slots, n, st);
    }
    Slots<double> s{(double*)p.scalar, (double*)p.motor, (double*)p.twist, (double*)p.wrench, (double*)p.pair,
                    (double*)p.jac, (double*)p.block, (double*)p.diag, (double*)p.full, p.row, p.fatal};
    return launch_traced_t(kind, s, slots, n, st);
}

}  // namespace exg
