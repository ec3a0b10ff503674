// Row-driven kernels of the reference's accelerator contract, one per MessageKind.
// Each row is the reference's baked [u32; 128] layout (crates/newton/src/accelerator/rows/*.rs,
// crates/newton/build.rs kernel preambles); slot indices address the typed storages.
// Scalar row words (gemm sign, assemble mass_scale) are f32 bit patterns as in the
// reference (crates/newton/src/accelerator/row.rs:29-36).
//
// Storage bytes are the reference's Pod layouts over the storage scalar T (f32 as
// `usual::<f32>()` registers them, or f64): a motor is its 8 even blades, a screw
// (twist, wrench, each wrench of a pair or of a Jacobian) its grade-2 store in
// ascending blade order with the markup signs (screw_to_pod / screw_from_pod,
// generated from the sign law; crates/newton/build.rs:106-152 probes the same map),
// a Block plain [[T; 6]; 6] in (force|torque) × (linear|angular) coordinates. Every
// kernel reads its slots into f64, computes in f64 and rounds once on store.
//
// Fatal marks (mark-and-continue, crates/newton/src/accelerator/shaders.rs:411-431): a
// row whose outputs are not all finite — the division guards of the traced kernels
// tripped — adds one to the dispatch's fatal counter; its outputs are written anyway.
//
//   Pre              rows/fixed.rs:14-22          Gather         rows/gather.rs:1-15, build.rs:1083-1154
//   connection       rows/fixed.rs:24-33          Gemm           rows/gemm.rs:1-23, build.rs:1239-1295
//   BodyPost*        rows/body_post.rs:1-20       AssembleBlock  rows/assemble.rs:1-31, build.rs:1310-1403
//   BodyPostGathered rows/body_post.rs:118-196    BlockMatVec    rows/matvec.rs, build.rs:1161-1232
//   BlockReduce      rows/reduce.rs               BlockCopy      rows/copy.rs
//   EvalSimpleSum    accelerator/simple_sum.glsl (a + b)
#include <hip/hip_runtime.h>
#include "physics_dev.h"
#include "rows.h"
#include "row_slots.h"

namespace exg {

template <class T>
__global__ void k_simple_sum(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    const double v = ld(s.scalar + w[0]) + ld(s.scalar + w[1]);
    st(s.scalar + w[2], v);
    mark(s, &v, 1);
}

template <class T>
__global__ void k_pre(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double e[8], mid[8], v[6], pose[8];
    ld_screw(s.twist + (size_t)w[0] * 6, v);
    ld_n(s.motor + (size_t)w[1] * 8, pose, 8);
    twist_exp(v, ld(s.scalar + w[4]), e);
    compose(pose, e, mid);
    st_n(s.motor + (size_t)w[2] * 8, mid, 8);
    st_screw(s.twist + (size_t)w[3] * 6, v);
    mark(s, mid, 8);
}

__device__ inline int n_params_of(int jk) {
    return (jk == SIMPLE_SPRING_DAMPER || jk == CRITICALLY_DAMPED_WARPED) ? 10 : (jk == PERPENDICULAR_DAMPER_WARPED ? 2 : 1);
}

template <class T>
__global__ void k_joint_plain(Slots<T> s, const uint32_t* slots, uint32_t n, int jk) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double prm[10] = {};
    const int np = n_params_of(jk);
    for (int p = 0; p < np; ++p) prm[p] = ld(s.scalar + w[7 + p]);
    double va[6], vb[6], ma[8], mb[8], out[12];
    ld_screw(s.twist + (size_t)w[0] * 6, va);
    ld_screw(s.twist + (size_t)w[1] * 6, vb);
    ld_n(s.motor + (size_t)w[2] * 8, ma, 8);
    ld_n(s.motor + (size_t)w[3] * 8, mb, 8);
    joint_wrench(jk, prm, ma, mb, va, vb, ld(s.scalar + w[5]), ld(s.scalar + w[6]), out, out + 6);
    T* conn = s.pair + (size_t)w[4] * 12;
    st_screw(conn, out);
    st_screw(conn + 6, out + 6);
    mark(s, out, 12);
}

// One thread per (row, axis): the 24 derivative directions of one connection.
template <class T>
__global__ void k_joint_jac(Slots<T> s, const uint32_t* slots, uint32_t n, int jk) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 24u) return;
    uint32_t r = t / 24u;
    int ax = t % 24u;
    const uint32_t* w = row_of(s, slots, r);
    double prm[10] = {};
    const int np = n_params_of(jk);
    for (int p = 0; p < np; ++p) prm[p] = ld(s.scalar + w[8 + p]);
    using D = Dual<double>;
    D pp[2][8], vv[2][6];
    for (int e = 0; e < 2; ++e) {
        double base[8], v[6];
        ld_n(s.motor + (size_t)w[2 + e] * 8, base, 8);
        if (ax < 12 && ax / 6 == e) seed_pose(base, ax % 6, pp[e]);
        else for (int k = 0; k < 8; ++k) pp[e][k] = D(base[k]);
        ld_screw(s.twist + (size_t)w[e] * 6, v);
        for (int k = 0; k < 6; ++k) vv[e][k] = D(v[k], (ax >= 12 && (ax - 12) == 6 * e + k) ? 1.0 : 0.0);
    }
    D wa[6], wb[6];
    joint_wrench(jk, prm, pp[0], pp[1], vv[0], vv[1], ld(s.scalar + w[6]), ld(s.scalar + w[7]), wa, wb);
    T* jac = s.jac + (size_t)w[4] * 288;
    double da[6], db[6];
    for (int k = 0; k < 6; ++k) {
        da[k] = wa[k].d;
        db[k] = wb[k].d;
    }
    st_screw(jac + (0 * 24 + ax) * 6, da);
    st_screw(jac + (1 * 24 + ax) * 6, db);
    mark(s, da, 6);
    mark(s, db, 6);
    if (ax == 0) {
        double val[12];
        for (int k = 0; k < 6; ++k) {
            val[k] = wa[k].v;
            val[6 + k] = wb[k].v;
        }
        T* conn = s.pair + (size_t)w[5] * 12;
        st_screw(conn, val);
        st_screw(conn + 6, val + 6);
    }
}

template <class T>
__global__ void k_gather(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    const double a1 = (double)w[2];
    const uint32_t nt = w[3];
    double acc[6], own[6], ext[6];
    ld_screw(s.wrench + (size_t)w[0] * 6, own);
    ld_screw(s.wrench + (size_t)w[1] * 6, ext);
    for (int k = 0; k < 6; ++k) acc[k] = a1 * own[k] + (1.0 - a1) * ext[k];
    for (uint32_t q = 0; q < nt; ++q) {
        const uint32_t t = w[4 + q];
        double pw[6];
        ld_screw(s.pair + (size_t)(t >> 1) * 12 + (t & 1u) * 6, pw);
        for (int k = 0; k < 6; ++k) acc[k] = acc[k] + pw[k];
    }
    st_screw(s.wrench + (size_t)w[0] * 6, acc);
    mark(s, acc, 6);
}

// one thread per output cell (row, r, c) of a block product
template <class T>
__global__ void k_gemm(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 36u) return;
    uint32_t r = t / 36u;
    int rc = t % 36u, rr = rc / 6, cc = rc % 6;
    const uint32_t* w = row_of(s, slots, r);
    const double a1 = (double)w[1];
    const uint32_t nt = w[2];
    const double sign = (double)bits_f32(w[3]);
    double acc = a1 * ld(s.block + (size_t)w[0] * 36 + rc);
    for (uint32_t q = 0; q < nt; ++q) {
        const T* a = s.block + (size_t)w[4 + 2 * q] * 36;
        const uint32_t bt = w[5 + 2 * q];
        const T* b = s.block + (size_t)(bt >> 1) * 36;
        const double diag = (double)(bt & 1u);
        double sum = 0.0;
        for (int k = 0; k < 6; ++k) {
            double right = sign * ld(b + k * 6 + cc);
            if (k == cc) right = right + diag * 2.0;
            sum = sum + ld(a + rr * 6 + k) * right;
        }
        acc = acc + sum;
    }
    // every cell of a row's output is read (accumulate rounds) and written by this thread only
    st(s.block + (size_t)w[0] * 36 + rc, acc);
    mark(s, &acc, 1);
}

template <class T>
__device__ void body_post_one(const Slots<T>& s, uint32_t mk, uint32_t vk, double mass, const double* t9,
                              uint32_t snap_k, const double* total, uint32_t mass_out, uint32_t rhs_out,
                              uint32_t scale_out, double half, double floor2) {
    double block[6][6], rhs[6], sc[2], vel[6], mid[8], snap[6];
    ld_n(s.motor + (size_t)mk * 8, mid, 8);
    ld_screw(s.twist + (size_t)vk * 6, vel);
    ld_screw(s.wrench + (size_t)snap_k * 6, snap);
    body_post(mid, vel, mass, t9, snap, total, half, floor2, block, rhs, sc);
    st_n(s.block + (size_t)mass_out * 36, &block[0][0], 36);
    st_screw(s.wrench + (size_t)rhs_out * 6, rhs);
    const double scale[6] = {sc[0], sc[0], sc[0], sc[1], sc[1], sc[1]};
    st_screw(s.wrench + (size_t)scale_out * 6, scale);
    mark(s, &block[0][0], 36);
    mark(s, rhs, 6);
}

template <class T>
__device__ inline void tensor_of(const Slots<T>& s, uint32_t ak, bool diagonal, double t9[9]) {
    if (diagonal) {
        for (int k = 0; k < 9; ++k) t9[k] = 0.0;
        for (int k = 0; k < 3; ++k) t9[4 * k] = ld(s.diag + (size_t)ak * 3 + k);
    } else {
        ld_n(s.full + (size_t)ak * 9, t9, 9);
    }
}

// one thread per (row, body): up to 13 bodies per row, 9 words each
template <class T>
__global__ void k_body_post(Slots<T> s, const uint32_t* slots, uint32_t n, int diagonal) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 13u) return;
    uint32_t r = t / 13u, b = t % 13u;
    const uint32_t* w = row_of(s, slots, r);
    if (b >= w[0]) return;
    const double half = ld(s.scalar + w[1]), floor2 = ld(s.scalar + w[2]);
    const uint32_t* o = w + 3 + 9 * b;
    double t9[9], total[6];
    tensor_of(s, o[3], diagonal != 0, t9);
    ld_screw(s.wrench + (size_t)o[5] * 6, total);
    body_post_one(s, o[0], o[1], ld(s.scalar + o[2]), t9, o[4], total, o[6], o[7], o[8], half, floor2);
}

template <class T>
__global__ void k_body_post_gathered(Slots<T> s, const uint32_t* slots, uint32_t n, int diagonal) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    const double half = ld(s.scalar + w[0]), floor2 = ld(s.scalar + w[1]);
    double acc[6];
    ld_screw(s.wrench + (size_t)w[7] * 6, acc);
    for (uint32_t q = 0; q < w[11]; ++q) {
        const uint32_t t = w[12 + q];
        double pw[6];
        ld_screw(s.pair + (size_t)(t >> 1) * 12 + (t & 1u) * 6, pw);
        for (int k = 0; k < 6; ++k) acc[k] = acc[k] + pw[k];
    }
    double t9[9];
    tensor_of(s, w[5], diagonal != 0, t9);
    body_post_one(s, w[2], w[3], ld(s.scalar + w[4]), t9, w[6], acc, w[8], w[9], w[10], half, floor2);
}

// one thread per output cell; a Jacobian column is a Pod wrench, read in coordinates
template <class T>
__global__ void k_assemble(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 36u) return;
    uint32_t r = t / 36u;
    int rc = t % 36u, rr = rc / 6, cc = rc % 6;
    const uint32_t* w = row_of(s, slots, r);
    const double mass_scale = (double)bits_f32(w[2]);
    const double pf = ld(s.scalar + w[3]), vf = ld(s.scalar + w[4]);
    const double a1 = (double)w[5];
    const uint32_t nt = w[6];
    double acc = a1 * ld(s.block + (size_t)w[0] * 36 + rc) + mass_scale * ld(s.block + (size_t)w[1] * 36 + rc);
    for (uint32_t q = 0; q < nt; ++q) {
        const uint32_t term = w[7 + q];
        const T* jac = s.jac + (size_t)(term >> 2) * 288;
        const int base = ((term >> 1) & 1u) * 144 + (term & 1u) * 36;
        double p[6], v[6];
        ld_screw(jac + base + cc * 6, p);
        ld_screw(jac + base + 72 + cc * 6, v);
        acc = acc + pf * p[rr] + vf * v[rr];
    }
    st(s.block + (size_t)w[0] * 36 + rc, acc);
    mark(s, &acc, 1);
}


// one thread per (row, output coordinate): each thread reads and writes only the Pod
// component of its own coordinate
template <class T>
__global__ void k_matvec(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 6u) return;
    uint32_t r = t / 6u;
    const int rr = t % 6u;
    const uint32_t* w = row_of(s, slots, r);
    const double a1 = (double)w[1];
    const uint32_t nt = w[2];
    int k;
    double sign;
    pod_slot_of(rr, k, sign);
    T* out = s.twist + (size_t)w[0] * 6 + k;
    double acc = a1 * (sign * ld(out));
    for (uint32_t q = 0; q < nt; ++q) {
        const T* x = s.block + (size_t)w[3 + 2 * q] * 36;
        double rhs[6];
        ld_screw(s.wrench + (size_t)w[4 + 2 * q] * 6, rhs);
        double sum = 0.0;
        for (int c = 0; c < 6; ++c) sum = sum + ld(x + rr * 6 + c) * rhs[c];
        acc = acc + sum;
    }
    st(out, sign * acc);
    mark(s, &acc, 1);
}

template <class T>
__global__ void k_reduce(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double acc = 0.0;
    for (uint32_t q = 0; q < w[1]; ++q) {
        const uint32_t t = w[2 + q];
        const T* b = s.block + (size_t)(t >> 1) * 36;
        const double diag = (double)(t & 1u);
        for (int rr = 0; rr < 6; ++rr)
            for (int c = 0; c < 6; ++c) {
                const double v = ld(b + rr * 6 + c);
                const double e = (rr == c) ? diag - v : -v;
                acc = acc + e * e;
            }
    }
    st(s.scalar + w[0], acc);
    mark(s, &acc, 1);
}

template <class T>
__global__ void k_copy(Slots<T> s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 36u) return;
    uint32_t r = t / 36u;
    const int rc = t % 36u;
    const uint32_t* w = row_of(s, slots, r);
    s.block[(size_t)w[0] * 36 + rc] = s.block[(size_t)w[1] * 36 + rc];   // exact: a byte copy
}

static inline dim3 grid_for(uint64_t threads) { return dim3((unsigned)((threads + RT - 1) / RT)); }

template <class T>
static void launch_rows_t(int kind, const Slots<T>& s, const uint32_t* slots, uint32_t n, hipStream_t st) {
    if (n == 0) return;
    switch (kind) {
        case K_SIMPLE_SUM: hipLaunchKernelGGL(k_simple_sum<T>, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_CDW_PLAIN: hipLaunchKernelGGL(k_joint_plain<T>, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)CRITICALLY_DAMPED_WARPED); break;
        case K_CDW_JAC: hipLaunchKernelGGL(k_joint_jac<T>, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)CRITICALLY_DAMPED_WARPED); break;
        case K_PERP_PLAIN: hipLaunchKernelGGL(k_joint_plain<T>, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)PERPENDICULAR_DAMPER_WARPED); break;
        case K_PERP_JAC: hipLaunchKernelGGL(k_joint_jac<T>, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)PERPENDICULAR_DAMPER_WARPED); break;
        case K_SSD_PLAIN: hipLaunchKernelGGL(k_joint_plain<T>, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)SIMPLE_SPRING_DAMPER); break;
        case K_SSD_JAC: hipLaunchKernelGGL(k_joint_jac<T>, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)SIMPLE_SPRING_DAMPER); break;
        case K_TORS_PLAIN: hipLaunchKernelGGL(k_joint_plain<T>, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)TORSIONAL_DAMPER_WARPED); break;
        case K_TORS_JAC: hipLaunchKernelGGL(k_joint_jac<T>, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)TORSIONAL_DAMPER_WARPED); break;
        case K_PRE: hipLaunchKernelGGL(k_pre<T>, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_GATHER: hipLaunchKernelGGL(k_gather<T>, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_GEMM: hipLaunchKernelGGL(k_gemm<T>, grid_for(36ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_BODY_POST_DIAG: hipLaunchKernelGGL(k_body_post<T>, grid_for(13ull * n), dim3(RT), 0, st, s, slots, n, 1); break;
        case K_BODY_POST_FULL: hipLaunchKernelGGL(k_body_post<T>, grid_for(13ull * n), dim3(RT), 0, st, s, slots, n, 0); break;
        case K_ASSEMBLE: hipLaunchKernelGGL(k_assemble<T>, grid_for(36ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_MATVEC: hipLaunchKernelGGL(k_matvec<T>, grid_for(6ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_REDUCE: hipLaunchKernelGGL(k_reduce<T>, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_COPY: hipLaunchKernelGGL(k_copy<T>, grid_for(36ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_BODY_POST_GATHERED_DIAG: hipLaunchKernelGGL(k_body_post_gathered<T>, grid_for(n), dim3(RT), 0, st, s, slots, n, 1); break;
        case K_BODY_POST_GATHERED_FULL: hipLaunchKernelGGL(k_body_post_gathered<T>, grid_for(n), dim3(RT), 0, st, s, slots, n, 0); break;
        default: break;
    }
}

void launch_rows(int kind, const StorePtrs& p, const uint32_t* slots, uint32_t n, hipStream_t st) {
    if (p.scalar_bytes == 4) {
        Slots<float> s{(float*)p.scalar, (float*)p.motor, (float*)p.twist, (float*)p.wrench, (float*)p.pair,
                       (float*)p.jac, (float*)p.block, (float*)p.diag, (float*)p.full, p.row, p.fatal};
        launch_rows_t(kind, s, slots, n, st);
    } else {
        Slots<double> s{(double*)p.scalar, (double*)p.motor, (double*)p.twist, (double*)p.wrench, (double*)p.pair,
                        (double*)p.jac, (double*)p.block, (double*)p.diag, (double*)p.full, p.row, p.fatal};
        launch_rows_t(kind, s, slots, n, st);
    }
}

}  // namespace exg
