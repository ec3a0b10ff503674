// Row-driven kernels of the reference's accelerator contract, one per MessageKind.
// Each row is the reference's baked [u32; 128] layout (crates/newton/src/accelerator/rows/*.rs,
// crates/newton/build.rs kernel preambles); slot indices address the typed storages.
// Scalar row words (gemm sign, assemble mass_scale) are f32 bit patterns as in the
// reference (crates/newton/src/accelerator/row.rs:29-36).
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

namespace exg {

constexpr int RT = 64;

__device__ inline const uint32_t* row_of(const StorePtrs& s, const uint32_t* slots, uint32_t r) {
    return s.row + (size_t)slots[r] * ROW_WORDS;
}
__device__ inline float bits_f32(uint32_t w) { return __uint_as_float(w); }

__global__ void k_simple_sum(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    s.scalar[w[2]] = s.scalar[w[0]] + s.scalar[w[1]];
}

__global__ void k_pre(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    const double* vel = s.twist + (size_t)w[0] * 6;
    const double* pose = s.motor + (size_t)w[1] * 8;
    double e[8], mid[8], v[6];
    for (int k = 0; k < 6; ++k) v[k] = vel[k];
    twist_exp(v, s.scalar[w[4]], e);
    compose(pose, e, mid);
    for (int k = 0; k < 8; ++k) s.motor[(size_t)w[2] * 8 + k] = mid[k];
    for (int k = 0; k < 6; ++k) s.twist[(size_t)w[3] * 6 + k] = v[k];
}

__device__ inline int n_params_of(int jk) {
    return (jk == SIMPLE_SPRING_DAMPER || jk == CRITICALLY_DAMPED_WARPED) ? 10 : (jk == PERPENDICULAR_DAMPER_WARPED ? 2 : 1);
}

__global__ void k_joint_plain(StorePtrs s, const uint32_t* slots, uint32_t n, int jk) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double prm[10] = {};
    int np = n_params_of(jk);
    for (int p = 0; p < np; ++p) prm[p] = s.scalar[w[7 + p]];
    double* conn = s.pair + (size_t)w[4] * 12;
    joint_wrench(jk, prm, s.motor + (size_t)w[2] * 8, s.motor + (size_t)w[3] * 8, s.twist + (size_t)w[0] * 6,
                 s.twist + (size_t)w[1] * 6, s.scalar[w[5]], s.scalar[w[6]], conn, conn + 6);
}

// One thread per (row, axis): the 24 derivative directions of one connection.
__global__ void k_joint_jac(StorePtrs s, const uint32_t* slots, uint32_t n, int jk) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 24u) return;
    uint32_t r = t / 24u;
    int ax = t % 24u;
    const uint32_t* w = row_of(s, slots, r);
    double prm[10] = {};
    int np = n_params_of(jk);
    for (int p = 0; p < np; ++p) prm[p] = s.scalar[w[8 + p]];
    using D = Dual<double>;
    D pp[2][8], vv[2][6];
    for (int e = 0; e < 2; ++e) {
        const double* base = s.motor + (size_t)w[2 + e] * 8;
        if (ax < 12 && ax / 6 == e) seed_pose(base, ax % 6, pp[e]);
        else for (int k = 0; k < 8; ++k) pp[e][k] = D(base[k]);
        const double* v = s.twist + (size_t)w[e] * 6;
        for (int k = 0; k < 6; ++k) vv[e][k] = D(v[k], (ax >= 12 && (ax - 12) == 6 * e + k) ? 1.0 : 0.0);
    }
    D wa[6], wb[6];
    joint_wrench(jk, prm, pp[0], pp[1], vv[0], vv[1], s.scalar[w[6]], s.scalar[w[7]], wa, wb);
    double* jac = s.jac + (size_t)w[4] * 288;
    for (int k = 0; k < 6; ++k) {
        jac[(0 * 24 + ax) * 6 + k] = wa[k].d;
        jac[(1 * 24 + ax) * 6 + k] = wb[k].d;
    }
    if (ax == 0) {
        double* conn = s.pair + (size_t)w[5] * 12;
        for (int k = 0; k < 6; ++k) {
            conn[k] = wa[k].v;
            conn[6 + k] = wb[k].v;
        }
    }
}

__global__ void k_gather(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double a1 = (double)w[2];
    uint32_t nt = w[3];
    double acc[6];
    for (int k = 0; k < 6; ++k) acc[k] = a1 * s.wrench[(size_t)w[0] * 6 + k] + (1.0 - a1) * s.wrench[(size_t)w[1] * 6 + k];
    for (uint32_t q = 0; q < nt; ++q) {
        uint32_t t = w[4 + q];
        const double* pw = s.pair + (size_t)(t >> 1) * 12 + (t & 1u) * 6;
        for (int k = 0; k < 6; ++k) acc[k] = acc[k] + pw[k];
    }
    for (int k = 0; k < 6; ++k) s.wrench[(size_t)w[0] * 6 + k] = acc[k];
}

// one thread per output cell (row, r, c) of a block product
__global__ void k_gemm(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 36u) return;
    uint32_t r = t / 36u;
    int rc = t % 36u, rr = rc / 6, cc = rc % 6;
    const uint32_t* w = row_of(s, slots, r);
    double a1 = (double)w[1];
    uint32_t nt = w[2];
    double sign = (double)bits_f32(w[3]);
    double acc = a1 * s.block[(size_t)w[0] * 36 + rc];
    for (uint32_t q = 0; q < nt; ++q) {
        const double* a = s.block + (size_t)w[4 + 2 * q] * 36;
        uint32_t bt = w[5 + 2 * q];
        const double* b = s.block + (size_t)(bt >> 1) * 36;
        double diag = (double)(bt & 1u);
        double sum = 0.0;
        for (int k = 0; k < 6; ++k) {
            double right = sign * b[k * 6 + cc];
            if (k == cc) right = right + diag * 2.0;
            sum = sum + a[rr * 6 + k] * right;
        }
        acc = acc + sum;
    }
    // every cell of a row's output is read (accumulate rounds) and written by this thread only
    s.block[(size_t)w[0] * 36 + rc] = acc;
}

__device__ void body_post_one(const StorePtrs& s, uint32_t mk, uint32_t vk, double mass, const double* t9,
                              uint32_t snap_k, const double* total, uint32_t mass_out, uint32_t rhs_out,
                              uint32_t scale_out, double half, double floor2) {
    double block[6][6], rhs[6], sc[2], vel[6], mid[8], snap[6];
    for (int k = 0; k < 8; ++k) mid[k] = s.motor[(size_t)mk * 8 + k];
    for (int k = 0; k < 6; ++k) {
        vel[k] = s.twist[(size_t)vk * 6 + k];
        snap[k] = s.wrench[(size_t)snap_k * 6 + k];
    }
    body_post(mid, vel, mass, t9, snap, total, half, floor2, block, rhs, sc);
    for (int r = 0; r < 6; ++r)
        for (int c = 0; c < 6; ++c) s.block[(size_t)mass_out * 36 + r * 6 + c] = block[r][c];
    for (int k = 0; k < 6; ++k) s.wrench[(size_t)rhs_out * 6 + k] = rhs[k];
    for (int k = 0; k < 3; ++k) {
        s.wrench[(size_t)scale_out * 6 + k] = sc[0];
        s.wrench[(size_t)scale_out * 6 + 3 + k] = sc[1];
    }
}

__device__ inline void tensor_of(const StorePtrs& s, uint32_t ak, bool diagonal, double t9[9]) {
    if (diagonal) {
        for (int k = 0; k < 9; ++k) t9[k] = 0.0;
        for (int k = 0; k < 3; ++k) t9[4 * k] = s.diag[(size_t)ak * 3 + k];
    } else {
        for (int k = 0; k < 9; ++k) t9[k] = s.full[(size_t)ak * 9 + k];
    }
}

// one thread per (row, body): up to 13 bodies per row, 9 words each
__global__ void k_body_post(StorePtrs s, const uint32_t* slots, uint32_t n, int diagonal) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 13u) return;
    uint32_t r = t / 13u, b = t % 13u;
    const uint32_t* w = row_of(s, slots, r);
    if (b >= w[0]) return;
    double half = s.scalar[w[1]], floor2 = s.scalar[w[2]];
    const uint32_t* o = w + 3 + 9 * b;
    double t9[9], total[6];
    tensor_of(s, o[3], diagonal != 0, t9);
    for (int k = 0; k < 6; ++k) total[k] = s.wrench[(size_t)o[5] * 6 + k];
    body_post_one(s, o[0], o[1], s.scalar[o[2]], t9, o[4], total, o[6], o[7], o[8], half, floor2);
}

__global__ void k_body_post_gathered(StorePtrs s, const uint32_t* slots, uint32_t n, int diagonal) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double half = s.scalar[w[0]], floor2 = s.scalar[w[1]];
    double acc[6];
    for (int k = 0; k < 6; ++k) acc[k] = s.wrench[(size_t)w[7] * 6 + k];
    for (uint32_t q = 0; q < w[11]; ++q) {
        uint32_t t = w[12 + q];
        const double* pw = s.pair + (size_t)(t >> 1) * 12 + (t & 1u) * 6;
        for (int k = 0; k < 6; ++k) acc[k] = acc[k] + pw[k];
    }
    double t9[9];
    tensor_of(s, w[5], diagonal != 0, t9);
    body_post_one(s, w[2], w[3], s.scalar[w[4]], t9, w[6], acc, w[8], w[9], w[10], half, floor2);
}

__global__ void k_assemble(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 36u) return;
    uint32_t r = t / 36u;
    int rc = t % 36u, rr = rc / 6, cc = rc % 6;
    const uint32_t* w = row_of(s, slots, r);
    double mass_scale = (double)bits_f32(w[2]);
    double pf = s.scalar[w[3]], vf = s.scalar[w[4]];
    double a1 = (double)w[5];
    uint32_t nt = w[6];
    double acc = a1 * s.block[(size_t)w[0] * 36 + rc] + mass_scale * s.block[(size_t)w[1] * 36 + rc];
    for (uint32_t q = 0; q < nt; ++q) {
        uint32_t term = w[7 + q];
        const double* jac = s.jac + (size_t)(term >> 2) * 288;
        int base = ((term >> 1) & 1u) * 144 + (term & 1u) * 36;
        double p = jac[base + cc * 6 + rr];
        double v = jac[base + 72 + cc * 6 + rr];
        acc = acc + pf * p + vf * v;
    }
    s.block[(size_t)w[0] * 36 + rc] = acc;
}

__global__ void k_matvec(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 6u) return;
    uint32_t r = t / 6u;
    int rr = t % 6u;
    const uint32_t* w = row_of(s, slots, r);
    double a1 = (double)w[1];
    uint32_t nt = w[2];
    double acc = a1 * s.twist[(size_t)w[0] * 6 + rr];
    for (uint32_t q = 0; q < nt; ++q) {
        const double* x = s.block + (size_t)w[3 + 2 * q] * 36;
        const double* rhs = s.wrench + (size_t)w[4 + 2 * q] * 6;
        double sum = 0.0;
        for (int c = 0; c < 6; ++c) sum = sum + x[rr * 6 + c] * rhs[c];
        acc = acc + sum;
    }
    s.twist[(size_t)w[0] * 6 + rr] = acc;
}

__global__ void k_reduce(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t r = blockIdx.x * RT + threadIdx.x;
    if (r >= n) return;
    const uint32_t* w = row_of(s, slots, r);
    double acc = 0.0;
    for (uint32_t q = 0; q < w[1]; ++q) {
        uint32_t t = w[2 + q];
        const double* b = s.block + (size_t)(t >> 1) * 36;
        double diag = (double)(t & 1u);
        for (int rr = 0; rr < 6; ++rr)
            for (int c = 0; c < 6; ++c) {
                double v = b[rr * 6 + c];
                double e = (rr == c) ? diag - v : -v;
                acc = acc + e * e;
            }
    }
    s.scalar[w[0]] = acc;
}

__global__ void k_copy(StorePtrs s, const uint32_t* slots, uint32_t n) {
    uint32_t t = blockIdx.x * RT + threadIdx.x;
    if (t >= n * 36u) return;
    uint32_t r = t / 36u;
    int rc = t % 36u;
    const uint32_t* w = row_of(s, slots, r);
    s.block[(size_t)w[0] * 36 + rc] = s.block[(size_t)w[1] * 36 + rc];
}

static inline dim3 grid_for(uint64_t threads) { return dim3((unsigned)((threads + RT - 1) / RT)); }

void launch_rows(int kind, const StorePtrs& s, const uint32_t* slots, uint32_t n, hipStream_t st) {
    if (n == 0) return;
    switch (kind) {
        case K_SIMPLE_SUM: hipLaunchKernelGGL(k_simple_sum, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_CDW_PLAIN: hipLaunchKernelGGL(k_joint_plain, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)CRITICALLY_DAMPED_WARPED); break;
        case K_CDW_JAC: hipLaunchKernelGGL(k_joint_jac, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)CRITICALLY_DAMPED_WARPED); break;
        case K_PERP_PLAIN: hipLaunchKernelGGL(k_joint_plain, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)PERPENDICULAR_DAMPER_WARPED); break;
        case K_PERP_JAC: hipLaunchKernelGGL(k_joint_jac, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)PERPENDICULAR_DAMPER_WARPED); break;
        case K_SSD_PLAIN: hipLaunchKernelGGL(k_joint_plain, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)SIMPLE_SPRING_DAMPER); break;
        case K_SSD_JAC: hipLaunchKernelGGL(k_joint_jac, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)SIMPLE_SPRING_DAMPER); break;
        case K_TORS_PLAIN: hipLaunchKernelGGL(k_joint_plain, grid_for(n), dim3(RT), 0, st, s, slots, n, (int)TORSIONAL_DAMPER_WARPED); break;
        case K_TORS_JAC: hipLaunchKernelGGL(k_joint_jac, grid_for(24ull * n), dim3(RT), 0, st, s, slots, n, (int)TORSIONAL_DAMPER_WARPED); break;
        case K_PRE: hipLaunchKernelGGL(k_pre, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_GATHER: hipLaunchKernelGGL(k_gather, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_GEMM: hipLaunchKernelGGL(k_gemm, grid_for(36ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_BODY_POST_DIAG: hipLaunchKernelGGL(k_body_post, grid_for(13ull * n), dim3(RT), 0, st, s, slots, n, 1); break;
        case K_BODY_POST_FULL: hipLaunchKernelGGL(k_body_post, grid_for(13ull * n), dim3(RT), 0, st, s, slots, n, 0); break;
        case K_ASSEMBLE: hipLaunchKernelGGL(k_assemble, grid_for(36ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_MATVEC: hipLaunchKernelGGL(k_matvec, grid_for(6ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_REDUCE: hipLaunchKernelGGL(k_reduce, grid_for(n), dim3(RT), 0, st, s, slots, n); break;
        case K_COPY: hipLaunchKernelGGL(k_copy, grid_for(36ull * n), dim3(RT), 0, st, s, slots, n); break;
        case K_BODY_POST_GATHERED_DIAG: hipLaunchKernelGGL(k_body_post_gathered, grid_for(n), dim3(RT), 0, st, s, slots, n, 1); break;
        case K_BODY_POST_GATHERED_FULL: hipLaunchKernelGGL(k_body_post_gathered, grid_for(n), dim3(RT), 0, st, s, slots, n, 0); break;
        default: break;
    }
}

}  // namespace exg
