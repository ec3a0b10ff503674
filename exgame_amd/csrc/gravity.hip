// All-pairs gravity with the reference's size-core law (crates/newton/src/gravity.rs:110-146,
// 530-601): for the unordered pair (lo, hi), d = |c_hi − c_lo|, pairs closer than 1e-9
// are skipped, s = size of the HIGHER slot, and
//     g = (G m_lo/r)(G m_hi/r)/r,  r = d if d ≥ s else s,
// so the force on lo is g·(c_hi − c_lo)/G and the force on hi its negation — i.e. on
// body i,  F_i = (G m_i / G) · Σ_j (G m_j) · r_ij⁻³ · (c_j − c_i).
// The moment of each target is taken about its island origin from the summed force.
//
// Layout: the charge of body i is 8 doubles [x y z | lx ly lz | G·m | size] (position
// relative to the world frame, lever arm from the island origin, charge, core size).
// Slots are body indices: bodies of one world are contiguous and in slot order.
//
// Work split: a table of entries, each one (tile of ≤ 256 targets, chunk of
// sources of the same world). A workgroup takes one entry: each thread owns a
// target and walks the chunk's sources, staged through LDS in blocks of 256 with
// s² and 1/s of each source, and writes its partial sum to partial[entry][axis][lane]
// (coalesced). A second table lists, per tile, its consecutive entries: the reduce
// kernel adds them in ascending chunk order, scales by the target's charge / G and
// forms the moment. A world's chunking depends only on its own body count, so its
// forces are bit-identical whatever other worlds share the launch. About four
// workgroups per CU fill all 256 CUs from a few thousand bodies up.
//
// f64 pairs: one hardware reciprocal square root refined by two Newton steps
// (y ← y + y·(1 − d²y²)/2) replaces the reference's sqrt and three divisions; the
// result agrees with the correctly rounded form to a few ulp.
// f32c pairs (mixed precision): sources and target are narrowed to f32 as they are
// staged; a pair costs one v_rsq_f32 and a dozen f32 operations, and each force
// component is accumulated as an unevaluated sum hi + lo — after every addition the
// rounding error of hi + x is recovered exactly (Knuth's two-sum) and added to lo.
// The chunk sums are widened to f64 for the reduction, scaling and moment.
#include <hip/hip_runtime.h>
#include "island.h"

namespace exg {

constexpr int GRAV_T = 256;
constexpr int GRAV_MIN_CHUNK = 64;

__device__ inline double rsqrt_refined(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double e = fma(-(x * y), y, 1.0);
    y = fma(y * 0.5, e, y);
    e = fma(-(x * y), y, 1.0);
    y = fma(y * 0.5, e, y);
    return y;
}

__device__ inline void two_sum_acc(float& hi, float& lo, float x) {
    const float t = hi + x;
    const float bp = t - hi;
    const float e = (hi - (t - bp)) + (x - bp);
    hi = t;
    lo = lo + e;
}

__global__ __launch_bounds__(GRAV_T) void gravity_partial_f64_kernel(const double* __restrict__ ch,
                                                                     const GravEntry* __restrict__ entries,
                                                                     double* __restrict__ partial) {
    __shared__ double sx[GRAV_T], sy[GRAV_T], sz[GRAV_T], sg[GRAV_T], ss2[GRAV_T], sinv[GRAV_T];
    const GravEntry en = entries[blockIdx.x];
    const int i = en.tgt0 + threadIdx.x;
    const bool live = (int)threadIdx.x < en.ntgt;
    double xi = 0.0, yi = 0.0, zi = 0.0, si2 = 1.0, sinv_i = 1.0;
    if (live) {
        xi = ch[(size_t)i * 8 + 0];
        yi = ch[(size_t)i * 8 + 1];
        zi = ch[(size_t)i * 8 + 2];
        double si = ch[(size_t)i * 8 + 7];
        si2 = si * si;
        sinv_i = 1.0 / si;
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int base = en.src0; base < en.src1; base += GRAV_T) {
        const int j = base + threadIdx.x;
        if (j < en.src1) {
            sx[threadIdx.x] = ch[(size_t)j * 8 + 0];
            sy[threadIdx.x] = ch[(size_t)j * 8 + 1];
            sz[threadIdx.x] = ch[(size_t)j * 8 + 2];
            sg[threadIdx.x] = ch[(size_t)j * 8 + 6];
            double sj = ch[(size_t)j * 8 + 7];
            ss2[threadIdx.x] = sj * sj;
            sinv[threadIdx.x] = 1.0 / sj;
        }
        __syncthreads();
        const int cnt = min(GRAV_T, en.src1 - base);
        if (live) {
            for (int q = 0; q < cnt; ++q) {
                const int jj = base + q;
                double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
                double d2 = dx * dx + dy * dy + dz * dz;
                bool hi_is_j = jj > i;
                double s2 = hi_is_j ? ss2[q] : si2;
                double inv_s = hi_is_j ? sinv[q] : sinv_i;
                double inv_r = (d2 >= s2) ? rsqrt_refined(d2) : inv_s;
                double w = sg[q] * (inv_r * inv_r * inv_r);
                // the pair itself and coincident pairs (d < 1e-9) contribute nothing
                w = (jj == i || d2 < 1e-18) ? 0.0 : w;
                fx = fma(w, dx, fx);
                fy = fma(w, dy, fy);
                fz = fma(w, dz, fz);
            }
        }
        __syncthreads();
    }
    if (live) {
        double* p = partial + (size_t)blockIdx.x * 3 * GRAV_T + threadIdx.x;
        p[0] = fx;
        p[GRAV_T] = fy;
        p[2 * GRAV_T] = fz;
    }
}

__global__ __launch_bounds__(GRAV_T) void gravity_partial_f32c_kernel(const double* __restrict__ ch,
                                                                      const GravEntry* __restrict__ entries,
                                                                      double* __restrict__ partial) {
    __shared__ float4 sp[GRAV_T];
    __shared__ float2 sc[GRAV_T];
    const GravEntry en = entries[blockIdx.x];
    const int i = en.tgt0 + threadIdx.x;
    const bool live = (int)threadIdx.x < en.ntgt;
    float4 me = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 mc = make_float2(1.f, 1.f);
    if (live) {
        const double* c = ch + (size_t)i * 8;
        me = make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[6]);
        const float s = (float)c[7];
        mc = make_float2(s * s, 1.0f / s);
    }
    float fx = 0.f, fy = 0.f, fz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
    for (int base = en.src0; base < en.src1; base += GRAV_T) {
        const int j = base + threadIdx.x;
        if (j < en.src1) {
            const double* c = ch + (size_t)j * 8;
            sp[threadIdx.x] = make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[6]);
            const float s = (float)c[7];
            sc[threadIdx.x] = make_float2(s * s, 1.0f / s);
        }
        __syncthreads();
        const int cnt = min(GRAV_T, en.src1 - base);
        if (live) {
            for (int q = 0; q < cnt; ++q) {
                const int jj = base + q;
                const float4 o = sp[q];
                const float dx = o.x - me.x, dy = o.y - me.y, dz = o.z - me.z;
                const float d2 = dx * dx + dy * dy + dz * dz;
                const bool hi_is_j = jj > i;
                const float2 cr = hi_is_j ? sc[q] : mc;
                const float inv_r = (d2 >= cr.x) ? __builtin_amdgcn_rsqf(d2) : cr.y;
                float w = o.w * (inv_r * inv_r * inv_r);
                w = (jj == i || d2 < 1e-18f) ? 0.f : w;
                two_sum_acc(fx, ex, w * dx);
                two_sum_acc(fy, ey, w * dy);
                two_sum_acc(fz, ez, w * dz);
            }
        }
        __syncthreads();
    }
    if (live) {
        double* p = partial + (size_t)blockIdx.x * 3 * GRAV_T + threadIdx.x;
        p[0] = (double)fx + (double)ex;
        p[GRAV_T] = (double)fy + (double)ey;
        p[2 * GRAV_T] = (double)fz + (double)ez;
    }
}

__global__ __launch_bounds__(GRAV_T) void gravity_reduce_kernel(const double* __restrict__ ch,
                                                                const GravTile* __restrict__ tiles,
                                                                const double* __restrict__ partial,
                                                                double* __restrict__ ext) {
    const GravTile tl = tiles[blockIdx.x];
    if ((int)threadIdx.x >= tl.ntgt) return;
    const int i = tl.tgt0 + threadIdx.x;
    double f[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < tl.ne; ++c)
        for (int k = 0; k < 3; ++k) f[k] += partial[((size_t)(tl.e0 + c) * 3 + k) * GRAV_T + threadIdx.x];
    const double a = ch[(size_t)i * 8 + 6] * tl.g_inv;
    for (int k = 0; k < 3; ++k) f[k] *= a;
    double lx = ch[(size_t)i * 8 + 3], ly = ch[(size_t)i * 8 + 4], lz = ch[(size_t)i * 8 + 5];
    double* o = ext + (size_t)i * 6;
    o[0] = f[0];
    o[1] = f[1];
    o[2] = f[2];
    o[3] = ly * f[2] - lz * f[1];
    o[4] = lz * f[0] - lx * f[2];
    o[5] = lx * f[1] - ly * f[0];
}

// Bodies of worlds without gravity get a zero wrench (a fill kernel, so a step
// can be captured in a graph).
__global__ __launch_bounds__(GRAV_T) void zero_wrench_kernel(double* __restrict__ ext, int b0, int n) {
    const int t = blockIdx.x * GRAV_T + threadIdx.x;
    if (t < n * 6) ext[(size_t)b0 * 6 + t] = 0.0;
}

void gravity_add_world(std::vector<GravEntry>& entries, std::vector<GravTile>& tiles, int b0, int n, double G,
                       int compute_units) {
    if (n <= 0) return;
    const int ntiles = (n + GRAV_T - 1) / GRAV_T;
    const int want = 4 * (compute_units > 0 ? compute_units : 256);
    int chunks = (want + ntiles - 1) / ntiles;
    const int max_chunks = (n + GRAV_MIN_CHUNK - 1) / GRAV_MIN_CHUNK;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    gravity_add_world_chunked(entries, tiles, b0, n, G, (n + chunks - 1) / chunks);
}

void gravity_add_world_chunked(std::vector<GravEntry>& entries, std::vector<GravTile>& tiles, int b0, int n, double G,
                               int chunk_len) {
    if (n <= 0) return;
    const int ntiles = (n + GRAV_T - 1) / GRAV_T;
    if (chunk_len < 1) chunk_len = n;
    const int chunks = (n + chunk_len - 1) / chunk_len;
    for (int t = 0; t < ntiles; ++t) {
        GravTile tl{};
        tl.tgt0 = b0 + t * GRAV_T;
        tl.ntgt = std::min(GRAV_T, n - t * GRAV_T);
        tl.e0 = (int)entries.size();
        tl.ne = chunks;
        tl.g_inv = 1.0 / G;
        for (int c = 0; c < chunks; ++c) {
            GravEntry e{};
            e.tgt0 = tl.tgt0;
            e.ntgt = tl.ntgt;
            e.src0 = b0 + c * chunk_len;
            e.src1 = b0 + std::min(n, (c + 1) * chunk_len);
            entries.push_back(e);
        }
        tiles.push_back(tl);
    }
}

void launch_gravity(const double* charges, const GravEntry* entries, int n_entries, const GravTile* tiles,
                    int n_tiles, int precision, double* partial, double* ext, hipStream_t stream) {
    if (n_entries <= 0 || n_tiles <= 0) return;
    if (precision == 1)
        hipLaunchKernelGGL(gravity_partial_f32c_kernel, dim3(n_entries), dim3(GRAV_T), 0, stream, charges, entries,
                           partial);
    else
        hipLaunchKernelGGL(gravity_partial_f64_kernel, dim3(n_entries), dim3(GRAV_T), 0, stream, charges, entries,
                           partial);
    hipLaunchKernelGGL(gravity_reduce_kernel, dim3(n_tiles), dim3(GRAV_T), 0, stream, charges, tiles, partial, ext);
}

void launch_zero_wrench(double* ext, int b0, int n, hipStream_t stream) {
    if (n <= 0) return;
    hipLaunchKernelGGL(zero_wrench_kernel, dim3((n * 6 + GRAV_T - 1) / GRAV_T), dim3(GRAV_T), 0, stream, ext, b0, n);
}

}  // namespace exg
