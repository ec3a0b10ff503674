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
//
// Work split: a 2-D grid of (target tile of 256, source chunk). Each thread owns one
// target and walks its chunk's sources, staged through LDS in tiles of 256 together
// with s² and 1/s of each source; partial sums go to partial[chunk][axis][body]
// (coalesced) and a second kernel adds the chunks in ascending order, scales by the
// target's charge and forms the moment. The chunk count is chosen so the grid holds
// about four workgroups per CU, which fills all 256 CUs at a few thousand bodies.
//
// Per pair: one hardware reciprocal square root refined by two Newton steps
// (y ← y + y·(1 − d²y²)/2) replaces the reference's sqrt and three divisions; the
// result agrees with the correctly rounded form to a few ulp.
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

__global__ __launch_bounds__(GRAV_T) void gravity_partial_kernel(const double* __restrict__ ch, int n, int chunk_len,
                                                                 double* __restrict__ partial) {
    __shared__ double sx[GRAV_T], sy[GRAV_T], sz[GRAV_T], sg[GRAV_T], ss2[GRAV_T], sinv[GRAV_T];
    const int i = blockIdx.x * GRAV_T + threadIdx.x;
    const int chunk = blockIdx.y;
    const int c0 = chunk * chunk_len;
    const int c1 = min(n, c0 + chunk_len);
    const bool live = i < n;
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
    for (int base = c0; base < c1; base += GRAV_T) {
        const int j = base + threadIdx.x;
        if (j < c1) {
            sx[threadIdx.x] = ch[(size_t)j * 8 + 0];
            sy[threadIdx.x] = ch[(size_t)j * 8 + 1];
            sz[threadIdx.x] = ch[(size_t)j * 8 + 2];
            sg[threadIdx.x] = ch[(size_t)j * 8 + 6];
            double sj = ch[(size_t)j * 8 + 7];
            ss2[threadIdx.x] = sj * sj;
            sinv[threadIdx.x] = 1.0 / sj;
        }
        __syncthreads();
        const int cnt = min(GRAV_T, c1 - base);
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
        double* p = partial + (size_t)chunk * 3 * n;
        p[i] = fx;
        p[(size_t)n + i] = fy;
        p[(size_t)2 * n + i] = fz;
    }
}

__global__ __launch_bounds__(GRAV_T) void gravity_reduce_kernel(const double* __restrict__ ch, int n, int chunks,
                                                                double g_inv, const double* __restrict__ partial,
                                                                double* __restrict__ ext) {
    const int i = blockIdx.x * GRAV_T + threadIdx.x;
    if (i >= n) return;
    double f[3] = {0.0, 0.0, 0.0};
    for (int c = 0; c < chunks; ++c)
        for (int k = 0; k < 3; ++k) f[k] += partial[((size_t)c * 3 + k) * n + i];
    const double a = ch[(size_t)i * 8 + 6] * g_inv;
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

// ───── mixed precision: FP32 pair arithmetic, two-sum compensated FP32 accumulators ─────
// Each step the f64 charges are narrowed once to (x, y, z, G·m) and (s², 1/s) in
// f32; a pair costs one v_rsq_f32 and a dozen f32 operations, and each force
// component is accumulated as an unevaluated sum hi + lo: after every addition
// the rounding error of hi + x is recovered exactly (Knuth's two-sum) and added to
// lo. The per-chunk sums are widened to f64 for the ordered chunk reduction, the
// scaling by the target's charge and the moment, which stay f64.
__device__ inline void two_sum_acc(float& hi, float& lo, float x) {
    const float t = hi + x;
    const float bp = t - hi;
    const float e = (hi - (t - bp)) + (x - bp);
    hi = t;
    lo = lo + e;
}

__global__ __launch_bounds__(GRAV_T) void gravity_narrow_kernel(const double* __restrict__ ch, int n,
                                                                float4* __restrict__ pos_g,
                                                                float2* __restrict__ core) {
    const int i = blockIdx.x * GRAV_T + threadIdx.x;
    if (i >= n) return;
    const double* c = ch + (size_t)i * 8;
    pos_g[i] = make_float4((float)c[0], (float)c[1], (float)c[2], (float)c[6]);
    const float sz = (float)c[7];
    core[i] = make_float2(sz * sz, 1.0f / sz);
}

__global__ __launch_bounds__(GRAV_T) void gravity_partial_f32_kernel(const float4* __restrict__ pos_g,
                                                                     const float2* __restrict__ core, int n,
                                                                     int chunk_len, double* __restrict__ partial) {
    __shared__ float4 sp[GRAV_T];
    __shared__ float2 sc[GRAV_T];
    const int i = blockIdx.x * GRAV_T + threadIdx.x;
    const int chunk = blockIdx.y;
    const int c0 = chunk * chunk_len;
    const int c1 = min(n, c0 + chunk_len);
    const bool live = i < n;
    float4 me = make_float4(0.f, 0.f, 0.f, 0.f);
    float2 mc = make_float2(1.f, 1.f);
    if (live) {
        me = pos_g[i];
        mc = core[i];
    }
    float fx = 0.f, fy = 0.f, fz = 0.f, ex = 0.f, ey = 0.f, ez = 0.f;
    for (int base = c0; base < c1; base += GRAV_T) {
        const int j = base + threadIdx.x;
        if (j < c1) {
            sp[threadIdx.x] = pos_g[j];
            sc[threadIdx.x] = core[j];
        }
        __syncthreads();
        const int cnt = min(GRAV_T, c1 - base);
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
        double* p = partial + (size_t)chunk * 3 * n;
        p[i] = (double)fx + (double)ex;
        p[(size_t)n + i] = (double)fy + (double)ey;
        p[(size_t)2 * n + i] = (double)fz + (double)ez;
    }
}

GravityPlan gravity_plan(int n, int compute_units) {
    GravityPlan p{};
    if (n <= 0) return p;
    const int tiles = (n + GRAV_T - 1) / GRAV_T;
    const int want = 4 * (compute_units > 0 ? compute_units : 256);
    int chunks = (want + tiles - 1) / tiles;
    const int max_chunks = (n + GRAV_MIN_CHUNK - 1) / GRAV_MIN_CHUNK;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    p.chunk_len = (n + chunks - 1) / chunks;
    p.chunks = (n + p.chunk_len - 1) / p.chunk_len;
    p.tiles = tiles;
    p.partial_doubles = (size_t)p.chunks * 3 * n;
    return p;
}

void launch_gravity_f32(const double* charges, int n, double g_inv, const GravityPlan& plan, double* partial,
                        float* narrow, double* ext, hipStream_t stream) {
    if (n <= 0) return;
    float4* pos_g = reinterpret_cast<float4*>(narrow);
    float2* core = reinterpret_cast<float2*>(narrow + (size_t)4 * n);
    hipLaunchKernelGGL(gravity_narrow_kernel, dim3(plan.tiles), dim3(GRAV_T), 0, stream, charges, n, pos_g, core);
    hipLaunchKernelGGL(gravity_partial_f32_kernel, dim3(plan.tiles, plan.chunks), dim3(GRAV_T), 0, stream, pos_g,
                       core, n, plan.chunk_len, partial);
    hipLaunchKernelGGL(gravity_reduce_kernel, dim3(plan.tiles), dim3(GRAV_T), 0, stream, charges, n, plan.chunks,
                       g_inv, partial, ext);
}

void launch_gravity_f64(const double* charges, int n, double g_inv, const GravityPlan& plan, double* partial,
                        double* ext, hipStream_t stream) {
    if (n <= 0) return;
    hipLaunchKernelGGL(gravity_partial_kernel, dim3(plan.tiles, plan.chunks), dim3(GRAV_T), 0, stream, charges, n,
                       plan.chunk_len, partial);
    hipLaunchKernelGGL(gravity_reduce_kernel, dim3(plan.tiles), dim3(GRAV_T), 0, stream, charges, n, plan.chunks,
                       g_inv, partial, ext);
}

}  // namespace exg
