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
// Work split: a table of entries, each one (tile of ≤ 64 targets, chunk of sources
// of the same world). A workgroup of four waves takes one entry: every wave holds
// the tile's 64 targets (one per lane) and walks a quarter of the chunk (its slice),
// staging 64 sources at a time in its own LDS rows with 1/s of each; source
// blocks wholly below or above the tile's slots skip the per-pair slot test. Each
// lane keeps two accumulators (even and odd sources, added at the end); the four
// slice sums are added in slice order through LDS and written to
// partial[entry][axis][lane]. A second table lists, per tile, its consecutive
// entries: the reduce kernel adds them in ascending chunk order, scales by the
// target's charge / G and forms the moment. A world's chunking depends only on its
// own body count (and the CU count), so its forces are bit-identical whatever other
// worlds share the launch.
//
// f64 pairs: one hardware reciprocal square root refined by a third-order step
// replaces the reference's sqrt and three divisions; the result is within 2 ulp of
// the correctly rounded form.
// f32 pairs (mixed precision): positions are staged as double-single hi + lo pairs of
// f32 (see f32ds_span), a pair costs one v_rsq_f32 and about fifteen f32 operations,
// and the f32 sums of groups of 16 sources are added into f64 accumulators; the
// reductions, scaling and moment are f64. Pairs closer than three core sizes of the
// target (the near field) are evaluated in f64 from the double-single records, and
// pairs within a tile's own 64 slots in f64 by the reduce kernel.
#include <hip/hip_runtime.h>
#include "island.h"

namespace exg {

static_assert(GRAV_TILE == 64, "one target per lane of a wave");
constexpr int GRAV_WAVES = 4;     // source slices per workgroup
constexpr int GRAV_MIN_CHUNK = 64;


// Wave-local staging: a wave's LDS rows are written and read by that wave only.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The tile's slice [s0, s1) of a chunk, for one wave: block position MODE 0 = every
// source below every target slot (the higher slot is the target), 1 = every source
// above (the source), 2 = mixed (tested per pair, the pair itself skipped).
// The core law r = max(d, s) enters as 1/r = min(1/d, 1/s): branch-free, and for a
// coincident pair (d = 0, whose refined 1/d is not a number) min returns 1/s.
template <int MODE>
__device__ __forceinline__ void f64_block(const double2* sxy, const double2* szg, const double* sinv, int cnt, int base,
                                          int i, double xi, double yi, double zi, double sinv_i, double f[2][3]) {
    auto pair = [&](int q, double* acc) {
        const int jj = base + q;
        const double2 a = sxy[q], b = szg[q];
        const double dx = a.x - xi, dy = a.y - yi, dz = b.x - zi;
        const double d2 = dx * dx + dy * dy + dz * dz;
        double inv_s = sinv_i;
        if (MODE == 1) inv_s = sinv[q];
        if (MODE == 2) inv_s = jj > i ? sinv[q] : sinv_i;
        const double inv_r = fmin(rsqrt_refined(d2), inv_s);
        double w = b.y * (inv_r * inv_r * inv_r);
        // the pair itself and coincident pairs (d < 1e-9) contribute nothing
        w = ((MODE == 2 && jj == i) || d2 < 1e-18) ? 0.0 : w;
        acc[0] = fma(w, dx, acc[0]);
        acc[1] = fma(w, dy, acc[1]);
        acc[2] = fma(w, dz, acc[2]);
    };
    int q = 0;
    for (; q + 2 <= cnt; q += 2) {
        pair(q, f[0]);
        pair(q + 1, f[1]);
    }
    if (q < cnt) pair(q, f[0]);
}

// f32 pairs on double-single positions: each coordinate is carried as hi + lo, two
// f32 values with hi = f32(x) and lo = f32(x − hi), so a difference is
// (x_hi − y_hi) + (x_lo − y_lo) — exact in its leading part for close pairs (the hi
// parts agree in their leading bits), rounded once — instead of the difference of
// two positions rounded to f32 (which loses the bits below 2^-24·|x|, a relative error
// of |x|/d·6e-8 on a pair at distance d). 1/d is the hardware estimate refined by one
// Newton step (its error enters 1/d³ three times over), the core law is
// min(1/d, 1/s), and the force is summed in f32 over GROUP sources at a time (two
// interleaved sums), each group's sum added to f64 accumulators.
constexpr int F32_GROUP = 16;
__device__ inline float rsqrt_refined_f32(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float e = fmaf(-(x * y), y, 1.0f);
    return fmaf(y * e, 0.5f, y);
}
// slice sums added in slice order; wave 0 writes the entry's partial
__device__ inline void write_partial(double fx, double fy, double fz, bool live, double* __restrict__ partial) {
    __shared__ double red[3][GRAV_WAVES - 1][GRAV_TILE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w > 0) {
        red[0][w - 1][lane] = fx;
        red[1][w - 1][lane] = fy;
        red[2][w - 1][lane] = fz;
    }
    __syncthreads();
    if (w == 0 && live) {
        for (int k = 0; k < GRAV_WAVES - 1; ++k) {
            fx += red[0][k][lane];
            fy += red[1][k][lane];
            fz += red[2][k][lane];
        }
        double* p = partial + (size_t)blockIdx.x * GRAV_PARTIAL_PER_ENTRY + lane;
        p[0] = fx;
        p[GRAV_TILE] = fy;
        p[2 * GRAV_TILE] = fz;
    }
}

__global__ __launch_bounds__(64 * GRAV_WAVES) void gravity_partial_f64_kernel(const double* __restrict__ ch,
                                                                              const GravEntry* __restrict__ entries,
                                                                              double* __restrict__ partial) {
    __shared__ double2 sxy[GRAV_WAVES][64], szg[GRAV_WAVES][64];
    __shared__ double sinv[GRAV_WAVES][64];
    const GravEntry en = entries[blockIdx.x];
    // the wave's index is wave-uniform: its slice bounds and loop counts live in scalar registers
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int i = en.tgt0 + lane;
    const bool live = lane < en.ntgt;
    double xi = 0.0, yi = 0.0, zi = 0.0, sinv_i = 1.0;
    if (live) {
        xi = ch[(size_t)i * 8 + 0];
        yi = ch[(size_t)i * 8 + 1];
        zi = ch[(size_t)i * 8 + 2];
        sinv_i = 1.0 / ch[(size_t)i * 8 + 7];
    }
    const int len = en.src1 - en.src0, per = (len + GRAV_WAVES - 1) / GRAV_WAVES;
    const int s0 = en.src0 + min(len, w * per), s1 = en.src0 + min(len, (w + 1) * per);
    const int t_lo = en.tgt0, t_hi = en.tgt0 + en.ntgt - 1;
    double f[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
    for (int base = s0; base < s1; base += 64) {
        const int j = base + lane;
        if (j < s1) {
            const double* c = ch + (size_t)j * 8;
            sxy[w][lane] = make_double2(c[0], c[1]);
            szg[w][lane] = make_double2(c[2], c[6]);
            sinv[w][lane] = 1.0 / c[7];
        }
        wave_lds_sync();
        const int cnt = min(64, s1 - base);
        if (base + cnt - 1 < t_lo) f64_block<0>(sxy[w], szg[w], sinv[w], cnt, base, i, xi, yi, zi, sinv_i, f);
        else if (base > t_hi) f64_block<1>(sxy[w], szg[w], sinv[w], cnt, base, i, xi, yi, zi, sinv_i, f);
        else f64_block<2>(sxy[w], szg[w], sinv[w], cnt, base, i, xi, yi, zi, sinv_i, f);
        wave_lds_sync();
    }
    write_partial(f[0][0] + f[1][0], f[0][1] + f[1][1], f[0][2] + f[1][2], live, partial);
}

// hi + lo split of a coordinate (lo = f32 of the remainder)
__device__ inline void split_ds(double x, float& hi, float& lo) {
    hi = (float)x;
    lo = (float)(x - (double)hi);
}

// The f32 source records of bodies [0, n): stage[2j] = x_hi y_hi z_hi G·m_hi,
// stage[2j+1] = x_lo y_lo z_lo (1/s)_hi, and after the 2n records of the pair loop,
// stage[2n + j] = G·m_lo (1/s)_lo 0 0, read only for near pairs.
__global__ __launch_bounds__(256) void gravity_stage_f32_kernel(const double* __restrict__ ch, int n,
                                                                float4* __restrict__ stage) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double* c = ch + (size_t)j * 8;
    float4 h, l, e;
    split_ds(c[0], h.x, l.x);
    split_ds(c[1], h.y, l.y);
    split_ds(c[2], h.z, l.z);
    split_ds(c[6], h.w, e.x);
    split_ds(1.0 / c[7], l.w, e.y);
    e.z = 0.f;
    e.w = 0.f;
    stage[2 * (size_t)j] = h;
    stage[2 * (size_t)j + 1] = l;
    stage[2 * (size_t)n + j] = e;
}

// Sources are read from the staged records with scalar loads (every lane of a wave
// reads the same source), so each pair's source operands sit in scalar registers.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4v* ConstF4;

// The near field. A pair closer than three core sizes of its target (d² < 9 s_i², on
// the f32 distance) is evaluated in f64 from the double-single records (positions,
// G·m and 1/s to 2^-48) instead of in f32: the strongest pulls carry the largest f32
// rounding. The pairs among a tile's own 64 slots are all left to the reduce kernel, in
// f64: an island's bodies hold consecutive slots, and the f32 forces of its bodies on
// each other would not cancel in the island's net force. Elsewhere near pairs are rare
// (one in twenty thousand in the 4,096-body chains world, where they and the island
// pairs hold three quarters of the f32 error of the islands' net forces), so the pair
// loop stays branch-free: it skips them and notes, per lane, the least d² − 9 s_i² of
// the group, and a group where some lane went below zero is passed over again for its
// near pairs.
constexpr float F32_NEAR_CORES2 = 9.f;

struct F32Target {
    float4 mh, ml;              // the target's records: x y z hi | G·m ; x y z lo | 1/s
    float near2;                // 9 s², 0 for a lane without a target
    double x, y, z, sinv;       // position and 1/s, hi + lo
};

// The f32 difference and squared distance of a pair, written out with explicit fused
// multiply-adds so that the pair loop and the near-field pass compute the same bits.
__device__ __forceinline__ float f32ds_d2(const f4v& h, const f4v& l, const F32Target& t, float& dx, float& dy,
                                          float& dz) {
    dx = (h.x - t.mh.x) + (l.x - t.ml.x);
    dy = (h.y - t.mh.y) + (l.y - t.ml.y);
    dz = (h.z - t.mh.z) + (l.z - t.ml.z);
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

// The f64 pair (source jj on target i) from the double-single records.
template <int MODE>
__device__ __forceinline__ void f64_near_pair(int jj, int i, const f4v& h, const f4v& l, const f4v& e,
                                              const F32Target& t, double f[3]) {
    const double dx = ((double)h.x + (double)l.x) - t.x;
    const double dy = ((double)h.y + (double)l.y) - t.y;
    const double dz = ((double)h.z + (double)l.z) - t.z;
    const double d2 = dx * dx + dy * dy + dz * dz;
    double inv_s = t.sinv;
    if (MODE == 1) inv_s = (double)l.w + (double)e.y;
    const double inv_r = fmin(rsqrt_refined(d2), inv_s);
    const double w = d2 < 1e-18 ? 0.0 : ((double)h.w + (double)e.x) * (inv_r * inv_r * inv_r);
    f[0] = fma(w, dx, f[0]);
    f[1] = fma(w, dy, f[1]);
    f[2] = fma(w, dz, f[2]);
}

// MODE 0: every source below every target (the higher slot, whose core size counts,
// is the target); MODE 1: every source above.
template <int MODE>
__device__ __forceinline__ void f32ds_span(ConstF4 src, ConstF4 src_e, int j0, int cnt, int i, const F32Target& t,
                                           double f[3]) {
    static_assert(MODE == 0 || MODE == 1, "the tile's own slots are summed by the reduce kernel");
    int q = 0;
    while (q < cnt) {
        const int qe = min(cnt, q + F32_GROUP);
        float a[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        // least d² − 9 s_i² of the group, in a vector register (a running lane mask
        // would take scalar registers the batched source loads need)
        float near_margin = 1.f;
        auto pair = [&](int k, float* acc) {
            const int jj = j0 + k;
            const f4v h = src[2 * jj], l = src[2 * jj + 1];
            float dx, dy, dz;
            const float d2 = f32ds_d2(h, l, t, dx, dy, dz);
            const float inv_r = fminf(rsqrt_refined_f32(d2), MODE == 1 ? l.w : t.ml.w);
            float w = h.w * (inv_r * inv_r * inv_r);
            w = d2 < t.near2 ? 0.f : w;          // near (and coincident) pairs: not here
            near_margin = fminf(near_margin, d2 - t.near2);
            acc[0] = fmaf(w, dx, acc[0]);
            acc[1] = fmaf(w, dy, acc[1]);
            acc[2] = fmaf(w, dz, acc[2]);
        };
        int k = q;
#pragma unroll 4
        for (; k + 2 <= qe; k += 2) {
            pair(k, a[0]);
            pair(k + 1, a[1]);
        }
        if (k < qe) pair(k, a[0]);
        for (int c = 0; c < 3; ++c) f[c] += (double)a[0][c] + (double)a[1][c];
        if (near_margin < 0.f) {   // rare: the group's near pairs in f64
            // four sources at a time, loaded and tested before any branch so that a
            // batch's loads are in flight together
            for (int kk = q; kk < qe; kk += 4) {
                f4v h[4], l[4];
                bool nearp[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int jj = j0 + min(kk + u, qe - 1);
                    h[u] = src[2 * jj];
                    l[u] = src[2 * jj + 1];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    float dx, dy, dz;
                    const float d2 = f32ds_d2(h[u], l[u], t, dx, dy, dz);
                    nearp[u] = kk + u < qe && d2 < t.near2;
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    if (nearp[u]) {
                        const f4v e = src_e[j0 + kk + u];
                        f64_near_pair<MODE>(j0 + kk + u, i, h[u], l[u], e, t, f);
                    }
            }
        }
        q = qe;
    }
}

__global__ __launch_bounds__(64 * GRAV_WAVES) void gravity_partial_f32c_kernel(const double* __restrict__ ch,
                                                                               const GravEntry* __restrict__ entries,
                                                                               const float4* __restrict__ stage,
                                                                               int n_stage,
                                                                               double* __restrict__ partial) {
    const GravEntry en = entries[blockIdx.x];
    // the wave's index is wave-uniform: its slice bounds and loop counts live in scalar registers
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int i = en.tgt0 + lane;
    const bool live = lane < en.ntgt;
    F32Target t;
    t.mh = make_float4(0.f, 0.f, 0.f, 0.f);
    t.ml = make_float4(0.f, 0.f, 0.f, 1.f);
    t.near2 = 0.f;
    t.x = t.y = t.z = 0.0;
    t.sinv = 1.0;
    if (live) {
        const float4 h = stage[2 * (size_t)i], l = stage[2 * (size_t)i + 1], e = stage[2 * (size_t)n_stage + i];
        t.mh = h;
        t.ml = l;
        t.x = (double)h.x + (double)l.x;
        t.y = (double)h.y + (double)l.y;
        t.z = (double)h.z + (double)l.z;
        t.sinv = (double)l.w + (double)e.y;
        const float sz = 1.f / l.w;
        t.near2 = F32_NEAR_CORES2 * sz * sz;
    }
    const int len = en.src1 - en.src0, per = (len + GRAV_WAVES - 1) / GRAV_WAVES;
    const int s0 = en.src0 + min(len, w * per), s1 = en.src0 + min(len, (w + 1) * per);
    const int t_lo = en.tgt0, t_hi = en.tgt0 + en.ntgt - 1;
    ConstF4 src = (ConstF4)stage, src_e = src + 2 * (size_t)n_stage;
    double f[3] = {0.0, 0.0, 0.0};
    // sources in blocks of 64, each clipped around the tile's own slots [t_lo, t_hi]
    // (those pairs are the reduce kernel's)
    for (int base = s0; base < s1; base += 64) {
        const int end = min(base + 64, s1);
        const int below = min(end, t_lo);
        if (base < below) f32ds_span<0>(src, src_e, base, below - base, i, t, f);
        const int above = max(base, t_hi + 1);
        if (above < end) f32ds_span<1>(src, src_e, above, end - above, i, t, f);
    }
    write_partial(f[0], f[1], f[2], live, partial);
}

// A tile's entries added in ascending chunk order (one target per lane).
__device__ inline void sum_partials(const GravTile& tl, const double* __restrict__ partial, int lane, double f[3]) {
    int c = 0;
    for (; c + 4 <= tl.ne; c += 4) {  // loads issued four chunks at a time, added in order
        double v[4][3];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < 3; ++k)
                v[u][k] = partial[(size_t)(tl.e0 + c + u) * GRAV_PARTIAL_PER_ENTRY + k * GRAV_TILE + lane];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int k = 0; k < 3; ++k) f[k] += v[u][k];
    }
    for (; c < tl.ne; ++c)
        for (int k = 0; k < 3; ++k) f[k] += partial[(size_t)(tl.e0 + c) * GRAV_PARTIAL_PER_ENTRY + k * GRAV_TILE + lane];
}

// Scale by the target's charge / G and write the wrench with the moment local × F.
__device__ inline void write_wrench(const double* __restrict__ ch, const GravTile& tl, int i, double f[3],
                                    double* __restrict__ ext) {
    gravity_wrench(ch + (size_t)i * 8, tl.g_inv, f, ext + (size_t)i * 6);
}

__global__ __launch_bounds__(GRAV_TILE) void gravity_reduce_kernel(const double* __restrict__ ch,
                                                                   const GravTile* __restrict__ tiles,
                                                                   const double* __restrict__ partial,
                                                                   double* __restrict__ ext) {
    const GravTile tl = tiles[blockIdx.x];
    const int lane = threadIdx.x;
    if (lane >= tl.ntgt) return;
    const int i = tl.tgt0 + lane;
    if (tl.ne == 0) {  // a world without gravity
        for (int k = 0; k < 6; ++k) ext[(size_t)i * 6 + k] = 0.0;
        return;
    }
    double f[3] = {0.0, 0.0, 0.0};
    sum_partials(tl, partial, lane, f);
    write_wrench(ch, tl, i, f, ext);
}

// The f32 pairs' reduce: the entries in ascending chunk order, then the pairs among the
// tile's own slots (which the f32 pair kernel leaves out) in f64 with the f64 kernel's
// pair arithmetic. Wave 0 sums the entries while waves 1-3 take the tile's sources
// [0, 21), [21, 42), [42, 64) from LDS; their sums are added in wave order.
__global__ __launch_bounds__(GRAV_TILE * GRAV_WAVES) void gravity_reduce_f32c_kernel(
    const double* __restrict__ ch, const GravTile* __restrict__ tiles, const double* __restrict__ partial,
    double* __restrict__ ext) {
    static_assert(GRAV_WAVES == 4 && GRAV_TILE == 64, "three waves of tile sources");
    __shared__ double tc[GRAV_TILE][5];                  // x y z G·m 1/s of the tile's bodies
    __shared__ double dsum[GRAV_WAVES - 1][3][GRAV_TILE];
    const GravTile tl = tiles[blockIdx.x];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int i = tl.tgt0 + lane;
    if (tl.ne == 0) {  // a world without gravity
        if (w == 0 && lane < tl.ntgt)
            for (int k = 0; k < 6; ++k) ext[(size_t)i * 6 + k] = 0.0;
        return;
    }
    if (w == 1 && lane < tl.ntgt) {
        const double* c = ch + (size_t)i * 8;
        tc[lane][0] = c[0];
        tc[lane][1] = c[1];
        tc[lane][2] = c[2];
        tc[lane][3] = c[6];
        tc[lane][4] = 1.0 / c[7];
    }
    double f[3] = {0.0, 0.0, 0.0};
    if (w == 0 && lane < tl.ntgt) sum_partials(tl, partial, lane, f);
    __syncthreads();
    if (w > 0) {
        double g[3] = {0.0, 0.0, 0.0};
        if (lane < tl.ntgt) {
            const double xi = tc[lane][0], yi = tc[lane][1], zi = tc[lane][2], sinv_i = tc[lane][4];
            const int ja = (w - 1) * 21, jb = w == 3 ? tl.ntgt : min(w * 21, tl.ntgt);
#pragma unroll 4
            for (int j = ja; j < jb; ++j) {
                const double dx = tc[j][0] - xi, dy = tc[j][1] - yi, dz = tc[j][2] - zi;
                const double d2 = dx * dx + dy * dy + dz * dz;
                const double inv_s = j > lane ? tc[j][4] : sinv_i;
                const double inv_r = fmin(rsqrt_refined(d2), inv_s);
                double wt = tc[j][3] * (inv_r * inv_r * inv_r);
                wt = (j == lane || d2 < 1e-18) ? 0.0 : wt;
                g[0] = fma(wt, dx, g[0]);
                g[1] = fma(wt, dy, g[1]);
                g[2] = fma(wt, dz, g[2]);
            }
        }
        for (int k = 0; k < 3; ++k) dsum[w - 1][k][lane] = g[k];
    }
    __syncthreads();
    if (w != 0 || lane >= tl.ntgt) return;
    for (int v = 0; v < GRAV_WAVES - 1; ++v)
        for (int k = 0; k < 3; ++k) f[k] += dsum[v][k][lane];
    write_wrench(ch, tl, i, f, ext);
}

// Bodies of worlds without gravity get a zero wrench (a fill kernel, so a step
// can be captured in a graph).
constexpr int ZERO_T = 256;
__global__ __launch_bounds__(ZERO_T) void zero_wrench_kernel(double* __restrict__ ext, int b0, int n) {
    const int t = blockIdx.x * ZERO_T + threadIdx.x;
    if (t < n * 6) ext[(size_t)b0 * 6 + t] = 0.0;
}

int gravity_chunk_len(int n, int compute_units) {
    if (n <= 0) return 1;
    // enough workgroups to fill the chip (4 per CU), chunks of at most 1,024 sources
    // (long chunks leave the later waves idle), at most 64 workgroups per CU (the
    // partial sums stay small), at least 64 sources per chunk
    const int ntiles = (n + GRAV_TILE - 1) / GRAV_TILE;
    const int cus = compute_units > 0 ? compute_units : 256;
    int chunks = (4 * cus + ntiles - 1) / ntiles;
    chunks = std::max(chunks, (n + 1023) / 1024);
    chunks = std::min(chunks, std::max(1, 64 * cus / ntiles));
    const int max_chunks = (n + GRAV_MIN_CHUNK - 1) / GRAV_MIN_CHUNK;
    if (chunks > max_chunks) chunks = max_chunks;
    if (chunks < 1) chunks = 1;
    return (n + chunks - 1) / chunks;
}

void gravity_add_world(std::vector<GravEntry>& entries, std::vector<GravTile>& tiles, int b0, int n, double G,
                       int compute_units) {
    if (n <= 0) return;
    gravity_add_world_chunked(entries, tiles, b0, n, G, gravity_chunk_len(n, compute_units));
}

void gravity_add_world_chunked(std::vector<GravEntry>& entries, std::vector<GravTile>& tiles, int b0, int n, double G,
                               int chunk_len) {
    if (n <= 0) return;
    const int ntiles = (n + GRAV_TILE - 1) / GRAV_TILE;
    if (chunk_len < 1) chunk_len = n;
    const int chunks = (n + chunk_len - 1) / chunk_len;
    for (int t = 0; t < ntiles; ++t) {
        GravTile tl{};
        tl.tgt0 = b0 + t * GRAV_TILE;
        tl.ntgt = std::min(GRAV_TILE, n - t * GRAV_TILE);
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
                    int n_tiles, int precision, double* partial, double* ext, void* stage, int n_bodies,
                    hipStream_t stream, hipEvent_t start, hipEvent_t stop, bool reduce) {
    if (n_tiles <= 0) return;
    // tiles without entries (worlds without gravity) still get their zero wrench from
    // the reduce, also when no world of the launch has gravity
    if (n_entries <= 0) {
        launch_k(gravity_reduce_kernel, dim3(n_tiles), dim3(GRAV_TILE), stream, start, stop, charges, tiles,
                 (const double*)partial, ext);
        return;
    }
    if (precision == 1) {
        launch_k(gravity_stage_f32_kernel, dim3((n_bodies + 255) / 256), dim3(256), stream, start, nullptr, charges,
                 n_bodies, (float4*)stage);
        launch_k(gravity_partial_f32c_kernel, dim3(n_entries), dim3(64 * GRAV_WAVES), stream, nullptr, nullptr, charges,
                 entries, (const float4*)stage, n_bodies, partial);
    } else {
        launch_k(gravity_partial_f64_kernel, dim3(n_entries), dim3(64 * GRAV_WAVES), stream, start,
                 reduce ? nullptr : stop, charges, entries, partial);
        if (!reduce) return;   // the islands add the partial sums
    }
    if (precision == 1)
        launch_k(gravity_reduce_f32c_kernel, dim3(n_tiles), dim3(GRAV_TILE * GRAV_WAVES), stream, nullptr, stop, charges,
                 tiles, (const double*)partial, ext);
    else
        launch_k(gravity_reduce_kernel, dim3(n_tiles), dim3(GRAV_TILE), stream, nullptr, stop, charges, tiles,
                 (const double*)partial, ext);
}

void launch_zero_wrench(double* ext, int b0, int n, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    if (n <= 0) return;
    launch_k(zero_wrench_kernel, dim3((n * 6 + ZERO_T - 1) / ZERO_T), dim3(ZERO_T), stream, start, stop, ext, b0, n);
}

}  // namespace exg
