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
// f32 (see f32ds_block), a pair costs one v_rsq_f32 and about fifteen f32 operations,
// and the f32 sums of groups of 16 sources are added into f64 accumulators; the
// reductions, scaling and moment are f64.
#include <hip/hip_runtime.h>
#include "island.h"

namespace exg {

static_assert(GRAV_TILE == 64, "one target per lane of a wave");
constexpr int GRAV_WAVES = 4;     // source slices per workgroup
constexpr int GRAV_MIN_CHUNK = 64;

// The hardware estimate (≈ 2^-24 relative) refined by one third-order step,
// y = y0 + y0·e·(1/2 + 3/8·e) with e = 1 − x·y0²: within 2 ulp of 1/sqrt(x) over
// 1e-40 … 2e30 (the bound of two Newton steps, in five operations instead of eight).
__device__ inline double rsqrt_refined(double x) {
    const double y0 = __builtin_amdgcn_rsq(x);
    const double e = fma(-(x * y0), y0, 1.0);
    const double c = fma(0.375, e, 0.5);
    return fma(y0 * e, c, y0);
}

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

// The f32 source records of bodies [0, n): x_hi y_hi z_hi G·m | x_lo y_lo z_lo 1/s.
__global__ __launch_bounds__(256) void gravity_stage_f32_kernel(const double* __restrict__ ch, int n,
                                                                float4* __restrict__ stage) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= n) return;
    const double* c = ch + (size_t)j * 8;
    float4 h, l;
    split_ds(c[0], h.x, l.x);
    split_ds(c[1], h.y, l.y);
    split_ds(c[2], h.z, l.z);
    h.w = (float)c[6];
    l.w = (float)(1.0 / c[7]);
    stage[2 * (size_t)j] = h;
    stage[2 * (size_t)j + 1] = l;
}

// Sources are read from the staged records with scalar loads (every lane of a wave
// reads the same source), so each pair's source operands sit in scalar registers.
typedef float f4v __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(4))) f4v* ConstF4;

template <int MODE>
__device__ __forceinline__ void f32ds_span(ConstF4 src, int j0, int cnt, int i, const float4& mh, const float4& ml,
                                           float minv, double f[3]) {
    int q = 0;
    while (q < cnt) {
        const int qe = min(cnt, q + F32_GROUP);
        float a[2][3] = {{0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}};
        auto pair = [&](int k, float* acc) {
            const int jj = j0 + k;
            const f4v h = src[2 * jj], l = src[2 * jj + 1];
            const float dx = (h.x - mh.x) + (l.x - ml.x);
            const float dy = (h.y - mh.y) + (l.y - ml.y);
            const float dz = (h.z - mh.z) + (l.z - ml.z);
            const float d2 = dx * dx + dy * dy + dz * dz;
            float inv_s = minv;
            if (MODE == 1) inv_s = l.w;
            if (MODE == 2) inv_s = jj > i ? l.w : minv;
            const float inv_r = fminf(rsqrt_refined_f32(d2), inv_s);
            float w = h.w * (inv_r * inv_r * inv_r);
            w = ((MODE == 2 && jj == i) || d2 < 1e-18f) ? 0.f : w;
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
        q = qe;
    }
}

__global__ __launch_bounds__(64 * GRAV_WAVES) void gravity_partial_f32c_kernel(const double* __restrict__ ch,
                                                                               const GravEntry* __restrict__ entries,
                                                                               const float4* __restrict__ stage,
                                                                               double* __restrict__ partial) {
    const GravEntry en = entries[blockIdx.x];
    // the wave's index is wave-uniform: its slice bounds and loop counts live in scalar registers
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int i = en.tgt0 + lane;
    const bool live = lane < en.ntgt;
    float4 mh = make_float4(0.f, 0.f, 0.f, 0.f), ml = make_float4(0.f, 0.f, 0.f, 0.f);
    float minv = 1.f;
    if (live) {
        const float4 h = stage[2 * (size_t)i], l = stage[2 * (size_t)i + 1];
        mh = h;
        ml = l;
        minv = l.w;
    }
    const int len = en.src1 - en.src0, per = (len + GRAV_WAVES - 1) / GRAV_WAVES;
    const int s0 = en.src0 + min(len, w * per), s1 = en.src0 + min(len, (w + 1) * per);
    const int t_lo = en.tgt0, t_hi = en.tgt0 + en.ntgt - 1;
    ConstF4 src = (ConstF4)stage;
    double f[3] = {0.0, 0.0, 0.0};
    for (int base = s0; base < s1; base += 64) {
        const int cnt = min(64, s1 - base);
        if (base + cnt - 1 < t_lo) f32ds_span<0>(src, base, cnt, i, mh, ml, minv, f);
        else if (base > t_hi) f32ds_span<1>(src, base, cnt, i, mh, ml, minv, f);
        else f32ds_span<2>(src, base, cnt, i, mh, ml, minv, f);
    }
    write_partial(f[0], f[1], f[2], live, partial);
}

__global__ __launch_bounds__(GRAV_TILE) void gravity_reduce_kernel(const double* __restrict__ ch,
                                                                   const GravTile* __restrict__ tiles,
                                                                   const double* __restrict__ partial,
                                                                   double* __restrict__ ext) {
    const GravTile tl = tiles[blockIdx.x];
    const int lane = threadIdx.x;
    if (lane >= tl.ntgt) return;
    const int i = tl.tgt0 + lane;
    double* o = ext + (size_t)i * 6;
    if (tl.ne == 0) {  // a world without gravity
        for (int k = 0; k < 6; ++k) o[k] = 0.0;
        return;
    }
    double f[3] = {0.0, 0.0, 0.0};
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
    const double a = ch[(size_t)i * 8 + 6] * tl.g_inv;
    for (int k = 0; k < 3; ++k) f[k] *= a;
    const double lx = ch[(size_t)i * 8 + 3], ly = ch[(size_t)i * 8 + 4], lz = ch[(size_t)i * 8 + 5];
    o[0] = f[0];
    o[1] = f[1];
    o[2] = f[2];
    o[3] = ly * f[2] - lz * f[1];
    o[4] = lz * f[0] - lx * f[2];
    o[5] = lx * f[1] - ly * f[0];
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
                    hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
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
                 entries, (const float4*)stage, partial);
    } else
        launch_k(gravity_partial_f64_kernel, dim3(n_entries), dim3(64 * GRAV_WAVES), stream, start, nullptr, charges,
                 entries, partial);
    launch_k(gravity_reduce_kernel, dim3(n_tiles), dim3(GRAV_TILE), stream, nullptr, stop, charges, tiles,
             (const double*)partial, ext);
}

void launch_zero_wrench(double* ext, int b0, int n, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    if (n <= 0) return;
    launch_k(zero_wrench_kernel, dim3((n * 6 + ZERO_T - 1) / ZERO_T), dim3(ZERO_T), stream, start, stop, ext, b0, n);
}

}  // namespace exg
