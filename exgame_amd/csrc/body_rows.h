// body_post (crates/newton/src/functions/src/lib.rs:170-247) shared by an aligned
// group of eight lanes: lane r (< 6; lanes 6 and 7 mirror row 5) forms row r of the
// world mass block Ad*_M ∘ I ∘ Ad_{M⁻¹}, each entry as body_post forms it, and the
// momentum P = block·V gathers its rows from the group, so every lane of the group
// ends with the same residual and self-scale. Device code: uses lane shuffles.
#pragma once
#include <hip/hip_runtime.h>
#include "physics_dev.h"

namespace exg {

// `v` as held by lane j of this lane's aligned group of eight. With j known at compile
// time (the unrolled row loops) a ds_swizzle in bitmask mode broadcasts it — lane
// ((lane & 0x18) | j) within each 32-lane half, no address to compute — otherwise a
// general shuffle.
template <int J>
__device__ __forceinline__ double group8_const(double v) {
    constexpr int pattern = 0x18 | (J << 5);   // and_mask 0b11000, or_mask J, xor_mask 0
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(bits & 0xffffffffll), pattern);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(bits >> 32), pattern);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// `v` from the lane whose index differs from this lane's by `Off` (a power of two
// below 32): a ds_swizzle in bitmask mode (xor_mask Off) on the two halves.
template <int Off>
__device__ __forceinline__ double xor_lane(double v) {
    static_assert(Off > 0 && Off < 32 && (Off & (Off - 1)) == 0, "xor offset within a 32-lane half");
    constexpr int pattern = 0x1f | (Off << 10);   // and_mask 0b11111, or_mask 0, xor_mask Off
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_swizzle((int)(bits & 0xffffffffll), pattern);
    const int hi = __builtin_amdgcn_ds_swizzle((int)(bits >> 32), pattern);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// Butterfly sum over the 64 lanes of a wave, every lane ending with the same value
// (partners at distance 32, 16, …, 1 in that order).
__device__ __forceinline__ double wave_sum(double v) {
    v += __shfl_xor(v, 32, 64);
    v += xor_lane<16>(v);
    v += xor_lane<8>(v);
    v += xor_lane<4>(v);
    v += xor_lane<2>(v);
    v += xor_lane<1>(v);
    return v;
}
__device__ __forceinline__ double group8(double v, int j) {
    if (__builtin_constant_p(j)) {
        switch (j) {
            case 0: return group8_const<0>(v);
            case 1: return group8_const<1>(v);
            case 2: return group8_const<2>(v);
            case 3: return group8_const<3>(v);
            case 4: return group8_const<4>(v);
            case 5: return group8_const<5>(v);
            case 6: return group8_const<6>(v);
            case 7: return group8_const<7>(v);
            default: break;
        }
    }
    return __shfl(v, j, 8);
}

// entry `row` of six values, by selects on the values (passed by value, so the
// selects cannot become a dynamically indexed load from private memory)
__device__ __forceinline__ double pick6(double v0, double v1, double v2, double v3, double v4, double v5, int row) {
    double out = v0;
    out = row == 1 ? v1 : out;
    out = row == 2 ? v2 : out;
    out = row == 3 ? v3 : out;
    out = row == 4 ? v4 : out;
    out = row == 5 ? v5 : out;
    return out;
}

// The first part of body_post_row, which does not depend on the applied wrench: row
// `row` of the world mass block and the momentum P = block·V (every lane of the group
// ends with all six entries).
__device__ __forceinline__ void body_post_row_momentum(const double pose_mid[8], const double vel[6], double mass,
                                                       const double t9[9], int row, double brow[6], double p[6]) {
    double inv[8], adi[6][6], cad[6][6];
    motor_rev(pose_mid, inv);
    ad_matrix(inv, adi);
    coad_matrix(pose_mid, cad);
    double cr[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) cr[k] = pick6(cad[0][k], cad[1][k], cad[2][k], cad[3][k], cad[4][k], cad[5][k], row);
#pragma unroll
    for (int col = 0; col < 6; ++col) {
        double x[6];
#pragma unroll
        for (int i = 0; i < 3; ++i) x[i] = adi[i][col] * mass;
#pragma unroll
        for (int r = 0; r < 3; ++r)
            x[3 + r] = t9[3 * r] * adi[3][col] + t9[3 * r + 1] * adi[4][col] + t9[3 * r + 2] * adi[5][col];
        double acc = cr[0] * x[0];
#pragma unroll
        for (int k = 1; k < 6; ++k) acc = acc + cr[k] * x[k];
        brow[col] = acc;
    }
    double pr = 0.0;
#pragma unroll
    for (int c = 0; c < 6; ++c) pr = pr + brow[c] * vel[c];
#pragma unroll
    for (int r = 0; r < 6; ++r) p[r] = group8(pr, r);
}
// The rest: residual −(P − snap − half·total) and the self-scale.
__device__ __forceinline__ void body_post_row_residual(const double p[6], const double snap[6], const double total[6],
                                                       double half, double floor2, double rhs[6], double scale2[2]) {
    double hh[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
        hh[r] = total[r] * half;
        rhs[r] = -(p[r] - snap[r] - hh[r]);
    }
#pragma unroll
    for (int o = 0; o < 2; ++o) {
        const int b = 3 * o;
        const double sp_ = p[b] * p[b] + p[b + 1] * p[b + 1] + p[b + 2] * p[b + 2];
        const double sn = snap[b] * snap[b] + snap[b + 1] * snap[b + 1] + snap[b + 2] * snap[b + 2];
        const double sh = hh[b] * hh[b] + hh[b + 1] * hh[b + 1] + hh[b + 2] * hh[b + 2];
        scale2[o] = fmax(fmax(fmax(sp_, sn), sh), floor2);
    }
}

__device__ __forceinline__ void body_post_row(const double pose_mid[8], const double vel[6], double mass,
                                              const double t9[9], const double snap[6], const double total[6],
                                              double half, double floor2, int row, double brow[6], double rhs[6],
                                              double scale2[2]) {
    double p[6];
    body_post_row_momentum(pose_mid, vel, mass, t9, row, brow, p);
    body_post_row_residual(p, snap, total, half, floor2, rhs, scale2);
}

}  // namespace exg
