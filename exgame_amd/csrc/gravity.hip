// All-pairs gravity with the reference's size-core law (crates/newton/src/gravity.rs:110-146,
// 530-601): for the unordered pair (lo, hi), d = |c_hi − c_lo|, pairs closer than 1e-9
// are skipped, s = size of the HIGHER slot, and
//     g = (G m_lo/r)(G m_hi/r)/r,  r = d if d ≥ s else s,
// so the force on lo is g·(c_hi − c_lo)/G and the force on hi its negation — i.e. on
// body i, g·(c_j − c_i)/G for every j. Each target sums its sources in ascending
// slot order, the order of the reference's row walk, and takes its moment about
// its island origin from the summed force.
//
// Layout: sources are staged through LDS in tiles of one workgroup; each thread
// owns one target and keeps its accumulators in registers. The work is
// VALU-bound (sqrt + reciprocal per pair), so the kernel is launched with many
// workgroups to fill all 256 CUs.
#include <hip/hip_runtime.h>
#include "island.h"

namespace exg {

constexpr int GRAV_TILE = 256;

__global__ __launch_bounds__(GRAV_TILE) void gravity_f64_kernel(const double* __restrict__ ch, int n, double g_inv,
                                                                 double* __restrict__ ext) {
    __shared__ double sx[GRAV_TILE], sy[GRAV_TILE], sz[GRAV_TILE], sg[GRAV_TILE], ss[GRAV_TILE];
    const int i = blockIdx.x * GRAV_TILE + threadIdx.x;
    const bool live = i < n;
    double xi = 0, yi = 0, zi = 0, gi = 0, si = 1;
    if (live) {
        xi = ch[(size_t)i * 8 + 0];
        yi = ch[(size_t)i * 8 + 1];
        zi = ch[(size_t)i * 8 + 2];
        gi = ch[(size_t)i * 8 + 6];
        si = ch[(size_t)i * 8 + 7];
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int base = 0; base < n; base += GRAV_TILE) {
        int j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = ch[(size_t)j * 8 + 0];
            sy[threadIdx.x] = ch[(size_t)j * 8 + 1];
            sz[threadIdx.x] = ch[(size_t)j * 8 + 2];
            sg[threadIdx.x] = ch[(size_t)j * 8 + 6];
            ss[threadIdx.x] = ch[(size_t)j * 8 + 7];
        }
        __syncthreads();
        int cnt = min(GRAV_TILE, n - base);
        if (live) {
            for (int q = 0; q < cnt; ++q) {
                int jj = base + q;
                double dx = sx[q] - xi, dy = sy[q] - yi, dz = sz[q] - zi;
                double d = sqrt(dx * dx + dy * dy + dz * dz);
                if (jj == i || d < 1e-9) continue;
                double s_hi = (jj > i) ? ss[q] : si;
                double r = (d >= s_hi) ? d : s_hi;
                double g = ((gi / r) * (sg[q] / r)) / r * g_inv;
                fx += g * dx;
                fy += g * dy;
                fz += g * dz;
            }
        }
        __syncthreads();
    }
    if (live) {
        double lx = ch[(size_t)i * 8 + 3], ly = ch[(size_t)i * 8 + 4], lz = ch[(size_t)i * 8 + 5];
        double* o = ext + (size_t)i * 6;
        o[0] = fx;
        o[1] = fy;
        o[2] = fz;
        o[3] = ly * fz - lz * fy;
        o[4] = lz * fx - lx * fz;
        o[5] = lx * fy - ly * fx;
    }
}

void launch_gravity_f64(const double* charges, int n, double g_inv, double* ext, hipStream_t stream) {
    if (n <= 0) return;
    int blocks = (n + GRAV_TILE - 1) / GRAV_TILE;
    hipLaunchKernelGGL(gravity_f64_kernel, dim3(blocks), dim3(GRAV_TILE), 0, stream, charges, n, g_inv, ext);
}

}  // namespace exg
