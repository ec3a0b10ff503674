// Pairwise gravity that allocates nothing: for callers that cannot call hipMalloc
// (code under HIP graph capture, tight serving loops, harnesses that forbid it).
//
// It computes what exg_gravity_accumulate_precision computes, with the same work
// split and kernels, so the two agree bit for bit; but the caller passes device
// buffers for the charges and the wrenches, and the work table and the per-chunk
// partial sums (and the per-source records of the f32 kernel) live in `scratch`, a
// fixed device array of this library (one copy
// per GPU, held until the process ends). The array's device address is taken once
// per device with hipGetSymbolAddress and only that address is ever handed to
// kernels. Every launch rewrites the parts of `scratch` it reads (the table fill
// kernel writes every table entry; the partial kernel writes every partial sum the
// reduce reads), so nothing needs clearing between calls. Calls that overlap on one
// GPU share `scratch` and race; serialise them (one stream, or events).
#include <hip/hip_runtime.h>
#include <mutex>
#include <string>
#include "../../include/exgame.h"
#include "island.h"

namespace exg {
namespace {

constexpr int NOALLOC_MAX_BODIES = EXG_GRAVITY_NOALLOC_MAX_BODIES;
constexpr int NOALLOC_MAX_TILES = NOALLOC_MAX_BODIES / GRAV_TILE;
constexpr int NOALLOC_MAX_ENTRIES = 16384;   // the work split's cap: 64 workgroups per CU of 256
constexpr size_t PARTIAL_BYTES = (size_t)NOALLOC_MAX_ENTRIES * GRAV_PARTIAL_PER_ENTRY * sizeof(double);
constexpr size_t ENTRY_BYTES = (size_t)NOALLOC_MAX_ENTRIES * sizeof(GravEntry);
constexpr size_t TILE_BYTES = (size_t)NOALLOC_MAX_TILES * sizeof(GravTile);
constexpr size_t STAGE_BYTES = (size_t)NOALLOC_MAX_BODIES * GRAV_STAGE_BYTES;
constexpr size_t SCRATCH_BYTES = PARTIAL_BYTES + STAGE_BYTES + ENTRY_BYTES + TILE_BYTES;
static_assert(PARTIAL_BYTES % 256 == 0 && STAGE_BYTES % 256 == 0 && ENTRY_BYTES % 256 == 0,
              "scratch sections stay 256-byte aligned");
static_assert(SCRATCH_BYTES == EXG_GRAVITY_NOALLOC_SCRATCH_BYTES, "the header states the scratch size");

// partial sums | per-source records | work entries | work tiles
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up once per device.
int scratch_of_current_device(char** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error(EXG_ERR_BACKEND, (std::string("exg_gravity_accumulate_noalloc: ") +
                                                            hipGetErrorString(e)).c_str());
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_error(EXG_ERR_UNSUPPORTED, "exg_gravity_accumulate_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return set_error(EXG_ERR_BACKEND, (std::string("exg_gravity_accumulate_noalloc: hipGetSymbolAddress: ") +
                                               hipGetErrorString(e)).c_str());
        g_scratch[dev] = p;
    }
    *out = static_cast<char*>(g_scratch[dev]);
    return EXG_OK;
}

// The work table of one world of n bodies in chunks of `chunk` sources: tile t holds
// entries [t·chunks, (t+1)·chunks) in ascending chunk order (as gravity_add_world_chunked
// lays it out).
__global__ __launch_bounds__(256) void noalloc_table_kernel(GravEntry* __restrict__ entries, GravTile* __restrict__ tiles,
                                                            int n, int chunk, int chunks, double g_inv) {
    const int ntiles = (n + GRAV_TILE - 1) / GRAV_TILE;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= ntiles * chunks) return;
    const int t = e / chunks, c = e - t * chunks;
    const int tgt0 = t * GRAV_TILE, ntgt = min(GRAV_TILE, n - tgt0);
    GravEntry en;
    en.tgt0 = tgt0;
    en.ntgt = ntgt;
    en.src0 = c * chunk;
    en.src1 = min(n, (c + 1) * chunk);
    entries[e] = en;
    if (c == 0) {
        GravTile tl;
        tl.tgt0 = tgt0;
        tl.ntgt = ntgt;
        tl.e0 = e;
        tl.ne = chunks;
        tl.g_inv = g_inv;
        tiles[t] = tl;
    }
}

}  // namespace
}  // namespace exg

extern "C" int exg_gravity_accumulate_noalloc(int n, const double* charges, double G, int precision, double* wrench,
                                              void* stream) {
    using namespace exg;
    if (!charges || !wrench) return set_error(EXG_ERR_INVALID, "exg_gravity_accumulate_noalloc: null buffer");
    if (n <= 0 || n > NOALLOC_MAX_BODIES)
        return set_error(EXG_ERR_INVALID, ("exg_gravity_accumulate_noalloc: n = " + std::to_string(n) +
                                           " is outside 1 .. " + std::to_string(NOALLOC_MAX_BODIES)).c_str());
    if (precision != EXG_PREC_F64 && precision != EXG_PREC_F32_COMPENSATED)
        return set_error(EXG_ERR_INVALID, "exg_gravity_accumulate_noalloc: unknown precision");
    char* base = nullptr;
    if (int rc = scratch_of_current_device(&base)) return rc;
    int dev = 0, cus = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e != hipSuccess)
        return set_error(EXG_ERR_BACKEND, (std::string("exg_gravity_accumulate_noalloc: ") + hipGetErrorString(e)).c_str());
    const int chunk = gravity_chunk_len(n, cus);
    const int chunks = (n + chunk - 1) / chunk;
    const int ntiles = (n + GRAV_TILE - 1) / GRAV_TILE;
    if ((long long)ntiles * chunks > NOALLOC_MAX_ENTRIES)
        return set_error(EXG_ERR_UNSUPPORTED, "exg_gravity_accumulate_noalloc: the work split exceeds the scratch table "
                                              "(a device with more than 256 compute units)");
    double* partial = reinterpret_cast<double*>(base);
    void* stage = base + PARTIAL_BYTES;
    GravEntry* entries = reinterpret_cast<GravEntry*>(base + PARTIAL_BYTES + STAGE_BYTES);
    GravTile* tiles = reinterpret_cast<GravTile*>(base + PARTIAL_BYTES + STAGE_BYTES + ENTRY_BYTES);
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int n_entries = ntiles * chunks;
    hipLaunchKernelGGL(noalloc_table_kernel, dim3((n_entries + 255) / 256), dim3(256), 0, st, entries, tiles, n, chunk,
                       chunks, 1.0 / G);
    launch_gravity(charges, entries, n_entries, tiles, ntiles, precision, partial, wrench, stage, n, st);
    e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(EXG_ERR_BACKEND, (std::string("exg_gravity_accumulate_noalloc: ") + hipGetErrorString(e)).c_str());
    return EXG_OK;
}
