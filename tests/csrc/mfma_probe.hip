// Test harness: the lane layout of the FP64 MFMA builtin on gfx950, in bounds.
// Block p is one wave of 64 lanes; lane l feeds av[p·64 + l] as its A operand and
// bv[p·64 + l] as its B operand and returns its four accumulator values in
// out[p·256 + l·4 + j]. The host arranges the operands and reads the layout back.
#include <hip/hip_runtime.h>
typedef double d4 __attribute__((ext_vector_type(4)));

__global__ void mfma_f64_probe(const double* av, const double* bv, double* out) {
    int p = blockIdx.x, l = threadIdx.x;
    d4 c = {0.0, 0.0, 0.0, 0.0};
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(av[p * 64 + l], bv[p * 64 + l], c, 0, 0, 0);
    for (int j = 0; j < 4; ++j) out[p * 256 + l * 4 + j] = c[j];
}

extern "C" int mfma_probe_run(const double* a, const double* b, double* out, int problems) {
    if (problems < 1 || problems > 4096) return 3;
    size_t n_in = (size_t)problems * 64 * 8, n_out = (size_t)problems * 256 * 8;
    double *da = nullptr, *db = nullptr, *dout = nullptr;
    if (hipMalloc(&da, n_in) || hipMalloc(&db, n_in) || hipMalloc(&dout, n_out)) return 1;
    hipMemcpy(da, a, n_in, hipMemcpyHostToDevice);
    hipMemcpy(db, b, n_in, hipMemcpyHostToDevice);
    hipLaunchKernelGGL(mfma_f64_probe, dim3(problems), dim3(64), 0, 0, da, db, dout);
    hipError_t e = hipMemcpy(out, dout, n_out, hipMemcpyDeviceToHost);
    hipFree(da);
    hipFree(db);
    hipFree(dout);
    return e == hipSuccess ? 0 : 2;
}
