// Slot access of the row kernels (rows.hip, rows_traced.hip): the typed storages of
// one store as device pointers over the storage scalar T, each row's [u32; 128]
// words, loads into f64 and rounding stores, screws through the grade-2 Pod map,
// and the fatal mark of a row with non-finite outputs.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "physics_dev.h"
#include "rows.h"

namespace exg {

constexpr int RT = 64;

template <class T>
struct Slots {
    T* scalar;
    T* motor;
    T* twist;
    T* wrench;
    T* pair;
    T* jac;
    T* block;
    T* diag;
    T* full;
    const uint32_t* row;
    unsigned* fatal;
};

template <class T>
__device__ inline const uint32_t* row_of(const Slots<T>& s, const uint32_t* slots, uint32_t r) {
    return s.row + (size_t)slots[r] * ROW_WORDS;
}
__device__ inline float bits_f32(uint32_t w) { return __uint_as_float(w); }

// slot access in f64: screws through the Pod map, everything else component-wise
template <class T>
__device__ inline double ld(const T* p) { return (double)*p; }
template <class T>
__device__ inline void st(T* p, double v) { *p = (T)v; }
template <class T>
__device__ inline void ld_n(const T* p, double* v, int n) { for (int k = 0; k < n; ++k) v[k] = (double)p[k]; }
template <class T>
__device__ inline void st_n(T* p, const double* v, int n) { for (int k = 0; k < n; ++k) p[k] = (T)v[k]; }
template <class T>
__device__ inline void ld_screw(const T* p, double c[6]) {
    double q[6];
    ld_n(p, q, 6);
    screw_from_pod(q, c);
}
template <class T>
__device__ inline void st_screw(T* p, const double c[6]) {
    double q[6];
    screw_to_pod(c, q);
    st_n(p, q, 6);
}
template <class T>
__device__ inline void mark(const Slots<T>& s, const double* v, int n) {
    bool bad = false;
    for (int k = 0; k < n; ++k) bad |= !isfinite(v[k]);
    if (bad) atomicAdd(s.fatal, 1u);
}

// the Pod component that carries screw coordinate c, and its sign (a signed permutation)
__device__ inline void pod_slot_of(int c, int& k, double& sign) {
    double probe[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, q[6];
    probe[c] = 1.0;
    screw_to_pod(probe, q);
    k = 0;
    sign = 0.0;
    for (int j = 0; j < 6; ++j)
        if (q[j] != 0.0) {
            k = j;
            sign = q[j];
        }
}

}  // namespace exg
