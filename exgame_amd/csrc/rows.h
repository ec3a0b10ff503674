// Row-driven kernels: one launch per MessageKind over a list of baked [u32; 128]
// rows, each row addressing slots of the typed storages (the reference's
// accelerator contract, crates/newton/src/accelerator/rows/*.rs).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace exg {

constexpr int ROW_WORDS = 128;

// Storage types (element widths in doubles; ROW is 128 u32 words).
enum StoreType : int {
    T_SCALAR = 0, T_MOTOR = 1, T_TWIST = 2, T_WRENCH = 3, T_WRENCH_PAIR = 4, T_JACOBIAN = 5,
    T_BLOCK = 6, T_DIAG_INERTIA = 7, T_FULL_INERTIA = 8, T_ROW = 9, N_STORE_TYPES = 10
};

// MessageKind numbering (crates/newton/src/accelerator/mod.rs:55-98).
enum Kind : int {
    K_SIMPLE_SUM = 0,
    K_CDW_PLAIN = 1, K_CDW_JAC = 2,
    K_PERP_PLAIN = 3, K_PERP_JAC = 4,
    K_SSD_PLAIN = 5, K_SSD_JAC = 6,
    K_TORS_PLAIN = 7, K_TORS_JAC = 8,
    K_PRE = 9, K_GATHER = 10, K_GEMM = 11,
    K_BODY_POST_DIAG = 12, K_BODY_POST_FULL = 13,
    K_ASSEMBLE = 14, K_MATVEC = 15, K_REDUCE = 16, K_COPY = 17,
    K_BODY_POST_GATHERED_DIAG = 18, K_BODY_POST_GATHERED_FULL = 19,
    N_KINDS = 20
};

// Device storages of one store; every non-row storage holds the same scalar type
// (scalar_bytes 4: f32, 8: f64). `fatal` counts rows with non-finite outputs.
struct StorePtrs {
    void* scalar;
    void* motor;
    void* twist;
    void* wrench;
    void* pair;
    void* jac;
    void* block;
    void* diag;
    void* full;
    const uint32_t* row;
    unsigned* fatal;
    int scalar_bytes;
};

void launch_rows(int kind, const StorePtrs& s, const uint32_t* row_slots, uint32_t count, hipStream_t stream);
// The joint and Pre kinds on the traced kernels (rows_traced.hip); false for the
// other kinds, which have no traced form.
bool launch_rows_traced(int kind, const StorePtrs& s, const uint32_t* row_slots, uint32_t count, hipStream_t stream);

}  // namespace exg
