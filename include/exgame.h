/*
 * exgame.h — C ABI of the MI355X-native per-step dynamics library (libexgame_hip.so).
 *
 * The library replaces the device half of the reference's accelerator boundary:
 * what `newton::Accelerator<T>` dispatches (crates/newton/src/accelerator/mod.rs:557-965)
 * on the `rembrandt` Vulkan device layer (crates/rembrandt/src/lib.rs:77-243) over
 * `aristotle::World` storage (crates/aristotle/src/world.rs:395-660). Every entry
 * point is plain C: pointers, sizes and integer status codes; no torch or HIP types.
 *
 * Data conventions (shared by every entry point):
 *   motor  8 doubles — even blades ascending {1, e12, e13, e23, e01, e02, e03, e0123}
 *          (crates/clifford/src/pga3/motor.rs:119-152, to_components/from_components)
 *   screw  6 doubles — (linear|force x y z, angular|torque x y z), the coordinate
 *          order of Twist::new / Wrench::new (crates/clifford/src/pga3/screw.rs:235-290)
 *   charge 8 doubles — com xyz (absolute), local xyz (arm about the island origin),
 *          G·m, effective size (crates/newton/src/gravity.rs:67-76)
 *   joint params 10 doubles — the joint's canonical order (crates/joints/src/lib.rs:273-299):
 *          axial [a0 a1 a2 b0 b1 b2 rest k c eps], perpendicular [c eps], torsional [c]
 *
 * Status codes map 1:1 onto the reference's EvalError (crates/newton/src/error.rs:11-43)
 * plus two for argument and capability errors; exg_last_error() describes the last one.
 * One world is driven from one host thread at a time.
 */
#ifndef EXGAME_H
#define EXGAME_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum exg_status {
    EXG_OK = 0,
    EXG_ERR_NO_SHADER = 1,    /* EvalError::NoShader — no kernel for a joint kind / message kind */
    EXG_ERR_TRAP = 2,         /* EvalError::Trap — a kernel trapped */
    EXG_ERR_FATAL = 3,        /* EvalError::Fatal — guarded fatal slots fired */
    EXG_ERR_WORKER_GONE = 4,  /* EvalError::WorkerGone — the device context is gone */
    EXG_ERR_BACKEND = 5,      /* EvalError::Backend — a HIP runtime call failed */
    EXG_ERR_INVALID = 6,      /* bad arguments (sizes, indices, topology) */
    EXG_ERR_UNSUPPORTED = 7,  /* a shape this build does not handle (e.g. island too large) */
    EXG_ERR_FULL = 8          /* an arena has no room for the world now (retire one and retry) */
};

/* Joint kinds — the four force laws of crates/joints (MessageKind families,
 * crates/newton/src/accelerator/mod.rs:55-77). */
enum exg_joint_kind {
    EXG_SIMPLE_SPRING_DAMPER = 0,
    EXG_CRITICALLY_DAMPED_WARPED = 1,
    EXG_PERPENDICULAR_DAMPER_WARPED = 2,
    EXG_TORSIONAL_DAMPER_WARPED = 3
};

/* Integrators of a world: the variants of ImplicitIntegrator
 * (crates/newton/src/integrator/implicit/mod.rs:43-48); the explicit ones run through
 * the bridge (bridge.rs:19-63). */
enum exg_integrator {
    EXG_EXPLICIT_EULER = 0,
    EXG_SYMPLECTIC_EULER = 1,
    EXG_NEWTON = 2,
    EXG_LIE_EULER = 3      /* crates/newton/src/integrator/explicit/lie_euler.rs:26-56 */
};

/* Device description — the analogue of rembrandt's GpuAccelerator::new + gpu_in_flight
 * (crates/rembrandt/src/lib.rs:77-243). */
typedef struct exg_device_info {
    int compute_units;
    int wave_size;
    int64_t lds_bytes_per_cu;
    int64_t hbm_bytes;
    char name[64];
} exg_device_info;

/* Last error message of the calling thread (empty string if none). */
const char* exg_last_error(void);

/* Fills `out` for device 0. */
int exg_device_info_get(exg_device_info* out);

typedef struct exg_world exg_world;

/* A world of islands of rigid bodies connected by joints, plus N-body gravity —
 * the device-resident state a `Mechanism` (crates/newton/src/mechanism/mechanism.rs:526-600)
 * steps, laid out structure-of-arrays in HBM.
 *
 *   n_bodies B: pose[B*8], momentum[B*6] (body frame), mass[B], inertia[B*9] (body-frame
 *               tensor, row-major), size[B] (effective size, crates/newton/src/body.rs:70-77)
 *   n_islands I: island_start[I+1] (bodies of island i are [start[i], start[i+1])),
 *               origin[I*3] (island anchor, absolute frame)
 *   n_joints J: kind[J], a[J], b[J] (global body indices inside one island, joints
 *               grouped island by island in declaration order), params[J*10]
 *   G: gravitational constant (gravity off when gravity_on == 0)
 * Returns EXG_ERR_UNSUPPORTED for an island beyond this build's on-chip size limits. */
/* The state is validated: EXG_ERR_INVALID for a negative or non-finite mass, an
 * inertia tensor of a body with mass that is not finite and positive definite, a
 * non-finite pose, momentum, size, origin, joint parameter or G.
 * Mass 0 is a kinematic body (Inertia::Kinematic, crates/newton/src/inertia.rs:103-151):
 * its momentum slot holds its body-frame velocity, it is never moved, and its joints
 * act on their other ends with that velocity (newton.rs:806-830). Its tensor is not
 * read (finite values required). An island's kinematic bodies come after its dynamic
 * ones (EXG_ERR_INVALID otherwise); islands of the mechanism-sized class (more than 8
 * bodies or 16 joints) with a kinematic body give EXG_ERR_UNSUPPORTED. */
int exg_world_create(int n_bodies, const double* pose, const double* momentum, const double* mass,
                     const double* inertia, const double* size, int n_islands, const int* island_start,
                     const double* origin, int n_joints, const int* kind, const int* a, const int* b,
                     const double* params, double G, int gravity_on, exg_world** out);
int exg_world_destroy(exg_world* w);

/* Solver settings: integrator, Newton tolerance² on the worst residual ratio
 * (newton.rs:139-151, 802-805), Newton–Schulz warm/cold budgets (newton.rs:58-61). */
int exg_world_set_solver(exg_world* w, int integrator, double tol2, int ns_warm, int ns_seed);

/* How the Newton–Schulz products run (no counterpart in the reference, whose GEMM rows
 * are one GLSL kernel, build.rs:1156-1403): EXG_NS_MFMA (default) uses FP64 MFMA
 * tiles; EXG_NS_VALU_ORDERED uses scalar FMAs in the reference's block summation order,
 * so results match the reference's arithmetic term for term. Both are f64. */
enum { EXG_NS_MFMA = 0, EXG_NS_VALU_ORDERED = 1 };
int exg_world_set_ns_engine(exg_world* w, int engine);

/* How the joint work of an 8-body island's step is spread over its waves (no
 * counterpart in the reference): EXG_JOINT_MAPPING_SPLIT (default) evaluates three
 * line-search candidates a round, one wave each, beside a wave that evaluates their
 * joints, and forms end b's velocity Jacobian columns by negation of end a's;
 * EXG_JOINT_MAPPING_FLAT evaluates four candidates a round with the joints on every
 * wave, one Jacobian column per lane. The same arithmetic in another order of
 * instructions: results agree to rounding (the compiler fuses multiply-adds
 * differently), not bit for bit. */
enum { EXG_JOINT_MAPPING_SPLIT = 0, EXG_JOINT_MAPPING_FLAT = 1 };
int exg_world_set_joint_mapping(exg_world* w, int mapping);

/* Advance `n_steps` steps of length dt > 0 with time warp `warp` ≥ 1 (aristotle Epoch dt/warp,
 * crates/aristotle/src/epoch.rs:238-248): per step, gravity from the charges published
 * by the previous step (zero on the first, gravity.rs:620-626), the island solve, the
 * re-anchoring and the new charges. Asynchronous on the world's stream. */
int exg_world_step(exg_world* w, int n_steps, double dt, double warp);

/* Replay steps from a captured HIP graph of two consecutive steps (re-captured when
 * dt, warp or a solver setting changes); exg_world_last_timing then reports only the
 * total time (out[0] = -1). No counterpart in the reference. */
int exg_world_set_graph(exg_world* w, int enable);

/* Block until the world's queued work has finished. Returns EXG_ERR_FATAL when island
 * steps since the previous sync ended with a non-finite state (mark-and-continue: the
 * state is kept as computed; exg_world_stats slot 6 counts them per island) — the
 * analogue of a message failing on its fatal slots (crates/newton/src/accelerator/shaders.rs:411-431). */
int exg_world_sync(exg_world* w);

/* Island-steps since creation that ended with a non-finite state (synchronizes). */
int exg_world_fatals(exg_world* w, int64_t* count);

/* Copy the state back: pose[B*8], momentum[B*6], origin[I*3]; any pointer may be NULL. */
int exg_world_download(exg_world* w, double* pose, double* momentum, double* origin);

/* Charges the next step's gravity will read, B*8. */
int exg_world_charges(exg_world* w, double* charges);

/* External wrench each body received in the last step (gravity), B*6, island frame. */
int exg_world_external(exg_world* w, double* ext);

/* Per-island counters since creation, I*8: spans, Newton iterations, Newton–Schulz
 * iterations, reseeds, line-search rejections, unconverged commits, steps that ended
 * with a non-finite state (fatal), spans found non-finite at the depth limit and not
 * committed (newton.rs:654-660). */
int exg_world_stats(exg_world* w, int32_t* stats);

/* Device time of the last exg_world_step call, split by kernel, in milliseconds:
 * out[0] gravity, out[1] island solve (sums over the steps of that call). */
int exg_world_last_timing(exg_world* w, double* out2);

/* ───────────── typed slot storage and row-driven dispatch ─────────────
 * The storage half of the boundary: one generational slot map per Pod type in HIP
 * device memory, the analogue of aristotle::World with `usual` storages
 * (crates/aristotle/src/world.rs:515-660). Keys are bit-exact with RawMap
 * (world.rs:300-385): insert pops the lowest free slot or appends, growing the
 * buffer ×2 (indices kept, storage generation +1); remove checks the generation,
 * splices the slot into the address-ordered free list and bumps its generation. */
enum exg_store_type {
    EXG_T_SCALAR = 0,        /* T (f64)                         */
    EXG_T_MOTOR = 1,         /* Motor: 8                        */
    EXG_T_TWIST = 2,         /* Twist: 6                        */
    EXG_T_WRENCH = 3,        /* Wrench: 6                       */
    EXG_T_WRENCH_PAIR = 4,   /* [Wrench; 2]: 12                 */
    EXG_T_JACOBIAN = 5,      /* [[Wrench; 24]; 2]: 288          */
    EXG_T_BLOCK = 6,         /* Block = [[T; 6]; 6]: 36         */
    EXG_T_DIAG_INERTIA = 7,  /* [T; 3]                          */
    EXG_T_FULL_INERTIA = 8,  /* [[T; 3]; 3]                     */
    EXG_T_ROW = 9            /* Row = [u32; 128] (rows/*.rs)    */
};

/* MessageKind numbering (crates/newton/src/accelerator/mod.rs:55-98). */
enum exg_kind {
    EXG_K_SIMPLE_SUM = 0,
    EXG_K_CRITICALLY_DAMPED_WARPED_PLAIN = 1, EXG_K_CRITICALLY_DAMPED_WARPED_JACOBIAN = 2,
    EXG_K_PERPENDICULAR_DAMPER_WARPED_PLAIN = 3, EXG_K_PERPENDICULAR_DAMPER_WARPED_JACOBIAN = 4,
    EXG_K_SIMPLE_SPRING_DAMPER_PLAIN = 5, EXG_K_SIMPLE_SPRING_DAMPER_JACOBIAN = 6,
    EXG_K_TORSIONAL_DAMPER_WARPED_PLAIN = 7, EXG_K_TORSIONAL_DAMPER_WARPED_JACOBIAN = 8,
    EXG_K_PRE = 9, EXG_K_GATHER = 10, EXG_K_GEMM = 11,
    EXG_K_BODY_POST_DIAGONAL = 12, EXG_K_BODY_POST_FULL = 13,
    EXG_K_ASSEMBLE_BLOCK = 14, EXG_K_BLOCK_MATVEC = 15, EXG_K_BLOCK_REDUCE = 16, EXG_K_BLOCK_COPY = 17,
    EXG_K_BODY_POST_GATHERED_DIAGONAL = 18, EXG_K_BODY_POST_GATHERED_FULL = 19
};

typedef struct exg_store exg_store;

/* Scalar type T of a store's storages (World::builder().usual::<T>(); the reference's
 * device path registers f32, crates/newton/src/accelerator/mod.rs:1015-1019). */
enum { EXG_SCALAR_F32 = 0, EXG_SCALAR_F64 = 1 };

/* World::builder().capacity(c).usual::<T>() — every storage starts with `capacity`
 * slots of the reference's Pod layouts over T:
 *   Motor   8 T  — even blades ascending {1, e12, e13, e23, e01, e02, e03, e0123}
 *   Twist / Wrench 6 T — the grade-2 store, blades ascending {e12, e13, e23, e01, e02,
 *           e03} with the markup signs (crates/clifford/src/pga3/screw.rs:102-130; the
 *           slot/sign map crates/newton/build.rs:106-152 probes): for a screw with
 *           constructor coordinates (linear|force l, angular|torque a) the store is
 *           (−a_z, a_y, −a_x, −l_x, −l_y, −l_z)
 *   [Wrench; 2] 12 T, [[Wrench; 24]; 2] 288 T (offset end·144 + axis·6 + component)
 *   Block [[T; 6]; 6] in coordinates (row: force|torque, column: linear|angular)
 *   inertia [T; 3] / [[T; 3]; 3], scalars T, Row [u32; 128].
 * The kernels compute in f64 and round once when they store. */
int exg_store_create_scalar(uint64_t capacity, int scalar, exg_store** out);
/* exg_store_create_scalar(capacity, EXG_SCALAR_F64, out). */
int exg_store_create(uint64_t capacity, exg_store** out);
int exg_store_scalar(exg_store* s, int* scalar);
int exg_store_destroy(exg_store* s);
/* WriteView::add → RawMap::insert: returns the slot index and its generation. */
int exg_slot_insert(exg_store* s, int type, const void* value, uint32_t* index, uint32_t* generation);
/* WorldKey::read / write / drop. `found` is 0 (and nothing is touched) for a stale,
 * free or foreign key — RawMap::get/set/remove returning None. */
int exg_slot_read(exg_store* s, int type, uint32_t index, uint32_t generation, void* out, int* found);
int exg_slot_write(exg_store* s, int type, uint32_t index, uint32_t generation, const void* value, int* found);
int exg_slot_remove(exg_store* s, int type, uint32_t index, uint32_t generation, void* value_out, int* found);
/* Capacity, high-water length and storage generation (World::generation, world.rs:627-634). */
int exg_storage_info(exg_store* s, int type, uint64_t* capacity, uint64_t* len, uint64_t* generation);
int exg_storage_device_ptr(exg_store* s, int type, void** ptr);
/* One round of rows of one MessageKind (Accelerator::submit_rounds → Shaders::dispatch,
 * accelerator/mod.rs:900-964, shaders.rs:369-480). `row_slots` are indices into the
 * ROW storage. Every slot a row names is validated first; two rows writing one slot
 * are refused with EXG_ERR_BACKEND (the Ledger, shaders.rs:287-390). Completes before
 * returning. Rows whose outputs are not all finite (a division guard tripped) are
 * counted and the call returns EXG_ERR_FATAL after writing every output
 * (mark-and-continue, shaders.rs:411-431). */
int exg_dispatch(exg_store* s, int kind, const uint32_t* row_slots, uint32_t count);
/* Kernels of the joint (EXG_K_*_PLAIN / _JACOBIAN) and Pre rows: EXG_CODEGEN_HANDWRITTEN
 * (default), or EXG_CODEGEN_TRACED — device functions generated at build time by
 * tracing the same scalar-generic joint laws over a symbolic scalar (the reference's
 * build.rs + viete route to its connection kernels, crates/newton/build.rs:673-821,
 * crates/viete/src/glsl.rs:12-80), one thread per row. Same rows, slots and results
 * to rounding. */
enum { EXG_CODEGEN_HANDWRITTEN = 0, EXG_CODEGEN_TRACED = 1 };
int exg_store_set_codegen(exg_store* s, int codegen);
/* Wait for queued work (the fence wait of a flush). */
int exg_flush(exg_store* s);
/* Rows of the last dispatch of `kind` that produced non-finite outputs. */
int exg_fatals(exg_store* s, int kind, uint32_t* count);

/* Stand-alone pairwise gravity on the device (gravity.rs:110-146, 530-601):
 * charges[n*8] in, wrench[n*6] out (force, local × force), host pointers. */
int exg_gravity_accumulate(int n, const double* charges, double G, double* wrench);

/* Arithmetic of a path (no counterpart in the reference, whose kernels are f32-only,
 * crates/newton/src/accelerator/mod.rs:1015-1019): EXG_PREC_F64 throughout, or
 * EXG_PREC_F32_COMPENSATED — FP32 pair/element arithmetic with two-sum compensated
 * FP32 accumulation, widened to f64 at the end; residuals, Newton–Schulz and the
 * state update stay f64 either way. */
enum { EXG_PREC_F64 = 0, EXG_PREC_F32_COMPENSATED = 1 };
int exg_gravity_accumulate_precision(int n, const double* charges, double G, int precision, double* wrench);
/* Pairwise gravity that allocates no device memory — no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). The same arithmetic and work split as
 * exg_gravity_accumulate_precision (results bit-identical), but `charges` (n*8) and
 * `wrench` (n*6) are DEVICE buffers the caller owns, and the work table and partial
 * sums live in a fixed device array of the library (EXG_GRAVITY_NOALLOC_SCRATCH_BYTES
 * per GPU, held until the process ends). n ≤ EXG_GRAVITY_NOALLOC_MAX_BODIES, else
 * EXG_ERR_INVALID. Asynchronous on `stream` (a hipStream_t; NULL for the default
 * stream). Calls that overlap on one GPU share the array and race: serialise them. */
#define EXG_GRAVITY_NOALLOC_MAX_BODIES 65536
#define EXG_GRAVITY_NOALLOC_SCRATCH_BYTES 29646848
int exg_gravity_accumulate_noalloc(int n, const double* charges, double G, int precision, double* wrench, void* stream);

/* Per-world precision of gravity and of the Jacobian assembly: with
 * EXG_PREC_F32_COMPENSATED the joint Jacobians are evaluated in FP32 forward mode and
 * assembled in FP32 with two-sum accumulation; the residual's wrenches, Newton–Schulz
 * and the state update stay f64, so a converged step meets the same f64 tolerance. */
int exg_world_set_precision(exg_world* w, int gravity, int jacobian);

/* Who adds the gravity pass's partial sums (f64 pairs): with on = 1 (the default) the
 * island kernels add their own bodies' partial sums in the reduce kernel's order and
 * arithmetic (the same bits), and the reduce kernel is not launched; with on = 0 the
 * reduce kernel runs. Takes effect with f64 gravity for worlds stepped by the explicit
 * integrators, and for worlds whose islands all have 2 to 8 bodies stepped by the
 * Newton integrator; the others run the reduce kernel either way. No counterpart in the reference (a schedule of GravityPropagator's
 * accumulate, crates/newton/src/gravity.rs:530-601). */
int exg_world_set_gravity_island_reduce(exg_world* w, int on);

/* A uniform force field on the world (UniformField, crates/newton/src/field.rs:16-44,
 * a ForceField beside the N-body gravity): a constant world acceleration a adds
 * F = m·a and a zero torque to every body's wrench after gravity, in every integrator;
 * kinematic bodies (mass 0) get nothing. (0, 0, 0), the default, is off. Finite values
 * only (EXG_ERR_INVALID otherwise). */
int exg_world_set_uniform_field(exg_world* w, double ax, double ay, double az);

/* ───────────── arenas: many independent worlds stepped together ─────────────
 * The batched / served form of a Driver epoch stepping several mechanisms
 * (crates/newton/src/driver.rs; accelerator/mod.rs:161-277, 613-630 flush their rows
 * together). An arena owns device arrays of fixed capacity (bodies, islands, joints,
 * and doubles of accepted inverses, Σ (6m)² per island); worlds are admitted into
 * first-fit ranges of them and retired to free them. A step advances every admitted
 * world by one implicit Newton step (gravity within each world only). A world's
 * result is bit-identical alone or among any others. Newton integrator only. */
typedef struct exg_arena exg_arena;
int exg_arena_create(int body_capacity, int island_capacity, int joint_capacity, long long inverse_capacity,
                     exg_arena** out);
int exg_arena_destroy(exg_arena* r);
/* Solver of every world: tolerance², Newton–Schulz budgets and engine, gravity precision. */
int exg_arena_set_solver(exg_arena* r, double tol2, int ns_warm, int ns_seed, int ns_engine, int gravity_precision);
/* Precision of gravity and of the Jacobian assembly for every world (EXG_PREC_*). */
int exg_arena_set_precision(exg_arena* r, int gravity, int jacobian);
/* Gravity work split of worlds admitted from now on: sources per chunk (≥ 64), or 0
 * for the single-world rule (chunks sized to fill the device for that world's body
 * count). A world keeps the chunking it was admitted with, so its forces do not
 * depend on the other worlds. */
int exg_arena_set_gravity_chunk(exg_arena* r, int chunk_len);
/* Replay steps from a captured HIP graph of two consecutive steps of every live
 * world (re-captured when dt, warp, a setting or the set of worlds changes). No
 * counterpart in the reference (its Driver records command buffers per flush). */
int exg_arena_set_graph(exg_arena* r, int enable);
/* Time the gravity pass of the live worlds with every world chunked by 0 (the
 * single-world rule), 128 … 8192, and make the fastest the setting for worlds
 * admitted from now on (written to *chosen_chunk when not null); the live worlds keep
 * their chunking. No counterpart in the reference. */
int exg_arena_autotune_gravity(exg_arena* r, int* chosen_chunk);
/* The joint mapping (EXG_JOINT_MAPPING_*) of worlds admitted from now on; a world
 * keeps the mapping it was admitted with, so its results do not depend on the others. */
int exg_arena_set_joint_mapping(exg_arena* r, int mapping);
/* Time one step of the live worlds with all of them on each joint mapping (every
 * trial from the same saved state, which is put back: the worlds do not advance and
 * keep their own mappings) and make the fastest the mapping of later admissions
 * (written to *chosen_mapping when not null). Allocates a copy of the state for the
 * duration of the call. No counterpart in the reference. */
int exg_arena_autotune_joints(exg_arena* r, int* chosen_mapping);
/* Admit a world (arrays as for exg_world_create); EXG_ERR_FULL when no range fits. */
int exg_arena_admit(exg_arena* r, int n_bodies, const double* pose, const double* momentum, const double* mass,
                    const double* inertia, const double* size, int n_islands, const int* island_start,
                    const double* origin, int n_joints, const int* kind, const int* a, const int* b,
                    const double* params, double G, int gravity_on, int* world_id);
int exg_arena_retire(exg_arena* r, int world_id);
/* Advance every admitted world n_steps steps of dt (asynchronous). */
int exg_arena_step(exg_arena* r, int n_steps, double dt, double warp);
/* As exg_world_sync, EXG_ERR_FATAL for non-finite island-steps of any world. */
int exg_arena_sync(exg_arena* r);
/* Completion marks, so that a host loop can keep steps queued while it waits for
 * earlier ones (no counterpart in the reference): exg_arena_mark puts a mark behind
 * the work enqueued so far and returns its ticket (consecutive from 0);
 * exg_arena_mark_done tells without blocking whether that work has finished,
 * exg_arena_mark_wait blocks until it has. Marks do not check fatals (exg_arena_sync
 * does). */
int exg_arena_mark(exg_arena* r, uint64_t* ticket);
int exg_arena_mark_done(exg_arena* r, uint64_t ticket, int* done);
int exg_arena_mark_wait(exg_arena* r, uint64_t ticket);
/* One world's state: pose[n*8], momentum[n*6], origin[islands*3]; any may be NULL. */
int exg_arena_download(exg_arena* r, int world_id, double* pose, double* momentum, double* origin);
/* One world's per-island solver counters, islands*8 (as exg_world_stats). */
int exg_arena_world_stats(exg_arena* r, int world_id, int32_t* stats);
/* Device time of the last exg_arena_step call, milliseconds. */
int exg_arena_last_ms(exg_arena* r, double* ms);
/* Used and total capacity: bodies, islands, joints, inverse doubles. */
int exg_arena_usage(exg_arena* r, int64_t* used4, int64_t* capacity4);
/* HBM bytes a world of this shape occupies in an arena with the given gravity chunk. */
int exg_arena_footprint(int n_bodies, int n_islands, int n_joints, long long inverse_doubles, int gravity_chunk,
                        int64_t* bytes);

#ifdef __cplusplus
}
#endif

#endif /* EXGAME_H */
