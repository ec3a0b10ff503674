// The implicit-midpoint Newton step of one island per workgroup, entirely on the
// device: Jacobian evaluation, residual, block assembly, Newton–Schulz inverse,
// the 8-lane line search, dt subdivision, commit, X checkpoint/rollback, and the
// island re-anchoring + gravity-charge publication that close a Mechanism step.
//
// The control flow is the reference's host loop moved into the kernel:
//   Newton::solve_step / newton_solve       crates/newton/src/integrator/implicit/newton.rs:334-683
//   run_budget / refine_inverse / seed      newton.rs:209-327
//   signal_log2 / commit_span               newton.rs:688-753
//   baked-row arithmetic (a1·out + terms)   crates/newton/build.rs:1156-1403
//   Island::recompute_centroid              crates/newton/src/mechanism/island.rs:206-244
//   GravityPropagator::charge_of + publish  crates/newton/src/gravity.rs:485-496, 627-638
//
// Everything an island touches during the step lives in LDS (A, X, X', R, the
// joint Jacobians, the 8 line-search lanes): one workgroup owns its island for
// the whole step and HBM is touched only to load the state and write it back,
// plus the accepted inverse X_acc on every commit/rollback. Every control
// decision is taken redundantly by all threads from values broadcast through LDS,
// so the workgroup stays converged.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "island.h"
#include "physics_dev.h"
#include "body_rows.h"

namespace exg {

typedef double d4 __attribute__((ext_vector_type(4)));

// Wave-uniform copies of values every lane already holds (LDS broadcasts, wave
// indices): they live in scalar registers, so branches on them are scalar branches
// instead of exec-masked vector ones.
__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline unsigned uni(unsigned v) { return (unsigned)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline double uni(double v) {
    long long b = __double_as_longlong(v);
    unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xffffffffll));
    unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

constexpr int NT = ISLAND_THREADS;
constexpr int LANES = 8;   // line-search step lengths α = 2^-l, l < LANES (newton.rs:469-522)
constexpr int LS = 4;      // candidates evaluated side by side: lanes 0-3, then 4-7 if all four fail
constexpr int MAX_ITERATIONS = 32;
constexpr int MAX_DEPTH = 8;
constexpr int THRESHOLD_LOG2 = 8;
constexpr int PER_HALVING_LOG2 = 4;

// Line-search candidates 1..LS-1 (candidate 0 is the Newton iterate's own storage).
template <int MAXM, int MAXJ>
struct LaneSlots {
    static constexpr int JD = MAXJ > 0 ? MAXJ : 1;
    double vmid[LS - 1][MAXM][6];
    double mid[LS - 1][MAXM][8];
    double total[LS - 1][MAXM][6];
    double rhs[LS - 1][MAXM][6];
    double scale[LS - 1][MAXM][2];
    double conn[LS - 1][JD][2][6];
};

// The n×n system matrices (n = 6m) are stored flat, row-major, padded to n16 =
// ⌈n/16⌉·16 rows and columns for the 16×16 MFMA tiles, with row stride LDM =
// MAXN + 1 (odd in doubles): the sixteen rows an MFMA operand column spans then fall
// on sixteen different LDS bank pairs. The padding stays zero.
//
// Three things are never live at once and share one region: the joint Jacobians
// (evaluation → assembly), T = 2I − A·X (inside one Newton–Schulz iteration) and
// line-search candidates 1..3 (line search → adoption). With MAXJ = 14 an 8-body
// island fits in 80 KB, so two islands share a CU.
template <int MAXM, int MAXJ, int NTH, bool JF>
struct IslandShared {
    static constexpr int MAXN = ((6 * MAXM + 15) / 16) * 16;
    static constexpr int LDM = MAXN + 1;
    static constexpr int NB = MAXN * LDM;
    static constexpr int JD = MAXJ > 0 ? MAXJ : 1;   // array extent (islands without joints: MAXJ = 0)
    using JT = typename std::conditional<JF, float, double>::type;   // Jacobian storage
    double A[NB];
    double X[NB];
    union {
        double T[NB];
        JT jac[JD][2][24][6];               // ∂wrench[end]/∂axis, per joint
        LaneSlots<MAXM, MAXJ> ls;
    } u;
    unsigned kmask[MAXN / 16];              // per 16-row tile of A: 4-column slabs that can be non-zero
    static constexpr int MAXT = MAXN / 16, WAVES = NTH / 64;
    static constexpr int PER_WAVE = (MAXT * MAXT + WAVES - 1) / WAVES;
    int tlist[2][WAVES][PER_WAVE];          // MFMA tiles of each wave: [0] sparse product, [1] dense
    int tcount[2][WAVES];
    double vmid0[MAXM][6], mid0[MAXM][8], total0[MAXM][6], rhs0[MAXM][6], scale0[MAXM][2];
    double conn0[JD][2][6];
    double wn[LS];
    double mass0[MAXM][36];                 // candidate 0's world mass block
    double pose[MAXM][8], mom[MAXM][6], massv[MAXM], tens[MAXM][9], ext[MAXM][6];
    double snap[MAXM][6], dv[MAXM][6], base[MAXM][6], best[MAXM][6];
    double jparams[JD][10];
    double red[NTH / 64];                    // wave partial sums: block_sum, and ‖I − A·X‖² inside
                                             // refine_inverse (written before the barrier that
                                             // publishes T, read before the iteration's last one)
    double seedsum[2 * MAXM];               // seed_transpose: block-row and block-column |A| sums
    double bcast[8];
    int jkind[JD], ja[JD], jb[JD];
    int jorder[JD];                         // joints grouped by kind (thread mapping only)
    int nbr_start[MAXM + 1], nbr[MAXM * MAXM];
    unsigned char nbr_row[MAXM * MAXM];     // block row of each neighbour-list entry
    int term_start[MAXM * MAXM + 1], terms[4 * JD];
    int inc_start[MAXM + 1], inc[2 * JD];
    int ibcast[8];                          // [0]: a non-zero Newton correction was written; [1]: seeded flag out
    int nd;                                 // dynamic bodies (the unknowns), the island's first nd
};

// Block-wide sum of one value per thread: a butterfly within each 64-lane wave,
// then the four wave sums added in wave order by every thread (the consumers compare
// the result against a ×4 margin or take ⌊log₂⌋, never test it for equality). The
// leading barrier protects s.red from the previous call's readers.
template <int MAXM, int MAXJ, int NTH, bool JF>
__device__ double block_sum(IslandShared<MAXM, MAXJ, NTH, JF>& s, double v) {
    v = wave_sum(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) s.red[threadIdx.x >> 6] = v;
    __syncthreads();
    double r = s.red[0];
#pragma unroll
    for (int w = 1; w < NTH / 64; ++w) r = r + s.red[w];
    return uni(r);
}

template <int MAXM, int MAXJ, int NTH, bool JF, bool FLAT = false>
struct IslandSolver {
    IslandShared<MAXM, MAXJ, NTH, JF>& s;
    const IslandArgs& a;
    static constexpr int LD = IslandShared<MAXM, MAXJ, NTH, JF>::LDM;
    using JT = typename IslandShared<MAXM, MAXJ, NTH, JF>::JT;
    // Line-search rounds with the joints on a wave of their own: with eight bodies per
    // candidate (one wave) and four or more waves, three candidates take a wave each
    // and the fourth wave evaluates their joints (3·MAXJ ≤ 64 items); otherwise LS
    // candidates per round in three phases (pre_candidates, joints_plain, gather_post).
    // (FLAT keeps LS candidates per round on any shape and every velocity column its
    // own lane: the same results, another mapping of the joint work onto the waves,
    // for the joint-mapping autotuner)
    static constexpr bool SPLIT = !FLAT && NTH / 64 >= 4 && 8 * MAXM == 64 && 3 * MAXJ <= 64;
    static constexpr int LSR = SPLIT ? 3 : LS;
    int m, J, isl, first_body;
    // dynamic bodies, the island's first md (the rest are kinematic references), read
    // from LDS at each use rather than held in a register across the solve
    __device__ int md() const { return uni(s.nd); }
    int n, n16;               // system size 6·md and its 16-padded size
    bool mfma;                // Newton–Schulz products on MFMA (else VALU in the oracle's summation order)
    double dt, warp, half, span;
    long long xoff;
    int st_spans, st_newton, st_ns, st_reseed, st_reject, st_uncommit, st_nonfinite;
    bool seeded;

    // the thread index as an opaque value at each use, so that index arithmetic is
    // redone where it is needed instead of being kept live across the whole solve
    static __device__ __forceinline__ int tid() {
        int t = (int)threadIdx.x;
        asm volatile("" : "+v"(t));
        return t;
    }
    __device__ double* A() { return s.A; }
    __device__ double* X() { return s.X; }
    __device__ double* R() { return s.u.T; }
    // candidate l's arrays (l = 0 is the Newton iterate)
    static __device__ int sl(int l) { return l > 0 ? l - 1 : 0; }
    __device__ double* VMID(int l, int i) { return l == 0 ? s.vmid0[i] : s.u.ls.vmid[sl(l)][i]; }
    __device__ double* MID(int l, int i) { return l == 0 ? s.mid0[i] : s.u.ls.mid[sl(l)][i]; }
    __device__ double* TOTAL(int l, int i) { return l == 0 ? s.total0[i] : s.u.ls.total[sl(l)][i]; }
    __device__ double* RHS(int l, int i) { return l == 0 ? s.rhs0[i] : s.u.ls.rhs[sl(l)][i]; }
    __device__ double* SCALE(int l, int i) { return l == 0 ? s.scale0[i] : s.u.ls.scale[sl(l)][i]; }
    __device__ double* CONN(int l, int j, int e) { return l == 0 ? s.conn0[j][e] : s.u.ls.conn[sl(l)][j][e]; }

    // ───── PRE: midpoint = pose ∘ exp(half·vmid) for lanes [l0, l1) ─────
    __device__ void pre(int l0, int l1) {
        for (int t = tid(); t < (l1 - l0) * m; t += NTH) {
            int lane = l0 + t / m, i = t % m;
            double e[8];
            twist_exp(VMID(lane, i), half, e);
            compose(s.pose[i], e, MID(lane, i));
        }
    }
    // line-search round `round`: candidate `slot` takes vmid = base + 2^-(round·LS + slot)·dv
    // (newton.rs:469-522) and its midpoint in the same pass
    __device__ void pre_candidates(int round) {
        for (int t = tid(); t < LS * m; t += NTH) {
            const int slot = t / m, i = t % m;
            const double alpha = ldexp(1.0, -(round * LS + slot));
            double v[6], e[8];
            for (int k = 0; k < 6; ++k) {
                v[k] = s.base[i][k] + s.dv[i][k] * alpha;
                VMID(slot, i)[k] = v[k];
            }
            twist_exp(v, half, e);
            compose(s.pose[i], e, MID(slot, i));
        }
    }
    // ───── Jacobian lanes: one thread per (joint, axis) ─────
    // item q < 12J: pose column q % 12 of joint q / 12 (in kind order); item 12J + q:
    // velocity column 12 + q % 12 of the same joint
    // With `mirrored`, velocity item q < 6J is column 12 + q % 6 of joint q / 6 (end a's
    // velocity axes), stored also as end b's column 18 + q % 6 negated.
    __device__ void jacobian_item(int item, bool mirrored = false) {
        const int pose_item = item < J * 12;
        const int q = pose_item ? item : item - J * 12;
        const int per = (pose_item || !mirrored) ? 12 : 6;
        const int j = s.jorder[q / per], ax = (pose_item ? 0 : 12) + q % per, ea = s.ja[j], eb = s.jb[j];
        // the f64 lanes also deliver the wrench values; with f32 lanes the residual's
        // wrenches come from an f64 value-only evaluation instead
        const bool values = !JF && ax == 0;
        JT da[6], db[6];
        joint_column(s.jkind[j], s.jparams[j], s.mid0[ea], s.mid0[eb], s.vmid0[ea], s.vmid0[eb], ax, dt, warp, da, db,
                     values ? s.conn0[j][0] : nullptr, values ? s.conn0[j][1] : nullptr);
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            s.u.jac[j][0][ax][c] = da[c];
            s.u.jac[j][1][ax][c] = db[c];
        }
        if (!pose_item && mirrored) {
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                s.u.jac[j][0][ax + 6][c] = -da[c];
                s.u.jac[j][1][ax + 6][c] = -db[c];
            }
        }
    }
    __device__ void joints_jacobian() {
        if constexpr (MAXJ == 0) return;
        const int P = J * 12, t = tid();
        if (!FLAT && P + 6 * J <= NTH) {
            // one column per thread: the pose columns, then the velocity columns of
            // end a, each also giving end b's column by negation (vrel = va − vb enters
            // the wrenches linearly, so ∂W/∂vb = −∂W/∂va exactly)
            if (t < P + 6 * J) jacobian_item(t, true);
            return;
        }
        int first, second;
        if (3 * P <= 2 * NTH) {
            // one pose column (tangents through the geometry) per thread; the cheaper
            // velocity columns, at most two per thread, on the remaining threads
            first = t < P ? t : (t - P < P ? t : -1);
            second = (t >= P && t + NTH - P < 2 * P) ? t + NTH - P : -1;
        } else {
            // pose columns first, then velocity columns: a thread that takes a second
            // column takes a cheap one (2P ≤ 2·NTH: at most two per thread)
            first = t < 2 * P ? t : -1;
            second = t + NTH < 2 * P ? t + NTH : -1;
        }
#pragma clang loop unroll(disable)
        for (int pass = 0; pass < 2; ++pass) {  // one call site: the column code is inlined once
            const int item = pass ? second : first;
            if (item >= 0) jacobian_item(item);
        }
    }
    // ───── value-only joints for lanes [l0, l1) ─────
    __device__ void joints_plain(int l0, int l1) {
        if constexpr (MAXJ == 0) return;
        const int nl = l1 - l0;
        for (int t = tid(); t < nl * J; t += NTH) {
            int lane = l0 + t % nl, j = s.jorder[t / nl];
            joint_wrench(s.jkind[j], s.jparams[j], MID(lane, s.ja[j]), MID(lane, s.jb[j]), VMID(lane, s.ja[j]),
                         VMID(lane, s.jb[j]), dt, warp, CONN(lane, j, 0), CONN(lane, j, 1));
        }
    }
    static __device__ double max2(double x, double y) { return x > y ? x : y; }
    // ───── gather (external + incident wrenches in joint order) + body post + wnorm ─────
    // Lanes [l0, l1), eight threads per (lane, body): lane l's bodies take threads
    // [l·LSTR, l·LSTR + 8m) with LSTR = 8·MAXM, so a wave holds whole lanes. The eight
    // threads of a body form one row each of its world mass block (body_rows.h) and
    // end with the same residual; the worst² residual ratio of each lane
    // (newton.rs:139-151) is folded across the wave in body order with max2 (NaN
    // handling as the sequential form) into s.wn (visible after the caller's barrier).
    // Candidate 0's world mass block goes to mass0 only when mass0 may take it: in
    // the second line-search round candidate 0 is lane 4, and mass0 must keep lane
    // 0's block unless a later lane is accepted (the oracle's lanes[0].mass).
    __device__ void gather_post(int l0, int l1, bool mass_of_candidate0 = true) {
        constexpr int LSTR = 8 * MAXM;
        static_assert(64 % LSTR == 0 && LS * LSTR <= NTH, "a wave holds whole line-search lanes");
        const int nl = l1 - l0, t = tid();
        const int ll = t / LSTR, i = (t % LSTR) >> 3, j = t & 7;
        double r0 = 0.0, r1 = 0.0;
        if (ll < nl && i < m) {   // kinematic bodies too (their rows are never read)
            const int lane = l0 + ll;
            double tot[6];
            for (int k = 0; k < 6; ++k) tot[k] = s.ext[i][k];
            for (int q = s.inc_start[i]; q < s.inc_start[i + 1]; ++q) {
                const int code = s.inc[q];
                const double* w = CONN(lane, code >> 1, code & 1);
                for (int k = 0; k < 6; ++k) tot[k] = tot[k] + w[k];
            }
            const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
            double brow[6], rhs[6], sc[2];
            body_post_row(MID(lane, i), VMID(lane, i), s.massv[i], s.tens[i], s.snap[i], tot, half, floor2,
                          j < 6 ? j : 5, brow, rhs, sc);
            if (j == 0) {
                double* tl = TOTAL(lane, i);
                double* rh = RHS(lane, i);
                double* scl = SCALE(lane, i);
                for (int k = 0; k < 6; ++k) {
                    tl[k] = tot[k];
                    rh[k] = rhs[k];
                }
                scl[0] = sc[0];
                scl[1] = sc[1];
            }
            if (lane == 0 && mass_of_candidate0 && j < 6)
                for (int c = 0; c < 6; ++c) s.mass0[i][j * 6 + c] = brow[c];
            const double lin2 = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
            const double ang2 = rhs[3] * rhs[3] + rhs[4] * rhs[4] + rhs[5] * rhs[5];
            r0 = lin2 / sc[0];
            r1 = ang2 / sc[1];
        }
        {   // every lane of every wave runs the shuffles
            double worst = 0.0;
            if constexpr (LSTR == 64) {
                // one candidate per wave: its bodies' ratios are read from fixed lanes
#pragma unroll
                for (int q = 0; q < MAXM; ++q)
                    if (q < md()) {
                        const double a0 = __longlong_as_double(
                            (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane(
                                             (int)(__double_as_longlong(r0) >> 32), q * 8) << 32) |
                                        (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(r0), q * 8)));
                        const double a1 = __longlong_as_double(
                            (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane(
                                             (int)(__double_as_longlong(r1) >> 32), q * 8) << 32) |
                                        (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(r1), q * 8)));
                        worst = max2(worst, a0);
                        worst = max2(worst, a1);
                    }
            } else {
                const int base = (t & 63) & ~(LSTR - 1);
                for (int q = 0; q < md(); ++q) {
                    const double a0 = __shfl(r0, base + q * 8, 64), a1 = __shfl(r1, base + q * 8, 64);
                    worst = max2(worst, a0);
                    worst = max2(worst, a1);
                }
            }
            if ((t % LSTR) == 0 && ll < nl) s.wn[l0 + ll] = worst;
        }
    }
    // ───── line-search round, joints beside the body rows (SPLIT classes) ─────
    // Phase 1: wave w < nl takes candidate w = round·LSR + w (vmid = base + 2^-w·dv),
    // eight threads per body; each forms the candidate's velocity and midpoint (the
    // same values on all eight; lane 0 of the group stores them), one row of the world
    // mass block and the momentum P = block·V, none of which depend on the joints. The
    // waves from nl on evaluate every (candidate, joint) wrench, forming the two ends'
    // candidate velocities and midpoints again from the same inputs by the same code
    // (so with the same bits). P waits in the candidate's rhs slot for phase 2 (no
    // registers held across the barrier); candidate 0's mass block goes to mass0 when
    // mass0 may take it (as in gather_post).
    __device__ void ls_round_bodies_joints(int round, int nl, bool mass_of_candidate0) {
        const int t = tid(), w = uni(t >> 6);
        if (w < nl) {
            const int i = (t & 63) >> 3, r = t & 7;
            if (i < m) {
                const double alpha = ldexp(1.0, -(round * LSR + w));
                double v[6], e[8], mid[8];
#pragma unroll
                for (int k = 0; k < 6; ++k) v[k] = s.base[i][k] + s.dv[i][k] * alpha;
                twist_exp(v, half, e);
                compose(s.pose[i], e, mid);
                if (r == 0) {
                    double* vm = VMID(w, i);
                    double* mw = MID(w, i);
#pragma unroll
                    for (int k = 0; k < 6; ++k) vm[k] = v[k];
#pragma unroll
                    for (int k = 0; k < 8; ++k) mw[k] = mid[k];
                }
                if (i < md()) {
                    double brow[6], p[6];
                    body_post_row_momentum(mid, v, s.massv[i], s.tens[i], r < 6 ? r : 5, brow, p);
                    if (r == 0) {
                        double* rh = RHS(w, i);
#pragma unroll
                        for (int k = 0; k < 6; ++k) rh[k] = p[k];
                    }
                    if (w == 0 && mass_of_candidate0 && r < 6)
                        for (int c = 0; c < 6; ++c) s.mass0[i][r * 6 + c] = brow[c];
                }
            }
        } else if constexpr (MAXJ > 0) {
            const int q = t - 64 * nl;
            if (q < nl * J) {
                const int lane = q % nl, j = s.jorder[q / nl];
                const double alpha = ldexp(1.0, -(round * LSR + lane));
                const int ends[2] = {s.ja[j], s.jb[j]};
                double v[2][6], mid[2][8];
#pragma unroll
                for (int en = 0; en < 2; ++en) {
                    const int b = ends[en];
                    double e[8];
#pragma unroll
                    for (int k = 0; k < 6; ++k) v[en][k] = s.base[b][k] + s.dv[b][k] * alpha;
                    twist_exp(v[en], half, e);
                    compose(s.pose[b], e, mid[en]);
                }
                joint_wrench(s.jkind[j], s.jparams[j], mid[0], mid[1], v[0], v[1], dt, warp, CONN(lane, j, 0),
                             CONN(lane, j, 1));
            }
        }
    }
    // Phase 2 (after a barrier): gather each candidate body's wrench (external +
    // incident in joint order), finish its residual and self-scale from P, and fold the
    // candidate's worst residual ratio across its wave into s.wn[w].
    __device__ void ls_round_gather(int nl) {
        const int t = tid(), w = uni(t >> 6);
        if (w >= nl) return;
        const int i = (t & 63) >> 3, r = t & 7;
        double r0 = 0.0, r1 = 0.0;
        if (i < md()) {
            double tot[6];
            for (int k = 0; k < 6; ++k) tot[k] = s.ext[i][k];
            for (int q = s.inc_start[i]; q < s.inc_start[i + 1]; ++q) {
                const int code = s.inc[q];
                const double* wr = CONN(w, code >> 1, code & 1);
                for (int k = 0; k < 6; ++k) tot[k] = tot[k] + wr[k];
            }
            const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
            double rhs[6], sc[2], p[6];
            const double* pp = RHS(w, i);
#pragma unroll
            for (int k = 0; k < 6; ++k) p[k] = pp[k];
            body_post_row_residual(p, s.snap[i], tot, half, floor2, rhs, sc);
            if (r == 0) {
                double* tl = TOTAL(w, i);
                double* rh = RHS(w, i);
                double* scl = SCALE(w, i);
                for (int k = 0; k < 6; ++k) {
                    tl[k] = tot[k];
                    rh[k] = rhs[k];
                }
                scl[0] = sc[0];
                scl[1] = sc[1];
            }
            const double lin2 = rhs[0] * rhs[0] + rhs[1] * rhs[1] + rhs[2] * rhs[2];
            const double ang2 = rhs[3] * rhs[3] + rhs[4] * rhs[4] + rhs[5] * rhs[5];
            r0 = lin2 / sc[0];
            r1 = ang2 / sc[1];
        }
        double worst = 0.0;
#pragma unroll
        for (int q = 0; q < MAXM; ++q)
            if (q < md()) {
                const double a0 = __longlong_as_double(
                    (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane(
                                     (int)(__double_as_longlong(r0) >> 32), q * 8) << 32) |
                                (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(r0), q * 8)));
                const double a1 = __longlong_as_double(
                    (long long)(((unsigned long long)(unsigned)__builtin_amdgcn_readlane(
                                     (int)(__double_as_longlong(r1) >> 32), q * 8) << 32) |
                                (unsigned)__builtin_amdgcn_readlane((int)__double_as_longlong(r1), q * 8)));
                worst = max2(worst, a0);
                worst = max2(worst, a1);
            }
        if ((t & 63) == 0) s.wn[w] = worst;
    }
    // ───── system matrix: A = 0·A + [i==j]·𝕀 + Σ terms (pf·∂W/∂pose + vf·∂W/∂vel) ─────
    __device__ void assemble() {
        double pf = -(half * half), vf = -half;
        double* Am = A();
        if (tid() == 0) s.ibcast[0] = 0;  // the correction's non-zero flag (set by matvec)
        // the non-zero blocks only (neighbour lists); the others stay as the launch
        // zeroed them, which is what 0·A + 0 would write
        for (int e = tid(); e < s.nbr_start[m] * 6; e += NTH) {
            const int q = e / 6, r = e - 6 * q;
            const int i = s.nbr_row[q];
            const int j = s.nbr[q], b = i * m + j;
            const int q0 = s.term_start[b], q1 = s.term_start[b + 1];
            const double dg = (i == j) ? 1.0 : 0.0;
            double* dst = Am + (6 * i + r) * LD + 6 * j;
            double acc[6];
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c] = 0.0 * dst[c] + dg * s.mass0[i][r * 6 + c];
            if constexpr (JF) {
                // FP32 terms, summed as an unevaluated hi + lo (two-sum), added in f64
                const float pf32 = (float)pf, vf32 = (float)vf;
                float hi[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, lo[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                for (int q = q0; q < q1; ++q) {
                    const int code = s.terms[q];
                    const int jn = code >> 2, re = (code >> 1) & 1, ce = code & 1;
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        const float x = pf32 * s.u.jac[jn][re][ce * 6 + c][r] + vf32 * s.u.jac[jn][re][12 + ce * 6 + c][r];
                        const float t = hi[c] + x, bp = t - hi[c];
                        lo[c] = lo[c] + ((hi[c] - (t - bp)) + (x - bp));
                        hi[c] = t;
                    }
                }
#pragma unroll
                for (int c = 0; c < 6; ++c) acc[c] = acc[c] + ((double)hi[c] + (double)lo[c]);
            } else {
                // two terms' loads in flight at a time; the sums keep the term order
#pragma unroll 2
                for (int q = q0; q < q1; ++q) {
                    const int code = s.terms[q];
                    const int jn = code >> 2, re = (code >> 1) & 1, ce = code & 1;
#pragma unroll
                    for (int c = 0; c < 6; ++c) {
                        double p = s.u.jac[jn][re][ce * 6 + c][r];
                        double v = s.u.jac[jn][re][12 + ce * 6 + c][r];
                        acc[c] = acc[c] + pf * p + vf * v;
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) dst[c] = acc[c];
        }
    }
    // ───── VALU form, the oracle's block order: out = 0·out + Σ_k a[i,k]·(sign·b[k,j] + diag·2I) ─────
    // With in_place, every cell is computed before any is stored (out may be bm).
    __device__ void gemm(const double* am, const double* bm, double* out, bool two_minus, bool sparse,
                         bool in_place) {
        constexpr int MAXCELLS = (IslandShared<MAXM, MAXJ, NTH, JF>::MAXN * IslandShared<MAXM, MAXJ, NTH, JF>::MAXN + NTH - 1) / NTH;
        double cell[MAXCELLS];
        double sign = two_minus ? -1.0 : 1.0;
        // constant cell indices (every loop unrolled), so that the cells are values of
        // this call alone rather than an array whose contents outlive it
#pragma unroll
        for (int u = 0; u < MAXCELLS; ++u) {
            const int e = tid() + u * NTH;
            if (e >= n * n) break;
            int row = e / n, col = e % n, i = row / 6, j = col / 6, c = col % 6;
            double acc = 0.0 * out[row * LD + col];
            int k0 = sparse ? s.nbr_start[i] : 0, k1 = sparse ? s.nbr_start[i + 1] : md();
            for (int q = k0; q < k1; ++q) {
                int k = sparse ? s.nbr[q] : q;
                double diag = (two_minus && k == j) ? 1.0 : 0.0;
                const double* ab = am + row * LD + 6 * k;
                const double* bb = bm + 6 * k * LD + col;
                double sum = 0.0;
                for (int kk = 0; kk < 6; ++kk) {
                    double right = sign * bb[kk * LD];
                    if (kk == c) right = right + diag * 2.0;
                    sum = sum + ab[kk] * right;
                }
                acc = acc + sum;
            }
            cell[u] = acc;
        }
        if (in_place) __syncthreads();
#pragma unroll
        for (int u = 0; u < MAXCELLS; ++u) {
            const int e = tid() + u * NTH;
            if (e >= n * n) break;
            out[(e / n) * LD + e % n] = cell[u];
        }
    }
    // Slab loop over NT tiles: operands of slab ks+1 are loaded while slab ks
    // multiplies (all loads stay inside the MAXN×LDM buffers; skipped slabs are loaded
    // and not used: loads under scalar branches measured 12 % slower, the waits no
    // longer overlap).
    template <int NT>
    __device__ __forceinline__ void mfma_slabs(const double* am, const double* bm, const int* ti, const int* tj,
                                               const unsigned* mk, d4* acc, int lr, int lk) {
        constexpr int MAXKS = IslandShared<MAXM, MAXJ, NTH, JF>::MAXN / 4;
        double an[NT], bn[NT];
#pragma unroll
        for (int u = 0; u < NT; ++u) {
            an[u] = am[(ti[u] * 16 + lr) * LD + lk];
            bn[u] = bm[lk * LD + tj[u] * 16 + lr];
        }
#pragma unroll
        for (int ks = 0; ks < MAXKS; ++ks) {
            double ac[NT], bc[NT];
#pragma unroll
            for (int u = 0; u < NT; ++u) {
                ac[u] = an[u];
                bc[u] = bn[u];
            }
            if (ks + 1 < MAXKS) {
#pragma unroll
                for (int u = 0; u < NT; ++u) {
                    an[u] = am[(ti[u] * 16 + lr) * LD + 4 * (ks + 1) + lk];
                    bn[u] = bm[(4 * (ks + 1) + lk) * LD + tj[u] * 16 + lr];
                }
            }
#pragma unroll
            for (int u = 0; u < NT; ++u)
                if (mk[u] & (1u << ks)) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[u], bc[u], acc[u], 0, 0, 0);
        }
    }
    // ───── MFMA form: out = a·b over 16×16 tiles, k in 4-column slabs ─────
    // Wave w computes tiles w, w+4, … of the n16×n16 product; a slab whose bit is
    // clear in kmask[tile row] is all-zero in `a` and skipped. With make_t the
    // epilogue stores T = 2I − a·b (what the second Newton–Schulz product needs) and
    // returns this thread's share of ‖I − a·b‖². Padding rows/columns are stored 0.
    // Lane layout of v_mfma_f64_16x16x4f64 (measured on gfx950): lane l supplies
    // a[l%16][l/16] and b[l/16][l%16]; accumulator j holds c[l/16 + 4j][l%16].
    // With in_place, all tiles are multiplied before any is stored (out may be am).
    // `which` selects the wave's tile list: 0 for the sparse product (tiles balanced by
    // their count of non-zero slabs), 1 for the dense one.
    __device__ double mfma_product(const double* am, const double* bm, double* out, const unsigned* kmask,
                                   bool make_t, bool in_place, int which) {
        const int wave = uni((int)(tid() >> 6)), lane = tid() & 63;
        const int lr = lane & 15, lk = lane >> 4;
        const int T = n16 >> 4, KS = n16 >> 2;
        const unsigned dense = (1u << KS) - 1u;
        constexpr int MAXT = IslandShared<MAXM, MAXJ, NTH, JF>::MAXN / 16;
        constexpr int PER_WAVE = (MAXT * MAXT + NTH / 64 - 1) / (NTH / 64);
        double part = 0.0;
        // this wave's tiles (wave-uniform), accumulated side by side
        int ti[PER_WAVE], tj[PER_WAVE];
        unsigned mk[PER_WAVE];
        d4 acc[PER_WAVE];
        const int ntiles = uni(s.tcount[which][wave]);
        // every slot is read (in bounds; unused ones hold 0) and selected on scalars, so
        // that tile indices and masks stay in scalar registers and each MFMA is guarded
        // by a scalar branch
#pragma unroll
        for (int u = 0; u < PER_WAVE; ++u) {
            const bool live = u < ntiles;
            const int t = uni(s.tlist[which][wave][u]);
            ti[u] = t & 255;
            tj[u] = t >> 8;
            const unsigned km = kmask ? uni(kmask[ti[u]]) : dense;
            mk[u] = uni(live ? km : 0u);
            acc[u] = d4{0.0, 0.0, 0.0, 0.0};
        }
        // the slab loop for exactly this wave's tile count (a uniform branch), so that
        // no operands are loaded for unused slots
        if constexpr (PER_WAVE >= 3) {
            if (ntiles >= 3) mfma_slabs<3>(am, bm, ti, tj, mk, acc, lr, lk);
            else if (ntiles == 2) mfma_slabs<2>(am, bm, ti, tj, mk, acc, lr, lk);
            else if (ntiles == 1) mfma_slabs<1>(am, bm, ti, tj, mk, acc, lr, lk);
        } else if constexpr (PER_WAVE == 2) {
            if (ntiles >= 2) mfma_slabs<2>(am, bm, ti, tj, mk, acc, lr, lk);
            else if (ntiles == 1) mfma_slabs<1>(am, bm, ti, tj, mk, acc, lr, lk);
        } else {
            if (ntiles >= 1) mfma_slabs<1>(am, bm, ti, tj, mk, acc, lr, lk);
        }
        if (in_place) __syncthreads();
        // epilogue; off the diagonal of an unpadded matrix T = −R and ‖I − R‖² takes v²
        const bool full = n == n16;
#pragma unroll
        for (int u = 0; u < PER_WAVE; ++u) {
            if (u >= ntiles) continue;
            const bool plain = full && ti[u] != tj[u];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int r = ti[u] * 16 + lk + 4 * j, c = tj[u] * 16 + lr;
                double v = acc[u][j];
                if (plain) {
                    if (make_t) {
                        part = fma(v, v, part);
                        v = -v;
                    }
                } else {
                    const bool real = r < n && c < n;
                    if (make_t) {
                        const double d = (r == c) ? 1.0 : 0.0;
                        const double x = d - v;
                        part += real ? x * x : 0.0;
                        v = 2.0 * d - v;
                    }
                    v = real ? v : 0.0;
                }
                out[r * LD + c] = v;
            }
        }
        return part;
    }
    // ───── Σ ‖diag·I − b‖² over the n×n entries of `mat` ─────
    __device__ double reduce_sq(const double* mat, bool identity_diag) {
        double part = 0.0;
        for (int e = tid(); e < n * n; e += NTH) {
            int row = e / n, col = e % n;
            double diag = (identity_diag && row == col) ? 1.0 : 0.0;
            double v = mat[row * LD + col];
            double x = diag - v;
            part = part + x * x;
        }
        return block_sum(s, part);
    }
    __device__ double refine_inverse() {  // newton.rs:255-279
        double before;
        if (mfma) {
            // T = 2I − A·X (A block-sparse: zero slabs skipped), then X ← X·T in place.
            // ‖I − A·X‖²: wave sums into s.red before the barrier that publishes T,
            // added in wave order after the second product (s.red's previous readers
            // finished before the previous iteration's closing barrier; the next
            // writers come after this one's)
            double part = mfma_product(A(), X(), R(), s.kmask, true, false, 0);
            part = wave_sum(part);
            if ((tid() & 63) == 0) s.red[tid() >> 6] = part;
            __syncthreads();
            mfma_product(X(), R(), X(), nullptr, false, true, 1);
            double r = s.red[0];
#pragma unroll
            for (int w = 1; w < NTH / 64; ++w) r = r + s.red[w];
            before = uni(r);
        } else {
            gemm(A(), X(), R(), false, true, false);
            __syncthreads();
            before = reduce_sq(R(), true);
            gemm(X(), R(), X(), true, false, true);
        }
        __syncthreads();
        st_ns++;
        return before;
    }
    __device__ void seed_transpose() {  // newton.rs:209-244
        // per block row i: Σ|A[i, ·]| and per block column i: Σ|A[·, i]|
        if (tid() < 2 * md()) {
            int i = tid() % md();
            bool col = tid() >= md();
            double sum = 0.0;
            const double* Am = A();
            for (int j = 0; j < md(); ++j)
                for (int r = 0; r < 6; ++r)
                    for (int c = 0; c < 6; ++c)
                        sum += fabs(col ? Am[(6 * j + r) * LD + 6 * i + c] : Am[(6 * i + r) * LD + 6 * j + c]);
            s.seedsum[tid()] = sum;
        }
        __syncthreads();
        if (tid() == 0) {
            double one = 0.0, inf = 0.0;
            for (int i = 0; i < md(); ++i) {
                inf = max2(inf, s.seedsum[i]);
                one = max2(one, s.seedsum[md() + i]);
            }
            s.bcast[0] = 1.0 / (one * inf);
        }
        __syncthreads();
        double sc = uni(s.bcast[0]);
        double* Xm = X();
        for (int e = tid(); e < n * n; e += NTH) {
            int row = e / n, col = e % n;
            Xm[row * LD + col] = A()[col * LD + row] * sc;
        }
        __syncthreads();
        seeded = true;
    }
    __device__ void run_budget(int budget, int seed_budget) {  // newton.rs:301-327
        bool has_prev = false, reseeded = false;
        double previous = 0.0;
        int remaining = budget;
        while (remaining > 0) {
            remaining -= 1;
            double before = refine_inverse();
            if (has_prev && !(before < 4.0 * previous) && !reseeded) {
                seed_transpose();
                reseeded = true;
                has_prev = false;
                remaining = seed_budget;
                st_reseed++;
                continue;
            }
            previous = before;
            has_prev = true;
        }
    }
    // ───── dv = 0·dv + Σ_j X[i,j]·rhs_j ─────
    // MFMA engine: four lanes per row, each summing every fourth block, combined by
    // two butterfly steps (all lanes of a wave take part in the shuffles). The ordered
    // engine keeps one lane per row and the reference's block order. A non-zero entry
    // sets s.ibcast[0] (cleared before the system is assembled; read after the barrier
    // that follows); each entry's writer also copies the iterate into the line
    // search's base.
    __device__ void matvec() {
        const double* Xm = X();
        static_assert(4 * 6 * MAXM <= NTH, "four lanes per row of the largest island");
        if (mfma) {
            const int t = tid(), row = t >> 2, part = t & 3;
            const bool live = t < 4 * n;
            double sum = 0.0;
            if (live)
                for (int j = part; j < md(); j += 4) {
                    const double* xb = Xm + row * LD + 6 * j;
                    double sb = 0.0;
                    for (int c = 0; c < 6; ++c) sb = sb + xb[c] * s.rhs0[j][c];
                    sum = sum + sb;
                }
            sum += xor_lane<1>(sum);
            sum += xor_lane<2>(sum);
            if (live && part == 0) {
                const double d = 0.0 * s.dv[row / 6][row % 6] + sum;
                s.dv[row / 6][row % 6] = d;
                s.base[row / 6][row % 6] = s.vmid0[row / 6][row % 6];
                if (d != 0.0) s.ibcast[0] = 1;
            }
            return;
        }
        for (int t = tid(); t < n; t += NTH) {
            int i = t / 6, r = t % 6;
            double acc = 0.0 * s.dv[i][r];
            for (int j = 0; j < md(); ++j) {
                const double* xb = Xm + t * LD + 6 * j;
                double sum = 0.0;
                for (int c = 0; c < 6; ++c) sum = sum + xb[c] * s.rhs0[j][c];
                acc = acc + sum;
            }
            s.dv[i][r] = acc;
            s.base[i][r] = s.vmid0[i][r];
            if (acc != 0.0) s.ibcast[0] = 1;
        }
    }
    __device__ void copy_vmid(int dst, const double (*src)[6]) {
        for (int t = tid(); t < m * 6; t += NTH) VMID(dst, t / 6)[t % 6] = src[t / 6][t % 6];
    }

    // ───── one Newton solve at the current `half` (newton.rs:334-530) ─────
    __device__ void newton_solve(double tol2, int ns_warm, int ns_seed, bool& converged, double& best,
                                 bool& assembled) {
        const double huge = 1000000000.0;
        // the iterate starts at each body's world velocity (the line-search candidates
        // are written from it before they are read)
        for (int t = tid(); t < m; t += NTH)
            world_velocity(s.pose[t], s.massv[t], s.tens[t], s.mom[t], s.vmid0[t]);
        __syncthreads();
        // kinematic bodies: the line search's base holds their prescribed velocity and
        // their dv is −0, the additive identity, so that every candidate base + α·dv is
        // that velocity bit for bit (the matvec writes the unknowns' rows only)
        for (int t = tid(); t < m * 6; t += NTH) {
            s.best[t / 6][t % 6] = s.vmid0[t / 6][t % 6];
            if (t >= md() * 6) {
                s.base[t / 6][t % 6] = s.vmid0[t / 6][t % 6];
                s.dv[t / 6][t % 6] = -0.0;
            }
        }
        __syncthreads();
        best = huge;
        converged = false;
        assembled = false;
        bool has_carried = false;
        double carried = 0.0;
        for (int iter = 0; iter < MAX_ITERATIONS; ++iter) {
            double rnorm;
            st_newton++;
            // the Jacobian at the iterate: after an accepted line-search lane the
            // iterate's midpoints, wrenches and residual are that lane's; otherwise
            // they are formed here (one call site of the Jacobian code)
            const bool fresh = !has_carried;
            has_carried = false;
            if (fresh) {
                pre(0, 1);
                __syncthreads();
            }
            joints_jacobian();
            if constexpr (JF)
                if (fresh) joints_plain(0, 1);
            __syncthreads();
            if (fresh) {
                gather_post(0, 1);
                __syncthreads();
                rnorm = uni(s.wn[0]);
            } else {
                rnorm = carried;
            }
            if (rnorm < best) {
                best = rnorm;
                for (int t = tid(); t < m * 6; t += NTH) s.best[t / 6][t % 6] = s.vmid0[t / 6][t % 6];
            }
            if (rnorm <= tol2) { converged = true; break; }
            if (!(rnorm < huge)) break;
            assemble();
            __syncthreads();
            assembled = true;
            int budget;
            if (seeded) budget = ns_warm;
            else { seed_transpose(); budget = ns_seed; }
            run_budget(budget, ns_seed);
            matvec();
            __syncthreads();
            // all-zero correction → hand over to subdivision
            if (uni(s.ibcast[0]) == 0) break;
            // backtracking line search from the same base, α = 2^-lane for lanes 0..7,
            // evaluated LS at a time; the first lane (in order) that lowers the
            // residual is taken, as when all eight are evaluated together
            // (the line search's base, a copy of the iterate, was written beside dv by
            // the matvec and published by its barrier)
            int accepted = -1, accepted_slot = -1;
            double tnorm = 0.0;
            for (int round = 0; round * LSR < LANES && accepted < 0; ++round) {
                const int nl = LANES - round * LSR < LSR ? LANES - round * LSR : LSR;
                if constexpr (SPLIT) {
                    ls_round_bodies_joints(round, nl, round == 0);
                    __syncthreads();
                    ls_round_gather(nl);
                    __syncthreads();
                } else {
                    pre_candidates(round);
                    __syncthreads();
                    joints_plain(0, LS);
                    __syncthreads();
                    gather_post(0, LS, round == 0);
                    __syncthreads();
                }
                for (int l = 0; l < nl; ++l)
                    if (uni(s.wn[l]) < rnorm) {
                        accepted = round * LSR + l;
                        accepted_slot = l;
                        tnorm = uni(s.wn[l]);
                        break;
                    }
            }
            if (accepted >= 0) {
                if (accepted > 0) {  // adopt_lane (cache.rs:1115-1136)
                    const int as = accepted_slot;
                    if (as > 0) {
                        for (int t = tid(); t < m * 6; t += NTH) {
                            s.vmid0[t / 6][t % 6] = VMID(as, t / 6)[t % 6];
                            s.rhs0[t / 6][t % 6] = RHS(as, t / 6)[t % 6];
                        }
                        for (int t = tid(); t < m * 8; t += NTH) s.mid0[t / 8][t % 8] = MID(as, t / 8)[t % 8];
                    }
                    // the accepted lane's mass block, recomputed bit-identically from its inputs
                    for (int t = tid(); t < md(); t += NTH) {
                        double block[6][6], rr[6], sc[2];
                        const double floor2 = (1.0 / 1000000000.0) * (1.0 / 1000000000.0);
                        body_post(MID(as, t), VMID(as, t), s.massv[t], s.tens[t], s.snap[t], TOTAL(as, t), half, floor2,
                                  block, rr, sc);
                        for (int r = 0; r < 6; ++r)
                            for (int c = 0; c < 6; ++c) s.mass0[t][r * 6 + c] = block[r][c];
                    }
                    __syncthreads();
                }
                has_carried = true;
                carried = tnorm;
            } else {
                for (int t = tid(); t < md() * 6; t += NTH) s.vmid0[t / 6][t % 6] = s.base[t / 6][t % 6] + s.dv[t / 6][t % 6] * 0.0;
                __syncthreads();
                st_reject++;
                break;
            }
        }
        __syncthreads();
        copy_vmid(0, s.best);
        __syncthreads();
    }

    static __device__ bool log2_floor(double v, int& e) {  // newton.rs:101-121
        double x = v;
        e = 0;
        if (!(x > 0.0)) return false;
        while (x >= 2.0) {
            x = x / 2.0;
            e += 1;
            if (e > 250) return false;
        }
        while (x < 1.0 && e > -250) {
            x = x * 2.0;
            e -= 1;
        }
        return true;
    }
    __device__ bool signal_log2(bool assembled, int& r) {
        if (!assembled) return false;
        double an = reduce_sq(A(), false);
        double mn = 0.0;
        {
            double part = 0.0;
            for (int e = tid(); e < md() * 36; e += NTH) {
                double v = s.mass0[e / 36][e % 36];
                part = part + v * v;
            }
            mn = block_sum(s, part);
        }
        double stiff = effective_zero(mn) ? 0.0 : an / mn;
        return log2_floor(stiff, r);
    }
    __device__ void restore_x() {
        double* Xm = X();
        const double* src = a.xacc + xoff;
        for (int e = tid(); e < n * n; e += NTH) Xm[(e / n) * LD + e % n] = src[e];
        __syncthreads();
    }
    __device__ void publish_x(bool& accepted_seeded) {
        const double* Xm = X();
        double* dst = a.xacc + xoff;
        // the copy and ‖X‖² in one pass (each thread's terms in reduce_sq's order)
        double part = 0.0;
        for (int e = tid(); e < n * n; e += NTH) {
            const double v = Xm[(e / n) * LD + e % n];
            dst[e] = v;
            const double x = 0.0 - v;
            part = part + x * x;
        }
        const double nx = block_sum(s, part);
        const double huge = 1000000000.0;
        if (nx < huge * huge) accepted_seeded = seeded;
        else { seeded = false; accepted_seeded = false; }
    }
    static __device__ int floor_div(int q, int d) { return (q >= 0) ? q / d : -((-q + d - 1) / d); }

    // ───── solve_step (newton.rs:543-683) ─────
    __device__ void solve_step(double tol2, int ns_warm, int ns_seed) {
        bool accepted_seeded = seeded;
        uint64_t remaining = 1ull << MAX_DEPTH;
        int depth = 0;
        while (remaining > 0) {
            uint64_t units = 1ull << (MAX_DEPTH - depth);
            while (units > remaining) { depth += 1; units >>= 1; }
            span = dt;
            for (int k = 0; k < depth; ++k) span = span / 2.0;
            half = span / 2.0;
            st_spans++;
            for (int t = tid(); t < m; t += NTH) world_momentum(s.pose[t], s.mom[t], s.snap[t]);
            __syncthreads();
            bool converged, assembled;
            double best;
            newton_solve(tol2, ns_warm, ns_seed, converged, best, assembled);
            bool finite = best < 1000000000.0;
            bool last = depth == MAX_DEPTH;
            if (!converged && !last) {
                restore_x();
                seeded = accepted_seeded;
                int r, down;
                if (signal_log2(assembled, r)) down = floor_div(r - THRESHOLD_LOG2, PER_HALVING_LOG2) + 1;
                else down = 1;
                int hi = MAX_DEPTH - depth;
                down = down < 1 ? 1 : (down > hi ? hi : down);
                depth += down;
                continue;
            }
            if (converged || finite) {
                if (!converged) st_uncommit++;
                for (int t = tid(); t < md(); t += NTH)   // the unknowns only (newton.rs:704-753)
                    commit_body(s.pose[t], s.mom[t], s.massv[t], s.tens[t], s.vmid0[t], s.snap[t], span, half,
                                converged);
                __syncthreads();
                publish_x(accepted_seeded);
            } else {
                restore_x();
                seeded = accepted_seeded;
                st_nonfinite++;
            }
            remaining -= units;
            auto aligned = [&](int d) {
                if (d <= 0) return false;
                uint64_t bigger = 1ull << (MAX_DEPTH - (d - 1));
                return bigger <= remaining && remaining % bigger == 0;
            };
            int r;
            if (remaining > 0 && converged && aligned(depth) && signal_log2(assembled, r)) {
                int up = floor_div(THRESHOLD_LOG2 - r, PER_HALVING_LOG2);
                while (up > 0 && aligned(depth)) { depth -= 1; up -= 1; }
            }
        }
    }
};

// ── the island's gravity wrenches from the gravity pass's partial sums ──
// In place of the gravity pass's reduce kernel, in two halves around the load phase's
// first barrier: thread 3i + k < 3m adds body i's axis-k partial sums over the tile's
// chunks in ascending order from 0 (loads issued four chunks at a time, beside the
// island's other loads) into s.ext[i][k]; after the barrier thread i < m forms the
// wrench with gravity_wrench. These are the reduce kernel's sums and arithmetic, so the
// wrenches are the same bits. Tiles are GRAV_TILE consecutive targets from body 0.
__device__ __forceinline__ void island_gravity_sum(const IslandArgs& a, int b0, int m, double (*ext)[6]) {
    const int t = threadIdx.x;
    if (t >= 3 * m) return;
    const int i = t / 3, k = t - 3 * i, gi = b0 + i;
    const GravTile& tl = a.grav_tiles[gi / GRAV_TILE];
    const double* __restrict__ p = a.grav_partial + (size_t)tl.e0 * GRAV_PARTIAL_PER_ENTRY + k * GRAV_TILE + (gi - tl.tgt0);
    double f = 0.0;
    int c = 0;
    for (; c + 4 <= tl.ne; c += 4) {
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = p[(size_t)(c + u) * GRAV_PARTIAL_PER_ENTRY];
#pragma unroll
        for (int u = 0; u < 4; ++u) f += v[u];
    }
    for (; c < tl.ne; ++c) f += p[(size_t)c * GRAV_PARTIAL_PER_ENTRY];
    ext[i][k] = f;
}
__device__ __forceinline__ void island_gravity_wrench(const IslandArgs& a, int b0, int m, double (*ext)[6]) {
    const int t = threadIdx.x;
    if (t >= m) return;
    double f[3] = {ext[t][0], ext[t][1], ext[t][2]};
    gravity_wrench(a.grav_charges + (size_t)(b0 + t) * 8, a.grav_tiles[(b0 + t) / GRAV_TILE].g_inv, f, ext[t]);
    for (int k = 0; k < 6; ++k) ext[t][k] = field_add(a, a.mass[b0 + t], k, ext[t][k]);
}

template <int MAXM, int MAXJ, int NTH, bool JF, bool FLAT = false>
__device__ __forceinline__ void island_step(const IslandArgs& a, IslandShared<MAXM, MAXJ, NTH, JF>& s) {
    if ((int)blockIdx.x >= a.n_islands) return;
    const int isl = a.island_ids ? a.island_ids[blockIdx.x] : (int)blockIdx.x;
    const int b0 = a.island_start[isl], m = a.island_end[isl] - b0;
    const int j0 = a.island_jstart[isl], J = a.island_jend[isl] - j0;
    if (m > MAXM || J > MAXJ || m <= 0) return;  // the host routes islands by size; never taken

    const bool island_reduce = a.grav_tiles != nullptr;
    if (island_reduce) island_gravity_sum(a, b0, m, s.ext);
    // ── load the island ──
    for (int t = threadIdx.x; t < m * 8; t += NTH) s.pose[t / 8][t % 8] = a.pose[(size_t)(b0 + t / 8) * 8 + t % 8];
    for (int t = threadIdx.x; t < m * 6; t += NTH) {
        s.mom[t / 6][t % 6] = a.mom[(size_t)(b0 + t / 6) * 6 + t % 6];
        if (!island_reduce)
            s.ext[t / 6][t % 6] = field_add(a, a.mass[b0 + t / 6], t % 6, a.ext[(size_t)(b0 + t / 6) * 6 + t % 6]);
        s.dv[t / 6][t % 6] = 0.0;
    }
    for (int t = threadIdx.x; t < m * 9; t += NTH) s.tens[t / 9][t % 9] = a.tens[(size_t)(b0 + t / 9) * 9 + t % 9];
    for (int t = threadIdx.x; t < m; t += NTH) s.massv[t] = a.mass[b0 + t];
    for (int t = threadIdx.x; t < J; t += NTH) {
        s.jkind[t] = a.jkind[j0 + t];
        s.ja[t] = a.ja[j0 + t] - b0;
        s.jb[t] = a.jb[j0 + t] - b0;
        for (int p = 0; p < 10; ++p) s.jparams[t][p] = a.jparams[(size_t)(j0 + t) * 10 + p];
    }
    const long long xoff = a.xacc_off[isl];
    constexpr int LD = IslandShared<MAXM, MAXJ, NTH, JF>::LDM;
    // the unknowns: the dynamic bodies, which come first (the host checks the order)
    int md = 0;
    for (int i = 0; i < m; ++i) md += a.mass[b0 + i] != 0.0 ? 1 : 0;
    md = uni(md);
    const int n = 6 * md;
    for (int e = threadIdx.x; e < IslandShared<MAXM, MAXJ, NTH, JF>::NB; e += NTH) {
        s.A[e] = 0.0;
        s.X[e] = 0.0;
        s.u.T[e] = 0.0;
    }
    __syncthreads();
    if (island_reduce) island_gravity_wrench(a, b0, m, s.ext);   // published by the barriers below
    for (int e = threadIdx.x; e < n * n; e += NTH) s.X[(e / n) * LD + e % n] = a.xacc[xoff + e];
    __syncthreads();
    // ── topology tables (built on the host from the joint graph, one record per island) ──
    {
        using SH = IslandShared<MAXM, MAXJ, NTH, JF>;
        static_assert(MAXM <= TOPO_MAXM && MAXJ <= TOPO_MAXJ && SH::WAVES <= TOPO_WAVES &&
                          SH::PER_WAVE <= TOPO_PER_WAVE && SH::MAXT <= TOPO_MAXT,
                      "island class exceeds the topology record");
        const IslandTopo& tp = a.topo[isl];
        const int t = threadIdx.x;
        if (t <= MAXM) {
            s.nbr_start[t] = tp.nbr_start[t];
            s.inc_start[t] = tp.inc_start[t];
        }
        if (t < MAXM * MAXM) s.nbr[t] = tp.nbr[t];
        if (t <= MAXM * MAXM) s.term_start[t] = tp.term_start[t];
        if (t < 4 * SH::JD) s.terms[t] = tp.terms[t];
        if (t < 2 * SH::JD) s.inc[t] = tp.inc[t];
        if (t < SH::JD) s.jorder[t] = tp.jorder[t];
        if (t < SH::MAXT) s.kmask[t] = tp.kmask[t];
        if (t == 0) s.nd = md;
        if (t < 2 * SH::WAVES * SH::PER_WAVE) {
            const int k = t / (SH::WAVES * SH::PER_WAVE), w = (t / SH::PER_WAVE) % SH::WAVES, u = t % SH::PER_WAVE;
            const int e = tp.tlist[k][w][u];
            s.tlist[k][w][u] = (e & 15) | ((e >> 4) << 8);   // row tile | column tile << 8
            if (u == 0) s.tcount[k][w] = tp.tcount[k][w];
        }
    }
    __syncthreads();
    // the block row of every neighbour-list entry (read by the assembly; published by
    // the barrier that follows the world momenta)
    if ((int)threadIdx.x < s.nbr_start[m]) {
        int i = 0;
        while (s.nbr_start[i + 1] <= (int)threadIdx.x) ++i;
        s.nbr_row[threadIdx.x] = (unsigned char)i;
    }

    IslandSolver<MAXM, MAXJ, NTH, JF, FLAT> sv{s, a};
    sv.m = m;
    sv.J = J;
    sv.n = n;
    sv.n16 = ((n + 15) / 16) * 16;
    sv.mfma = a.ns_engine == 0;
    sv.isl = isl;
    sv.first_body = b0;
    sv.dt = a.dt;
    sv.warp = a.warp;
    sv.xoff = xoff;
    sv.st_spans = sv.st_newton = sv.st_ns = sv.st_reseed = sv.st_reject = sv.st_uncommit = sv.st_nonfinite = 0;
    sv.seeded = a.seeded[isl] != 0;
    sv.solve_step(a.tol2, a.ns_warm, a.ns_seed);
    __syncthreads();

    // ── recompute_centroid (island.rs:206-244) and publish the charges ──
    // One thread per body forms its position (into snap, whose span is over); thread 0
    // adds them in body order and forms the shift; one thread per body composes its pose
    // with the shift and checks its state. (best holds the shift: the solve is over.)
    double* shift_sh = &s.best[0][0];
    if (threadIdx.x < m) body_position(s.pose[threadIdx.x], s.snap[threadIdx.x]);
    if (threadIdx.x == 0) {
        s.ibcast[1] = sv.seeded ? 1 : 0;
        s.ibcast[2] = 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total_m = 0.0, wx[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < m; ++i) {
            total_m += s.massv[i];
            for (int k = 0; k < 3; ++k) wx[k] += s.snap[i][k] * s.massv[i];
        }
        const bool shifted = !effective_zero(total_m);
        if (shifted) {
            double inv = 1.0 / total_m;
            double sh[6] = {-(wx[0] * inv), -(wx[1] * inv), -(wx[2] * inv), 0.0, 0.0, 0.0};
            twist_exp(sh, 1.0, shift_sh);
            for (int k = 0; k < 3; ++k) a.origin[isl * 3 + k] += wx[k] * inv;
        }
        s.ibcast[3] = shifted ? 1 : 0;
        for (int k = 0; k < 3; ++k) s.bcast[2 + k] = a.origin[isl * 3 + k];
    }
    __syncthreads();
    if (threadIdx.x < m) {
        const int i = threadIdx.x;
        if (s.ibcast[3]) {
            double np[8];
            compose(s.pose[i], shift_sh, np);
            for (int k = 0; k < 8; ++k) s.pose[i][k] = np[k];
        }
        bool bad = false;
        for (int k = 0; k < 8; ++k) bad |= !isfinite(s.pose[i][k]);
        for (int k = 0; k < 6; ++k) bad |= !isfinite(s.mom[i][k]);
        if (bad) s.ibcast[2] = 1;
    }
    __syncthreads();
    double org[3] = {s.bcast[2], s.bcast[3], s.bcast[4]};
    for (int t = threadIdx.x; t < m; t += NTH) {
        double p[3];
        body_position(s.pose[t], p);
        double* c = a.charges_out + (size_t)(b0 + t) * 8;
        for (int k = 0; k < 3; ++k) {
            c[k] = p[k] + org[k];
            c[3 + k] = p[k];
        }
        c[6] = (a.G_of_island ? a.G_of_island[isl] : a.G) * s.massv[t];
        c[7] = a.size[b0 + t];
    }
    // ── write the state back ──
    for (int t = threadIdx.x; t < m * 8; t += NTH) a.pose[(size_t)(b0 + t / 8) * 8 + t % 8] = s.pose[t / 8][t % 8];
    for (int t = threadIdx.x; t < m * 6; t += NTH) a.mom[(size_t)(b0 + t / 6) * 6 + t % 6] = s.mom[t / 6][t % 6];
    if (threadIdx.x == 0) {
        a.seeded[isl] = s.ibcast[1];
        int* st = a.stats + (size_t)isl * ISLAND_STATS;
        st[0] += sv.st_spans;
        st[1] += sv.st_newton;
        st[2] += sv.st_ns;
        st[3] += sv.st_reseed;
        st[4] += sv.st_reject;
        st[5] += sv.st_uncommit;
        st[STAT_NONFINITE_SPANS] += sv.st_nonfinite;
        if (s.ibcast[2]) mark_fatal(a, isl);
    }
}

// Kernel variants by island class (the host routes each island to the smallest
// that holds it; grids run over that class's island ids):
//   m1    single bodies without joints: one thread per island (singles.hip)
//   m4j10 up to 4 bodies / 10 joints (a 4-body rig: springs and torsional dampers
//         around the ring and across one diagonal): 128 threads
//   m8j14 up to 8 bodies / 14 joints: the shared state fits two workgroups per CU
//         (two islands in flight, 2 waves per SIMD, ≤ 256 registers per lane)
//   m8j16 up to 8 bodies / 16 joints: one workgroup holds the CU
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m4j10(IslandArgs a) {
    __shared__ IslandShared<4, 10, 128, false> s;
    island_step<4, 10, 128, false>(a, s);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m8j14(IslandArgs a) {
    __shared__ IslandShared<8, 14, 256, false> s;
    island_step<8, 14, 256, false>(a, s);
}
// the same step with four line-search candidates per round, the joints on every wave
// and one Jacobian column per lane (the same results to rounding: the fused
// multiply-adds fall differently; selected per world by its joint mapping)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m8j14_flat(IslandArgs a) {
    __shared__ IslandShared<8, 14, 256, false> s;
    island_step<8, 14, 256, false, true>(a, s);
}
__global__ __launch_bounds__(256) void island_step_kernel_m8j16(IslandArgs a) {
    __shared__ IslandShared<8, 16, 256, false> s;
    island_step<8, 16, 256, false>(a, s);
}
// Mixed precision: FP32 Jacobian lanes and compensated FP32 assembly (f64 residual,
// Newton–Schulz and state update).
__global__ __launch_bounds__(128) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m4j10_jf32(IslandArgs a) {
    __shared__ IslandShared<4, 10, 128, true> s;
    island_step<4, 10, 128, true>(a, s);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m8j14_jf32(IslandArgs a) {
    __shared__ IslandShared<8, 14, 256, true> s;
    island_step<8, 14, 256, true>(a, s);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m8j14_jf32_flat(IslandArgs a) {
    __shared__ IslandShared<8, 14, 256, true> s;
    island_step<8, 14, 256, true, true>(a, s);
}
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void island_step_kernel_m8j16_jf32(IslandArgs a) {
    __shared__ IslandShared<8, 16, 256, true> s;
    island_step<8, 16, 256, true>(a, s);
}

int island_class(int m, int j) {
    if (m == 1 && j == 0) return ISLAND_CLASS_M1;
    if (m <= 4 && j <= 10) return ISLAND_CLASS_M4J10;
    if (m <= 8 && j <= 14) return ISLAND_CLASS_M8J14;
    if (m <= 8 && j <= 16) return ISLAND_CLASS_M8J16;
    if (m <= LARGE_MAXM && j <= LARGE_MAXJ) return ISLAND_CLASS_LARGE;
    return -1;
}

void island_class_shape(int cls, int& waves, int& per_wave) {
    switch (cls) {
        case ISLAND_CLASS_M4J10:
            waves = IslandShared<4, 10, 128, false>::WAVES;
            per_wave = IslandShared<4, 10, 128, false>::PER_WAVE;
            break;
        case ISLAND_CLASS_M8J14:
        case ISLAND_CLASS_M8J14_FLAT:
            waves = IslandShared<8, 14, 256, false>::WAVES;
            per_wave = IslandShared<8, 14, 256, false>::PER_WAVE;
            break;
        case ISLAND_CLASS_M8J16:
            waves = IslandShared<8, 16, 256, false>::WAVES;
            per_wave = IslandShared<8, 16, 256, false>::PER_WAVE;
            break;
        default:
            waves = per_wave = 1;
    }
}

hipError_t ClassStreams::create() {
    hipError_t e = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
    for (int k = 0; k < ISLAND_CLASSES && e == hipSuccess; ++k) {
        e = hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&join[k], hipEventDisableTiming);
    }
    return e;
}
void ClassStreams::destroy() {
    for (int k = 0; k < ISLAND_CLASSES; ++k) {
        if (join[k]) (void)hipEventDestroy(join[k]);
        if (side[k]) (void)hipStreamDestroy(side[k]);
        join[k] = nullptr;
        side[k] = nullptr;
    }
    if (fork) (void)hipEventDestroy(fork);
    fork = nullptr;
}

hipError_t launch_classes(IslandArgs a, const ClassLaunch& c, hipEvent_t start, hipEvent_t stop) {
    static const int order[ISLAND_CLASSES] = {ISLAND_CLASS_LARGE, ISLAND_CLASS_M8J14, ISLAND_CLASS_M8J14_FLAT,
                                              ISLAND_CLASS_M8J16, ISLAND_CLASS_M4J10, ISLAND_CLASS_M1};
    int live[ISLAND_CLASSES], nl = 0;
    for (int k = 0; k < ISLAND_CLASSES; ++k)
        if (c.n[order[k]] > 0) live[nl++] = order[k];
    if (nl == 0) return hipSuccess;
    if (nl == 1) {
        a.island_ids = c.ids[live[0]];
        a.n_islands = c.n[live[0]];
        launch_island_class(a, live[0], c.main, start, stop);
        return hipGetLastError();
    }
    hipError_t e = hipEventRecord(c.fork, c.main);
    for (int k = 1; k < nl && e == hipSuccess; ++k) e = hipStreamWaitEvent(c.side[live[k]], c.fork, 0);
    if (e != hipSuccess) return e;
    for (int k = 0; k < nl; ++k) {
        a.island_ids = c.ids[live[k]];
        a.n_islands = c.n[live[k]];
        launch_island_class(a, live[k], k == 0 ? c.main : c.side[live[k]], k == 0 ? start : nullptr, nullptr);
    }
    e = hipGetLastError();
    for (int k = 1; k < nl && e == hipSuccess; ++k) {
        e = hipEventRecord(c.join[live[k]], c.side[live[k]]);
        if (e == hipSuccess) e = hipStreamWaitEvent(c.main, c.join[live[k]], 0);
    }
    if (e == hipSuccess && stop) e = hipEventRecord(stop, c.main);
    return e;
}

void launch_island_class(const IslandArgs& a, int cls, hipStream_t stream, hipEvent_t start, hipEvent_t stop) {
    if (a.n_islands <= 0) return;
    const dim3 grid(a.n_islands);
    switch (cls) {
        case ISLAND_CLASS_M1:
            launch_singles(a, stream, start, stop);
            break;
        case ISLAND_CLASS_M4J10:
            if (a.jac_f32) launch_k(island_step_kernel_m4j10_jf32, grid, dim3(128), stream, start, stop, a);
            else launch_k(island_step_kernel_m4j10, grid, dim3(128), stream, start, stop, a);
            break;
        case ISLAND_CLASS_M8J14:
            if (a.jac_f32) launch_k(island_step_kernel_m8j14_jf32, grid, dim3(256), stream, start, stop, a);
            else launch_k(island_step_kernel_m8j14, grid, dim3(256), stream, start, stop, a);
            break;
        case ISLAND_CLASS_M8J14_FLAT:
            if (a.jac_f32) launch_k(island_step_kernel_m8j14_jf32_flat, grid, dim3(256), stream, start, stop, a);
            else launch_k(island_step_kernel_m8j14_flat, grid, dim3(256), stream, start, stop, a);
            break;
        case ISLAND_CLASS_LARGE:
            launch_island_large(a, stream, start, stop);
            break;
        default:
            if (a.jac_f32) launch_k(island_step_kernel_m8j16_jf32, grid, dim3(256), stream, start, stop, a);
            else launch_k(island_step_kernel_m8j16, grid, dim3(256), stream, start, stop, a);
            break;
    }
}

}  // namespace exg
