// mdn_kernels.h -- gfx950 kernels of the mechanical-equilibration path.
//
// The interaction kernels are templates over the arithmetic type R of the per-interaction math
// (double: the FP64 path; float: the mixed path). Coordinates, search directions, forces after
// the gather, every reduction and all CG control scalars are FP64 in both instantiations; the
// slots hold R.
//
// Layout in HBM (all device buffers owned by the context):
//   x[2][3*n_beads] (current / next coordinates), F, F_prev, d[2] (search direction, current
//   / next), x_min[2] (minimum-energy backup, committed / candidate): the reference's flat AoS
//   coord vectors (DofSerializer.hpp:64-187), double.
//   Term tables: structure of arrays per force field (bead indices int32, parameters double).
//   slots[S]: one 3-vector of R (24 B in FP64, 12 B in the mixed path) per (interaction, bead
//   of the interaction) of the linkers, motors, excluded volume and boundary, laid out
//   bead-sorted: bead b owns slots [bstart[b], bstart[b+1]) in ascending order of (force field,
//   interaction, bead of the interaction), which is the order the reference adds into `force`.
//   Interaction kernels write through pos[] (slot id -> bead-sorted position). The filament
//   stretching and bending forces, which come first in that order, go to bead-major pieces
//   piece[5][n_beads] instead (coalesced writes: consecutive cylinders have consecutive beads),
//   or are recomputed per bead by bead blocks and summed in the same order into bsum[n_beads]
//   (the bonded sums, DevSys::bonded_gather); the gather starts from a bead's pieces or bonded
//   sum, then streams its contiguous slot range, so it reproduces the reference's per-bead
//   summation order exactly and needs no atomics.
//
// Grid layout of the interaction kernels: blocks [blk[k], blk[k+1]) hold force field dff[k]'s
// interactions (each range padded to whole blocks, so a block belongs to one force field; the
// longest-running force fields first, the filament terms last), then, where the bonded sums are
// used, bead blocks.
//
// Reductions are deterministic: block tree reduction, block partials in fixed slots, then a
// one-block finalize kernel reduces the partials with all its threads in a fixed order and runs
// the scalar control logic of CGMethod::minimize on an LDS copy of the state. (A last-block
// election would need a device-scope fence per block, which on gfx950 writes back the L2 each
// time; one kernel boundary is far cheaper.)
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdn_terms.h"

namespace mdn {

constexpr int BLOCK = 256;
constexpr int FIN = 1024;        // threads of the one-block finalize kernels
constexpr int NFF = 6;
constexpr int MAXELEM = 8;       // boundary elements held in the kernel argument
constexpr int NTRIAL = 2;        // line-search step lengths evaluated per launch (lambda, lambda/2)
constexpr int PSTRIDE = 4;       // partial slots per block
// beads per gather block (GB / 64 waves sum, all 256 threads stage slots): 64 below
// GB_LARGE_BEADS beads, 128 from there on (half the blocks to dispatch and to finalize)
constexpr int GB_LARGE_BEADS = 100000;
constexpr int DEFER_CAP = 256;   // excluded-volume energies deferred to the Simpson pass per launch

// An excluded-volume pair whose closed-form energy is non-finite (or < -1): the reference then
// integrates numerically (energyN). The pass runs block-parallel in the finalize kernel.
struct EnDefer {
    int code;      // trial / report slot the energy belongs to
    int t;         // pair index
    double ks;     // kRepScaled
    double c[12];  // the four bead positions
    double U;      // result
};

enum { FF_STRETCH = 0, FF_BEND, FF_LINKER, FF_MOTOR, FF_EXVOL, FF_BOUNDARY };

struct DevSys {
    int n_beads;
    int n[NFF];                 // interactions per force field
    // interaction blocks in dispatch order: range k = [blk[k], blk[k + 1]) holds force field
    // dff[k]; blk[NFF] = all interaction blocks, blk[NFF + 1] = + one block per BLOCK beads.
    // fblk[ff] = first block of force field ff.
    int blk[NFF + 2];
    int dff[NFF];
    int fblk[NFF];
    long slot0[NFF + 1];        // first slot id of each force field, total
    // stretching
    const int* str_b;  const double* str_k; const double* str_eql;
    // bending
    const int* bend_b; const double* bend_k; const double* bend_eqt;
    // linkers / motors
    const int* lnk_b; const double* lnk_k; const double* lnk_eql; const double* lnk_p1; const double* lnk_p2;
    double* lnk_sf;
    const int* mot_b; const double* mot_k; const double* mot_eql; const double* mot_p1; const double* mot_p2;
    double* mot_sf;
    // The linker / motor tables (and the tensions *_sf) are stored in the order of the first bead
    // of each crosslink's first cylinder, so that a wave's table reads, first-cylinder coordinates
    // and slot writes are close together; *_ord[i] is the crosslink's own index, which places its
    // slots (so the results do not depend on the order)
    const int* lnk_ord; const int* mot_ord;
    // excluded volume: beads (4 per pair) and kRepScaled = krep * L0_i * L0_j
    const int* ev_b;   const double* ev_ks;
    // boundary: bead, element, krep, slen
    const int* bnd_b;  const int* bnd_e; const double* bnd_krep; const double* bnd_slen;
    int n_elem;
    BElem elem[MAXELEM];
    // filament bonded forces, bead-major: piece[k][bead] (3 values of the evaluation's arithmetic
    // type) for k = 0: the stretching term with the bead as bead 2, 1: as bead 1, 2..4: the bending
    // term with the bead as bead 3, 2, 1; pmask[bead] bit k marks the pieces the bead has
    void* piece;
    const unsigned char* pmask;
    // Bonded sums: instead of the piece arrays, bead blocks appended to the terms (or
    // line-search) kernel's grid recompute each bead's bonded pieces from the coordinates
    // (bonded_at_bead) and store their sum in piece order, bsum[bead] (FP64, 3 values), which the
    // gather then starts from; the filament terms' own blocks leave their forces out.
    // bonded_gather selects this for the minimizer's kernels (force evaluations choose on the
    // host). bnbr[j][bead], j = 0..3: the beads two before, one before, one after and two after
    // it along its filament (the bead itself where there is none); bterm[0][bead]: the stretching
    // term with the bead as bead 1, bterm[1][bead]: the bending term with the bead in the middle
    // (their neighbours along the filament are the adjacent term indices; out of range where the
    // term does not exist). blk_nb = the interaction blocks of the other force fields, which are
    // dispatched first.
    // bcontig: the filaments' beads are numbered consecutively, filament after filament (the
    // bead order of serializeDof), so a bead's neighbours are the adjacent indices and its
    // bending term is 2 bterm[0][bead] - bead - 1: the bead blocks then read bterm[0] alone
    const int* bnbr;
    const int* bterm;
    int bcontig;
    double* bsum;
    int bonded_gather;
    int blk_nb;
    // bead blocks (ceil(n_beads / BLOCK)) and the columns of a line-search slot's partials:
    // [0, blk[NFF]) interaction blocks, [blk[NFF], + nbb) the line search's bead blocks,
    // [blk[NFF] + nbb, + nbb) the bead kernel's (k_cg_pos)
    int nbb;
    int lsp;
    // with bonded_gather, where the line search takes the filament terms' energies: 0 their own
    // blocks (energies only), 1 the bead blocks (each bead the terms it starts), whose grid then
    // leaves the filament terms' blocks out
    int bead_energy;
    // bead-sorted slot layout (linkers, motors, excluded volume, boundary)
    const int* pos;        // [S] slot id -> position
    const int* bstart;     // [n_beads + 1]
    const int* sorted_id;  // [S] position -> slot id
    // deferred Simpson energies
    EnDefer* defer;
    int* defer_n;
};

// ---------------------------------------------------------------------------------------
// CG device state (CGMethod::minimize locals + MinimizationResult)
enum { PH_LS = 0, PH_MOVE = 1, PH_GATHER = 2 };
// CgState.status beyond the MDN_ERR_* codes: the minimizer paused for the numerical
// excluded-volume force fallback of the current step
constexpr int ST_NEED_FORCEN = 100;

struct CgState {
    // parameters
    int method, try_recover, force_call_limit, energy_call_limit;
    double max_dist, lambda_max, grad_tol, energy_rel_tol;
    // control
    int phase, parity, dpar, bkpar, fpar, done, status;
    int num_iter, num_energy_call, num_force_call, num_safe_mode, num_recovery;
    unsigned error_flags;
    int safe_mode, recovery_move;
    double grad_inf_norm, energy_rel_change, prev_energy, last_fdotf, fdotf;
    // pending shiftSearchDir(beta): d <- F + beta d; dir_reset marks resetCGParams
    int dir_pending, dir_reset;
    double dir_beta;
    // line search
    int ls_active, ecur_valid, ls_trials, spec_ok;
    double ls_lambda, ls_cur, ls_slope, ls_reduced, lambda, new_energy, rec_step;
    // minimum-energy backup (CGMethod.hpp:794-814)
    double total_energy, minimum_e;
    int have_min_e;
    // the quadratic line search (MDN_LS_QUADRATIC): line_search = ls_algo; q_phase marks a launch
    // that evaluates the parabola's vertex; the recorded (lambda, energy) list as its first entry
    // and the largest later energy; the parabola's points; the backtracking step and energy;
    // q_stale: the step is not the accepted trial, so E(x_new) must be evaluated
    int ls_algo, q_phase, q_nrec, q_iter, q_stale, q_pad;
    double q_lv0, q_ev0, q_emax, q_x0, q_x1, q_x2, q_y0, q_y1, q_y2, q_bt, q_ebt;
    // launches that did work (statistics)
    int n_ls_launch, n_move_launch, n_gather_launch;
    int reserved[2];
};

__device__ inline int ff_of_block(const DevSys& s, int b) {
    int k = 0;
#pragma unroll
    for (int q = 1; q <= NFF; q++) k += (b >= s.blk[q]) ? 1 : 0;
    return k < NFF ? s.dff[k] : NFF;   // NFF means a bead block
}

// ---- deterministic reduction helpers (64-wide waves) ----------------------------------
// A wave reduction in six DPP steps on the VALU (no LDS traffic): row shifts by 1, 2, 4 and 8
// leave each 16-lane row's total in its lane 15, row_bcast:15 adds row 0's total into row 1 and
// row 2's into row 3, row_bcast:31 adds lane 31 into rows 2 and 3; lane 63 then holds the
// reduction of all 64 lanes, read back into every lane. Lanes without a source keep the
// identity. The whole wave must be active.
template <int CTRL, int ROWS>
__device__ inline double dpp_or(double v, double ident) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(ident), __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(ident), __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ inline double lane63(double v) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), 63),
                            __builtin_amdgcn_readlane(__double2loint(v), 63));
}
template <int OP>   // 0 sum, 1 max, 2 min
__device__ inline double comb(double a, double b) { return OP == 0 ? a + b : OP == 1 ? rmax(a, b) : rmin(a, b); }
template <int OP>
__device__ inline double wave_reduce(double v) {
    const double id = OP == 0 ? 0.0 : OP == 1 ? -INFINITY : INFINITY;
    v = comb<OP>(v, dpp_or<0x111, 0xF>(v, id));   // row_shr:1
    v = comb<OP>(v, dpp_or<0x112, 0xF>(v, id));   // row_shr:2
    v = comb<OP>(v, dpp_or<0x114, 0xF>(v, id));   // row_shr:4
    v = comb<OP>(v, dpp_or<0x118, 0xF>(v, id));   // row_shr:8
    v = comb<OP>(v, dpp_or<0x142, 0xA>(v, id));   // row_bcast:15 -> rows 1, 3
    v = comb<OP>(v, dpp_or<0x143, 0xC>(v, id));   // row_bcast:31 -> rows 2, 3
    return lane63(v);
}
// The first three steps of wave_reduce: lane 8k + 7 then holds the reduction of lanes 8k .. 8k + 7.
template <int OP>
__device__ inline double wave_fold8(double v) {
    const double id = OP == 0 ? 0.0 : OP == 1 ? -INFINITY : INFINITY;
    v = comb<OP>(v, dpp_or<0x111, 0xF>(v, id));   // row_shr:1
    v = comb<OP>(v, dpp_or<0x112, 0xF>(v, id));   // row_shr:2
    v = comb<OP>(v, dpp_or<0x114, 0xF>(v, id));   // row_shr:4
    return v;
}
__device__ inline double wave_sum(double v) { return wave_reduce<0>(v); }
__device__ inline double wave_max(double v) { return wave_reduce<1>(v); }
__device__ inline double wave_min(double v) { return wave_reduce<2>(v); }
// op: 0 sum, 1 max, 2 min, for NV values at once over a block of NTH threads. Results valid in
// thread 0.
template <int NV, int NTH = BLOCK>
__device__ inline void block_reduce_n(double* v, const int* op, double* lds /* >= NV * NTH/64 */) {
    constexpr int NW = NTH / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++)
        v[i] = op[i] == 0 ? wave_sum(v[i]) : op[i] == 1 ? wave_max(v[i]) : wave_min(v[i]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) lds[i * NW + w] = v[i];
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) {
            double r = lds[i * NW];
            for (int q = 1; q < NW; q++) {
                double u = lds[i * NW + q];
                r = op[i] == 0 ? r + u : op[i] == 1 ? rmax(r, u) : rmin(r, u);
            }
            v[i] = r;
        }
}

// The same for a large block whose waves would otherwise each run NV full wave reductions on
// shared SIMDs: every wave folds groups of 8 lanes (three DPP row shifts), the NTH / 8 group
// results go through LDS, and wave i combines value i (NV <= NTH / 64) from its NTH / 8 group
// results with one full wave reduction, leaving it in res[i]. Fixed order, so deterministic;
// every thread returns with all NV results.
template <int NV, int NTH>
__device__ inline void block_reduce_fold(double* v, const int* op, double* lds /* >= NV * NTH / 8 */,
                                             double* res /* NV */) {
    static_assert(NV <= NTH / 64, "one wave per value");
    constexpr int NP = NTH / 8;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < NV; i++)
        v[i] = op[i] == 0 ? wave_fold8<0>(v[i]) : op[i] == 1 ? wave_fold8<1>(v[i]) : wave_fold8<2>(v[i]);
    if ((lane & 7) == 7)
#pragma unroll
        for (int i = 0; i < NV; i++) lds[i * NP + w * 8 + (lane >> 3)] = v[i];
    __syncthreads();
    if (w < NV) {   // wave-uniform
        const int o = op[w];
        const double id = o == 0 ? 0.0 : o == 1 ? -INFINITY : INFINITY;
        double r = lane < NP ? lds[w * NP + lane] : id;
        for (int q = lane + 64; q < NP; q += 64) {
            const double u = lds[w * NP + q];
            r = o == 0 ? r + u : o == 1 ? rmax(r, u) : rmin(r, u);
        }
        r = o == 0 ? wave_sum(r) : o == 1 ? wave_max(r) : wave_min(r);
        if (lane == 0) res[w] = r;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) v[i] = res[i];
}

// Pins values loaded at a kernel's entry before its first branch: without a use there, the
// compiler sinks each load into the branch that needs it, which turns one round trip to the
// state into a chain of dependent ones. An empty asm statement, no instruction.
__device__ inline void keep_loaded(int a) { asm volatile("" ::"s"(a)); }
__device__ inline void keep_loaded(double a) { asm volatile("" ::"s"(a)); }

// Stage the whole CG state through LDS: one parallel round trip instead of a chain of
// dependent global accesses by the thread that runs the scalar logic.
template <typename S>
__device__ inline void state_load(S* lds, const S* g) {
    const int n = sizeof(S) / sizeof(int);
    const int* src = reinterpret_cast<const int*>(g);
    int* dst = reinterpret_cast<int*>(lds);
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
    __syncthreads();
}
template <typename S>
__device__ inline void state_store(S* g, const S* lds) {
    __syncthreads();
    const int n = sizeof(S) / sizeof(int);
    const int* src = reinterpret_cast<const int*>(lds);
    int* dst = reinterpret_cast<int*>(g);
    for (int i = threadIdx.x; i < n; i += blockDim.x) dst[i] = src[i];
}

// ---- positions x + lambda * d_eff, with d_eff = d or the pending F + beta d ------------
struct Pos {
    const double* __restrict__ x;
    const double* __restrict__ d;
    const double* __restrict__ f;
    double beta, lam;
    int mode;   // 0: x; 1: x + lam d; 2: x + lam (F + beta d)
};

__device__ inline D3 pos_of(const Pos& p, int b) {
    const double* x = p.x + 3 * b;
    if (p.mode == 0) return D3{x[0], x[1], x[2]};
    const double* d = p.d + 3 * b;
    if (p.mode == 1) return D3{x[0] + p.lam * d[0], x[1] + p.lam * d[1], x[2] + p.lam * d[2]};
    const double* f = p.f + 3 * b;
    // shiftSearchDir (CGMethod.cpp:356-360) then calcCoordLineSearch (:43-46)
    double d0 = f[0] + p.beta * d[0], d1 = f[1] + p.beta * d[1], d2 = f[2] + p.beta * d[2];
    return D3{x[0] + p.lam * d0, x[1] + p.lam * d1, x[2] + p.lam * d2};
}

// Slots hold the per-(interaction, bead) forces in the arithmetic type R of the evaluation:
// 3 values per slot, bead-sorted position p at [3p, 3p + 3). Pieces: 3 values at
// [3 (k n_beads + bead), + 3).
template <typename R>
__device__ inline void put_piece(const DevSys& s, int k, int bead, D3 v) {
    R* q = static_cast<R*>(s.piece) + 3L * ((long)k * s.n_beads + bead);
    q[0] = (R)v.x; q[1] = (R)v.y; q[2] = (R)v.z;
}
template <typename R>
__device__ inline void put_slot(const DevSys& s, R* __restrict__ slots, long id, D3 v) {
    R* q = slots + 3L * s.pos[id];
    q[0] = (R)v.x; q[1] = (R)v.y; q[2] = (R)v.z;
}

// Energy (and, when slots != null, forces into the bead-sorted slots) of interaction t of
// force field ff at positions p. *bad: excluded-volume energy non-finite after the Simpson
// fallback (the reference's -1 sentinel). *fbad: excluded-volume analytic forces non-finite
// (the reference switches to forceN).
__device__ inline void push_defer(const DevSys& s, int code, int t, D3 c1, D3 c2, D3 c3, D3 c4) {
    const int k = atomicAdd(s.defer_n, 1);
    if (k >= DEFER_CAP) return;   // overflow is reported by the finalize kernel
    EnDefer& e = s.defer[k];
    e.code = code;
    e.t = t;
    e.ks = s.ev_ks[t];
    const D3 cs[4] = {c1, c2, c3, c4};
    for (int q = 0; q < 4; q++) { e.c[3 * q] = cs[q].x; e.c[3 * q + 1] = cs[q].y; e.c[3 * q + 2] = cs[q].z; }
}

// code: which energy the value belongs to (recorded with a deferred Simpson pair), or -1 when
// only the forces are wanted. *bad_t receives the excluded-volume pair whose forces need the
// numerical fallback. Excluded volume is evaluated in FP64 in both instantiations: its closed
// form divides a bracket that vanishes like the squared triple product of the pair by that
// square, so FP32 rounding is amplified like 1/tau^2 for near-coplanar pairs (tau the normalized
// triple product), and FP32 gains nothing on gfx950 for this latency-bound term.
// SPLIT: an excluded-volume pair's energy and forces by two calls, each forming the pair's
// invariants: more arithmetic, fewer registers live at once, for kernels whose register budget
// the shared form overflows. The FP64 path's values are the same bit for bit; the mixed path's
// FMA-contracted forms may fuse differently in the other context (differences at rounding level).
template <typename R, bool SPLIT = false>
__device__ inline double term_ef(const DevSys& s, int ff, int t, const Pos& p, R* __restrict__ slots,
                                 int code, int* fbad, int* bad_t = nullptr) {
    D3 out[4];
    switch (ff) {
    // filament stretching / bending: forces into the bead-major pieces
    case FF_STRETCH: {
        const int b1 = s.str_b[2 * t], b2 = s.str_b[2 * t + 1];
        D3 c1 = pos_of(p, b1), c2 = pos_of(p, b2);
        if (slots) {
            stretch_forces<R>(c1, c2, s.str_k[t], s.str_eql[t], out);
            put_piece<R>(s, 1, b1, out[0]);
            put_piece<R>(s, 0, b2, out[1]);
        }
        return (double)stretch_energy<R>(c1, c2, s.str_k[t], s.str_eql[t]);
    }
    case FF_BEND: {
        const int b1 = s.bend_b[3 * t], b2 = s.bend_b[3 * t + 1], b3 = s.bend_b[3 * t + 2];
        D3 c1 = pos_of(p, b1), c2 = pos_of(p, b2), c3 = pos_of(p, b3);
        if (slots) {
            bend_forces<R>(c1, c2, c3, s.bend_k[t], s.bend_eqt[t], out);
            put_piece<R>(s, 4, b1, out[0]);
            put_piece<R>(s, 3, b2, out[1]);
            put_piece<R>(s, 2, b3, out[2]);
        }
        return (double)bend_energy<R>(c1, c2, c3, s.bend_k[t], s.bend_eqt[t]);
    }
    case FF_LINKER:
    case FF_MOTOR: {
        const bool lk = ff == FF_LINKER;
        const int* b = lk ? s.lnk_b : s.mot_b;
        const double k = lk ? s.lnk_k[t] : s.mot_k[t], eql = lk ? s.lnk_eql[t] : s.mot_eql[t];
        const double p1 = lk ? s.lnk_p1[t] : s.mot_p1[t], p2 = lk ? s.lnk_p2[t] : s.mot_p2[t];
        D3 c1 = pos_of(p, b[4 * t]), c2 = pos_of(p, b[4 * t + 1]), c3 = pos_of(p, b[4 * t + 2]), c4 = pos_of(p, b[4 * t + 3]);
        if (slots) {
            double sf = xlink_forces<R>(c1, c2, c3, c4, k, eql, p1, p2, out);
            (lk ? s.lnk_sf : s.mot_sf)[t] = sf;
            const long id = s.slot0[ff] + 4L * (lk ? s.lnk_ord : s.mot_ord)[t];
            for (int q = 0; q < 4; q++) put_slot(s, slots, id + q, out[q]);
        }
        return (double)xlink_energy<R>(c1, c2, c3, c4, k, eql, p1, p2);
    }
    case FF_EXVOL: {
        const int tp = t;
        D3 c1 = pos_of(p, s.ev_b[4 * tp]), c2 = pos_of(p, s.ev_b[4 * tp + 1]), c3 = pos_of(p, s.ev_b[4 * tp + 2]),
           c4 = pos_of(p, s.ev_b[4 * tp + 3]);
        const long id = s.slot0[FF_EXVOL] + 4L * tp;
        // FP64 in both paths: the reference's expression order (FP64 path), the reduced-operation
        // form of the same closed form (mixed path)
        constexpr bool fast = sizeof(R) == 4;
        int deferred = 0;
        double U = 0.0;
        if (slots) {
            bool ok = true;
            if (fast) {
                if (code < 0) {
                    ok = exvol_forces_fast(c1, c2, c3, c4, s.ev_ks[tp], out);
                } else if (SPLIT) {
                    U = exvol_energy_fast(c1, c2, c3, c4, s.ev_ks[tp], &deferred);
                    ok = exvol_forces_fast(c1, c2, c3, c4, s.ev_ks[tp], out);
                } else {
                    U = exvol_energy_forces_fast(c1, c2, c3, c4, s.ev_ks[tp], &deferred, &ok, out);
                }
            } else if (code < 0) {
                ok = exvol_forces<double>(c1, c2, c3, c4, s.ev_ks[tp], out);
            } else if (SPLIT) {
                U = exvol_energy<double>(c1, c2, c3, c4, s.ev_ks[tp], &deferred);
                ok = exvol_forces<double>(c1, c2, c3, c4, s.ev_ks[tp], out);
            } else {
                U = exvol_energy_forces<double>(c1, c2, c3, c4, s.ev_ks[tp], &deferred, &ok, out);
            }
            if (!ok) {
                out[0] = out[1] = out[2] = out[3] = D3{0.0, 0.0, 0.0};
                *fbad = 1;
                if (bad_t) *bad_t = tp;
            }
            for (int q = 0; q < 4; q++) put_slot(s, slots, id + q, out[q]);
        } else {
            U = fast ? exvol_energy_fast(c1, c2, c3, c4, s.ev_ks[tp], &deferred)
                     : exvol_energy<double>(c1, c2, c3, c4, s.ev_ks[tp], &deferred);
        }
        if (deferred && code >= 0) push_defer(s, code, tp, c1, c2, c3, c4);
        return U;
    }
    default: {
        D3 c = pos_of(p, s.bnd_b[t]);
        const BElem& e = s.elem[s.bnd_e[t]];
        if (slots) put_slot(s, slots, s.slot0[FF_BOUNDARY] + t, boundary_force<R>(e, c, s.bnd_krep[t], s.bnd_slen[t]));
        return (double)boundary_energy<R>(e, c, s.bnd_krep[t], s.bnd_slen[t]);
    }
    }
}

// Gather of GB beads per block, in the reference's per-bead `+=` order:
//  1. the bead's filament bonded pieces (pmask bits; bmask bit 0 selects the stretching pieces
//     0-1, bit 1 the bending pieces 2-4): independent coalesced loads;
//  2. the bead's slots (linkers, motors, excluded volume, boundary), which own the contiguous
//     range [bstart[b0], bstart[b0 + GB]): all 256 threads stream it into LDS with coalesced,
//     unrolled loads (the first chunk's loads issued before the pieces are added), then the first
//     wave adds each bead's slots in ascending order.
// Only slots whose id is in [lo, hi) are added when `all` is false. The first wave accumulates in
// FP64 and returns the sum for threads < GB.
constexpr int GCHUNK_BYTES = 12288;   // LDS bytes per chunk (512 FP64 slots, 1024 FP32 slots)
template <typename R> constexpr int gchunk() { return GCHUNK_BYTES / (3 * (int)sizeof(R)); }
template <typename R> constexpr int gunroll() { return 3 * gchunk<R>() / BLOCK; }
constexpr int NBOND = 5;              // bonded pieces per bead

// The slot buffer is padded by GCHUNK_PAD values past its last slot and the piece arrays cover
// every bead, so the loads below are unconditional (the results are selected, not branched on).
constexpr long GCHUNK_PAD = GCHUNK_BYTES / 4;   // values (>= 3 * gchunk<R>() for R = float, double)

// The bonded pieces of bead b (the bits of m, in piece order k = 0..4: 0 / 1 the stretching term
// with the bead as bead 2 / 1, 2 / 3 / 4 the bending term with the bead as bead 3 / 2 / 1)
// recomputed from the coordinates x and added to acc: the same term functions on the same inputs
// as in term_ef, so the values the piece arrays would hold (rounded to R, accumulated in FP64).
// Every load is issued up front and every term evaluated (absent pieces see in-range beads and
// terms and are dropped by the select), so no load waits behind a branch. With E, *e also
// receives the energies of the filament terms bead b starts, as bead_bonded_energy forms them
// (the same term functions on the same coordinates and constants, already loaded here).
template <typename R, bool E = false>
__device__ inline D3 bonded_at_bead(const DevSys& s, const double* __restrict__ x, int b, unsigned m, D3 acc,
                                    double* e = nullptr) {
    const long nb = s.n_beads;
    const int ns = s.n[FF_STRETCH], nd = s.n[FF_BEND];
    int id[5], t1, tc;   // t1: stretching term (bead, next); tc: bending term (previous, bead, next)
    if (s.bcontig) {     // absent neighbours: in-range adjacent beads, dropped by the mask
        const int hi = (int)nb - 1;
        id[0] = max(b - 2, 0); id[1] = max(b - 1, 0); id[2] = b; id[3] = min(b + 1, hi); id[4] = min(b + 2, hi);
        t1 = s.bterm[b];
        tc = 2 * t1 - b - 1;
    } else {
        id[0] = s.bnbr[b]; id[1] = s.bnbr[nb + b]; id[2] = b; id[3] = s.bnbr[2 * nb + b]; id[4] = s.bnbr[3 * nb + b];
        t1 = s.bterm[b];
        tc = s.bterm[nb + b];
    }
    D3 c[5];   // along the filament: two before .. two after
#pragma unroll
    for (int j = 0; j < 5; j++) c[j] = D3{x[3L * id[j]], x[3L * id[j] + 1], x[3L * id[j] + 2]};
    double sk[2] = {0.0, 0.0}, sl[2] = {0.0, 0.0}, bk[3] = {0.0, 0.0, 0.0}, bt[3] = {0.0, 0.0, 0.0};
    if (ns > 0)
#pragma unroll
        for (int j = 0; j < 2; j++) {
            const int t = min(max(t1 - 1 + j, 0), ns - 1);
            sk[j] = s.str_k[t];
            sl[j] = s.str_eql[t];
        }
    if (nd > 0)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int t = min(max(tc - 1 + j, 0), nd - 1);
            bk[j] = s.bend_k[t];
            bt[j] = s.bend_eqt[t];
        }
    D3 o[5][3];
    stretch_forces<R>(c[1], c[2], sk[0], sl[0], o[0]);       // k = 0: as bead 2 -> o[0][1]
    stretch_forces<R>(c[2], c[3], sk[1], sl[1], o[1]);       // k = 1: as bead 1 -> o[1][0]
    bend_forces<R>(c[0], c[1], c[2], bk[0], bt[0], o[2]);    // k = 2: as bead 3 -> o[2][2]
    bend_forces<R>(c[1], c[2], c[3], bk[1], bt[1], o[3]);    // k = 3: as bead 2 -> o[3][1]
    bend_forces<R>(c[2], c[3], c[4], bk[2], bt[2], o[4]);    // k = 4: as bead 1 -> o[4][0]
    const D3 v[5] = {o[0][1], o[1][0], o[2][2], o[3][1], o[4][0]};
#pragma unroll
    for (int k = 0; k < NBOND; k++) {
        const bool on = (m >> k) & 1;
        acc.x = on ? acc.x + v[k].x : acc.x;
        acc.y = on ? acc.y + v[k].y : acc.y;
        acc.z = on ? acc.z + v[k].z : acc.z;
    }
    if constexpr (E) {   // the stretching term (b, next) and the bending term centred on b
        const double es = (double)stretch_energy<R>(c[2], c[3], sk[1], sl[1]);
        const double eb = (double)bend_energy<R>(c[1], c[2], c[3], bk[1], bt[1]);
        double en = 0.0;
        en = (m & 0x02u) ? en + es : en;   // bead 1 of a stretching term
        en = (m & 0x08u) ? en + eb : en;   // bead 2 of a bending term
        *e = en;
    }
    return acc;
}

// A bead block: bsum[b] = base[b] (or 0) + bead b's bonded pieces selected by bmask, in piece
// order, recomputed at the coordinates x. Bead index b = k-th bead block * BLOCK + thread.
template <typename R>
__device__ inline void bead_bonded_sum(const DevSys& s, const double* __restrict__ x, int k, int bmask,
                                       const double* __restrict__ base) {
    const int b = k * BLOCK + threadIdx.x;
    if (b >= s.n_beads) return;
    const unsigned m = bmask ? s.pmask[b] & ((bmask & 1 ? 0x3u : 0u) | (bmask & 2 ? 0x1cu : 0u)) : 0u;
    D3 acc = base ? D3{base[3L * b], base[3L * b + 1], base[3L * b + 2]} : D3{0.0, 0.0, 0.0};
    acc = bonded_at_bead<R>(s, x, b, m, acc);
    s.bsum[3L * b] = acc.x;
    s.bsum[3L * b + 1] = acc.y;
    s.bsum[3L * b + 2] = acc.z;
}

// BG: each bead starts from its bonded sum bsum (base and the bonded pieces, written by the
// preceding kernel's bead blocks) instead of acc plus the piece arrays.
template <typename R, int GB, bool BG = false>
__device__ inline D3 gather_block(const DevSys& s, int bmask, const R* __restrict__ slots, long lo, long hi,
                                  bool all, D3 acc, double* lds_raw /* GCHUNK_BYTES */) {
    constexpr int CH = gchunk<R>(), UN = gunroll<R>();
    static_assert(3 * CH <= GCHUNK_PAD, "slot padding");
    R* lds = reinterpret_cast<R*>(lds_raw);
    const int b0 = blockIdx.x * GB;
    const int b = b0 + threadIdx.x;
    const int bend = min(b0 + GB, s.n_beads);
    const bool mine = threadIdx.x < GB && b < s.n_beads;
    const int bc = mine ? b : b0;   // an in-range bead for the unconditional loads
    // bonded pieces in flight first
    const unsigned m = (!BG && mine && bmask) ? s.pmask[bc] & ((bmask & 1 ? 0x3u : 0u) | (bmask & 2 ? 0x1cu : 0u)) : 0u;
    R pv[NBOND][3] = {};
    if (BG && threadIdx.x < GB) {   // the summing waves (uniform branch)
        const double* q = s.bsum + 3L * bc;
        acc = mine ? D3{q[0], q[1], q[2]} : D3{0.0, 0.0, 0.0};
    }
    if (!BG && threadIdx.x < GB) {   // the summing waves (uniform branch)
        const R* pc = static_cast<const R*>(s.piece);
#pragma unroll
        for (int k = 0; k < NBOND; k++)
#pragma unroll
            for (int q = 0; q < 3; q++) pv[k][q] = pc[3L * ((long)k * s.n_beads + bc) + q];
    }
    // then the first slot chunk
    const int r0 = s.bstart[b0], r1 = s.bstart[bend];
    const int my0 = mine ? s.bstart[b] : 0, my1 = mine ? s.bstart[b + 1] : 0;
    int c0 = r0;
    R v[UN];
    {
        const int n = 3 * (min(c0 + CH, r1) - c0);   // block-uniform: whole rows of loads are skipped
#pragma unroll
        for (int j = 0; j < UN; j++)
            if (j * BLOCK < n) v[j] = slots[3L * c0 + threadIdx.x + j * BLOCK];
    }
    if (!BG) {
#pragma unroll
        for (int k = 0; k < NBOND; k++) {
            const bool on = (m >> k) & 1;
            acc.x = on ? acc.x + (double)pv[k][0] : acc.x;
            acc.y = on ? acc.y + (double)pv[k][1] : acc.y;
            acc.z = on ? acc.z + (double)pv[k][2] : acc.z;
        }
    }
    while (c0 < r1) {   // uniform over the block
        const int c1 = min(c0 + CH, r1);
        const int n = 3 * (c1 - c0);
#pragma unroll
        for (int j = 0; j < UN; j++)
            if (j * BLOCK < n) lds[threadIdx.x + j * BLOCK] = v[j];
        __syncthreads();
        const int q0 = max(my0, c0), q1 = min(my1, c1);
        if (all) {
            // four slots' LDS reads in flight at once; the sums stay in slot order
            int q = q0;
            for (; q + 4 <= q1; q += 4) {
                const R* w = lds + 3 * (q - c0);
                R u[12];
#pragma unroll
                for (int j = 0; j < 12; j++) u[j] = w[j];
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    acc.x += (double)u[3 * j]; acc.y += (double)u[3 * j + 1]; acc.z += (double)u[3 * j + 2];
                }
            }
            if (q < q1) {   // the last one to three slots, their reads in flight together
                const R* w = lds + 3 * (q - c0);
                const int nr = q1 - q;
                R u[9];
#pragma unroll
                for (int j = 0; j < 3; j++) {
                    const bool on = j < nr;
                    u[3 * j] = on ? w[3 * j] : (R)0;
                    u[3 * j + 1] = on ? w[3 * j + 1] : (R)0;
                    u[3 * j + 2] = on ? w[3 * j + 2] : (R)0;
                }
#pragma unroll
                for (int j = 0; j < 3; j++)
                    if (j < nr) {
                        acc.x += (double)u[3 * j]; acc.y += (double)u[3 * j + 1]; acc.z += (double)u[3 * j + 2];
                    }
            }
        } else {
            for (int q = q0; q < q1; q++) {
                const long id = s.sorted_id[q];
                if (id < lo || id >= hi) continue;
                const R* w = lds + 3 * (q - c0);
                acc.x += (double)w[0]; acc.y += (double)w[1]; acc.z += (double)w[2];
            }
        }
        c0 = c1;
        if (c0 < r1) {
            const int n2 = 3 * (min(c0 + CH, r1) - c0);
#pragma unroll
            for (int j = 0; j < UN; j++)
                if (j * BLOCK < n2) v[j] = slots[3L * c0 + threadIdx.x + j * BLOCK];
        }
        __syncthreads();   // the chunk is consumed before the next one is stored
    }
    return acc;
}

// =======================================================================================
// Plain (non-CG) kernels
// =======================================================================================

// serializeDof from the device coordinates: out[3i + q] = xo[3 g[i] + q] for a current bead
// (g[i] >= 0, below the current bead count) or xn[3 (-1 - g[i]) + q] for a new one (the host
// checked both ranges). One thread per coordinate.
__global__ void __launch_bounds__(BLOCK) k_dof_gather(int n, const int* __restrict__ g, const double* __restrict__ xo,
                                                      const double* __restrict__ xn, double* __restrict__ out) {
    const long e = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (e >= 3L * n) return;
    const int i = (int)(e / 3), q = (int)(e - 3L * i);
    const int s = g[i];
    out[e] = s >= 0 ? xo[3L * s + q] : xn[3L * (-1 - s) + q];
}

// Slots of every interaction at x (ForceFieldManager::computeForces before the gather).
// Excluded-volume pairs needing the numerical fallback are listed in bad_list.
// Grid: the interaction blocks. The mixed instantiation is held to 4 waves per SIMD (126 VGPRs,
// no spills; measured 1-2 % faster per evaluation at 42,000 and 420,000 beads), the FP64 one
// keeps 3 (at 4 it spills and was slower at 42,000 beads).
template <typename R>
__global__ void __launch_bounds__(BLOCK, sizeof(R) == 4 ? 4 : 1) k_forces_terms(DevSys s, const double* __restrict__ x,
                                                        R* __restrict__ slots, int* bad_count, int* bad_list,
                                                        int bead0, int bmask, const double* __restrict__ base) {
    // blocks from bead0 on (when bead0 >= 0) are bead blocks: the bonded sums (bead_bonded_sum)
    if (bead0 >= 0 && (int)blockIdx.x >= bead0) {
        bead_bonded_sum<R>(s, x, blockIdx.x - bead0, bmask, base);
        return;
    }
    const int ff = ff_of_block(s, blockIdx.x);
    if (ff >= NFF) return;
    const int t = (blockIdx.x - s.fblk[ff]) * BLOCK + threadIdx.x;
    if (t >= s.n[ff]) return;
    Pos p{x, nullptr, nullptr, 0.0, 0.0, 0};
    int fbad = 0, bad_t = -1;
    (void)term_ef<R>(s, ff, t, p, slots, -1, &fbad, &bad_t);   // forces only: the energy half is dead code
    if (fbad) {
        const int k = atomicAdd(bad_count, 1);
        bad_list[k] = bad_t;
    }
}

// force[b] = base[b] + the bonded pieces selected by bmask + b's slots with id in [lo, hi)
// (base may be null: zero-fill). BG: the first two terms come as the bonded sums of the terms
// kernel's bead blocks (which took base and bmask); else from base and the piece arrays. Grid:
// gather blocks (GB beads each).
template <typename R, int GB, bool BG>
__global__ void __launch_bounds__(BLOCK) k_gather_plain(DevSys s, int bmask, const R* __restrict__ slots, long lo,
                                                        long hi, const double* __restrict__ base, double* __restrict__ f) {
    __shared__ double L[GCHUNK_BYTES / sizeof(double)];
    const int b = blockIdx.x * GB + threadIdx.x;
    const bool mine = threadIdx.x < GB && b < s.n_beads;
    D3 acc = (!BG && base && mine) ? D3{base[3 * b], base[3 * b + 1], base[3 * b + 2]} : D3{0.0, 0.0, 0.0};
    acc = gather_block<R, GB, BG>(s, bmask, slots, lo, hi, lo == 0 && hi == s.slot0[NFF], acc, L);
    if (!mine) return;
    f[3 * b] = acc.x; f[3 * b + 1] = acc.y; f[3 * b + 2] = acc.z;
}

// The FP64 gathers of several tangent products in one launch (k_gather_plain's gather of the
// piece arrays and every slot, from zero): product k = blockIdx.y from its slots at
// slots + k slot_stride and its pieces at piece + k piece_stride into f + k n_var.
template <int GB>
__global__ void __launch_bounds__(BLOCK) k_gather_tangents(DevSys s, const double* __restrict__ slots, long slot_stride,
                                                           double* piece, long piece_stride, double* __restrict__ f,
                                                           long n_var) {
    __shared__ double L[GCHUNK_BYTES / sizeof(double)];
    const long k = blockIdx.y;
    DevSys sk = s;
    sk.piece = piece + k * piece_stride;
    const int b = blockIdx.x * GB + threadIdx.x;
    const bool mine = threadIdx.x < GB && b < s.n_beads;
    const D3 acc = gather_block<double, GB, false>(sk, 3, slots + k * slot_stride, 0L, s.slot0[NFF], true,
                                                   D3{0.0, 0.0, 0.0}, L);
    if (!mine) return;
    double* o = f + k * n_var;
    o[3 * b] = acc.x; o[3 * b + 1] = acc.y; o[3 * b + 2] = acc.z;
}

// Block-parallel Simpson energies (energyN) of the pairs deferred by the preceding kernel.
// Thread 0 receives, per code, the sum of the energies in (code, pair index) order and a flag for
// a pair that stays non-finite (the reference's -1 sentinel); the deferred list is reset.
// Returns the number of deferred pairs (DEFER_CAP exceeded when larger). The count is read and
// reset with atomics: in the batched minimizer the pairs were pushed by other waves of the same
// workgroup within the kernel.
// run_defers_n: the same with the count n already read (by every thread).
template <int NTH = BLOCK>
__device__ inline int run_defers_n(const DevSys& s, int n, double* lds, double* sum, int* badc, int ncode) {
    const int m = min(n, DEFER_CAP);
    const int op1[1] = {0};
    for (int k = 0; k < m; k++) {
        EnDefer& e = s.defer[k];
        const D3 c1{e.c[0], e.c[1], e.c[2]}, c2{e.c[3], e.c[4], e.c[5]}, c3{e.c[6], e.c[7], e.c[8]},
            c4{e.c[9], e.c[10], e.c[11]};
        double r = exvol_energyN_share(c1, c2, c3, c4, threadIdx.x, NTH);
        block_reduce_n<1, NTH>(&r, op1, lds);
        if (threadIdx.x == 0) e.U = exvol_energyN_finish(c1, c2, c3, c4, e.ks, r);
    }
    if (threadIdx.x == 0) {
        for (int c = 0; c < ncode; c++) {
            sum[c] = 0.0;
            badc[c] = 0;
            int last_t = -1;
            for (;;) {
                int best = -1;
                for (int k = 0; k < m; k++)
                    if (s.defer[k].code == c && s.defer[k].t > last_t && (best < 0 || s.defer[k].t < s.defer[best].t))
                        best = k;
                if (best < 0) break;
                const double U = s.defer[best].U;
                if (energy_bad(U)) badc[c] = 1;
                else sum[c] += U;
                last_t = s.defer[best].t;
            }
        }
        __hip_atomic_store(s.defer_n, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return n;
}
template <int NTH = BLOCK>
__device__ inline int run_defers(const DevSys& s, double* lds, double* sum, int* badc, int ncode) {
    __shared__ int nsh;
    __syncthreads();
    if (threadIdx.x == 0) nsh = __hip_atomic_load(s.defer_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    return run_defers_n<NTH>(s, nsh, lds, sum, badc, ncode);
}

// Seeded Gaussian displacement, counter-based so every coordinate's value depends only on
// (seed, stream, index): SplitMix64's output function (Steele, Lea and Flood, OOPSLA 2014) turns
// the counters into two uniforms per coordinate, u1 in (0, 1] and u2 in [0, 1), and the
// Box-Muller transform gives x[i] += sigma sqrt(-2 ln u1) cos(2 pi u2).
//   key = mix(seed + G (stream + 1)),  r(j) = mix(key + G (j + 1)),  G = 0x9E3779B97F4A7C15,
//   u1 = ((r(2i) >> 11) + 1) 2^-53,     u2 = (r(2i + 1) >> 11) 2^-53.
__host__ __device__ inline unsigned long long splitmix_mix(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__global__ void __launch_bounds__(BLOCK) k_perturb(double* __restrict__ x, long n, unsigned long long seed,
                                                  unsigned long long stream, double sigma) {
    const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    const unsigned long long G = 0x9E3779B97F4A7C15ULL;
    const unsigned long long key = splitmix_mix(seed + G * (stream + 1));
    const unsigned long long r1 = splitmix_mix(key + G * (2ULL * (unsigned long long)i + 1));
    const unsigned long long r2 = splitmix_mix(key + G * (2ULL * (unsigned long long)i + 2));
    const double u1 = (double)((r1 >> 11) + 1) * 0x1p-53, u2 = (double)(r2 >> 11) * 0x1p-53;
    x[i] += sigma * (sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

// x[i] += d[i], i < n (a displacement of the device coordinates)
__global__ void __launch_bounds__(BLOCK) k_axpy_inplace(double* __restrict__ x, const double* __restrict__ d, long n) {
    const long i = (long)blockIdx.x * BLOCK + threadIdx.x;
    if (i < n) x[i] += d[i];
}

// v.v and max|v| of a vector (the minimizer's start: F.F and maxF): grid-stride block partials,
// value-major, then one block in a fixed order into out[0..1].
__global__ void __launch_bounds__(BLOCK) k_vec_stats(const double* __restrict__ v, long n, double* __restrict__ partial) {
    __shared__ double lds[2 * (BLOCK / 64)];
    double a[2] = {0.0, 0.0};
    for (long i = (long)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (long)gridDim.x * BLOCK) {
        const double x = v[i];
        a[0] += x * x;
        a[1] = rmax(a[1], fabs(x));
    }
    const int op[2] = {0, 1};
    block_reduce_n<2>(a, op, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = a[0];
        partial[gridDim.x + blockIdx.x] = a[1];
    }
}
__global__ void __launch_bounds__(BLOCK) k_vec_stats_fin(const double* __restrict__ partial, int nb,
                                                         double* __restrict__ out) {
    __shared__ double lds[2 * (BLOCK / 64)];
    double a[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nb; b += BLOCK) {
        a[0] += partial[b];
        a[1] = rmax(a[1], partial[nb + b]);
    }
    const int op[2] = {0, 1};
    block_reduce_n<2>(a, op, lds);
    if (threadIdx.x == 0) {
        out[0] = a[0];
        out[1] = a[1];
    }
}

// Per-FF energies at x, block partials. Grid: interaction blocks only.
template <typename R>
__global__ void __launch_bounds__(BLOCK) k_energy_report(DevSys s, const double* __restrict__ x,
                                                         double* __restrict__ partial) {
    __shared__ double lds[BLOCK / 64];
    const int ff = ff_of_block(s, blockIdx.x);
    double v = 0.0;
    if (ff < NFF) {
        const int t = (blockIdx.x - s.fblk[ff]) * BLOCK + threadIdx.x;
        int fbad = 0;
        Pos p{x, nullptr, nullptr, 0.0, 0.0, 0};
        if (t < s.n[ff]) v = term_ef<R>(s, ff, t, p, (R*)nullptr, 0, &fbad);
    }
    const int op1[1] = {0};
    block_reduce_n<1>(&v, op1, lds);
    if (threadIdx.x == 0) partial[PSTRIDE * blockIdx.x] = v;
}

// out[0..5] per force field (excluded volume = -1 on the sentinel), out[6] =
// computeIndividualEnergies total (plain sum), out[7] = computeEnergy total (+inf on a sentinel),
// out[8] = deferred Simpson pairs beyond capacity. One block.
__global__ void __launch_bounds__(BLOCK) k_energy_report_fin(DevSys s, const double* __restrict__ partial,
                                                             double* __restrict__ out) {
    __shared__ double lds[NFF * (BLOCK / 64)];
    double acc[NFF];
    int opn[NFF];
#pragma unroll
    for (int f2 = 0; f2 < NFF; f2++) {
        acc[f2] = 0.0;
        opn[f2] = 0;
        const int b1 = s.fblk[f2] + (s.n[f2] + BLOCK - 1) / BLOCK;
        for (int b = s.fblk[f2] + threadIdx.x; b < b1; b += BLOCK) acc[f2] += partial[PSTRIDE * b];
    }
    block_reduce_n<NFF>(acc, opn, lds);
    double dsum[1];
    int dbad[1];
    const int n = run_defers(s, lds, dsum, dbad, 1);
    if (threadIdx.x == 0) {
        acc[FF_EXVOL] += dsum[0];
        double tot = 0.0, tot_s = 0.0;
        bool failed = false;
        for (int f2 = 0; f2 < NFF; f2++) {
            double ef = (f2 == FF_EXVOL && dbad[0]) ? -1.0 : acc[f2];
            out[f2] = ef;
            tot += ef;
            if (!failed) {
                if (ef <= -1) failed = true;
                else tot_s += ef;
            }
        }
        out[6] = tot;
        out[7] = failed ? INFINITY : tot_s;
        out[8] = n > DEFER_CAP ? (double)n : 0.0;
    }
}

// Numerical excluded-volume force fallback (forceN, CylinderExclVolRepulsion.cpp:984-1254)
// for one pair: block b handles rows i = b, b + grid, ... of the 10^4 x 10^4 Simpson grid
// and writes 6 partials per block; k_exvol_forceN_finish adds them in block order.
__device__ inline void exvol_force_integrand(double a, double b, double c, double d, double e, double F,
                                             double s, double t, double* out) {
    double r_2 = c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;
    double x = (1.0 / (r_2 * r_2 * r_2));
    out[0] = x; out[1] = s * x; out[2] = s * s * x; out[3] = s * t * x; out[4] = t * x; out[5] = t * t * x;
}

struct PairGeom { D3 c1, c2, c3, c4; };

__device__ inline PairGeom exvol_pair_geom(const DevSys& s, const double* x, int t, int movebeads) {
    PairGeom g{ld3(x, s.ev_b[4 * t]), ld3(x, s.ev_b[4 * t + 1]), ld3(x, s.ev_b[4 * t + 2]), ld3(x, s.ev_b[4 * t + 3])};
    if (movebeads && are_in_plane(g.c1, g.c2, g.c3, g.c4)) {
        D3 n2 = move_out_of_plane(g.c1, g.c2, g.c3, 2, 0.01);
        D3 n1 = move_out_of_plane(g.c1, g.c2, g.c3, 1, 0.01);
        g.c2 = n2;
        g.c1 = n1;
    }
    return g;
}

__global__ void __launch_bounds__(BLOCK) k_exvol_forceN_rows(DevSys s, const double* __restrict__ x, int t,
                                                             int movebeads, double* __restrict__ partial) {
    __shared__ double lds[6 * (BLOCK / 64)];
    PairGeom g = exvol_pair_geom(s, x, t, movebeads);
    const SixInv v6 = six_inv(g.c1, g.c2, g.c3, g.c4);
    const double a = v6.a, b = v6.b, c = v6.c, d = v6.d, e = v6.e, F = v6.F;
    const int N = 10000;
    const double delta = 1.0 / ((double)N);
    double acc[6] = {0, 0, 0, 0, 0, 0}, in[6];
    // rows i = 0..N with s = (i+1)*delta (the reference's loop advances s before use)
    for (int i = blockIdx.x; i <= N; i += gridDim.x) {
        const double sv = (i + 1) * delta;
        const double fs = (i % 2 == 0) ? 2.0 : 4.0;
        for (int j = 1 + threadIdx.x; j < N; j += BLOCK) {
            const double tv = (j + 1) * delta;
            const double ft = (j % 2 == 0) ? 2.0 : 4.0;
            if (i == 0) {
                exvol_force_integrand(a, b, c, d, e, F, 0.0, tv, in);
                for (int k = 0; k < 6; k++) acc[k] += ft * in[k];
                exvol_force_integrand(a, b, c, d, e, F, 1.0, tv, in);
                for (int k = 0; k < 6; k++) acc[k] += ft * in[k];
            } else if (i < N) {
                exvol_force_integrand(a, b, c, d, e, F, sv, tv, in);
                for (int k = 0; k < 6; k++) acc[k] += fs * ft * in[k];
            }
        }
        if (threadIdx.x == 0 && i > 0 && i < N) {
            for (int l = 0; l < 2; l++) {
                exvol_force_integrand(a, b, c, d, e, F, sv, (double)l, in);
                for (int k = 0; k < 6; k++) acc[k] += fs * in[k];
            }
        }
    }
    const int op6[6] = {0, 0, 0, 0, 0, 0};
    block_reduce_n<6>(acc, op6, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) partial[6 * blockIdx.x + k] = acc[k];
}

template <typename R>
__global__ void k_exvol_forceN_finish(DevSys s, const double* __restrict__ x, int t, int movebeads, int nblk,
                                      const double* __restrict__ partial, R* __restrict__ slots, int* ok_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    PairGeom g = exvol_pair_geom(s, x, t, movebeads);
    const SixInv v6 = six_inv(g.c1, g.c2, g.c3, g.c4);
    const double a = v6.a, b = v6.b, c = v6.c, d = v6.d, e = v6.e, F = v6.F;
    double T1[6] = {0, 0, 0, 0, 0, 0}, in[6], rest[6] = {0, 0, 0, 0, 0, 0};
    const double lim[2] = {0.0, 1.0};
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2; q++) {
            exvol_force_integrand(a, b, c, d, e, F, lim[p], lim[q], in);
            for (int k = 0; k < 6; k++) T1[k] += in[k];
        }
    for (int blk = 0; blk < nblk; blk++)
        for (int k = 0; k < 6; k++) rest[k] += partial[6 * blk + k];
    const double delta = 1.0 / 10000.0;
    double I6[6];
    for (int k = 0; k < 6; k++) I6[k] = (delta * delta / 9.0) * (T1[k] + rest[k]);
    const double ks = s.ev_ks[t];
    double vA[3] = {g.c1.x - g.c2.x, g.c1.y - g.c2.y, g.c1.z - g.c2.z};
    double vB[3] = {g.c3.x - g.c4.x, g.c3.y - g.c4.y, g.c3.z - g.c4.z};
    double vC[3] = {g.c3.x - g.c1.x, g.c3.y - g.c1.y, g.c3.z - g.c1.z};
    double f[4][3];
    bool ok = true;
    for (int k = 0; k < 3; k++) {
        f[0][k] = 4.0 * ks * (-I6[0] * vC[k] + I6[1] * (vC[k] - vA[k]) + I6[2] * vA[k] + I6[4] * vB[k] - I6[3] * vB[k]);
        f[1][k] = 4.0 * ks * (-I6[1] * vC[k] - I6[2] * vA[k] + I6[3] * vB[k]);
        f[2][k] = 4.0 * ks * (I6[0] * vC[k] + I6[1] * vA[k] - I6[4] * (vB[k] + vC[k]) - I6[3] * vA[k] + I6[5] * vB[k]);
        f[3][k] = 4.0 * ks * (I6[4] * vC[k] + I6[3] * vA[k] - I6[5] * vB[k]);
        for (int q = 0; q < 4; q++) ok = ok && !bad_component<double>(f[q][k]);
    }
    *ok_out = ok ? 1 : 0;
    if (ok) {
        const long id = s.slot0[FF_EXVOL] + 4L * t;
        for (int q = 0; q < 4; q++) put_slot(s, slots, id + q, D3{f[q][0], f[q][1], f[q][2]});
    }
}

// =======================================================================================
// CG kernels. One "slot" of the device-driven minimizer is four kernels:
//     k_cg_pos -> k_cg_ls -> k_cg_gather -> k_cg_fin
// Each reads the device state first and returns at once when it has nothing to do.
// A typical CG iteration is one slot: k_cg_pos writes the trial coordinates x + lambda0 d (the
// pending direction update d = F + beta d applied on the fly), which are also the speculative
// next coordinates; k_cg_ls evaluates the trial energy there with its forces; k_cg_gather sums
// those forces into the next force buffer, speculatively; k_cg_fin runs the line-search
// decision and, when lambda0 is accepted, the post-step control at once. After a rejection a
// launch evaluates the next two halvings; when the accepted step is not a speculative one (or
// for the recovery move) the next slot is the move (k_cg_pos: x_next at the accepted step,
// k_cg_ls: its forces), its gather real, and its k_cg_fin the post-step control.
// Precomputing the trial coordinates once per bead lets every interaction read one coordinate
// vector instead of three (x, d, F) per bead it touches.
// =======================================================================================

struct CgBufs {
    double* x[2];       // current / next coordinates (parity)
    double* d[2];       // current / next search direction (dpar)
    double* xmin[2];    // committed / candidate minimum-energy backup (bkpar)
    double* f[2];       // current / next force (fpar)
    double* fp;         // previous force (the batched minimizer's; here F_prev is f[fpar])
    double* xh;         // x + (lambda / 2) d, the second trial of a launch after a rejection
    void* slots;        // 3 values of the evaluation's arithmetic type per slot
    double* partial;    // line-search partials, value-major: [value][interaction blocks + bead blocks]
    double* gpartial;   // gather partials, value-major: [value][gather blocks]
    int* bad_count;     // excluded-volume pairs needing the numerical fallback in a CG move
};

__device__ inline bool call_limit(const CgState* st) {
    return (st->energy_call_limit != 0 && st->num_energy_call > st->energy_call_limit) ||
           (st->force_call_limit != 0 && st->num_force_call > st->force_call_limit);
}
__device__ inline bool converged(const CgState* st) {
    return st->grad_inf_norm < st->grad_tol || st->energy_rel_change < st->energy_rel_tol;
}

// ---- line search, bead part: trial coordinates, d.F, the recovery step ------------------
// Grid: bead blocks. Phase PH_LS: x_next = x + lambda0 d_eff (trial 1, and the speculative next
// coordinates), x + (lambda0 / 2) d after a rejection (trial 2), and on the first launch of a
// search the d.F and max-step partials and the materialized direction. Phase PH_MOVE:
// x_next = x + lambda d (moveAlongSearchDir, CGMethod.cpp:350-354; the recovery move,
// MController.h:318-321). The partials go after the interaction blocks' (index T + block).
__global__ void __launch_bounds__(BLOCK) k_cg_pos(DevSys s, const CgState* __restrict__ st, CgBufs B) {
    __shared__ double lds[2 * (BLOCK / 64)];
    // every state field this kernel reads, loaded before the first branch: one round trip to the
    // state instead of a chain of dependent ones
    const int done = st->done, phase = st->phase, par = st->parity, dpar = st->dpar, fpar = st->fpar;
    const int ls_active = st->ls_active, dir_pending = st->dir_pending;
    const double lambda = st->lambda, lam0 = st->ls_lambda, beta = st->dir_beta;
    keep_loaded(done); keep_loaded(phase); keep_loaded(par); keep_loaded(dpar); keep_loaded(fpar);
    keep_loaded(ls_active); keep_loaded(dir_pending);
    keep_loaded(lambda); keep_loaded(lam0); keep_loaded(beta);
    if (done) return;
    if (phase != PH_LS && phase != PH_MOVE) return;
    // block b covers the 3 * BLOCK coordinates of beads [BLOCK b, BLOCK (b + 1)), coalesced
    const long e0 = 3L * BLOCK * blockIdx.x + threadIdx.x, ne = 3L * s.n_beads;
    const double* __restrict__ xc = B.x[par];
    const double* __restrict__ dc = B.d[dpar];
    double* __restrict__ xn = B.x[par ^ 1];
    if (phase == PH_MOVE) {
        const double lam = lambda;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const long e = e0 + (long)k * BLOCK;
            if (e < ne) xn[e] = xc[e] + lam * dc[e];
        }
        return;
    }
    const bool first = ls_active == 0;
    const bool pend = first && dir_pending;
    const double* __restrict__ fcur = B.f[fpar];
    double v[2] = {0.0, 0.0};
    double* __restrict__ dn = B.d[dpar ^ 1];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const long e = e0 + (long)k * BLOCK;
        if (e >= ne) continue;
        const double fv = fcur[e];
        const double de = pend ? fv + beta * dc[e] : dc[e];
        xn[e] = xc[e] + lam0 * de;   // trial 1 = speculative x_next
        if (!first) B.xh[e] = xc[e] + (lam0 * 0.5) * de;
        if (first) {
            v[0] += de * fv;                          // d.F
            v[1] = rmax(v[1], fabs(de));              // max|d|, for the recovery step
            if (pend) dn[e] = de;                     // shiftSearchDir materialized
        }
    }
    if (!first) return;   // the partials are read on the first launch of a search only
    const int op2[2] = {0, 1};
    block_reduce_n<2>(v, op2, lds);
    if (threadIdx.x == 0) {
        const long np = s.lsp, k = s.blk[NFF] + s.nbb + blockIdx.x;   // value-major
        B.partial[k] = v[0];
        B.partial[np + k] = v[1];
    }
}

// The energies of the filament terms bead b starts (the stretching term (b, next) and the
// bending term centred on b), so that over all beads every term counts once; both evaluated and
// selected (absent terms see in-range indices and are dropped). The same term functions as
// term_ef.
template <typename R>
__device__ inline double bead_bonded_energy(const DevSys& s, const double* __restrict__ x, int b) {
    const long nb = s.n_beads;
    const int ns = s.n[FF_STRETCH], nd = s.n[FF_BEND];
    const unsigned m = s.pmask[b];
    const int t1r = s.bterm[b];
    const int ip = s.bcontig ? max(b - 1, 0) : s.bnbr[nb + b];
    const int in = s.bcontig ? min(b + 1, (int)nb - 1) : s.bnbr[2 * nb + b];
    const int tcr = s.bcontig ? 2 * t1r - b - 1 : s.bterm[nb + b];
    const int t1 = min(max(t1r, 0), max(ns - 1, 0)), tc = min(max(tcr, 0), max(nd - 1, 0));
    auto X = [x](int i) { return D3{x[3L * i], x[3L * i + 1], x[3L * i + 2]}; };
    const D3 cp = X(ip), c0 = X(b), cn = X(in);
    const double sk = ns > 0 ? s.str_k[t1] : 0.0, sl = ns > 0 ? s.str_eql[t1] : 0.0;
    const double bk = nd > 0 ? s.bend_k[tc] : 0.0, bt = nd > 0 ? s.bend_eqt[tc] : 0.0;
    const double es = (double)stretch_energy<R>(c0, cn, sk, sl);
    const double eb = (double)bend_energy<R>(cp, c0, cn, bk, bt);
    double e = 0.0;
    e = (m & 0x02u) ? e + es : e;   // bead 1 of a stretching term
    e = (m & 0x08u) ? e + eb : e;   // bead 2 of a bending term
    return e;
}

// A bead block of the line search (bonded_gather): the bonded sums at x_next (the speculative or
// the move's forces) and, in phase PH_LS with energies, the block partials of the filament terms' energies at
// the trial points the interaction blocks evaluate (x_next; x + (lambda / 2) d after a
// rejection; x when the current energy is not known), in the columns after the interaction
// blocks'.
template <typename R>
__device__ inline void bead_ls(const DevSys& s, const CgBufs& B, int j, bool move, bool first, bool need_e0, int par,
                               bool energies, double* lds) {
    const int b = j * BLOCK + threadIdx.x;
    const bool on = b < s.n_beads;
    const double* __restrict__ xn = B.x[par ^ 1];
    const bool with_e = !move && energies;   // block-uniform
    double v[3] = {0.0, 0.0, 0.0};
    if (on) {
        const D3 acc = with_e ? bonded_at_bead<R, true>(s, xn, b, s.pmask[b], D3{0.0, 0.0, 0.0}, &v[0])
                              : bonded_at_bead<R>(s, xn, b, s.pmask[b], D3{0.0, 0.0, 0.0});
        s.bsum[3L * b] = acc.x;
        s.bsum[3L * b + 1] = acc.y;
        s.bsum[3L * b + 2] = acc.z;
    }
    if (!with_e) return;
    if (on) {
        if (!first) v[1] = bead_bonded_energy<R>(s, B.xh, b);
        if (need_e0) v[2] = bead_bonded_energy<R>(s, B.x[par], b);
    }
    const int op3[3] = {0, 0, 0};
    block_reduce_n<3>(v, op3, lds);
    if (threadIdx.x == 0) {
        const long np = s.lsp, k = s.blk[NFF] + j;   // value-major
        B.partial[k] = v[0];
        B.partial[np + k] = v[1];
        B.partial[2 * np + k] = v[2];
    }
}

// ---- line search, interaction part: trial energies, speculative forces -----------------
// Grid: the interaction blocks. Phase PH_LS: E(lambda) with its forces at x_next (into the
// slots), E(lambda / 2) at x_half after a rejection, E(x) when the current energy is not known;
// block partials E(lambda), E(lambda / 2), E(x). Phase PH_MOVE: the forces at x_next.
// 3 waves per SIMD (at most 168 VGPRs) measured faster than the unconstrained 2.
template <typename R>
__global__ void __launch_bounds__(BLOCK, 3) k_cg_ls(DevSys s, const CgState* __restrict__ st, CgBufs B) {
    __shared__ double lds[3 * (BLOCK / 64)];
    // the state fields, loaded before the first branch (one round trip)
    const int done = st->done, phase = st->phase, par = st->parity, ls_active = st->ls_active;
    const int ecur_valid = st->ecur_valid;
    keep_loaded(done); keep_loaded(phase); keep_loaded(par); keep_loaded(ls_active); keep_loaded(ecur_valid);
    if (done) return;
    if (phase != PH_LS && phase != PH_MOVE) return;
    const bool first = ls_active == 0;
    const bool need_e0 = first && !ecur_valid;
    // with bonded_gather bead blocks follow the interaction blocks and carry the filament terms'
    // forces (and, with bead_energy, their energies: then the filament terms' blocks are left out)
    const int nint = (s.bonded_gather && s.bead_energy) ? s.blk_nb : s.blk[NFF];
    if ((int)blockIdx.x >= nint) {   // (or the one block of a grid without interactions)
        if (s.bonded_gather)
            bead_ls<R>(s, B, blockIdx.x - nint, phase == PH_MOVE, first, need_e0, par, s.bead_energy != 0, lds);
        return;
    }
    const int ff = ff_of_block(s, blockIdx.x);
    const int t = (blockIdx.x - s.fblk[ff]) * BLOCK + threadIdx.x;
    const double* __restrict__ xn = B.x[par ^ 1];
    const Pos p1{xn, nullptr, nullptr, 0.0, 0.0, 0};
    // with bonded_gather (and not bead_energy) the filament terms' blocks take energies only
    const bool bonded_e = s.bonded_gather && (ff == FF_STRETCH || ff == FF_BEND);
    R* const tslots = bonded_e ? nullptr : static_cast<R*>(B.slots);
    if (phase == PH_MOVE) {
        if (t < s.n[ff] && !bonded_e) {
            int fbad = 0;
            (void)term_ef<R>(s, ff, t, p1, tslots, -1, &fbad);
            if (fbad) atomicAdd(B.bad_count, 1);
        }
        return;
    }
    double v[3] = {0.0, 0.0, 0.0};
    if (t < s.n[ff]) {
        int fbad = 0;
        v[0] = term_ef<R, true>(s, ff, t, p1, tslots, 0, &fbad);   // split: no register spills at 3 waves
        if (fbad) atomicAdd(B.bad_count, 1);
        if (!first) {   // the first launch of a search evaluates lambda0 alone (see ls_control)
            const Pos p2{B.xh, nullptr, nullptr, 0.0, 0.0, 0};
            v[1] = term_ef<R>(s, ff, t, p2, (R*)nullptr, 1, &fbad);
        }
        if (need_e0) {
            const Pos p0{B.x[par], nullptr, nullptr, 0.0, 0.0, 0};
            v[2] = term_ef<R>(s, ff, t, p0, (R*)nullptr, 2, &fbad);
        }
    }
    const int op3[3] = {0, 0, 0};
    block_reduce_n<3>(v, op3, lds);
    if (threadIdx.x == 0) {
        const long np = s.lsp, k = blockIdx.x;   // value-major
        B.partial[k] = v[0];
        B.partial[np + k] = v[1];
        B.partial[2 * np + k] = v[2];
    }
}

__device__ inline bool call_limit_s(const CgState& S) {
    return (S.energy_call_limit != 0 && S.num_energy_call > S.energy_call_limit) ||
           (S.force_call_limit != 0 && S.num_force_call > S.force_call_limit);
}
__device__ inline bool converged_s(const CgState& S) {
    return S.grad_inf_norm < S.grad_tol || S.energy_rel_change < S.energy_rel_tol;
}

// Scalar logic after the trial energies: backtrackingLineSearch (CGMethod.hpp:238-373) over the
// evaluated step lengths, and the failure handling of CGMethod::minimize (:1337-1367).
// acc = {E(lambda), E(lambda/2), bad, bad2, E(x), bad0, d.F, recovery step}
// A search that ends with lambda = 0 (CGMethod.hpp:1335-1365): recovery on -> safe mode, or the
// recovery move when already in safe mode; recovery off -> the line-search error.
__device__ inline void ls_failed(CgState& S) {
    if (S.try_recover) {
        if (S.safe_mode) {
            if (S.rec_step == INFINITY) {   // MController.h:312-316 throws
                S.status = 7;               // MDN_ERR_RECOVERY
                S.done = 1;
                return;
            }
            S.lambda = S.rec_step;
            S.recovery_move = 1;
            S.phase = PH_MOVE;
            return;
        }
        S.safe_mode = 1;
        S.num_safe_mode += 1;
        S.dir_pending = 1;             // resetCGParams before the next line search
        S.dir_beta = 0.0;
        S.dir_reset = 1;
        // the next backtracking search starts again from lambda0 = min(lambdaMax, maxDist / maxF)
        // (CGMethod.hpp:258-262); maxF is unchanged, no force evaluation happened
        S.ls_lambda = rmin(S.lambda_max, S.max_dist / S.grad_inf_norm);
        if (call_limit_s(S) || converged_s(S)) S.done = 1;
        return;
    }
    S.error_flags |= 1u;
    S.done = 1;
}

// ---- the quadratic line search (quadraticLineSearchV2 + quadraticoptimization,
// CGMethod.hpp:522-764), every trial at x + lambda d; the same guards as the oracle's: a
// non-finite vertex returns x1 unevaluated, and at most QUAD_ITER_CAP parabola fits.
constexpr int QUAD_ITER_CAP = 64;
constexpr double LAMBDAQUADTOL = 1e-3;   // CGMethod.hpp:106

// the search's end at step lam (spec: this launch's speculative forces are at lam)
__device__ inline void quad_finish(CgState& S, double lam, bool spec) {
    S.q_phase = 0;
    S.ls_active = 0;
    if (lam <= 0.0 || lam <= 1e-8) lam = 0.0;
    S.total_energy = lam > 0.0 ? S.q_ebt : S.ls_cur;   // TotalEnergy (CGMethod.hpp:649-652)
    if (lam == 0.0) {
        ls_failed(S);
        return;
    }
    S.lambda = lam;
    S.new_energy = S.q_ebt;        // res.newEnergy = the backtracking phase's energy (:655)
    S.recovery_move = 0;
    S.q_stale = 1;
    S.phase = spec ? PH_GATHER : PH_MOVE;
}

// The next parabola fit through (x0, y0), (x1, y1), (x2, y2): ends the search, or sets up the
// launch that evaluates the vertex.
__device__ inline void quad_next(CgState& S) {
    if (++S.q_iter > QUAD_ITER_CAP) {
        quad_finish(S, S.q_x1, false);
        return;
    }
    const double x0 = S.q_x0, x1 = S.q_x1, x2 = S.q_x2, y0 = S.q_y0, y1 = S.q_y1, y2 = S.q_y2;
    const double d = (x0 - x1) * (x0 - x2) * (x1 - x2);
    const double a = y0 * (x1 - x2) - y1 * (x0 - x2) + y2 * (x0 - x1);
    const double b = -(y0 * (x1 + x2) * (x1 - x2) - y1 * (x0 + x2) * (x0 - x2) + y2 * (x0 + x1) * (x0 - x1));
    const double lq = -b / (2 * a);
    if (!isfinite(lq)) { quad_finish(S, x1, false); return; }
    if (fabs(lq - x1) / x1 < LAMBDAQUADTOL) { quad_finish(S, lq, false); return; }
    if (a / d < 0) { quad_finish(S, S.q_bt, false); return; }   // the model has no minimum
    S.ls_lambda = lq;   // trial 1 of the next launch
    S.q_phase = 1;
}

// The vertex's energy eq (trial 1 of this launch): accept it, or replace a bracket point.
__device__ inline void quad_eval(CgState& S, double eq, bool spec) {
    const double lq = S.ls_lambda;
    S.num_energy_call += 1;
    if (fabs(eq - S.q_y1) / S.q_y1 < LAMBDAQUADTOL) { quad_finish(S, lq, spec); return; }
    if (lq < S.q_x1) {
        if (eq < S.q_y0 && eq < S.q_y1) { S.q_x1 = lq; S.q_y1 = eq; }
        else { S.q_x0 = lq; S.q_y0 = eq; }
    } else if (S.q_x1 < lq) {
        if (eq < S.q_y1 && eq < S.q_y2) { S.q_x1 = lq; S.q_y1 = eq; }
        else { S.q_x2 = lq; S.q_y2 = eq; }
    } else {
        quad_finish(S, S.q_x1, false);
        return;
    }
    quad_next(S);
}

// An upper bound on the minimizer slots per force call (the guard on its exit): the first launch
// of a search, one launch per two of the at most 60 halvings of lambda0 <= 1e10 down to 1e-8,
// the parabola fits of the quadratic search, a move.
MDN_HD inline long slots_per_force_call(const CgState& S) { return 32 + (S.ls_algo == 1 ? QUAD_ITER_CAP : 0); }

__device__ inline void ls_control(CgState& S, const double* acc, int& bad_count) {
    const bool first = S.ls_active == 0;
    const double E1 = acc[2] > 0.0 ? INFINITY : acc[0];   // computeEnergy: -1 sentinel -> +inf
    const double E2 = acc[3] > 0.0 ? INFINITY : acc[1];
    const double lam0 = S.ls_lambda;
    S.n_ls_launch += 1;
    if (S.q_phase) {   // a launch that evaluated the parabola's vertex (trial 1)
        const bool spec = bad_count == 0;
        bad_count = 0;
        quad_eval(S, E1, spec);
        return;
    }
    if (first) {
        if (S.dir_pending) {
            if (!S.dir_reset && acc[6] <= 0) {
                // d.F <= 0 after shiftSearchDir(beta): resetCGParams (CGMethod.hpp:1505-1507);
                // redo this line-search launch with d = F
                S.dir_beta = 0.0;
                S.dir_reset = 1;
                bad_count = 0;
                return;
            }
            S.dpar ^= 1;
            S.dir_pending = 0;
            if (S.dir_reset) S.last_fdotf = S.fdotf;   // lastFDotF = F.F
            S.dir_reset = 0;
        }
        S.num_iter += 1;             // numIter++ at the top of the loop (CGMethod.hpp:1297)
        S.rec_step = acc[7];
        S.ls_trials = 0;
        S.q_nrec = 0;
        S.q_emax = -INFINITY;
        if (S.grad_inf_norm != 0.0) {   // maxForce == 0 returns lambda = 0 at once (:252-256)
            if (!S.ecur_valid) S.ls_cur = acc[5] > 0.0 ? INFINITY : acc[4];
            S.num_energy_call += 1;
            S.ls_slope = acc[6];
            S.ls_reduced = (S.safe_mode ? 0.0 : 0.4) * acc[6];
            S.ls_active = 1;
        }
    }
    // QUADRATIC outside safe mode (CGMethod.hpp:1322-1325): the backtracking phase of
    // quadraticLineSearchV2, its ideal change -BACKTRACKSLOPE * lambda * d.F in that order
    const bool quad = S.ls_algo == 1 && !S.safe_mode;
    bool conv = false;
    int acc_trial = -1;
    double lambda = lam0, energyLambda = 0.0;
    if (first && S.grad_inf_norm == 0.0) {
        conv = true;
        lambda = 0.0;
    }
    // The first launch of a search evaluates lambda0 alone (usually accepted: its forces are
    // the speculative ones); after a rejection each launch evaluates the next NTRIAL halvings.
    // The trial sequence is the reference's either way.
    const int ntrial = first ? 1 : NTRIAL;
    for (int k = 0; k < ntrial && !conv; k++) {
        energyLambda = k == 0 ? E1 : E2;
        S.num_energy_call += 1;
        S.ls_trials += 1;
        const double lt = lambda, et = energyLambda;
        const double ideal = quad ? -0.4 * lambda * S.ls_slope : -S.ls_reduced * lambda;
        if (lambda < 0 || ideal > 0) { conv = true; lambda = 0.0; }
        const double change = energyLambda - S.ls_cur;
        if (change <= ideal) { conv = true; acc_trial = k; }
        else lambda *= 0.5;
        if (lambda <= 0.0 || lambda <= 1e-8) {
            conv = true;
            acc_trial = -1;
            lambda = 0.0;
            energyLambda = S.ls_cur;
        }
        if (quad) {   // the recorded entry of this trial (CGMethod.hpp:601-613): the trial, or
                      // (0, E(x)) when lambda fell to the tolerance
            const double l = lambda == 0.0 ? 0.0 : lt, e = lambda == 0.0 ? S.ls_cur : et;
            if (S.q_nrec++ == 0) { S.q_lv0 = l; S.q_ev0 = e; }
            else S.q_emax = rmax(S.q_emax, e);
        }
    }
    const bool spec = bad_count == 0;
    bad_count = 0;
    if (!conv) {                 // keep halving in the next launch
        S.ls_lambda = lambda;
        return;
    }
    S.total_energy = energyLambda;   // TotalEnergy = energyLambda (CGMethod.hpp:368)
    S.ls_active = 0;
    if (lambda != 0.0) {
        if (quad && S.ls_trials > 1 && energyLambda < S.ls_cur && energyLambda < S.q_emax) {
            // the parabola phase (CGMethod.hpp:618-646): through (0, E(x)), the accepted trial
            // and the first entry (the reference's index loop ends at 0)
            S.ls_active = 1;
            S.q_x0 = 0.0; S.q_y0 = S.ls_cur;
            S.q_x1 = lambda; S.q_y1 = energyLambda;
            S.q_x2 = S.q_lv0; S.q_y2 = S.q_ev0;
            S.q_bt = lambda; S.q_ebt = energyLambda;
            S.q_iter = 0;
            quad_next(S);
            return;
        }
        S.lambda = lambda;
        S.new_energy = energyLambda;
        S.recovery_move = 0;
        // accepted at the first trial with usable speculative forces: gather directly
        S.phase = (acc_trial == 0 && spec) ? PH_GATHER : PH_MOVE;
        return;
    }
    ls_failed(S);
}

// ---- gather F into the next force buffer; F.F, F.F_prev, max|F| partials ---------------
// After a line-search launch the gather is speculative (k_cg_fin uses it when lambda0 is
// accepted); after a move it is the step's. Grid: bead blocks only.
template <typename R, int GB, bool BG>
__global__ void __launch_bounds__(BLOCK) k_cg_gather(DevSys s, const CgState* __restrict__ st, CgBufs B) {
    __shared__ double L[GCHUNK_BYTES / sizeof(double)];
    // the state fields, loaded before the first branch (one round trip)
    const int done = st->done, phase = st->phase, par = st->parity, bk = st->bkpar, fpar = st->fpar;
    const int num_iter = st->num_iter, force_call_limit = st->force_call_limit;
    keep_loaded(done); keep_loaded(phase); keep_loaded(par); keep_loaded(bk); keep_loaded(fpar);
    keep_loaded(num_iter); keep_loaded(force_call_limit);
    if (done) return;
    if (phase != PH_LS && phase != PH_MOVE) return;
    const int b = blockIdx.x * GB + threadIdx.x;
    // copycoordsifminimumE candidate (CGMethod.hpp:1564-1566): x_next goes to the candidate
    // buffer whenever the iteration part of the condition can hold after this step; the energy
    // and force parts decide in k_cg_fin whether it is committed
    const bool cand = num_iter + 1 > force_call_limit / 2;
    double* __restrict__ fnext = B.f[fpar ^ 1];
    // forcePrev = force after every step and reset (CGMethod.hpp:966-968, 1532-1535), and the
    // current force buffer is replaced only by a committed gather: F_prev is f[fpar]
    const double* __restrict__ fprev = B.f[fpar];
    D3 acc = gather_block<R, GB, BG>(s, 3, static_cast<const R*>(B.slots), 0, 0, true, D3{0.0, 0.0, 0.0}, L);
    constexpr int NSW = GB / 64;   // summing waves
    __shared__ double red[3 * NSW];
    if (threadIdx.x < GB) {   // whole waves
        double v[3] = {0.0, 0.0, 0.0};
        if (b < s.n_beads) {
            const double fv[3] = {acc.x, acc.y, acc.z};
            const double* xn = B.x[par ^ 1];
            for (int q = 0; q < 3; q++) {
                fnext[3 * b + q] = fv[q];
                v[0] += fv[q] * fv[q];
                v[1] += fv[q] * fprev[3 * b + q];
                v[2] = rmax(v[2], fabs(fv[q]));
                if (cand) B.xmin[bk ^ 1][3 * b + q] = xn[3 * b + q];
            }
        }
        v[0] = wave_sum(v[0]);
        v[1] = wave_sum(v[1]);
        v[2] = wave_max(v[2]);
        if ((threadIdx.x & 63) == 0)
            for (int q = 0; q < 3; q++) red[q * NSW + (threadIdx.x >> 6)] = v[q];
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // the summing waves' results in wave order
        double r[3] = {red[0], red[NSW], red[2 * NSW]};
        for (int k = 1; k < NSW; k++) {
            r[0] += red[k];
            r[1] += red[NSW + k];
            r[2] = rmax(r[2], red[2 * NSW + k]);
        }
        for (int q = 0; q < 3; q++) B.gpartial[(long)q * gridDim.x + blockIdx.x] = r[q];   // value-major
    }
}

// Scalar logic after a force evaluation: maxF, energyRelChange, beta, the minimum-energy
// backup and the loop condition (CGMethod.hpp:1406-1576).
__device__ inline void gather_control(CgState& S, const double* acc, int bad_count) {
    if (bad_count != 0) {
        // an excluded-volume pair needs the numerical force fallback (forceN): pause before any
        // state change; the host completes this step's forces and re-runs the gather
        S.status = ST_NEED_FORCEN;
        S.done = 1;
        return;
    }
    const double FF = acc[0], FP = acc[1], MX = acc[2];
    const bool cand = S.num_iter > S.force_call_limit / 2 && S.total_energy <= S.minimum_e;
    S.n_gather_launch += 1;
    S.parity ^= 1;
    S.fpar ^= 1;
    S.num_force_call += 1;
    S.grad_inf_norm = MX;
    S.fdotf = FF;
    if (!isfinite(MX)) {        // maxF error (CGMethod.hpp:202-212)
        S.status = 4;
        S.done = 1;
        return;
    }
    S.phase = PH_LS;
    S.ls_active = 0;
    S.ls_lambda = rmin(S.lambda_max, S.max_dist / MX);
    if (S.recovery_move) {
        // after the recovery move: forces, maxF, resetCGParams, safe mode off (:1342-1348)
        S.num_recovery += 1;
        S.recovery_move = 0;
        S.safe_mode = 0;
        S.ecur_valid = 0;
        S.dir_pending = 1;
        S.dir_beta = 0.0;
        S.dir_reset = 1;
    } else {
        S.energy_rel_change = fabs(S.new_energy - S.prev_energy) / fabs(S.prev_energy);
        S.prev_energy = S.new_energy;
        S.ls_cur = S.new_energy;      // E(x_new) = the accepted trial energy, bit for bit
        S.ecur_valid = S.q_stale ? 0 : 1;   // unless the quadratic search moved elsewhere
        S.q_stale = 0;
        double beta = 0.0;
        if (S.method == 1) beta = FF / S.last_fdotf;
        else if (S.method == 2) beta = rmax(0.0, (FF - FP) / S.last_fdotf);
        S.dir_pending = 1;
        S.dir_beta = beta;
        S.dir_reset = 0;
        if (cand && MX <= 10 * S.grad_tol) {   // commit the candidate backup
            S.bkpar ^= 1;
            S.minimum_e = S.total_energy;
            S.have_min_e = 1;
        }
        if (S.method != 0) S.last_fdotf = FF;   // lastFDotF = newGrad (CGMethod.hpp:1570-1572)
    }
    if (call_limit_s(S) || converged_s(S)) S.done = 1;
}

// One block after every slot: the line-search decision (phase PH_LS: the k_cg_ls and k_cg_pos
// partials in a fixed order (the columns of DevSys::lsp), the deferred Simpson energies,
// ls_control) and, when the step is accepted at the
// speculative trial or a move just ran (phase PH_MOVE), the post-step control on the gather
// partials (gather_control). The kernel is a chain of latencies, so every independent global
// read (the state, the deferred-pair and bad-pair counts, both sets of partials) is issued at
// the start, the eleven sums and extrema are reduced in one pass, and thread 0 runs the scalar
// control on a register copy of the state and writes it back.
__global__ void __launch_bounds__(FIN) k_cg_fin(DevSys s, CgState* __restrict__ st, CgBufs B, int ngb) {
    constexpr int NW = sizeof(CgState) / sizeof(int);
    static_assert(NW <= FIN, "the state is staged one word per thread");
    __shared__ CgState Sl;
    __shared__ double lds[8 * (FIN / 8)];
    __shared__ double res8[8];
    __shared__ int sh_ndef, sh_bad;
    if (threadIdx.x < NW) reinterpret_cast<int*>(&Sl)[threadIdx.x] = reinterpret_cast<const int*>(st)[threadIdx.x];
    if (threadIdx.x == 0) {
        sh_ndef = __hip_atomic_load(s.defer_n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        sh_bad = *B.bad_count;
    }
    const int T = s.blk[NFF], np = s.lsp;
    // energy columns: the interaction blocks in the line search's grid, then (bonded_gather) its
    // bead blocks at T; the bead kernel's columns at T + nbb
    const bool be = s.bonded_gather && s.bead_energy;
    const int ne0 = be ? s.blk_nb : T, neb = be ? s.nbb : 0;
    const int ne = ne0 + neb, p0c = T + s.nbb;
    // E(lambda), E(lambda/2), -, -, E(x), -, d.F, max|d| (then the recovery step) | F.F, F.F_prev, max|F|
    double acc[11] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    // E(lambda / 2) is read only after a launch that evaluated it (a search's later launches) and
    // E(x) only when the first launch of a search evaluated it: the common launch reads one energy
    // column; the flags are loaded up front and only those rarer loads wait for them
    const int ls_active = st->ls_active, ecur_valid = st->ecur_valid;
    // the other columns' loads in flight before any sum (one round trip when they fit)
    const int nc = ne > s.nbb ? ne : s.nbb;
    for (int base = 0; base < nc; base += 8 * FIN) {
        double p0[8], q0[8], q1[8];
#pragma unroll
        for (int j = 0; j < 8; j++) {
            const int c = base + threadIdx.x + j * FIN;
            const long col = c < ne0 ? c : T + (c - ne0);
            p0[j] = c < ne ? B.partial[col] : 0.0;
            q0[j] = c < s.nbb ? B.partial[p0c + c] : 0.0;
            q1[j] = c < s.nbb ? B.partial[np + p0c + c] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 8; j++) {
            acc[0] += p0[j];
            acc[6] += q0[j];
            acc[7] = rmax(acc[7], q1[j]);
        }
    }
    for (int base = 0; base < ngb; base += 4 * FIN) {
        double p[4][3];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int bk = base + threadIdx.x + j * FIN;
            for (int q = 0; q < 3; q++) p[j][q] = bk < ngb ? B.gpartial[(long)q * ngb + bk] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            acc[8] += p[j][0];
            acc[9] += p[j][1];
            acc[10] = rmax(acc[10], p[j][2]);
        }
    }
    const bool want_half = ls_active != 0, want_x = ls_active == 0 && ecur_valid == 0;
    if (want_half || want_x)   // uniform
        for (int base = 0; base < ne; base += 8 * FIN) {
            double p1[8], p2[8];
#pragma unroll
            for (int j = 0; j < 8; j++) {
                const int c = base + threadIdx.x + j * FIN;
                const long col = c < ne0 ? c : T + (c - ne0);
                p1[j] = want_half && c < ne ? B.partial[np + col] : 0.0;
                p2[j] = want_x && c < ne ? B.partial[2L * np + col] : 0.0;
            }
#pragma unroll
            for (int j = 0; j < 8; j++) {
                acc[1] += p1[j];
                acc[4] += p2[j];
            }
        }
    __syncthreads();
    const int phase0 = Sl.phase;
    if (Sl.done || (phase0 != PH_LS && phase0 != PH_MOVE)) return;
    // the eight reduced values (the three flags of acc are set from the deferred pass below)
    double r8[8] = {acc[0], acc[1], acc[4], acc[6], acc[7], acc[8], acc[9], acc[10]};
    const int op[8] = {0, 0, 0, 0, 1, 0, 0, 1};
    block_reduce_fold<8, FIN>(r8, op, lds, res8);
    acc[0] = r8[0]; acc[1] = r8[1]; acc[4] = r8[2]; acc[6] = r8[3];
    // the recovery step, min over components of maxDist / |d| (MController.h:293-305), equals
    // maxDist / max|d| under correctly rounded division: one division instead of one per component
    acc[7] = Sl.max_dist / r8[4];
    acc[8] = r8[5]; acc[9] = r8[6]; acc[10] = r8[7];
    double dsum[3] = {0.0, 0.0, 0.0};
    int dbad[3] = {0, 0, 0};
    const int ndef = sh_ndef;
    if (phase0 == PH_LS && ndef > 0) (void)run_defers_n<FIN>(s, ndef, lds, dsum, dbad, 3);
    if (threadIdx.x == 0) {
        CgState S = Sl;
        int bad = sh_bad;
        bool gather = phase0 == PH_MOVE;
        if (phase0 == PH_LS) {
            acc[0] += dsum[0]; acc[2] = dbad[0];
            acc[1] += dsum[1]; acc[3] = dbad[1];
            acc[4] += dsum[2]; acc[5] = dbad[2];
            if (ndef > DEFER_CAP) {
                S.status = 9;
                S.done = 1;
            } else {
                ls_control(S, acc, bad);
                gather = !S.done && S.phase == PH_GATHER;
            }
        } else {
            S.n_move_launch += 1;
        }
        if (gather) gather_control(S, acc + 8, bad);
        if (bad != sh_bad) *B.bad_count = bad;
        *st = S;
    }
}

// =======================================================================================
// Tangent (Hessian-vector product) and parameter-gradient kernels
// =======================================================================================
__device__ inline DualD3 pos_dual(const double* __restrict__ x, const double* __restrict__ v, int b) {
    return DualD3{ld3(x, b), ld3(v, b)};
}

// d/de of interaction t's forces at x + e v (forward-mode dual arithmetic) into the pieces /
// slots (FP64); their gather is (dF/dx) v = -H v. RT = DualD: FP64 throughout; RT = DualF: the
// mixed path's tangents, the stretching, bending and crosslink arithmetic in FP32 duals (bond
// vectors, lengths and extensions formed in FP64 as in the mixed forces), excluded volume and the
// boundary in FP64. *fbad: an excluded-volume pair whose analytic force is non-finite (the
// reference would switch to forceN, which has no tangent here).
template <typename RT>
__device__ inline void term_tangent(const DevSys& s, int ff, int t, const double* __restrict__ x,
                                    const double* __restrict__ v, double* __restrict__ slots,
                                    double* __restrict__ piece, int* fbad) {
    // the filament terms' tangents go to the bead-major pieces of this product: piece[k][bead]
    auto put = [&](int k, int bead, D3 d) {
        double* q = piece + 3L * ((long)k * s.n_beads + bead);
        q[0] = d.x; q[1] = d.y; q[2] = d.z;
    };
    DualD3 out[4];
    switch (ff) {
    case FF_STRETCH: {
        const int b1 = s.str_b[2 * t], b2 = s.str_b[2 * t + 1];
        stretch_forces<RT>(pos_dual(x, v, b1), pos_dual(x, v, b2), s.str_k[t], s.str_eql[t], out);
        put(1, b1, out[0].d);
        put(0, b2, out[1].d);
        return;
    }
    case FF_BEND: {
        const int b1 = s.bend_b[3 * t], b2 = s.bend_b[3 * t + 1], b3 = s.bend_b[3 * t + 2];
        bend_forces<RT>(pos_dual(x, v, b1), pos_dual(x, v, b2), pos_dual(x, v, b3), s.bend_k[t], s.bend_eqt[t], out);
        put(4, b1, out[0].d);
        put(3, b2, out[1].d);
        put(2, b3, out[2].d);
        return;
    }
    case FF_LINKER:
    case FF_MOTOR: {
        const bool lk = ff == FF_LINKER;
        const int* b = lk ? s.lnk_b : s.mot_b;
        const double k = lk ? s.lnk_k[t] : s.mot_k[t], eql = lk ? s.lnk_eql[t] : s.mot_eql[t];
        const double p1 = lk ? s.lnk_p1[t] : s.mot_p1[t], p2 = lk ? s.lnk_p2[t] : s.mot_p2[t];
        (void)xlink_forces<RT>(pos_dual(x, v, b[4 * t]), pos_dual(x, v, b[4 * t + 1]), pos_dual(x, v, b[4 * t + 2]),
                                  pos_dual(x, v, b[4 * t + 3]), k, eql, p1, p2, out);
        const long id = s.slot0[ff] + 4L * (lk ? s.lnk_ord : s.mot_ord)[t];
        for (int q = 0; q < 4; q++) put_slot(s, slots, id + q, out[q].d);
        return;
    }
    case FF_EXVOL: {
        const bool ok = exvol_forces<DualD>(pos_dual(x, v, s.ev_b[4 * t]), pos_dual(x, v, s.ev_b[4 * t + 1]),
                                            pos_dual(x, v, s.ev_b[4 * t + 2]), pos_dual(x, v, s.ev_b[4 * t + 3]),
                                            s.ev_ks[t], out);
        const long id = s.slot0[FF_EXVOL] + 4L * t;
        for (int q = 0; q < 4; q++) put_slot(s, slots, id + q, ok ? out[q].d : D3{0.0, 0.0, 0.0});
        if (!ok) *fbad = 1;
        return;
    }
    default: {
        const BElem& e = s.elem[s.bnd_e[t]];
        put_slot(s, slots, s.slot0[FF_BOUNDARY] + t,
                 boundary_force_dual(e, pos_dual(x, v, s.bnd_b[t]), s.bnd_krep[t], s.bnd_slen[t]).d);
        return;
    }
    }
}

// Grid: (interaction blocks, products). Product k = blockIdx.y takes its tangent from
// v + k v_stride and writes its slots and pieces at slots + k slot_stride, piece + k piece_stride.
template <typename RT>
__global__ void __launch_bounds__(BLOCK) k_tangent_terms(DevSys s, const double* __restrict__ x,
                                                         const double* __restrict__ v, long v_stride,
                                                         double* __restrict__ slots, long slot_stride,
                                                         double* __restrict__ piece, long piece_stride,
                                                         int* bad_count) {
    const int ff = ff_of_block(s, blockIdx.x);
    if (ff >= NFF) return;
    const int t = (blockIdx.x - s.fblk[ff]) * BLOCK + threadIdx.x;
    if (t >= s.n[ff]) return;
    const long k = blockIdx.y;
    int fbad = 0;
    term_tangent<RT>(s, ff, t, x, v + k * v_stride, slots + k * slot_stride, piece + k * piece_stride, &fbad);
    if (fbad) atomicAdd(bad_count, 1);
}

// Per-pair excluded-volume energies (closed form, FP64). A pair whose closed form is non-finite
// or < -1 (the reference then integrates numerically: k_ev_pair_energyN) gets NaN and is listed
// in bad_list (bad_count of them, in no particular order).
__global__ void k_ev_pair_energy(DevSys s, const double* __restrict__ x, double* __restrict__ U, int* bad_count,
                                 int* bad_list) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= s.n[FF_EXVOL]) return;
    int deferred = 0;
    const double u = exvol_energy<double>(ld3(x, s.ev_b[4 * t]), ld3(x, s.ev_b[4 * t + 1]), ld3(x, s.ev_b[4 * t + 2]),
                                          ld3(x, s.ev_b[4 * t + 3]), s.ev_ks[t], &deferred);
    U[t] = deferred ? NAN : u;
    if (deferred) bad_list[atomicAdd(bad_count, 1)] = t;
}
// The Simpson energy (energyN) of pair t into U[t] (-1 when it stays non-finite). One block.
__global__ void __launch_bounds__(BLOCK) k_ev_pair_energyN(DevSys s, const double* __restrict__ x, int t,
                                                           double* __restrict__ U) {
    __shared__ double lds[BLOCK / 64];
    const D3 c1 = ld3(x, s.ev_b[4 * t]), c2 = ld3(x, s.ev_b[4 * t + 1]), c3 = ld3(x, s.ev_b[4 * t + 2]),
             c4 = ld3(x, s.ev_b[4 * t + 3]);
    double r = exvol_energyN_share(c1, c2, c3, c4, threadIdx.x, BLOCK);
    const int op1[1] = {0};
    block_reduce_n<1>(&r, op1, lds);
    if (threadIdx.x == 0) {
        const double u = exvol_energyN_finish(c1, c2, c3, c4, s.ev_ks[t], r);
        U[t] = energy_bad(u) ? -1.0 : u;
    }
}

// dE/dtheta per cylinder c, theta = (stretching k, rest length L0, bending k):
//   stretching term c (FilamentStretchingHarmonic.cpp:185-199): dE/dk = (L - L0)^2 / 2,
//     dE/dL0 = -k (L - L0);
//   excluded volume (kRepScaled = krep L0_i L0_j, CylinderExclVolume.cpp:73-138): every pair with
//     cylinder c adds U / L0_c to dE/dL0_c; those pairs are the ones in the incidence list of c's
//     first bead with the bead as the pair's bead 1 or bead 3;
//   bending term whose second cylinder is c (its k, FilamentBending.cpp:29-60): dE/dk = U / k.
// out[c], out[n_cyl + c], out[2 n_cyl + c]. R = float: the mixed path's arithmetic (the
// extension formed in FP64 and the stretching derivatives rounded to FP32 as the mixed stretching
// energy is, FP32 bending energy, FP64 excluded-volume energies).
// PG_LANES lanes per cylinder (in one wave): they stride over the incidence list, their sums of
// pair energies meet in a fixed xor tree (the same on every run), and lane 0 writes the
// stretching derivatives, lane 1 the bending one. Grid: n_cyl * PG_LANES threads.
constexpr int PG_LANES = 8;
template <typename R>
__global__ void __launch_bounds__(BLOCK) k_param_grad(DevSys s, int n_cyl, const int* __restrict__ cyl_b1,
                                                      const int* __restrict__ cyl_bend, const double* __restrict__ x,
                                                      const double* __restrict__ U, double* __restrict__ out) {
    static_assert(64 % PG_LANES == 0, "a cylinder's lanes share one wave");
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int c = g / PG_LANES, lane = g % PG_LANES;
    double eu = 0.0;   // sum of U over c's pairs, this lane's share
    if (c < n_cyl) {
        const int b = cyl_b1[c];
        const long e0 = s.slot0[FF_EXVOL], e1 = s.slot0[FF_EXVOL + 1];
        for (int q = s.bstart[b] + lane; q < s.bstart[b + 1]; q += PG_LANES) {
            const long id = s.sorted_id[q];
            const long loc = id - e0;
            if (id >= e0 && id < e1 && (loc & 1) == 0) eu += U[loc >> 2];   // bead 1 or bead 3 of the pair
        }
    }
#pragma unroll
    for (int m = PG_LANES / 2; m >= 1; m /= 2) eu += __shfl_xor(eu, m, PG_LANES);
    if (c >= n_cyl) return;
    if (lane == 0) {
        const V3<double> r = dif<double>(ld3(x, s.str_b[2 * c + 1]), ld3(x, s.str_b[2 * c]));
        const double dl = sqrt(r.x * r.x + r.y * r.y + r.z * r.z) - s.str_eql[c];
        out[c] = (double)(R)(0.5 * dl * dl);
        out[n_cyl + c] = (double)(R)(-s.str_k[c] * dl) + eu / s.str_eql[c];
    } else if (lane == 1) {
        const int tb = cyl_bend[c];
        out[2 * n_cyl + c] = tb >= 0 ? (double)bend_energy<R>(ld3(x, s.bend_b[3 * tb]), ld3(x, s.bend_b[3 * tb + 1]),
                                                              ld3(x, s.bend_b[3 * tb + 2]), 1.0, s.bend_eqt[tb])
                                     : 0.0;
    }
}

// =======================================================================================
// Neighbor lists
// =======================================================================================
struct NLDev {
    int n_cyl;
    const int* cyl_b1; const int* cyl_b2; const int* cyl_fil; const int* cyl_pos; const int* cyl_type;
    const double* cyl_kexvol; const double* cyl_eql; const char* cyl_minus;
    double rmaxsq, rminsq;      // squares formed in float, widened to double
    double side;                // bin side
    int nb[3];
    double grid[3];
    int ta, tb;                 // the filament-type pair a list is restricted to (ta <= tb; -1: any)
    // bins are numbered tile by tile: tiles of T x T x T bins (T = 2, 4 or 8), nt[] tiles per
    // axis; bin id = tile * T^3 + lx + T (ly + T lz)
    int T, lt;                  // lt = log2 T
    int nt[3];
};

__device__ inline int nl_bin_id(const NLDev& n, int x, int y, int z) {
    const int m = n.T - 1;
    const int tile = (x >> n.lt) + n.nt[0] * ((y >> n.lt) + n.nt[1] * (z >> n.lt));
    return (tile << (3 * n.lt)) + (x & m) + ((y & m) << n.lt) + ((z & m) << (2 * n.lt));
}

// midpoints 0.5*(x1 + x2) (Cylinder.cpp:48) and bin index; flags out-of-grid cylinders
__global__ void k_nl_bin(NLDev n, const double* __restrict__ x, double* __restrict__ mid, int* __restrict__ bin,
                         int* __restrict__ bin_count, int* __restrict__ err) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n.n_cyl) return;
    int b1 = n.cyl_b1[c], b2 = n.cyl_b2[c];
    int ix[3];
    bool ok = true;
    for (int q = 0; q < 3; q++) {
        double m = 0.5 * (x[3 * b1 + q] + x[3 * b2 + q]);
        mid[3 * c + q] = m;
        ok = ok && (m >= 0 && m < n.grid[q]);
        int i = ok ? (int)(m / n.side) : 0;
        if (i >= n.nb[q]) i = n.nb[q] - 1;
        ix[q] = i;
    }
    if (!ok) {
        atomicAdd(err, 1);
        bin[c] = 0;
        return;
    }
    int bb = nl_bin_id(n, ix[0], ix[1], ix[2]);
    bin[c] = bb;
    atomicAdd(&bin_count[bb], 1);
}

// the members of each bin, with their bin and midpoint in the same (bin) order
__global__ void k_nl_fill(int n_cyl, const int* __restrict__ bin, const int* __restrict__ bin_start,
                          int* __restrict__ cursor, const double* __restrict__ mid, int* __restrict__ members,
                          int* __restrict__ mbin, double* __restrict__ mmid) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cyl) return;
    int bb = bin[c];
    int k = atomicAdd(&cursor[bb], 1);
    const int m = bin_start[bb] + k;
    members[m] = c;
    mbin[m] = bb;
    for (int q = 0; q < 3; q++) mmid[3 * (long)m + q] = mid[3 * (long)c + q];
}

// HybridNeighborListImpl.cpp:411-439 predicate, on explicit midpoints (a candidate's from LDS)
__device__ inline bool nl_keep_at(const NLDev& n, int c, int cn, const double* mc, const double* mn) {
    if (c <= cn) return false;
    if (n.cyl_fil[c] == n.cyl_fil[cn]) {
        int sep = abs(n.cyl_pos[c] - n.cyl_pos[cn]);
        if (sep <= 2) return false;
    }
    if (n.ta >= 0) {
        const int t1 = n.cyl_type[c], t2 = n.cyl_type[cn];
        if (min(t1, t2) != n.ta || max(t1, t2) != n.tb) return false;
    }
    double d2 = 0.0;
    for (int q = 0; q < 3; q++) {
        double v = mn[q] - mc[q];
        d2 += v * v;
    }
    return !(d2 < n.rminsq || d2 > n.rmaxsq);
}

// Cell-list pairs, one workgroup per tile of T^3 bins (HybridCylinderCylinderNL::updateNeighborsbin,
// HybridNeighborListImpl.cpp:332-471, over the 27 bins around each cylinder's bin):
//   1. each thread takes halo bins (the tile's bins and one bin around them, (T + 2)^3) and
//      reads their member ranges; a block scan of the counts places every halo bin's members
//      contiguously, in halo-bin order, in LDS with their midpoints,
//   2. every cylinder of the tile tests the members of its 27 bins, read from LDS as nine
//      x-rows of three consecutive halo bins.
// A halo with more than NL_CAP members is read from global memory instead (the same 27 bins). The
// bins are wider than the range, so the pair set is the predicate's whatever the bin size.
// PASS 0: count per cylinder; PASS 1: emit into [off[c], off[c + 1]) and sort ascending.
constexpr int NL_CAP = 768;
constexpr int NL_HMAX = 10 * 10 * 10;   // (T + 2)^3 for T <= 8

template <int PASS>
__global__ void __launch_bounds__(BLOCK) k_nl_tile(NLDev n, const int* __restrict__ bin_start,
                                                   const int* __restrict__ members, const int* __restrict__ mbin,
                                                   const double* __restrict__ mmid, int* __restrict__ count,
                                                   const int* __restrict__ off, int* __restrict__ pairs) {
    __shared__ int s_a[NL_HMAX], s_start[NL_HMAX + 1], s_part[BLOCK];
    __shared__ int s_sc[NL_CAP];
    __shared__ double s_mid[3 * NL_CAP];
    const int T = n.T, lt = n.lt, T3 = 1 << (3 * lt), H = T + 2, H3 = H * H * H;
    const int tile = blockIdx.x;
    const int i0 = bin_start[(long)tile * T3], i1 = bin_start[(long)(tile + 1) * T3];
    if (i0 == i1) return;   // an empty tile (the whole workgroup)
    const int tx = tile % n.nt[0], ty = (tile / n.nt[0]) % n.nt[1], tz = tile / (n.nt[0] * n.nt[1]);
    const int hx0 = tx * T - 1, hy0 = ty * T - 1, hz0 = tz * T - 1;
    // 1. the halo bins' member ranges (bins outside the grid are empty)
    const int run = (H3 + BLOCK - 1) / BLOCK, r0 = threadIdx.x * run, r1 = min(r0 + run, H3);
    int tot = 0;
    for (int h = r0; h < r1; h++) {
        const int X = hx0 + h % H, Y = hy0 + (h / H) % H, Z = hz0 + h / (H * H);
        int a = 0, c = 0;
        if (X >= 0 && Y >= 0 && Z >= 0 && X < n.nb[0] && Y < n.nb[1] && Z < n.nb[2]) {
            const int g = nl_bin_id(n, X, Y, Z);
            a = bin_start[g];
            c = bin_start[g + 1] - a;
        }
        s_a[h] = a;
        s_start[h] = c;
        tot += c;
    }
    s_part[threadIdx.x] = tot;
    __syncthreads();
    for (int step = 1; step < BLOCK; step *= 2) {   // inclusive scan of the runs' totals
        const int v = threadIdx.x >= step ? s_part[threadIdx.x - step] : 0;
        __syncthreads();
        s_part[threadIdx.x] += v;
        __syncthreads();
    }
    {
        int acc = s_part[threadIdx.x] - tot;   // exclusive prefix of this run
        for (int h = r0; h < r1; h++) {
            const int v = s_start[h];
            s_start[h] = acc;
            acc += v;
        }
        if (threadIdx.x == BLOCK - 1) s_start[H3] = s_part[BLOCK - 1];
    }
    __syncthreads();
    const int nh = s_start[H3];
    const bool staged = nh <= NL_CAP;   // the same for the whole workgroup
    if (staged) {
        int h = 0;
        for (int j = threadIdx.x; j < nh; j += BLOCK) {
            // the halo bin holding position j: the last h with s_start[h] <= j (binary search)
            int lo = 0, hi = H3 - 1;
            while (lo < hi) {
                const int md = (lo + hi + 1) >> 1;
                if (s_start[md] <= j) lo = md;
                else hi = md - 1;
            }
            h = lo;
            const int m = s_a[h] + (j - s_start[h]);
            s_sc[j] = members[m];
            for (int q = 0; q < 3; q++) s_mid[3 * j + q] = mmid[3 * (long)m + q];
        }
        __syncthreads();
    }
    // 2. the tile's cylinders
    for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
        const int c = members[i];
        const double mc[3] = {mmid[3 * (long)i], mmid[3 * (long)i + 1], mmid[3 * (long)i + 2]};
        const int loc = mbin[i] & (T3 - 1);
        const int lx = (loc & (T - 1)) + 1, ly = ((loc >> lt) & (T - 1)) + 1, lz = (loc >> (2 * lt)) + 1;   // halo coords
        int cnt = 0;
        const long base = PASS ? off[c] : 0;
        if (staged) {
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++) {
                    const int h = (lx - 1) + H * ((ly + dy) + H * (lz + dz));
                    for (int k = s_start[h]; k < s_start[h + 3]; k++) {
                        const int cn = s_sc[k];
                        if (nl_keep_at(n, c, cn, mc, &s_mid[3 * k])) {
                            if (PASS) pairs[2 * (base + cnt) + 1] = cn;
                            cnt++;
                        }
                    }
                }
        } else {
            for (int dz = -1; dz <= 1; dz++)
                for (int dy = -1; dy <= 1; dy++)
                    for (int dx = -1; dx <= 1; dx++) {
                        const int X = hx0 + lx + dx, Y = hy0 + ly + dy, Z = hz0 + lz + dz;
                        if (X < 0 || Y < 0 || Z < 0 || X >= n.nb[0] || Y >= n.nb[1] || Z >= n.nb[2]) continue;
                        const int nbin = nl_bin_id(n, X, Y, Z);
                        for (int m = bin_start[nbin]; m < bin_start[nbin + 1]; m++) {
                            const int cn = members[m];
                            if (nl_keep_at(n, c, cn, mc, &mmid[3 * (long)m])) {
                                if (PASS) pairs[2 * (base + cnt) + 1] = cn;
                                cnt++;
                            }
                        }
                    }
        }
        if (!PASS) {
            count[c] = cnt;
            continue;
        }
        // insertion sort of the neighbour ids, then the owning cylinder id
        for (int a = 1; a < cnt; a++) {
            const int v = pairs[2 * (base + a) + 1];
            int j = a - 1;
            while (j >= 0 && pairs[2 * (base + j) + 1] > v) {
                pairs[2 * (base + j + 1) + 1] = pairs[2 * (base + j) + 1];
                j--;
            }
            pairs[2 * (base + j + 1) + 1] = v;
        }
        for (int a = 0; a < cnt; a++) pairs[2 * (base + a)] = c;
    }
}

// HybridBindingSearchManager::updateAllPossibleBindingsstencilHYBD (HybridBindingSearchManager.cpp:
// 759-966) for one list pair (c, cn): the distance between site point p1 = x1 + s1 (x2 - x1) and
// x3 + t (x4 - x3) is a quadratic A t^2 + B t + C in t; its roots at rmin^2 and rmax^2 bound the
// sites s2 of cn in range. The expression types are the reference's: site positions and squared
// bounds in float, the rest in FP64.
struct BindDev {
    int n_sites, shift, step, ftype1, ftype2;
    const float* site;
    const float* free;    // [n_cyl * n_sites] or null
    float rminsq, rmaxsq, p2_first, p2_last;
};

template <int PASS>
__global__ void k_bind_sites(NLDev n, BindDev b, const double* __restrict__ x, long n_pairs,
                             const int* __restrict__ pairs, int* __restrict__ count, const int* __restrict__ off,
                             uint32_t* __restrict__ out) {
    const long pi = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pi >= n_pairs) return;
    const int c = pairs[2 * pi], cn = pairs[2 * pi + 1];
    int cnt = 0;
    const long base = PASS ? off[pi] : 0;
    // same filament, and the reference's type test (it compares c's own type with the
    // complementary one, so only a pair of equal types passes)
    const int tc = n.cyl_type[c];
    const bool typed = tc == b.ftype1 || tc == b.ftype2;
    const int compl_t = tc == b.ftype1 ? b.ftype2 : b.ftype1;
    if (n.cyl_fil[c] != n.cyl_fil[cn] && typed && tc == compl_t) {
        const D3 x1 = ld3(x, n.cyl_b1[c]), x2 = ld3(x, n.cyl_b2[c]);
        const D3 x3 = ld3(x, n.cyl_b1[cn]), x4 = ld3(x, n.cyl_b2[cn]);
        const double a12[3] = {x2.x - x1.x, x2.y - x1.y, x2.z - x1.z};
        const double a13[3] = {x3.x - x1.x, x3.y - x1.y, x3.z - x1.z};
        const double a34[3] = {x4.x - x3.x, x4.y - x3.y, x4.z - x3.z};
        auto dot = [](const double* u, const double* v) { return u[0] * v[0] + u[1] * v[1] + u[2] * v[2]; };
        const double s13 = dot(a13, a13), s12 = dot(a12, a12), d13_12 = dot(a13, a12);
        const double s34 = dot(a34, a34), d13_34 = dot(a13, a34), d34_12 = dot(a34, a12);
        for (int p1 = 0; p1 < b.n_sites; p1++) {
            if (b.step > 0 && p1 % b.step != 0) continue;
            if (b.free && b.free[(long)c * b.n_sites + p1] != 1.0f) continue;
            const float mp1 = b.site[p1];
            const double A = s34;
            const double B = 2.0 * d13_34 - 2.0 * mp1 * d34_12;
            const double C = s13 + (double)(mp1 * mp1) * s12 - 2.0 * mp1 * d13_12;
            const double Bsq = B * B;
            const double C1 = C - (double)b.rminsq, C2 = C - (double)b.rmaxsq;
            const double q1 = Bsq - 4 * A * C1, q2 = Bsq - 4 * A * C2;
            const bool st1 = q1 < 0, st2 = q2 < 0;
            if (st1 && st2) continue;
            double mn[2] = {0.0, 0.0}, mx[2] = {0.0, 0.0};
            if (!st1) {
                const double r1 = (-B + sqrt(q1)) / (2 * A), r2 = (-B - sqrt(q1)) / (2 * A);
                mn[0] = r1 < r2 ? r1 : r2;
                mn[1] = r1 < r2 ? r2 : r1;
                if (mn[0] < b.p2_first && mn[1] > b.p2_last) continue;   // every site too close
            }
            if (!st2) {
                const double r1 = (-B + sqrt(q2)) / (2 * A), r2 = (-B - sqrt(q2)) / (2 * A);
                mx[0] = r1 < r2 ? r1 : r2;
                mx[1] = r1 < r2 ? r2 : r1;
                if (mx[0] > b.p2_last || mx[1] < b.p2_first) continue;   // every site too far
            }
            for (int p2 = 0; p2 < b.n_sites; p2++) {
                if (b.step > 0 && p2 % b.step != 0) continue;
                if (b.free && b.free[(long)cn * b.n_sites + p2] != 1.0f) continue;
                const float mp2 = b.site[p2];
                if (!st2 && (mp2 < mx[0] || mp2 > mx[1])) continue;
                if (!st1 && (mp2 > mn[0] && mp2 < mn[1])) continue;
                if (PASS) {
                    out[2 * (base + cnt)] = ((uint32_t)c << b.shift) | (uint32_t)p1;
                    out[2 * (base + cnt) + 1] = ((uint32_t)cn << b.shift) | (uint32_t)p2;
                }
                cnt++;
            }
        }
    }
    if (!PASS) count[pi] = cnt;
}

// CylinderExclVolume::vectorize (CylinderExclVolume.cpp:73-138)
__global__ void k_ev_vectorize(NLDev n, long n_pairs, const int* __restrict__ pairs, int* __restrict__ ev_b,
                               double* __restrict__ ev_ks) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    int ci = pairs[2 * p], cn = pairs[2 * p + 1];
    ev_b[4 * p] = n.cyl_b1[ci];
    ev_b[4 * p + 1] = n.cyl_b2[ci];
    ev_b[4 * p + 2] = n.cyl_b1[cn];
    ev_b[4 * p + 3] = n.cyl_b2[cn];
    double ki = n.cyl_kexvol[ci], kn = n.cyl_kexvol[cn];
    double krep = (n.cyl_type[ci] != n.cyl_type[cn]) ? ((ki < kn) ? kn : ki) : ki;
    ev_ks[p] = krep * n.cyl_eql[ci] * n.cyl_eql[cn];
}

// BoundaryCylinderNL::updateNeighbors (NeighborListImpl.cpp:738-752): per (element, cylinder)
// interaction count (minus end adds its first bead, every cylinder its second bead)
__global__ void k_bnd_count(NLDev n, DevSys s, double cutoff, const double* __restrict__ mid, int* __restrict__ cnt) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)s.n_elem * n.n_cyl) return;
    int e = (int)(i / n.n_cyl), c = (int)(i % n.n_cyl);
    D3 m{mid[3 * c], mid[3 * c + 1], mid[3 * c + 2]};
    bool in = boundary_distance(s.elem[e], m) < cutoff;
    cnt[i] = in ? (n.cyl_minus[c] ? 2 : 1) : 0;
}
__global__ void k_bnd_emit(NLDev n, DevSys s, const int* __restrict__ cnt, const int* __restrict__ off,
                           const double* __restrict__ krep_e, const double* __restrict__ slen_e,
                           int* __restrict__ bnd_b, int* __restrict__ bnd_e, double* __restrict__ krep,
                           double* __restrict__ slen) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)s.n_elem * n.n_cyl) return;
    int k = cnt[i];
    if (!k) return;
    int e = (int)(i / n.n_cyl), c = (int)(i % n.n_cyl);
    int o = off[i];
    if (k == 2) {
        bnd_b[o] = n.cyl_b1[c]; bnd_e[o] = e; krep[o] = krep_e[e]; slen[o] = slen_e[e];
        o++;
    }
    bnd_b[o] = n.cyl_b2[c]; bnd_e[o] = e; krep[o] = krep_e[e]; slen[o] = slen_e[e];
}

// ---- boundary load forces on the filament ends ---------------------------------------
// BoundaryCylinderRepulsion::computeLoadForces (BoundaryCylinderRepulsion.cpp:122-187, the
// helper :125-162) after the reset of ForceFieldManager::computeLoadForces
// (ForceFieldManager.cpp:292-310): for the plus-end cylinder of every filament (bead bd = its
// second bead, bo = its first) and the minus-end cylinder (bd = its first bead, bo = its
// second), loadForces[i] = sum over the boundary elements whose neighbor list holds the cylinder,
// in element order, of max(-n(p_i).dir, 0) * kRep exp(-d(p_i) / lambda) / lambda
// (BoundaryCylinderRepulsionExp::loadForces, BoundaryCylinderRepulsionExp.cpp:113-118) at
// p_i = bd + (i * monomerSize) * dir, dir = normalizedVector(bd - bo), i < cylinderNumMon of the
// filament's type. FP64. One thread per (end, filament, monomer); out[end][filament][stride].
constexpr int MAX_FIL_TYPES = 8;
struct LoadParams {
    int n_types, stride;
    double mon_size[MAX_FIL_TYPES];
    int num_mon[MAX_FIL_TYPES];
};
__global__ void k_load_forces(NLDev n, DevSys s, const double* __restrict__ x, const int* __restrict__ fil_ends,
                              const int* __restrict__ bnd_cnt, const double* __restrict__ krep_e,
                              const double* __restrict__ slen_e, int n_fil, LoadParams P, double* __restrict__ out) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2L * n_fil * P.stride) return;
    const int end = (int)(i / ((long)n_fil * P.stride));          // 0: plus, 1: minus
    const int f = (int)((i / P.stride) % n_fil), k = (int)(i % P.stride);
    const int c = fil_ends[2 * f + (end == 0 ? 1 : 0)];             // last / first cylinder
    const int t = n.cyl_type[c];
    double acc = 0.0;
    if (t >= 0 && t < P.n_types && k < P.num_mon[t]) {
        const D3 xd = ld3(x, end == 0 ? n.cyl_b2[c] : n.cyl_b1[c]);
        const D3 xo = ld3(x, end == 0 ? n.cyl_b1[c] : n.cyl_b2[c]);
        double dir[3] = {xd.x - xo.x, xd.y - xo.y, xd.z - xo.z};
        const double norm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
        for (int q = 0; q < 3; q++) dir[q] /= norm;
        const double step = k * P.mon_size[t];
        const D3 p{xd.x + step * dir[0], xd.y + step * dir[1], xd.z + step * dir[2]};
        for (int e = 0; e < s.n_elem; e++) {
            if (bnd_cnt[(long)e * n.n_cyl + c] == 0) continue;   // not in element e's neighbor list
            const V3<double> nn = boundary_normal<double>(s.elem[e], p);
            const double dotnd = nn.x * dir[0] + nn.y * dir[1] + nn.z * dir[2];
            const double proj = rmax(-dotnd, 0.0);
            const double R = -boundary_distance(s.elem[e], p) / slen_e[e];
            acc += proj * (krep_e[e] * exp(R) / slen_e[e]);
        }
    }
    out[i] = acc;
}

// dynamic incidence lists: per bead, the excluded-volume and boundary slot ids, ascending
__global__ void k_inc_count(DevSys s, int* __restrict__ cnt) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nev = 4L * s.n[FF_EXVOL], nb = s.n[FF_BOUNDARY];
    if (i < nev) atomicAdd(&cnt[s.ev_b[i]], 1);
    else if (i < nev + nb) atomicAdd(&cnt[s.bnd_b[i - nev]], 1);
}
__global__ void k_inc_fill(DevSys s, const int* __restrict__ start, int* __restrict__ cursor, int* __restrict__ out) {
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long nev = 4L * s.n[FF_EXVOL], nb = s.n[FF_BOUNDARY];
    int bead;
    long slot;
    if (i < nev) { bead = s.ev_b[i]; slot = s.slot0[FF_EXVOL] + i; }
    else if (i < nev + nb) { bead = s.bnd_b[i - nev]; slot = s.slot0[FF_BOUNDARY] + (i - nev); }
    else return;
    int k = atomicAdd(&cursor[bead], 1);
    out[start[bead] + k] = (int)slot;
}
__global__ void k_inc_sort(int n_beads, const int* __restrict__ start, int* __restrict__ list) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_beads) return;
    int lo = start[b], hi = start[b + 1];
    for (int i = lo + 1; i < hi; i++) {
        int v = list[i], j = i - 1;
        while (j >= lo && list[j] > v) { list[j + 1] = list[j]; j--; }
        list[j + 1] = v;
    }
}
// per-bead slot count of the merged (static + dynamic) layout
__global__ void k_inc_total(int n_beads, const int* __restrict__ s_start, const int* __restrict__ d_start,
                            int* __restrict__ cnt) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_beads) return;
    cnt[b] = (s_start[b + 1] - s_start[b]) + (d_start[b + 1] - d_start[b]);
}
// bead-sorted layout: static slots then dynamic slots of each bead (static ids are smaller, so
// the bead's range stays ascending); sorted_id[p] = slot id at p, pos[slot id] = p
__global__ void k_inc_merge(int n_beads, const int* __restrict__ s_start, const int* __restrict__ s_list,
                            const int* __restrict__ d_start, const int* __restrict__ d_list,
                            const int* __restrict__ bstart, int* __restrict__ sorted_id, int* __restrict__ pos) {
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n_beads) return;
    int p = bstart[b];
    for (int q = s_start[b]; q < s_start[b + 1]; q++, p++) {
        int id = s_list[q];
        sorted_id[p] = id;
        pos[id] = p;
    }
    for (int q = d_start[b]; q < d_start[b + 1]; q++, p++) {
        int id = d_list[q];
        sorted_id[p] = id;
        pos[id] = p;
    }
}

}  // namespace mdn
