// mdn_batched.h -- persistent batched minimizer: many independent systems minimized by one
// kernel launch. Each workgroup pulls a system from a device queue (in the order the host gives,
// e.g. largest first) and runs CGMethod::minimize on it to completion with workgroup barriers
// only; then it pulls the next one. Every system keeps its own tables and vectors (the device
// buffers of its context, described by a DevSys and CgBufs record).
//
// The work of one CG step is the same as in the multi-kernel minimizer (k_cg_ls, k_cg_gather,
// k_cg_fin, with the move when it is needed) and the scalar control is the same code
// (ls_control, gather_control), so a system follows the reference's CGMethod::minimize
// (CGMethod.hpp:826-1743) exactly as the single-system path does; only the order of the
// block-level energy and dot-product reductions differs.
//
// Citations are relative to medyan-5.4.0/ in the reference tree.
#pragma once

#include "mdn_kernels.h"

namespace mdn {

struct BatchItem {
    DevSys s;         // the system's tables
    CgBufs B;         // its CG vectors (B.x[0] holds the start coordinates)
    CgState* st;      // out: final CG state
    double* report;   // out: [0..8] FP64 energy report at the start, [9..17] at the result
};

// values per thread per step of the workgroup's streaming passes over a system's vectors
constexpr int BU = 8;

struct BatchParams {
    int method, try_recover, force_call_limit, energy_call_limit, report_on_call_limit;
    int limit_override;   // > 0: force-call limit of every system (schedule tuning runs)
    int line_search;      // MDN_LS_*
    double max_dist, lambda_max, grad_tol, energy_rel_tol;
};

// Per-force-field energies at positions p summed over this thread's interactions; code tags
// pairs deferred to the Simpson pass.
template <typename R, int NT>
__device__ inline void blk_energies(const DevSys& s, const Pos& p, int code, double* e) {
    for (int ff = 0; ff < NFF; ff++) {
        double a = 0.0;
        for (int t = threadIdx.x; t < s.n[ff]; t += NT) {
            int fb = 0;
            a += term_ef<R>(s, ff, t, p, (R*)nullptr, code, &fb);
        }
        e[ff] = a;
    }
}

// The energy report of k_energy_report / k_energy_report_fin at x (arithmetic R) into
// rep[0..8] (thread 0).
template <typename R, int NT>
__device__ inline void blk_report(const DevSys& s, const double* x, double* rep, double* red) {
    double acc[NFF];
    const Pos p{x, nullptr, nullptr, 0.0, 0.0, 0};
    blk_energies<R, NT>(s, p, 0, acc);
    const int opn[NFF] = {0, 0, 0, 0, 0, 0};
    block_reduce_n<NFF, NT>(acc, opn, red);
    double dsum[1];
    int dbad[1];
    const int n = run_defers<NT>(s, red, dsum, dbad, 1);
    if (threadIdx.x == 0) {
        acc[FF_EXVOL] += dsum[0];
        double tot = 0.0, tot_s = 0.0;
        bool failed = false;
        for (int f2 = 0; f2 < NFF; f2++) {
            const double ef = (f2 == FF_EXVOL && dbad[0]) ? -1.0 : acc[f2];
            rep[f2] = ef;
            tot += ef;
            if (!failed) {
                if (ef <= -1) failed = true;
                else tot_s += ef;
            }
        }
        rep[6] = tot;
        rep[7] = failed ? INFINITY : tot_s;
        rep[8] = n > DEFER_CAP ? (double)n : 0.0;
    }
    __syncthreads();
}

// The filament stretching and bending forces as per-bead bonded sums at x (bonded_at_bead into
// s.bsum: each bead's pieces recomputed from the coordinates, FP64 sums in piece order, the values
// the piece arrays would hold) and, with E, the energies of the terms each bead starts added to
// *e. Beads are adjacent or read from the neighbour tables: one round of loads per bead instead of
// an index load, then coordinate loads, per term.
// Two beads per step (b and b + NT, both beads' loads before either's stores, so the second
// bead's loads are in flight with the first's); each thread still visits its beads in order.
template <typename R, int NT, bool E>
__device__ inline void blk_bonded(const DevSys& s, const double* __restrict__ x, double* e) {
    const int nb = s.n_beads;
    for (int b0 = threadIdx.x; b0 < nb; b0 += 2 * NT) {
        const int b1 = b0 + NT, b1c = min(b1, nb - 1);   // an absent second bead: any bead, dropped
        double e0 = 0.0, e1 = 0.0;
        const D3 a0 = bonded_at_bead<R, E>(s, x, b0, s.pmask[b0], D3{0.0, 0.0, 0.0}, &e0);
        const D3 a1 = bonded_at_bead<R, E>(s, x, b1c, s.pmask[b1c], D3{0.0, 0.0, 0.0}, &e1);
        s.bsum[3L * b0] = a0.x;
        s.bsum[3L * b0 + 1] = a0.y;
        s.bsum[3L * b0 + 2] = a0.z;
        if constexpr (E) *e += e0;
        if (b1 < nb) {
            s.bsum[3L * b1] = a1.x;
            s.bsum[3L * b1 + 1] = a1.y;
            s.bsum[3L * b1 + 2] = a1.z;
            if constexpr (E) *e += e1;
        }
    }
}
// the energies of the filament terms each bead starts at x, added to *e (two beads per step)
template <typename R, int NT>
__device__ inline void blk_bonded_energy(const DevSys& s, const double* __restrict__ x, double* e) {
    const int nb = s.n_beads;
    for (int b0 = threadIdx.x; b0 < nb; b0 += 2 * NT) {
        const int b1 = b0 + NT;
        const double e0 = bead_bonded_energy<R>(s, x, b0);
        const double e1 = bead_bonded_energy<R>(s, x, min(b1, nb - 1));
        *e += e0;
        if (b1 < nb) *e += e1;
    }
}
__device__ inline bool is_bonded(int ff) { return ff == FF_STRETCH || ff == FF_BEND; }

// force[b] = the bead's bonded sum + its slots (the order of gather_block), with the partial
// dot products F.F, F.F_prev (fp may be null) and max|F| of this thread's beads, and the
// minimum-energy backup candidate xmin_dst = xn when cand.
template <typename R, int NT>
__device__ inline void blk_gather(const DevSys& s, const R* __restrict__ slots, double* __restrict__ f,
                                  const double* __restrict__ fp, bool cand, const double* __restrict__ xn,
                                  double* __restrict__ xmin_dst, double* v) {
    const int nb = s.n_beads;
    for (int b0 = threadIdx.x; b0 < nb; b0 += 2 * NT) {
        // two beads per step (b0, b0 + NT), every load of both before the first use; each
        // thread still adds its beads in order
        int bb[2] = {b0, min(b0 + NT, nb - 1)};
        int q0[2], q1[2];
        D3 acc[2];
        double fpv[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int b = bb[u];
            q0[u] = s.bstart[b];
            q1[u] = s.bstart[b + 1];
            acc[u] = D3{s.bsum[3L * b], s.bsum[3L * b + 1], s.bsum[3L * b + 2]};
            if (fp)
                for (int c = 0; c < 3; c++) fpv[u][c] = fp[3 * b + c];
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            if (u == 1 && b0 + NT >= nb) break;
            const int b = bb[u];
            D3 a = acc[u];
            for (int q = q0[u]; q < q1[u]; q++) {
                const R* w = slots + 3L * q;
                a.x += (double)w[0]; a.y += (double)w[1]; a.z += (double)w[2];
            }
            const double fv[3] = {a.x, a.y, a.z};
            for (int c = 0; c < 3; c++) {
                f[3 * b + c] = fv[c];
                v[0] += fv[c] * fv[c];
                if (fp) v[1] += fv[c] * fpv[u][c];
                v[2] = rmax(v[2], fabs(fv[c]));
                if (cand) xmin_dst[3 * b + c] = xn[3 * b + c];
            }
        }
    }
}

// The line search: k_cg_pos and k_cg_ls (the bead work; the trial energy at lambda with
// speculative forces, at lambda/2 too after a rejection, E(x) when needed) and the first half of
// k_cg_fin (reductions, Simpson energies, ls_control).
template <typename R, int NT>
__device__ inline void blk_ls(const DevSys& s, const CgBufs& B, CgState& S, double* red, int& sh_bad) {
    const bool first = S.ls_active == 0;
    const bool pend = first && S.dir_pending;
    const double lam0 = S.ls_lambda, beta = S.dir_beta;
    const bool need_e0 = first && !S.ecur_valid;
    const int par = S.parity, dpar = S.dpar;
    const double* __restrict__ xc = B.x[par];
    const double* __restrict__ dc = B.d[dpar];
    R* slots = static_cast<R*>(B.slots);
    const double* __restrict__ fcur = B.f[S.fpar];
    const double lamh = lam0 * 0.5;
    double acc[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, INFINITY};
    // the bead pass first (k_cg_pos): the trial coordinates x + lambda0 (F + beta d) (the
    // speculative x_next) and, after a rejection, x + (lambda0 / 2) d, written once per bead with
    // the expression pos_of forms on the fly, so the interactions below read one coordinate
    // vector per bead instead of three
    const double maxd = S.max_dist;
    double* __restrict__ xn = B.x[par ^ 1];
    double* __restrict__ dn = B.d[dpar ^ 1];
    // BU values per thread per step, every load issued before the first use (one workgroup
    // streams the whole system: latency-bound with one load in flight per thread); each thread
    // still visits i = tid, tid + NT, ... in order, so the d.F partial sums are unchanged
    const int n3 = 3 * s.n_beads;
    for (int i0 = threadIdx.x; i0 < n3; i0 += BU * NT) {
        double fv[BU], dv[BU], xv[BU];
#pragma unroll
        for (int u = 0; u < BU; u++) {
            const int i = i0 + u * NT;
            fv[u] = i < n3 ? fcur[i] : 0.0;
            dv[u] = i < n3 ? dc[i] : 0.0;
            xv[u] = i < n3 ? xc[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < BU; u++) {
            const int i = i0 + u * NT;
            if (i >= n3) break;
            const double de = pend ? fv[u] + beta * dv[u] : dv[u];
            xn[i] = xv[u] + lam0 * de;   // speculative x_next
            if (!first) B.xh[i] = xv[u] + lamh * de;
            if (first) {
                acc[6] += de * fv[u];                         // d.F
                acc[7] = rmin(acc[7], maxd / fabs(de));       // MController.h:293-305
                if (pend) {
                    dn[i] = de;                               // shiftSearchDir materialized
                    B.fp[i] = fv[u];                          // forcePrev = force
                }
            }
        }
    }
    __syncthreads();
    const Pos p1{xn, nullptr, nullptr, 0.0, 0.0, 0};
    const Pos p2{B.xh, nullptr, nullptr, 0.0, 0.0, 0};
    const Pos pz{xc, nullptr, nullptr, 0.0, 0.0, 0};
    int fbad = 0;
    // one pass per evaluation keeps a single interaction's working set live at a time; an
    // excluded-volume pair's energy and forces by two calls (the whole minimizer is one kernel,
    // whose register budget the shared form overflows; measured faster in the mixed path,
    // unchanged in FP64)
    for (int ff = 0; ff < NFF; ff++)
        if (!is_bonded(ff))
            for (int t = threadIdx.x; t < s.n[ff]; t += NT) acc[0] += term_ef<R, true>(s, ff, t, p1, slots, 0, &fbad);
    blk_bonded<R, NT, true>(s, xn, &acc[0]);
    if (fbad) atomicOr(&sh_bad, 1);
    if (!first) {   // the first launch of a search evaluates lambda0 alone (see ls_control)
        for (int ff = 0; ff < NFF; ff++)
            if (!is_bonded(ff))
                for (int t = threadIdx.x; t < s.n[ff]; t += NT) {
                    int fb2 = 0;
                    acc[1] += term_ef<R>(s, ff, t, p2, (R*)nullptr, 1, &fb2);
                }
        blk_bonded_energy<R, NT>(s, B.xh, &acc[1]);
    }
    if (need_e0) {
        for (int ff = 0; ff < NFF; ff++)
            if (!is_bonded(ff))
                for (int t = threadIdx.x; t < s.n[ff]; t += NT) {
                    int fb2 = 0;
                    acc[4] += term_ef<R>(s, ff, t, pz, (R*)nullptr, 2, &fb2);
                }
        blk_bonded_energy<R, NT>(s, xc, &acc[4]);
    }
    const int op8[8] = {0, 0, 1, 1, 0, 1, 0, 2};
    block_reduce_n<8, NT>(acc, op8, red);
    double dsum[3];
    int dbad[3];
    const int n = run_defers<NT>(s, red, dsum, dbad, 3);
    if (threadIdx.x == 0) {
        acc[0] += dsum[0]; acc[2] = dbad[0];
        acc[1] += dsum[1]; acc[3] = dbad[1];
        acc[4] += dsum[2]; acc[5] = dbad[2];
        if (n > DEFER_CAP) {
            S.status = 9;
            S.done = 1;
        } else {
            CgState Sr = S;   // the scalar control on a register copy (LDS latency per field access)
            ls_control(Sr, acc, sh_bad);
            S = Sr;
        }
    }
    __syncthreads();
}

// The move (k_cg_pos and k_cg_ls in phase PH_MOVE): forces and x_next at the accepted step (or the
// recovery step).
template <typename R, int NT>
__device__ inline void blk_move(const DevSys& s, const CgBufs& B, const CgState& S, int& sh_bad) {
    const double lam = S.lambda;
    const double* __restrict__ x = B.x[S.parity];
    const double* __restrict__ d = B.d[S.dpar];
    double* __restrict__ xn = B.x[S.parity ^ 1];
    // moveAlongSearchDir (CGMethod.cpp:350-354) / the recovery move (MController.h:318-321),
    // first, so that the interactions read the new coordinates
    const int n3 = 3 * s.n_beads;
    for (int i0 = threadIdx.x; i0 < n3; i0 += BU * NT) {
        double xv[BU], dv[BU];
#pragma unroll
        for (int u = 0; u < BU; u++) {
            const int i = i0 + u * NT;
            xv[u] = i < n3 ? x[i] : 0.0;
            dv[u] = i < n3 ? d[i] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < BU; u++)
            if (i0 + u * NT < n3) xn[i0 + u * NT] = xv[u] + lam * dv[u];
    }
    __syncthreads();
    const Pos p{xn, nullptr, nullptr, 0.0, 0.0, 0};
    R* slots = static_cast<R*>(B.slots);
    int fbad = 0;
    for (int ff = 0; ff < NFF; ff++)
        if (!is_bonded(ff))
            for (int t = threadIdx.x; t < s.n[ff]; t += NT) (void)term_ef<R>(s, ff, t, p, slots, -1, &fbad);
    blk_bonded<R, NT, false>(s, xn, nullptr);
    if (fbad) atomicOr(&sh_bad, 1);
    __syncthreads();
}

// The gather (k_cg_gather and the second half of k_cg_fin) into the next force buffer.
template <typename R, int NT>
__device__ inline void blk_gather_step(const DevSys& s, const CgBufs& B, CgState& S, double* red, int& sh_bad) {
    // copycoordsifminimumE candidate (CGMethod.hpp:1564-1566)
    const bool cand = S.num_iter > S.force_call_limit / 2 && S.total_energy <= S.minimum_e;
    const int par = S.parity, bk = S.bkpar;
    double v[3] = {0.0, 0.0, 0.0};
    blk_gather<R, NT>(s, static_cast<const R*>(B.slots), B.f[S.fpar ^ 1], B.fp, cand, B.x[par ^ 1], B.xmin[bk ^ 1], v);
    const int op3[3] = {0, 0, 1};
    block_reduce_n<3, NT>(v, op3, red);
    if (threadIdx.x == 0) {
        CgState Sr = S;
        gather_control(Sr, v, sh_bad);
        S = Sr;
    }
    __syncthreads();
}

// CGMethod::minimize of one system by one workgroup: the set-up of mdn_minimize
// (CGMethod.hpp:876-896: energies, forces, maxF; resetCGParams applied by the first line
// search), the CG slots until done, the minimum-energy restore on a call limit
// (CGMethod.hpp:1579-1590) and the final energy report.
template <typename R, int NT>
__device__ inline void blk_minimize(const BatchItem& it, const BatchParams& P, CgState& S, double* red,
                                    int& sh_bad) {
    const DevSys& s = it.s;
    const CgBufs& B = it.B;
    const int nb = s.n_beads;
    __shared__ double e_start;
    // energies before: the FP64 report; the minimizer's own start energy in R arithmetic
    blk_report<double, NT>(s, B.x[0], it.report, red);
    if constexpr (sizeof(R) == sizeof(double)) {
        if (threadIdx.x == 0) e_start = it.report[6];
    } else {
        blk_report<R, NT>(s, B.x[0], it.report + 18, red);   // scratch report
        if (threadIdx.x == 0) e_start = it.report[18 + 6];
    }
    // forces at the start, F.F and max|F|; searchDir.assign(numDof, 0)
    if (threadIdx.x == 0) sh_bad = 0;
    __syncthreads();
    {
        const Pos p0{B.x[0], nullptr, nullptr, 0.0, 0.0, 0};
        R* slots = static_cast<R*>(B.slots);
        int fbad = 0;
        for (int ff = 0; ff < NFF; ff++)
            if (!is_bonded(ff))
                for (int t = threadIdx.x; t < s.n[ff]; t += NT) (void)term_ef<R>(s, ff, t, p0, slots, -1, &fbad);
        blk_bonded<R, NT, false>(s, B.x[0], nullptr);
        if (fbad) atomicOr(&sh_bad, 1);
        for (int i = threadIdx.x; i < 3 * nb; i += NT) B.d[0][i] = 0.0;
    }
    __syncthreads();
    double v[3] = {0.0, 0.0, 0.0};
    blk_gather<R, NT>(s, static_cast<const R*>(B.slots), B.f[0], nullptr, false, nullptr, nullptr, v);   // fpar = 0
    const int op3[3] = {0, 0, 1};
    block_reduce_n<3, NT>(v, op3, red);
    if (threadIdx.x == 0) {
        S = CgState{};
        S.method = P.method;
        S.ls_algo = P.line_search;
        S.try_recover = P.try_recover;
        S.force_call_limit = P.limit_override > 0 ? P.limit_override : max(P.force_call_limit, 3 * nb);
        S.energy_call_limit = P.energy_call_limit;
        S.max_dist = P.max_dist;
        S.lambda_max = P.lambda_max;
        S.grad_tol = P.grad_tol;
        S.energy_rel_tol = P.energy_rel_tol;
        S.num_energy_call = 1;
        S.num_force_call = 1;
        S.grad_inf_norm = v[2];
        S.energy_rel_change = INFINITY;
        S.prev_energy = e_start;
        S.minimum_e = INFINITY;
        S.fdotf = v[0];
        S.dir_pending = 1;
        S.dir_beta = 0.0;
        S.dir_reset = 1;
        S.phase = PH_LS;
        S.ls_lambda = rmin(S.lambda_max, S.max_dist / S.grad_inf_norm);
        if (sh_bad) {
            // an excluded-volume pair needs the numerical force fallback, which this kernel does
            // not run (reported as MDN_ERR_UNSUPPORTED, like a pause for it later in the run)
            S.status = ST_NEED_FORCEN;
            S.done = 1;
        } else if (!isfinite(v[2])) {
            S.status = 4;   // MDN_ERR_NONFINITE: maxF not finite (CGMethod.hpp:202-212)
            S.done = 1;
        } else {
            const bool limit0 = S.force_call_limit != 0 && S.num_force_call > S.force_call_limit;
            S.done = (S.grad_inf_norm < S.grad_tol || S.energy_rel_change < S.energy_rel_tol || limit0) ? 1 : 0;
        }
        sh_bad = 0;
    }
    __syncthreads();
    // CG slots; a step costs at most one launch per two halvings from lambda0 to the tolerance,
    // the parabola fits of the quadratic search and a move, so the bound below is never reached
    // by a correct state machine (it guards the exit)
    const long max_slots = slots_per_force_call(S) * ((long)S.force_call_limit + 16) + 128;
    for (long k = 0; k < max_slots && !S.done; k++) {
        if (S.phase == PH_LS) blk_ls<R, NT>(s, B, S, red, sh_bad);
        if (!S.done && S.phase == PH_MOVE) blk_move<R, NT>(s, B, S, sh_bad);
        if (!S.done && (S.phase == PH_MOVE || S.phase == PH_GATHER)) blk_gather_step<R, NT>(s, B, S, red, sh_bad);
    }
    if (threadIdx.x == 0) {
        if (!S.done) {
            S.status = 3;   // MDN_ERR_STATE: did not terminate
            S.done = 1;
        }
        const bool limit = call_limit_s(S);
        S.reserved[0] = (S.status == 0 && S.error_flags == 0 && P.report_on_call_limit && limit && S.have_min_e) ? 1 : 0;
    }
    __syncthreads();
    if (S.reserved[0]) {   // copybackupcoordinates
        double* xr = B.x[S.parity];
        const double* xb = B.xmin[S.bkpar];
        for (int i = threadIdx.x; i < 3 * nb; i += NT) xr[i] = xb[i];
        __syncthreads();
    }
    if (S.status == 0 && S.error_flags == 0) blk_report<double, NT>(s, B.x[S.parity], it.report + 9, red);
    if (threadIdx.x == 0) *it.st = S;
    __syncthreads();   // S is reused by the next system
}

// Persistent kernel: workgroups pull systems order[k], k = 0, 1, ... from the queue counter until
// it passes n (every workgroup then returns).
template <typename R, int NT>
__global__ void __launch_bounds__(NT) k_cg_batched(const BatchItem* __restrict__ items, const int* __restrict__ order,
                                                   int n, int* __restrict__ queue, BatchParams P) {
    __shared__ CgState S;
    __shared__ double red[8 * (NT / 64)];
    __shared__ int sh_bad, sh_k;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) sh_k = atomicAdd(queue, 1);
        __syncthreads();
        const int k = sh_k;
        if (k >= n) return;
        blk_minimize<R, NT>(items[order[k]], P, S, red, sh_bad);
    }
}

// The final CG state and the energy report (27 values) of every system of a batch, gathered into
// contiguous buffers for one read-back: states[i], reports[27 i ...]. Grid-stride over
// (system, word).
__global__ void __launch_bounds__(256) k_batch_collect(const BatchItem* __restrict__ items, int n,
                                                       CgState* __restrict__ states, double* __restrict__ reports) {
    constexpr int NW = sizeof(CgState) / sizeof(int);
    static_assert(sizeof(CgState) % sizeof(int) == 0, "CgState in whole words");
    const long total = (long)n * (NW + 27);
    for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += (long)gridDim.x * blockDim.x) {
        const int i = (int)(k / (NW + 27)), w = (int)(k % (NW + 27));
        if (w < NW) reinterpret_cast<int*>(states + i)[w] = reinterpret_cast<const int*>(items[i].st)[w];
        else reports[27L * i + (w - NW)] = items[i].report[w - NW];
    }
}

// One launch for the start copies of a whole batch (each system's device coordinates into its
// x[0], the start's backup for the schedule tuning and its restores): copy i moves c[i].n doubles
// from c[i].src to c[i].dst. Grid (blocks per copy, systems), both strided.
struct BatchCopy {
    const double* src;
    double* dst;
    long n;
};
__global__ void __launch_bounds__(256) k_batch_copy(const BatchCopy* __restrict__ c, int n) {
    for (int i = blockIdx.y; i < n; i += gridDim.y) {
        const BatchCopy e = c[i];
        for (long k = (long)blockIdx.x * blockDim.x + threadIdx.x; k < e.n; k += (long)gridDim.x * blockDim.x)
            e.dst[k] = e.src[k];
    }
}

__global__ void k_fill_int(int* p, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

}  // namespace mdn
