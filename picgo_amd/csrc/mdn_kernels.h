// mdn_kernels.h -- gfx950 kernels of the mechanical-equilibration path (FP64 path).
//
// Layout in HBM (all device buffers owned by the context):
//   x[3*n_beads], F, F_prev, d (search direction), x_min: the reference's flat AoS coord
//   vectors (DofSerializer.hpp:64-187), double.
//   Term tables, structure of arrays per force field (bead indices int32, parameters double).
//   slots[S]: one D3 per (interaction, bead-of-interaction); slot index order = force-field
//   order x term order x bead order, i.e. the reference's += order into `force`.
//   Incidence CSR per bead (static part: stretching/bending/linker/motor, built at upload;
//   dynamic part: excluded volume/boundary, built after each neighbor-list rebuild), each bead's
//   slots ascending, so an owner-computes gather reproduces the reference's per-bead
//   summation order exactly and needs no atomics.
//
// Grid layout of the term / bead kernels: blocks [blk[ff], blk[ff+1]) hold force field ff's
// terms (each FF range padded to whole blocks, so a block belongs to one FF and per-FF energies
// come out of block partials without mixing), blocks [blk[6], blk[7]) hold the beads.
//
// Reductions are deterministic: block tree reduction, block partials in fixed slots, and the
// last block to finish (ticket counter + device-scope fences) reduces the partials in block
// order and runs the scalar control logic of CGMethod::minimize.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdn_terms.h"

namespace mdn {

constexpr int BLOCK = 256;
constexpr int NFF = 6;
constexpr int MAXK = 8;          // line-search trials evaluated per launch
constexpr int MAXELEM = 8;       // boundary elements held in the kernel argument

enum { FF_STRETCH = 0, FF_BEND, FF_LINKER, FF_MOTOR, FF_EXVOL, FF_BOUNDARY };

struct DevSys {
    int n_beads;
    int n[NFF];                 // interactions per force field
    int blk[NFF + 2];           // first block of each FF range, beads, end
    long slot0[NFF + 1];        // first slot of each force field, total
    // stretching
    const int* str_b;  const double* str_k; const double* str_eql;
    // bending
    const int* bend_b; const double* bend_k; const double* bend_eqt;
    // linkers / motors
    const int* lnk_b; const double* lnk_k; const double* lnk_eql; const double* lnk_p1; const double* lnk_p2;
    double* lnk_sf;
    const int* mot_b; const double* mot_k; const double* mot_eql; const double* mot_p1; const double* mot_p2;
    double* mot_sf;
    // excluded volume: beads (4 per pair) and kRepScaled = krep * L0_i * L0_j
    const int* ev_b;   const double* ev_ks;
    // boundary: bead, element, krep, slen
    const int* bnd_b;  const int* bnd_e; const double* bnd_krep; const double* bnd_slen;
    int n_elem;
    BElem elem[MAXELEM];
    // incidence CSR (static + dynamic), slot indices
    const int* inc_s_start; const int* inc_s;
    const int* inc_d_start; const int* inc_d;
};

// ---------------------------------------------------------------------------------------
// CG device state (CGMethod::minimize locals + MinimizationResult)
enum { PH_LS = 0, PH_MOVE = 1, PH_GATHERED = 2, PH_RESET = 3, PH_DONE = 4 };

struct CgState {
    // parameters
    int method, try_recover, force_call_limit, energy_call_limit;
    double max_dist, lambda_max, grad_tol, energy_rel_tol;
    // control
    int phase, parity, done, status;
    int num_iter, num_energy_call, num_force_call, num_safe_mode, num_recovery;
    unsigned error_flags;
    int safe_mode, recovery_move;
    double grad_inf_norm, energy_rel_change, prev_energy, last_fdotf, fdotf, beta;
    // line search
    int ls_active, ls_K, ecur_valid, ls_trials;
    double ls_lambda, ls_cur, ls_slope, ls_reduced, lambda, new_energy, rec_step;
    // minimum-energy backup (CGMethod.hpp:794-814)
    double total_energy, minimum_e;
    int have_min_e, do_backup;
    // ticket counters of the last-block reductions
    unsigned ticket[8];
};

__device__ inline int ff_of_block(const DevSys& s, int b) {
    int ff = 0;
#pragma unroll
    for (int k = 1; k <= NFF; k++) ff += (b >= s.blk[k]) ? 1 : 0;
    return ff;   // NFF means a bead block
}

// ---- deterministic block reduction helpers (64-wide waves) ----------------------------
__device__ inline double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}
__device__ inline double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = rmax(v, __shfl_down(v, o, 64));
    return v;
}
__device__ inline double wave_min(double v) {
    for (int o = 32; o > 0; o >>= 1) v = rmin(v, __shfl_down(v, o, 64));
    return v;
}
// op: 0 sum, 1 max, 2 min. Result valid in thread 0.
template <int OP>
__device__ inline double block_reduce(double v, double* lds) {
    v = OP == 0 ? wave_sum(v) : OP == 1 ? wave_max(v) : wave_min(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) lds[w] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        r = lds[0];
        for (int i = 1; i < BLOCK / 64; i++)
            r = OP == 0 ? r + lds[i] : OP == 1 ? rmax(r, lds[i]) : rmin(r, lds[i]);
    }
    return r;
}

// Last-block election. Every block must call it after writing its partials.
__device__ inline bool last_block(unsigned* ticket) {
    __shared__ int am_last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = atomicAdd(ticket, 1u);
        am_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (am_last) {
        __threadfence();
        if (threadIdx.x == 0) *ticket = 0u;
    }
    return am_last != 0;
}

// coherent (L2-bypassing) load of a partial written by another block
__device__ inline double ld_partial(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline int ld_partial_i(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------------------
// Energy of term t of force field ff at positions x + lambda*d. *bad is set for an excluded
// volume pair that stays non-finite (the reference's -1 sentinel).
__device__ inline double term_energy(const DevSys& s, int ff, int t, const double* __restrict__ x,
                                     const double* __restrict__ d, double lam, int* bad) {
    switch (ff) {
    case FF_STRETCH: {
        D3 c1 = ld3_at(x, d, lam, s.str_b[2 * t]), c2 = ld3_at(x, d, lam, s.str_b[2 * t + 1]);
        return stretch_energy(c1, c2, s.str_k[t], s.str_eql[t]);
    }
    case FF_BEND: {
        D3 c1 = ld3_at(x, d, lam, s.bend_b[3 * t]), c2 = ld3_at(x, d, lam, s.bend_b[3 * t + 1]),
           c3 = ld3_at(x, d, lam, s.bend_b[3 * t + 2]);
        return bend_energy(c1, c2, c3, s.bend_k[t], s.bend_eqt[t]);
    }
    case FF_LINKER:
    case FF_MOTOR: {
        const int* b = ff == FF_LINKER ? s.lnk_b : s.mot_b;
        D3 c1 = ld3_at(x, d, lam, b[4 * t]), c2 = ld3_at(x, d, lam, b[4 * t + 1]),
           c3 = ld3_at(x, d, lam, b[4 * t + 2]), c4 = ld3_at(x, d, lam, b[4 * t + 3]);
        if (ff == FF_LINKER)
            return xlink_energy(c1, c2, c3, c4, s.lnk_k[t], s.lnk_eql[t], s.lnk_p1[t], s.lnk_p2[t]);
        return xlink_energy(c1, c2, c3, c4, s.mot_k[t], s.mot_eql[t], s.mot_p1[t], s.mot_p2[t]);
    }
    case FF_EXVOL: {
        D3 c1 = ld3_at(x, d, lam, s.ev_b[4 * t]), c2 = ld3_at(x, d, lam, s.ev_b[4 * t + 1]),
           c3 = ld3_at(x, d, lam, s.ev_b[4 * t + 2]), c4 = ld3_at(x, d, lam, s.ev_b[4 * t + 3]);
        return exvol_energy(c1, c2, c3, c4, s.ev_ks[t], bad);
    }
    default: {
        D3 c = ld3_at(x, d, lam, s.bnd_b[t]);
        return boundary_energy(s.elem[s.bnd_e[t]], c, s.bnd_krep[t], s.bnd_slen[t]);
    }
    }
}

// Forces of term t into its slots. Returns false for an excluded-volume pair whose analytic
// forces are non-finite (the reference then integrates numerically, forceN).
__device__ inline bool term_forces(const DevSys& s, int ff, int t, const double* __restrict__ x,
                                   const double* __restrict__ d, double lam, D3* __restrict__ slots) {
    D3 out[4];
    switch (ff) {
    case FF_STRETCH: {
        D3 c1 = ld3_at(x, d, lam, s.str_b[2 * t]), c2 = ld3_at(x, d, lam, s.str_b[2 * t + 1]);
        stretch_forces(c1, c2, s.str_k[t], s.str_eql[t], out);
        D3* o = slots + s.slot0[FF_STRETCH] + 2 * (long)t;
        o[0] = out[0]; o[1] = out[1];
        return true;
    }
    case FF_BEND: {
        D3 c1 = ld3_at(x, d, lam, s.bend_b[3 * t]), c2 = ld3_at(x, d, lam, s.bend_b[3 * t + 1]),
           c3 = ld3_at(x, d, lam, s.bend_b[3 * t + 2]);
        bend_forces(c1, c2, c3, s.bend_k[t], s.bend_eqt[t], out);
        D3* o = slots + s.slot0[FF_BEND] + 3 * (long)t;
        o[0] = out[0]; o[1] = out[1]; o[2] = out[2];
        return true;
    }
    case FF_LINKER:
    case FF_MOTOR: {
        const int* b = ff == FF_LINKER ? s.lnk_b : s.mot_b;
        D3 c1 = ld3_at(x, d, lam, b[4 * t]), c2 = ld3_at(x, d, lam, b[4 * t + 1]),
           c3 = ld3_at(x, d, lam, b[4 * t + 2]), c4 = ld3_at(x, d, lam, b[4 * t + 3]);
        double sf;
        if (ff == FF_LINKER) {
            sf = xlink_forces(c1, c2, c3, c4, s.lnk_k[t], s.lnk_eql[t], s.lnk_p1[t], s.lnk_p2[t], out);
            s.lnk_sf[t] = sf;
        } else {
            sf = xlink_forces(c1, c2, c3, c4, s.mot_k[t], s.mot_eql[t], s.mot_p1[t], s.mot_p2[t], out);
            s.mot_sf[t] = sf;
        }
        D3* o = slots + s.slot0[ff] + 4 * (long)t;
        o[0] = out[0]; o[1] = out[1]; o[2] = out[2]; o[3] = out[3];
        return true;
    }
    case FF_EXVOL: {
        D3 c1 = ld3_at(x, d, lam, s.ev_b[4 * t]), c2 = ld3_at(x, d, lam, s.ev_b[4 * t + 1]),
           c3 = ld3_at(x, d, lam, s.ev_b[4 * t + 2]), c4 = ld3_at(x, d, lam, s.ev_b[4 * t + 3]);
        bool ok = exvol_forces(c1, c2, c3, c4, s.ev_ks[t], out);
        D3* o = slots + s.slot0[FF_EXVOL] + 4 * (long)t;
        if (!ok) out[0] = out[1] = out[2] = out[3] = D3{0.0, 0.0, 0.0};
        o[0] = out[0]; o[1] = out[1]; o[2] = out[2]; o[3] = out[3];
        return ok;
    }
    default: {
        D3 c = ld3_at(x, d, lam, s.bnd_b[t]);
        slots[s.slot0[FF_BOUNDARY] + t] = boundary_force(s.elem[s.bnd_e[t]], c, s.bnd_krep[t], s.bnd_slen[t]);
        return true;
    }
    }
}

// Sum of bead b's slots with slot index in [lo, hi), in ascending slot order, starting from
// `acc` (the reference's `+=` into force).
__device__ inline D3 gather_bead(const DevSys& s, int b, const D3* __restrict__ slots, long lo, long hi, D3 acc) {
    for (int q = s.inc_s_start[b]; q < s.inc_s_start[b + 1]; q++) {
        long sl = s.inc_s[q];
        if (sl < lo || sl >= hi) continue;
        D3 v = slots[sl];
        acc.x += v.x; acc.y += v.y; acc.z += v.z;
    }
    for (int q = s.inc_d_start[b]; q < s.inc_d_start[b + 1]; q++) {
        long sl = s.inc_d[q];
        if (sl < lo || sl >= hi) continue;
        D3 v = slots[sl];
        acc.x += v.x; acc.y += v.y; acc.z += v.z;
    }
    return acc;
}

// =======================================================================================
// Plain (non-CG) kernels
// =======================================================================================

// Slots of every term at x (forces of ForceFieldManager::computeForces before the gather).
// bad_count / bad_list collect excluded-volume pairs that need the numerical fallback.
__global__ void __launch_bounds__(BLOCK) k_forces_terms(DevSys s, const double* __restrict__ x,
                                                        D3* __restrict__ slots, int* bad_count, int* bad_list) {
    const int ff = ff_of_block(s, blockIdx.x);
    if (ff >= NFF) return;
    const int t = (blockIdx.x - s.blk[ff]) * BLOCK + threadIdx.x;
    if (t >= s.n[ff]) return;
    if (!term_forces(s, ff, t, x, nullptr, 0.0, slots)) {
        int k = atomicAdd(bad_count, 1);
        bad_list[k] = t;
    }
}

// force[b] = base[b] + sum of b's slots within [lo, hi) (base may be null: zero-fill).
__global__ void __launch_bounds__(BLOCK) k_gather_plain(DevSys s, const D3* __restrict__ slots, long lo, long hi,
                                                        const double* __restrict__ base, double* __restrict__ f) {
    const int b = blockIdx.x * BLOCK + threadIdx.x;   // grid: bead blocks only
    if (b >= s.n_beads) return;
    D3 acc = base ? ld3(base, b) : D3{0.0, 0.0, 0.0};
    acc = gather_bead(s, b, slots, lo, hi, acc);
    f[3 * b] = acc.x; f[3 * b + 1] = acc.y; f[3 * b + 2] = acc.z;
}

// Per-FF energies at x: out[0..5] per force field (exvol = -1 on the sentinel), out[6] =
// computeIndividualEnergies total (plain sum), out[7] = computeEnergy total (+inf on sentinel).
__global__ void __launch_bounds__(BLOCK) k_energy_report(DevSys s, const double* __restrict__ x,
                                                         double* __restrict__ partial, int* __restrict__ partial_bad,
                                                         unsigned* ticket, double* __restrict__ out) {
    __shared__ double lds[BLOCK / 64];
    const int ff = ff_of_block(s, blockIdx.x);
    double e = 0.0;
    int bad = 0;
    if (ff < NFF) {
        const int t = (blockIdx.x - s.blk[ff]) * BLOCK + threadIdx.x;
        if (t < s.n[ff]) e = term_energy(s, ff, t, x, nullptr, 0.0, &bad);
    }
    double bsum = block_reduce<0>(e, lds);
    double bbad = block_reduce<1>((double)bad, lds);
    if (threadIdx.x == 0) {
        partial[blockIdx.x] = bsum;
        partial_bad[blockIdx.x] = bbad > 0.0;
    }
    if (!last_block(ticket)) return;
    if (threadIdx.x == 0) {
        double tot = 0.0, tot_s = 0.0;
        bool failed = false;
        for (int f2 = 0; f2 < NFF; f2++) {
            double ef = 0.0;
            int bd = 0;
            for (int b = s.blk[f2]; b < s.blk[f2 + 1]; b++) {
                ef += ld_partial(&partial[b]);
                bd |= ld_partial_i(&partial_bad[b]);
            }
            if (bd) ef = -1.0;
            out[f2] = ef;
            tot += ef;
            if (!failed) {
                if (ef <= -1) failed = true;
                else tot_s += ef;
            }
        }
        out[6] = tot;
        out[7] = failed ? INFINITY : tot_s;
    }
}

// Numerical excluded-volume force fallback (forceN, CylinderExclVolRepulsion.cpp:984-1254)
// for one pair: block b handles rows i = b, b + grid, ... of the 10^4 x 10^4 Simpson grid
// and writes 6 row-sum partials per block; k_exvol_forceN_finish adds them in block order.
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
    __shared__ double lds[BLOCK / 64];
    PairGeom g = exvol_pair_geom(s, x, t, movebeads);
    double a = scalar_product(g.c1, g.c2, g.c1, g.c2), b = scalar_product(g.c3, g.c4, g.c3, g.c4);
    double c = scalar_product(g.c3, g.c1, g.c3, g.c1), d = scalar_product(g.c1, g.c2, g.c3, g.c4);
    double e = scalar_product(g.c1, g.c2, g.c3, g.c1), F = scalar_product(g.c3, g.c4, g.c3, g.c1);
    const int N = 10000;
    const double delta = 1.0 / ((double)N);
    double acc[6] = {0, 0, 0, 0, 0, 0}, in[6];
    // T2 / T3 / T4 rows: row i has s = (i+1)*delta (the loop advances s before use)
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
    for (int k = 0; k < 6; k++) {
        double r = block_reduce<0>(acc[k], lds);
        if (threadIdx.x == 0) partial[6 * blockIdx.x + k] = r;
    }
}

__global__ void k_exvol_forceN_finish(DevSys s, const double* __restrict__ x, int t, int movebeads, int nblk,
                                      const double* __restrict__ partial, D3* __restrict__ slots, int* ok_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    PairGeom g = exvol_pair_geom(s, x, t, movebeads);
    double a = scalar_product(g.c1, g.c2, g.c1, g.c2), b = scalar_product(g.c3, g.c4, g.c3, g.c4);
    double c = scalar_product(g.c3, g.c1, g.c3, g.c1), d = scalar_product(g.c1, g.c2, g.c3, g.c4);
    double e = scalar_product(g.c1, g.c2, g.c3, g.c1), F = scalar_product(g.c3, g.c4, g.c3, g.c1);
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
        for (int q = 0; q < 4; q++) ok = ok && !bad_component(f[q][k]);
    }
    *ok_out = ok ? 1 : 0;
    if (ok) {
        D3* o = slots + s.slot0[FF_EXVOL] + 4 * (long)t;
        for (int q = 0; q < 4; q++) o[q] = D3{f[q][0], f[q][1], f[q][2]};
    }
}

// =======================================================================================
// CG kernels: one "slot" of the device-driven minimizer is
//     k_cg_ls -> k_cg_move -> k_cg_gather -> k_cg_dir
// Each kernel reads the device state first and returns at once when its phase is not due.
// =======================================================================================

struct CgBufs {
    double* x[2];       // current / next coordinates (parity)
    double* f;          // force
    double* fp;         // previous force
    double* d;          // search direction
    double* xmin;       // minimum-energy backup
    D3* slots;
    double* partial;    // [blocks * 16]
    int* partial_i;     // [blocks * 16]
    int* bad_count;     // excluded-volume pairs needing the numerical fallback in a CG move
};

// ---- line-search trials: E(x + lambda_k d) for k = 0..K (k = 0 only when E(x) unknown) ----
__global__ void __launch_bounds__(BLOCK) k_cg_ls(DevSys s, CgState* __restrict__ st, CgBufs B) {
    __shared__ double lds[BLOCK / 64];
    if (st->done || st->phase != PH_LS) return;
    const int K = st->ls_K;
    const bool first = st->ls_active == 0;
    const bool need_e0 = first && !st->ecur_valid;
    const double lam0 = st->ls_lambda;
    const double* __restrict__ x = B.x[st->parity];
    const int ff = ff_of_block(s, blockIdx.x);
    double e[MAXK + 1];
    int bad[MAXK + 1];
#pragma unroll
    for (int k = 0; k <= MAXK; k++) { e[k] = 0.0; bad[k] = 0; }
    double slope = 0.0, step = INFINITY;
    if (ff < NFF) {
        const int t = (blockIdx.x - s.blk[ff]) * BLOCK + threadIdx.x;
        if (t < s.n[ff]) {
            if (need_e0) e[0] = term_energy(s, ff, t, x, nullptr, 0.0, &bad[0]);
            double lam = lam0;
            for (int k = 1; k <= K; k++) {
                e[k] = term_energy(s, ff, t, x, B.d, lam, &bad[k]);
                lam *= 0.5;
            }
        }
    } else if (first) {
        const int b = (blockIdx.x - s.blk[NFF]) * BLOCK + threadIdx.x;
        if (b < s.n_beads) {
            for (int q = 0; q < 3; q++) {
                double dv = B.d[3 * b + q], fv = B.f[3 * b + q];
                slope += dv * fv;
                // moveAlongDirectionLimitingBeadMove_ step (MController.h:293-305)
                double v = st->max_dist / fabs(dv);
                step = rmin(step, v);
            }
        }
    }
    double* P = B.partial + (long)blockIdx.x * 16;
    int* PI = B.partial_i + (long)blockIdx.x * 16;
    for (int k = 0; k <= K; k++) {
        double r = block_reduce<0>(e[k], lds);
        double rb = block_reduce<1>((double)bad[k], lds);
        if (threadIdx.x == 0) { P[k] = r; PI[k] = rb > 0.0; }
    }
    if (first) {
        double r = block_reduce<0>(slope, lds);
        double m = block_reduce<2>(step, lds);
        if (threadIdx.x == 0) { P[MAXK + 1] = r; P[MAXK + 2] = m; }
    }
    if (!last_block(&st->ticket[0])) return;
    if (threadIdx.x != 0) return;

    // ---- final reductions in block order, per force field then total (computeEnergy) ----
    double E[MAXK + 1];
    for (int k = 0; k <= K; k++) {
        double tot = 0.0;
        bool failed = false;
        for (int f2 = 0; f2 < NFF && !failed; f2++) {
            double ef = 0.0;
            int bd = 0;
            for (int b = s.blk[f2]; b < s.blk[f2 + 1]; b++) {
                ef += ld_partial(&B.partial[(long)b * 16 + k]);
                bd |= ld_partial_i(&B.partial_i[(long)b * 16 + k]);
            }
            if (bd) ef = -1.0;
            if (ef <= -1) failed = true;
            else tot += ef;
        }
        E[k] = failed ? INFINITY : tot;
    }
    double lambda = lam0;
    if (first) {
        double slope_t = 0.0, step_t = INFINITY;
        if (s.blk[NFF + 1] > s.blk[NFF]) {
            for (int b = s.blk[NFF]; b < s.blk[NFF + 1]; b++) {
                slope_t += ld_partial(&B.partial[(long)b * 16 + MAXK + 1]);
                step_t = rmin(step_t, ld_partial(&B.partial[(long)b * 16 + MAXK + 2]));
            }
        }
        st->num_iter += 1;           // numIter++ at the top of the loop (CGMethod.hpp:1297)
        st->rec_step = step_t;
        st->ls_trials = 0;
        if (st->grad_inf_norm == 0.0) {
            // backtrackingLineSearch returns lambda = 0 at once when maxForce == 0 (:252-256)
        } else {
            st->ls_cur = need_e0 ? E[0] : st->ls_cur;
            st->num_energy_call += 1;
            st->ls_slope = slope_t;
            st->ls_reduced = (st->safe_mode ? 0.0 : 0.4) * slope_t;
            st->ls_active = 1;
        }
    }
    // ---- backtrackingLineSearch loop (CGMethod.hpp:297-361) over the evaluated trials ----
    bool conv = false;
    double energyLambda = 0.0;
    if (first && st->grad_inf_norm == 0.0) {
        conv = true;
        lambda = 0.0;
        energyLambda = 0.0;
    }
    for (int k = 1; k <= K && !conv; k++) {
        energyLambda = E[k];
        st->num_energy_call += 1;
        st->ls_trials += 1;
        const double ideal = -st->ls_reduced * lambda;
        if (lambda < 0 || ideal > 0) { conv = true; lambda = 0.0; }
        const double change = energyLambda - st->ls_cur;
        if (change <= ideal) conv = true;
        else lambda *= 0.5;
        if (lambda <= 0.0 || lambda <= 1e-8) {
            conv = true;
            lambda = 0.0;
            energyLambda = st->ls_cur;
        }
    }
    if (!conv) {                 // keep halving in the next slot
        st->ls_lambda = lambda;
        st->ls_K = MAXK;
        return;
    }
    const double rec_step = st->rec_step;
    st->total_energy = energyLambda;   // TotalEnergy = energyLambda (CGMethod.hpp:368)
    st->ls_active = 0;
    {
        int k2 = st->ls_trials + 1;
        st->ls_K = k2 < 2 ? 2 : (k2 > MAXK ? MAXK : k2);
    }
    if (lambda != 0.0) {
        st->lambda = lambda;
        st->new_energy = energyLambda;
        st->recovery_move = 0;
        st->phase = PH_MOVE;
        return;
    }
    // ---- line-search failure (CGMethod.hpp:1337-1367) ----
    if (st->try_recover) {
        if (st->safe_mode) {
            if (rec_step == INFINITY) {   // MController.h:312-316 throws
                st->status = 7;           // MDN_ERR_RECOVERY
                st->done = 1;
                return;
            }
            st->lambda = rec_step;
            st->recovery_move = 1;
            st->phase = PH_MOVE;
        } else {
            st->safe_mode = 1;
            st->num_safe_mode += 1;
            st->phase = PH_RESET;
        }
    } else {
        st->error_flags |= 1u;
        st->done = 1;
    }
    // loop condition after an iteration that did not move
    if (!st->done && st->phase == PH_RESET) {
        bool limit = (st->energy_call_limit != 0 && st->num_energy_call > st->energy_call_limit) ||
                     (st->force_call_limit != 0 && st->num_force_call > st->force_call_limit);
        bool convd = st->grad_inf_norm < st->grad_tol || st->energy_rel_change < st->energy_rel_tol;
        if (limit || convd) st->done = 1;
    }
}

// ---- move + term forces at x + lambda d; writes x_next ----
__global__ void __launch_bounds__(BLOCK) k_cg_move(DevSys s, CgState* __restrict__ st, CgBufs B) {
    if (st->done || st->phase != PH_MOVE) return;
    const double lam = st->lambda;
    const double* __restrict__ x = B.x[st->parity];
    const int ff = ff_of_block(s, blockIdx.x);
    if (ff < NFF) {
        const int t = (blockIdx.x - s.blk[ff]) * BLOCK + threadIdx.x;
        if (t < s.n[ff])
            if (!term_forces(s, ff, t, x, B.d, lam, B.slots)) atomicAdd(B.bad_count, 1);
    } else {
        const int b = (blockIdx.x - s.blk[NFF]) * BLOCK + threadIdx.x;
        if (b < s.n_beads) {
            double* xn = B.x[st->parity ^ 1];
            for (int q = 0; q < 3; q++) xn[3 * b + q] = x[3 * b + q] + lam * B.d[3 * b + q];
        }
    }
}

// ---- gather F; reductions F.F, F.F_prev, max|F|; or materialize a CG reset ----
__global__ void __launch_bounds__(BLOCK) k_cg_gather(DevSys s, CgState* __restrict__ st, CgBufs B) {
    __shared__ double lds[BLOCK / 64];
    if (st->done) return;
    const int phase = st->phase;
    if (phase != PH_MOVE && phase != PH_RESET) return;
    const int b = blockIdx.x * BLOCK + threadIdx.x;   // grid: bead blocks only
    if (phase == PH_RESET) {
        // resetCGParams (CGMethod.hpp:963-972): d = F (shiftSearchDir(0)), F_prev = F
        if (b < s.n_beads)
            for (int q = 0; q < 3; q++) {
                double fv = B.f[3 * b + q];
                B.d[3 * b + q] = fv + 0 * B.d[3 * b + q];
                B.fp[3 * b + q] = fv;
            }
        if (!last_block(&st->ticket[1])) return;
        if (threadIdx.x == 0) {
            st->last_fdotf = st->fdotf;
            st->phase = PH_LS;
            st->ls_lambda = rmin(st->lambda_max, st->max_dist / st->grad_inf_norm);
        }
        return;
    }
    double ff = 0.0, fpd = 0.0, mx = 0.0;
    if (b < s.n_beads) {
        D3 acc = gather_bead(s, b, B.slots, 0, s.slot0[NFF], D3{0.0, 0.0, 0.0});
        double fv[3] = {acc.x, acc.y, acc.z};
        for (int q = 0; q < 3; q++) {
            B.f[3 * b + q] = fv[q];
            ff += fv[q] * fv[q];
            fpd += fv[q] * B.fp[3 * b + q];
            mx = rmax(mx, fabs(fv[q]));
        }
    }
    double r0 = block_reduce<0>(ff, lds), r1 = block_reduce<0>(fpd, lds), r2 = block_reduce<1>(mx, lds);
    double* P = B.partial + (long)blockIdx.x * 16;
    if (threadIdx.x == 0) { P[0] = r0; P[1] = r1; P[2] = r2; }
    if (!last_block(&st->ticket[2])) return;
    if (threadIdx.x != 0) return;
    double FF = 0.0, FP = 0.0, MX = 0.0;
    for (int k = 0; k < (int)gridDim.x; k++) {
        FF += ld_partial(&B.partial[(long)k * 16 + 0]);
        FP += ld_partial(&B.partial[(long)k * 16 + 1]);
        MX = rmax(MX, ld_partial(&B.partial[(long)k * 16 + 2]));
    }
    st->parity ^= 1;
    st->num_force_call += 1;
    st->grad_inf_norm = MX;
    st->fdotf = FF;
    if (!isfinite(MX)) {        // maxF error (CGMethod.hpp:202-212)
        st->status = 4;
        st->done = 1;
        return;
    }
    if (*B.bad_count != 0) {    // numerical fallback needed: handled by the host
        st->status = 9;
        st->done = 1;
        return;
    }
    if (st->recovery_move) {
        st->num_recovery += 1;
        st->recovery_move = 0;
        st->safe_mode = 0;
        st->ecur_valid = 0;
        st->phase = PH_RESET;
        bool limit = (st->energy_call_limit != 0 && st->num_energy_call > st->energy_call_limit) ||
                     (st->force_call_limit != 0 && st->num_force_call > st->force_call_limit);
        bool convd = st->grad_inf_norm < st->grad_tol || st->energy_rel_change < st->energy_rel_tol;
        if (limit || convd) st->done = 1;
        return;
    }
    st->energy_rel_change = fabs(st->new_energy - st->prev_energy) / fabs(st->prev_energy);
    st->prev_energy = st->new_energy;
    st->ls_cur = st->new_energy;      // E(x_new) = accepted trial energy, bit for bit
    st->ecur_valid = 1;
    double beta = 0.0;
    if (st->method == 1) beta = FF / st->last_fdotf;
    else if (st->method == 2) beta = rmax(0.0, (FF - FP) / st->last_fdotf);
    st->beta = beta;
    // copycoordsifminimumE condition (CGMethod.hpp:1564-1566)
    st->do_backup = (MX <= 10 * st->grad_tol && st->num_iter > st->force_call_limit / 2 &&
                     st->total_energy <= st->minimum_e) ? 1 : 0;
    if (st->do_backup) {
        st->minimum_e = st->total_energy;
        st->have_min_e = 1;
    }
    st->phase = PH_GATHERED;
}

// ---- d = F + beta d, F_prev = F, d.F check ----
__global__ void __launch_bounds__(BLOCK) k_cg_dir(DevSys s, CgState* __restrict__ st, CgBufs B) {
    __shared__ double lds[BLOCK / 64];
    if (st->done || st->phase != PH_GATHERED) return;
    const int b = blockIdx.x * BLOCK + threadIdx.x;   // grid: bead blocks only
    const double beta = st->beta;
    const bool pr = st->method == 2;
    double df = 0.0;
    if (b < s.n_beads) {
        const double* xc = B.x[st->parity];
        for (int q = 0; q < 3; q++) {
            double fv = B.f[3 * b + q];
            double dv = fv + beta * B.d[3 * b + q];
            B.d[3 * b + q] = dv;
            df += dv * fv;
            if (pr) B.fp[3 * b + q] = fv;
            if (st->do_backup) B.xmin[3 * b + q] = xc[3 * b + q];
        }
    }
    double r = block_reduce<0>(df, lds);
    if (threadIdx.x == 0) B.partial[(long)blockIdx.x * 16 + 3] = r;
    if (!last_block(&st->ticket[3])) return;
    if (threadIdx.x != 0) return;
    double DF = 0.0;
    for (int k = 0; k < (int)gridDim.x; k++) DF += ld_partial(&B.partial[(long)k * 16 + 3]);
    st->last_fdotf = st->fdotf;   // lastFDotF = newGrad (== F.F also after a reset)
    st->phase = (DF <= 0) ? PH_RESET : PH_LS;
    st->ls_lambda = rmin(st->lambda_max, st->max_dist / st->grad_inf_norm);
    bool limit = (st->energy_call_limit != 0 && st->num_energy_call > st->energy_call_limit) ||
                 (st->force_call_limit != 0 && st->num_force_call > st->force_call_limit);
    bool convd = st->grad_inf_norm < st->grad_tol || st->energy_rel_change < st->energy_rel_tol;
    if (limit || convd) st->done = 1;
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
};

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
    int bb = ix[0] + n.nb[0] * (ix[1] + n.nb[1] * ix[2]);
    bin[c] = bb;
    atomicAdd(&bin_count[bb], 1);
}

__global__ void k_nl_fill(int n_cyl, const int* __restrict__ bin, const int* __restrict__ bin_start,
                          int* __restrict__ cursor, int* __restrict__ members) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_cyl) return;
    int bb = bin[c];
    int k = atomicAdd(&cursor[bb], 1);
    members[bin_start[bb] + k] = c;
}

// HybridNeighborListImpl.cpp:411-439 predicate
__device__ inline bool nl_keep(const NLDev& n, const double* mid, int c, int cn) {
    if (c <= cn) return false;
    if (n.cyl_fil[c] == n.cyl_fil[cn]) {
        int sep = abs(n.cyl_pos[c] - n.cyl_pos[cn]);
        if (sep <= 2) return false;
    }
    double d2 = 0.0;
    for (int q = 0; q < 3; q++) {
        double v = mid[3 * cn + q] - mid[3 * c + q];
        d2 += v * v;
    }
    if (d2 < n.rminsq || d2 > n.rmaxsq) return false;
    return true;
}

// pass 0: count per cylinder; pass 1: emit into [off[c], off[c+1]) and sort ascending
template <int PASS>
__global__ void k_nl_pairs(NLDev n, const double* __restrict__ mid, const int* __restrict__ bin,
                           const int* __restrict__ bin_start, const int* __restrict__ members,
                           int* __restrict__ count, const int* __restrict__ off, int* __restrict__ pairs) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n.n_cyl) return;
    const int bb = bin[c];
    const int bx = bb % n.nb[0], by = (bb / n.nb[0]) % n.nb[1], bz = bb / (n.nb[0] * n.nb[1]);
    int cnt = 0;
    int base = PASS ? off[c] : 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                int X = bx + dx, Y = by + dy, Z = bz + dz;
                if (X < 0 || Y < 0 || Z < 0 || X >= n.nb[0] || Y >= n.nb[1] || Z >= n.nb[2]) continue;
                int nbin = X + n.nb[0] * (Y + n.nb[1] * Z);
                for (int m = bin_start[nbin]; m < bin_start[nbin + 1]; m++) {
                    int cn = members[m];
                    if (nl_keep(n, mid, c, cn)) {
                        if (PASS) pairs[2 * (long)(base + cnt) + 1] = cn;
                        cnt++;
                    }
                }
            }
    if (!PASS) {
        count[c] = cnt;
        return;
    }
    // insertion sort of the neighbour ids, then the owning cylinder id
    for (int i = 1; i < cnt; i++) {
        int v = pairs[2 * (long)(base + i) + 1];
        int j = i - 1;
        while (j >= 0 && pairs[2 * (long)(base + j) + 1] > v) {
            pairs[2 * (long)(base + j + 1) + 1] = pairs[2 * (long)(base + j) + 1];
            j--;
        }
        pairs[2 * (long)(base + j + 1) + 1] = v;
    }
    for (int i = 0; i < cnt; i++) pairs[2 * (long)(base + i)] = c;
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

// dynamic incidence CSR: per bead, the excluded-volume and boundary slots, ascending
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

}  // namespace mdn
