/*
 * oracle.h -- CPU FP64 restatement of MEDYAN 5.4.0's mechanical-equilibration path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the oracle the gfx950 library (libmdn_mech.so) is
 * tested against, and the single-core CPU baseline timed beside it. Only tests/, the smoke
 * entry point and the benchmark's cpu_baseline leg may load it; the product never does.
 *
 * It restates, from the mathematics and the cited reference sources read as text (the
 * reference itself cannot be compiled here: every translation unit needs spdlog / Eigen /
 * Boost, none of which exist in this image), the functions below with the reference's
 * expression order, IEEE double arithmetic and no FMA contraction (built with
 * -ffp-contract=off).
 *
 * Citations are relative to medyan-5.4.0/ in the reference tree.
 *
 * Pinning (what checks each function beyond retyping its formula):
 *   stretching / linker / motor : closed forms, Newton's third law, finite-difference
 *                                 energy-force consistency (TestFFCommon.hpp:106-134 recipe)
 *   bending                     : closed form k(1-cos) at 0/90/180 degrees, finite differences
 *   excluded volume             : the reference's three fixtures (TestCylinderExclVolume.cpp),
 *                                 brute-force Gauss-Legendre quadrature of the integral
 *   boundary                    : the reference fixture (TestBoundaryBeadRepulsion.cpp), closed form
 *   neighbor lists              : brute force O(N^2) vs cell list, identical pair sets
 *   CG                          : the reference harmonic-10D fixture (TestCGMethod.cpp), exact
 *                                 quadratic minima, forced line-search failures
 *   Parity unpinned (no reference vector exists and the reference cannot be built):
 *     exact rounding of the excluded-volume closed form on real networks, the coplanar-shift
 *     force branch, the Simpson fallbacks (energyN / forceN), and libm atan/exp ulps.
 */
#ifndef MDN_ORACLE_H
#define MDN_ORACLE_H

#include "../include/mdn_mech.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- interaction kernels, reference signatures (beadSet = flat offsets 3*bead) -------- */

/* FilamentStretchingHarmonic::energy / forces (FilamentStretchingHarmonic.cpp:185-275) */
double or_stretch_energy(const double* coord, const int* beadSet, const double* kstr,
                         const double* eql, int nint);
void or_stretch_forces(const double* coord, double* f, const int* beadSet, const double* kstr,
                       const double* eql, int nint);

/* FilamentBendingCosine::energy / forces (FilamentBendingCosine.cpp:199-398) */
double or_bend_energy(const double* coord, int nint, const int* beadSet, const double* kbend,
                      const double* eqt);
void or_bend_forces(const double* coord, double* f, int nint, const int* beadSet,
                    const double* kbend, const double* eqt);

/* LinkerStretchingHarmonic / MotorGhostStretchingHarmonic ::energy / forces
 * (LinkerStretchingHarmonic.cpp:191-336, MotorGhostStretchingHarmonic.cpp:234-389) */
double or_xlink_energy(const double* coord, const int* beadSet, const double* kstr,
                       const double* eql, const double* pos1, const double* pos2, int nint);
void or_xlink_forces(const double* coord, double* f, const int* beadSet, const double* kstr,
                     const double* eql, const double* pos1, const double* pos2,
                     double* stretchforce, int nint);

/* CylinderExclVolRepulsion::energy / forces (CylinderExclVolRepulsion.cpp:330-882).
 * energy returns -1 when a pair is non-finite even after the Simpson fallback. */
double or_exvol_energy(const double* coord, const int* beadSet, const double* krep,
                       const double* eqLengths, int nint);
void or_exvol_forces(const double* coord, double* f, const int* beadSet, const double* krep,
                     const double* eqLengths, int nint);
/* Simpson fallbacks energyN / forceN (CylinderExclVolRepulsion.cpp:884-1254), exposed for
 * tests. */
double or_exvol_energyN(const double* coord, const int* beadSet, const double* krep,
                        const double* eqLengths, int i, int movebeads);
void or_exvol_forceN(const double* coord, double* f, const int* beadSet, const double* krep,
                     const double* eqLengths, int i, int movebeads);
/* MathFunctions.h:1101-1135 areInPlane (|n.v| < 0.1 after normalization) */
int or_are_in_plane(const double* p1, const double* p2, const double* p3, const double* p4);
/* movePointOutOfPlane (MathFunctions.h:1221-1278): point i in {1, 2} moved by d along the unit
 * normal of (p2 - p1, p3 - p1), exposed for tests */
void or_move_out_of_plane(const double* p1, const double* p2, const double* p3, double* newp, int i,
                          double d);

/* BoundaryCylinderRepulsionExp::energy / forces (BoundaryCylinderRepulsionExp.cpp:27-118) */
double or_boundary_energy(const double* coord, const int* beadSet, const double* krep,
                          const double* slen, const int* nneighbors,
                          const mdn_boundary_element* elems, int nelem);
void or_boundary_forces(const double* coord, double* f, const int* beadSet, const double* krep,
                        const double* slen, const int* nneighbors,
                        const mdn_boundary_element* elems, int nelem);
/* PlaneBoundaryElement / SphereBoundaryElement distance and normal (BoundaryElementImpl.h) */
double or_boundary_distance(const mdn_boundary_element* e, const double* x);
void or_boundary_normal(const mdn_boundary_element* e, const double* x, double* n);
/* PlaneBoundaryElement constructor: (a,b,c) = normal, d = -a x0 - b y0 - c z0 */
mdn_boundary_element or_plane_element(const double* point, const double* normal, double krep,
                                      double slen);

/* ---- neighbor lists ----------------------------------------------------------------- */

/* Half list with the HybridCylinderCylinderNL predicate (HybridNeighborListImpl.cpp:411-439):
 * pair (c, c') kept iff c > c', not (same filament and |pos - pos'| <= 2) and
 * rMin^2 <= d^2(mid_c, mid_c') <= rMax^2 with the squares formed in float.
 * Output pairs ordered by c ascending, then c' ascending. Returns the pair count, or -1 when
 * cap is too small (the count is still written to *needed). */
long or_nl_brute(const double* mid, const int* fil, const int* pos, int n_cyl, float rmax,
                 float rmin, int* pairs, long cap, long* needed);
/* Same pair set through a cell list with bins of side 1.125 * rmax
 * (HybridNeighborListImpl.cpp:176-240). Returns -2 when a midpoint is outside
 * [0, grid_size). */
long or_nl_cells(const double* mid, const int* fil, const int* pos, int n_cyl, float rmax,
                 float rmin, const double* grid_size, int* pairs, long cap, long* needed);

/* ---- the system (vectorized tables) ------------------------------------------------- */
typedef struct or_system {
    int n_beads, n_cyl, n_fil;
    /* per cylinder */
    int *cyl_b1, *cyl_b2, *cyl_fil, *cyl_pos, *cyl_type;
    char* cyl_minus_end;
    double *cyl_kexvol, *cyl_eql;
    /* stretching */
    int* str_beadSet; double *str_k, *str_eql;
    /* bending */
    int n_bend; int* bend_beadSet; double *bend_k, *bend_eqt;
    /* linkers / motors */
    int n_link; int* lnk_beadSet; double *lnk_k, *lnk_eql, *lnk_p1, *lnk_p2, *lnk_sf;
    int n_mot; int* mot_beadSet; double *mot_k, *mot_eql, *mot_p1, *mot_p2, *mot_sf;
    /* excluded volume (from the NL) */
    long n_pairs; int* pairs; int* ev_beadSet; double *ev_krep, *ev_eqlen;
    /* boundary (from the boundary NL) */
    int n_elem; mdn_boundary_element* elems; double bnd_cutoff;
    int n_bint; int* bnd_beadSet; double *bnd_krep, *bnd_slen; int* bnd_nneighbors;
    char* bnd_member;   /* [n_elem * n_cyl]: cylinder c in element e's neighbor list */
    float nl_rmax, nl_rmin; double grid[3];
    int use_cells;
} or_system;

int or_system_build(const mdn_system_desc* d, or_system* s);
/* Rebuild both neighbor lists from coord and vectorize exvol + boundary. Returns MDN_OK or
 * MDN_ERR_NL_BOUNDS. */
int or_system_nl(or_system* s, const double* coord);
void or_system_free(or_system* s);

/* ForceFieldManager::computeEnergy: total with the <= -1 sentinel -> +inf; rep may be NULL. */
double or_ffm_energy(const or_system* s, const double* coord, mdn_energy_report* rep);
/* ForceFieldManager::computeForces: zero-fill then accumulate every FF in order. */
void or_ffm_forces(or_system* s, const double* coord, double* f);
/* One force field, accumulating. */
void or_ff_forces(or_system* s, int ff, const double* coord, double* f);
double or_ff_energy(const or_system* s, int ff, const double* coord);

/* ---- minimizer ---------------------------------------------------------------------- */
typedef double (*or_energy_fn)(void* user, const double* coord);
typedef void (*or_force_fn)(void* user, const double* coord, double* force, int nvar);
typedef void (*or_energy_ind_fn)(void* user, const double* coord, mdn_energy_report* rep);
/* returns MDN_OK or MDN_ERR_RECOVERY */
typedef int (*or_recovery_fn)(void* user, double* coord, const double* searchDir, int ndof);

typedef struct or_cg_callbacks {
    or_energy_fn energy;
    or_force_fn force;
    or_energy_ind_fn energy_individual;
    or_recovery_fn recovery;
    void* user;
    int n_beads;      /* Bead::numBeads() for the force-call limit (CGMethod.hpp:855-858) */
} or_cg_callbacks;

/* CGMethod::minimize<method> (CGMethod.hpp:826-1743) with backtrackingLineSearch
 * (CGMethod.hpp:238-373). Returns MDN_OK, MDN_ERR_NONFINITE (maxF error) or
 * MDN_ERR_RECOVERY. A line-search error with recovery disabled returns MDN_OK with
 * res->error_flags = MDN_MIN_ERR_LINESEARCH, as the reference returns the result. */
int or_cg_minimize(const mdn_cg_params* p, double* coord, int ndof, int nvar,
                   const or_cg_callbacks* cb, double* force_out, mdn_min_result* res);

/* MController::run wiring (MController.h:122-162) + moveAlongDirectionLimitingBeadMove_
 * (MController.h:273-339) as the recovery function. The neighbor lists must be built. */
int or_system_minimize(or_system* s, const mdn_cg_params* p, double* coord, double* force_out,
                       mdn_min_result* res);

/* Boundary load forces on the filament ends after a minimization
 * (BoundaryCylinderRepulsion::computeLoadForces, BoundaryCylinderRepulsion.cpp:122-187, with the
 * reset of ForceFieldManager::computeLoadForces, ForceFieldManager.cpp:292-310), from the
 * boundary neighbor list of the last or_system_nl. out_plus[f * stride + i] = loadForcesP[i] of
 * filament f's plus-end bead, out_minus likewise for its minus-end bead; i < num_mon[type]
 * (cylinderNumMon), zeros up to stride. Returns MDN_ERR_ARG on an unknown type or a stride
 * below num_mon. */
int or_system_load_forces(const or_system* s, const double* coord, int n_types, const double* mon_size,
                          const int* num_mon, int stride, double* out_plus, double* out_minus);

/* ---- handle API for language bindings ------------------------------------------------ */
or_system* or_system_create(const mdn_system_desc* d, int* status);
void or_system_destroy(or_system* s);
void or_system_set_cells(or_system* s, int use_cells);
long or_system_npairs(const or_system* s);
const int* or_system_pairs(const or_system* s);
int or_system_nbint(const or_system* s);
const int* or_system_bnd_beadset(const or_system* s);
const int* or_system_bnd_nneighbors(const or_system* s);
int or_system_ncyl(const or_system* s);
int or_system_nbend(const or_system* s);
const double* or_system_stretchforce(const or_system* s, int motor);

/* HybridBindingSearchManager::updateAllPossibleBindingsstencilHYBD (HybridBindingSearchManager.cpp:
 * 759-966) by brute force over all cylinder pairs at coord: the pairs (t1, t2) in the order
 * (cylinder 1, cylinder 2, site 1, site 2), t = (cylinder << shift_bits) | site. Returns the
 * count, or -1 when cap is too small (*needed holds the count). */
long or_binding_pairs(const or_system* s, const double* coord, const mdn_binding_params* p, uint32_t* out,
                      long cap, long* needed);

#ifdef __cplusplus
}
#endif
#endif
