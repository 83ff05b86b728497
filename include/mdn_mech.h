/*
 * mdn_mech.h -- C ABI of the MI355X (gfx950) mechanical-equilibration library.
 *
 * The library replaces the mechanical-equilibration path of MEDYAN 5.4.0:
 *   - ForceFieldManager energy / force evaluation over six force fields
 *     (FilamentStretching<Harmonic>, FilamentBending<Cosine>, LinkerStretching<Harmonic>,
 *      MotorGhostStretching<Harmonic>, CylinderExclVolume<CylinderExclVolRepulsion>,
 *      BoundaryCylinderRepulsion<BoundaryCylinderRepulsionExp>),
 *   - the cylinder-cylinder hybrid neighbor list and the boundary-cylinder neighbor list,
 *   - CGMethod::minimize (steepest descent, Fletcher-Reeves, Polak-Ribiere) with the
 *     backtracking line search, safe mode and recovery move.
 *
 * Conventions (all entry points):
 *   - every function returns an int status, MDN_OK (0) on success; mdn_last_error() gives
 *     a message for the last failing call on a context;
 *   - host buffers are owned by the caller; device buffers are owned by the context;
 *   - coordinates and forces are the reference's flat `coord` vector: 3*n_beads doubles,
 *     bead-major (x,y,z per bead), in bead index order (src/Structure/DofSerializer.hpp:64-187);
 *   - one context serves one caller thread (the reference path is single-threaded,
 *     src/Controller/MController.h:108-211); contexts are not thread-safe.
 *
 * Citations are relative to medyan-5.4.0/ in the reference tree.
 */
#ifndef MDN_MECH_H
#define MDN_MECH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDN_ABI_VERSION 1

/* ---- status codes ------------------------------------------------------------------ */
enum {
    MDN_OK = 0,
    MDN_ERR_ARG = 1,          /* invalid argument / inconsistent system description          */
    MDN_ERR_HIP = 2,          /* HIP runtime error (no device, launch failure, ...)          */
    MDN_ERR_STATE = 3,        /* call order violated (e.g. energy before upload)             */
    MDN_ERR_NONFINITE = 4,    /* maxF non-finite: CGMethod.hpp:202-212 throws "maxF error"   */
    MDN_ERR_LINESEARCH = 5,   /* line search failed with recovery disabled                   */
    MDN_ERR_NL_BOUNDS = 6,    /* cylinder midpoint outside the neighbor-list grid            */
    MDN_ERR_RECOVERY = 7,     /* recovery step infinite: MController.h:312-316               */
    MDN_ERR_NOMEM = 8,
    MDN_ERR_UNSUPPORTED = 9,
    MDN_ERR_CAPACITY = 10     /* caller buffer too small                                     */
};

/* ---- force-field slots, in ForceFieldManager order (src/Controller/MController.cpp:70-178) */
enum {
    MDN_FF_STRETCH = 0,   /* FilamentStretching<FilamentStretchingHarmonic>              */
    MDN_FF_BEND = 1,      /* FilamentBending<FilamentBendingCosine>                      */
    MDN_FF_LINKER = 2,    /* LinkerStretching<LinkerStretchingHarmonic>                  */
    MDN_FF_MOTOR = 3,     /* MotorGhostStretching<MotorGhostStretchingHarmonic>          */
    MDN_FF_EXVOL = 4,     /* CylinderExclVolume<CylinderExclVolRepulsion>                */
    MDN_FF_BOUNDARY = 5,  /* BoundaryCylinderRepulsion<BoundaryCylinderRepulsionExp>     */
    MDN_NUM_FF = 6
};

/* ---- system description (the data serializeDof + vectorize read) ----------------------- */

/* A linker or motor ghost: a spring between the point at relative position pos1 on
 * cylinder cyl1 and pos2 on cyl2 (Linker.cpp:60-71, LinkerStretching.cpp:30-57). */
typedef struct mdn_crosslink {
    int32_t cyl1, cyl2;   /* global cylinder indices */
    double pos1, pos2;    /* relative positions in [0,1] (adjustedrelativeposition = identity) */
    double kstr, eql;     /* stretching constant (pN/nm), rest length (nm) */
} mdn_crosslink;

enum { MDN_BND_PLANE = 0, MDN_BND_SPHERE = 1 };

/* A boundary element (src/Structure/BoundaryElementImpl.h:30-197).
 * plane : p = {a, b, c, d}, distance = (a x + b y + c z + d) / |(a,b,c)|, normal = (a,b,c)
 * sphere: p = {cx, cy, cz, radius}, distance = radius - |x - c|, normal = (c - x)/|c - x| */
typedef struct mdn_boundary_element {
    int32_t kind;
    int32_t reserved;
    double p[4];
    double krep;          /* BoundaryK (pN)            */
    double slen;          /* BScreenLength (nm)        */
} mdn_boundary_element;

/* The mechanical system. Cylinders are numbered filament-major: cylinder k (0-based) of
 * filament f joins beads fil_beads[off_f + k] and fil_beads[off_f + k + 1] with
 * off_f = fil_bead_offset[f]; its global index is off_f - f + k. The global index is also the
 * cylinder id used by the neighbor-list half-list rule (HybridNeighborListImpl.cpp:413), and k
 * its position on the filament (:415-418). Every filament has at least two beads. */
typedef struct mdn_system_desc {
    int32_t n_beads;
    int32_t n_filaments;
    const int32_t* fil_bead_offset;   /* [n_filaments + 1]                                   */
    const int32_t* fil_beads;         /* bead indices, minus end first                       */
    const int32_t* fil_type;          /* [n_filaments] filament type, or NULL (all type 0)   */
    const double* cyl_kstr;           /* [n_cyl] FStretchingK * cylinderSize / eqLength      */
    const double* cyl_eql;            /* [n_cyl] cylinder equilibrium length                 */
    const double* cyl_kbend;          /* [n_cyl] bending constant; the bending term whose
                                         second cylinder is c uses cyl_kbend[c]
                                         (FilamentBending.cpp:29-60)                         */
    const double* cyl_eqtheta;        /* [n_cyl] equilibrium bending angle                   */
    const double* cyl_kexvol;         /* [n_cyl] excluded-volume constant                    */
    int32_t n_linkers;
    const mdn_crosslink* linkers;
    int32_t n_motors;
    const mdn_crosslink* motors;
    int32_t n_boundary;
    const mdn_boundary_element* boundary;   /* registry order = interaction order           */
    double boundary_cutoff;           /* BoundaryCutoff (strict <, NeighborListImpl.cpp:748)  */
    double exvol_rmax;                /* VolumeCutoff, passed through float like the
                                         reference (HybridNeighborListImpl.h:74-80)          */
    double exvol_rmin;
    double grid_size[3];              /* neighbor-list grid covers [0, grid_size) per axis   */
} mdn_system_desc;

typedef struct mdn_energy_report {
    double total;                     /* sum of ff[] in order; +inf on a failure sentinel    */
    double ff[MDN_NUM_FF];
} mdn_energy_report;

enum { MDN_CG_STEEPEST = 0, MDN_CG_FLETCHER_REEVES = 1, MDN_CG_POLAK_RIBIERE = 2 };
/* ConjugateGradientParams::lineSearchAlgorithm (CGMethod.hpp:51, :1319-1331): BACKTRACKING, or
 * QUADRATIC (quadraticLineSearchV2 + quadraticoptimization, CGMethod.hpp:522-764, outside safe
 * mode; safe mode backtracks with slope 0 in both). The quadratic search here evaluates every
 * trial at coord + lambda * searchDir; the reference's version moves coord in place and evaluates
 * a line-search buffer it does not refresh, so its values depend on earlier calls. */
enum { MDN_LS_BACKTRACKING = 0, MDN_LS_QUADRATIC = 1 };
enum { MDN_PREC_FP64 = 0, MDN_PREC_MIXED = 1 };

/* ConjugateGradientParams (CGMethod.hpp:44-63) plus the precision selector. */
typedef struct mdn_cg_params {
    int32_t method;               /* MDN_CG_*                                               */
    int32_t try_recover;          /* tryToRecoverInLineSearchError                          */
    double max_dist;              /* maxDist                                                */
    double lambda_max;            /* lambdaMax                                              */
    double grad_tol;              /* gradTol (max component |F|)                            */
    double energy_rel_tol;        /* energyRelTol                                           */
    int32_t force_call_limit;     /* raised to max(limit, 3 * n_beads) (CGMethod.hpp:855-858) */
    int32_t energy_call_limit;    /* 0 = no limit                                           */
    int32_t report_on_call_limit; /* restore minimum-energy coordinates on call limit       */
    int32_t precision;            /* MDN_PREC_*                                             */
    int32_t line_search;          /* MDN_LS_* (0 = BACKTRACKING, the reference's default)   */
    int32_t reserved;             /* 0                                                      */
} mdn_cg_params;

enum { MDN_MIN_ERR_LINESEARCH = 1 };   /* MinimizationResult::errorLineSearch */

/* MinimizationResult (MinimizationTypes.hpp:10-64) plus path statistics. */
typedef struct mdn_min_result {
    mdn_energy_report energies_before;
    mdn_energy_report energies_after;
    uint32_t error_flags;
    int32_t num_energy_call;
    int32_t num_force_call;
    int32_t force_call_limit;
    int32_t energy_call_limit;
    int32_t num_iter;
    int32_t num_safe_mode;        /* line-search failures that entered safe mode            */
    int32_t num_recovery;         /* recovery moves                                         */
    double grad_inf_norm;
    double grad_tol;
    double energy_rel_change;
    double energy_rel_tol;
    int32_t converged_grad;       /* gradInfNorm < gradTol fired                            */
    int32_t converged_energy;     /* energyRelChange < energyRelTol fired                   */
    int32_t call_limit_exceeded;
    int32_t restored_min_energy;  /* copybackupcoordinates() restored coordMinE             */
} mdn_min_result;

/* ---- library ------------------------------------------------------------------------ */
typedef struct mdn_ctx mdn_ctx;

/* ABI version (MDN_ABI_VERSION). */
int mdn_abi_version(void);

/* Create a context bound to HIP device `device` with its own stream. */
int mdn_ctx_create(int device, mdn_ctx** out);
int mdn_ctx_destroy(mdn_ctx* ctx);
const char* mdn_last_error(const mdn_ctx* ctx);

/* Upload the system and its coordinates, and build the bonded and crosslink tables
 * (ForceFieldManager::vectorizeAllForceFields, ForceFieldManager.cpp:35-50, for the
 * stretching, bending, linker and motor force fields). coord NULL: the coordinates that
 * mdn_dof_serialize gathered on the device for sys->n_beads beads. */
int mdn_system_upload(mdn_ctx* ctx, const mdn_system_desc* sys, const double* coord);

/* Counts: n_beads, n_cylinders, and interaction counts per force field (after mdn_nl_build
 * for MDN_FF_EXVOL and MDN_FF_BOUNDARY). out must hold 2 + MDN_NUM_FF int64 values. */
int mdn_system_counts(mdn_ctx* ctx, int64_t* out);

/* Incremental re-vectorization of the crosslink force fields alone: replaces the linker and
 * motor-ghost tables of the uploaded system (same filaments, beads and cylinders) and keeps the
 * filament tables, the neighbor lists and the excluded-volume / boundary tables of the last
 * mdn_nl_build. Replaces the LinkerStretching / MotorGhostStretching part of
 * ForceFieldManager::vectorizeAllForceFields (ForceFieldManager.cpp:35-50 ->
 * LinkerStretching.cpp:30-57, MotorGhostStretching.cpp:30-57), which the reference reruns for
 * every force field before each minimization although between chemistry steps only the bound
 * linkers and motors change. Indices are validated before any state changes (MDN_ERR_ARG). */
int mdn_crosslinks_update(mdn_ctx* ctx, const mdn_crosslink* linkers, int32_t n_linkers,
                          const mdn_crosslink* motors, int32_t n_motors);

/* Rebuild the excluded-volume neighbor list (HybridCylinderCylinderNL::reset,
 * HybridNeighborListImpl.cpp:527-556) and the boundary neighbor list
 * (BoundaryCylinderNL::reset, NeighborListImpl.cpp:801-809) from `coord` (host, or NULL for
 * the context's current coordinates) and vectorize the two force fields
 * (CylinderExclVolume.cpp:33-138, BoundaryCylinderRepulsion.cpp:32-87).
 * n_pairs receives the number of excluded-volume pairs. */
int mdn_nl_build(mdn_ctx* ctx, const double* coord, int64_t* n_pairs);

/* Copy the excluded-volume pair list (cylinder index pairs (c, c'), c > c', ordered by c then
 * c') into out[2*cap]. */
int mdn_nl_pairs(mdn_ctx* ctx, int32_t* out, int64_t cap, int64_t* n_pairs);

/* Copy the boundary interaction list: bead index per interaction, element-major
 * (BoundaryCylinderRepulsion.cpp:56-84), and per-element counts (nneighbors). */
int mdn_boundary_list(mdn_ctx* ctx, int32_t* out_beads, int64_t cap, int32_t* out_nneighbors,
                      int64_t* n_interactions);

/* ForceFieldManager::computeIndividualEnergies (ForceFieldManager.cpp:209-227): per-FF
 * energies at `coord` (host; NULL = current coordinates). A force field that reports the
 * failure sentinel (energy <= -1, ForceFieldManager.cpp:150-161) makes total = +inf. */
int mdn_energy(mdn_ctx* ctx, const double* coord, mdn_energy_report* out);

/* ForceFieldManager::computeForces (ForceFieldManager.cpp:242-290): force[0..n_var) is
 * zero-filled, then every force field accumulates in order. n_var = 3 * n_beads. */
int mdn_forces(mdn_ctx* ctx, const double* coord, double* force, int32_t n_var);

/* ForceField::computeForces of one force field (ForceField.h:34-99): ACCUMULATES (+=) the
 * forces of force field `ff` (MDN_FF_*) into force[0..n_var). */
int mdn_ff_forces(mdn_ctx* ctx, int32_t ff, const double* coord, double* force, int32_t n_var);

/* Linker / motor stretchForce after the last force evaluation
 * (LinkerStretchingHarmonic.cpp:291, ForceField::assignforcemags). */
int mdn_crosslink_stretch_forces(mdn_ctx* ctx, double* linker_out, double* motor_out);

/* Boundary load forces on the filament ends after a minimization, for the chemistry's rate
 * changers (BoundaryCylinderRepulsion::computeLoadForces, BoundaryCylinderRepulsion.cpp:122-187,
 * reached through ForceFieldManager::computeLoadForces, ForceFieldManager.cpp:292-310, and
 * MController's afterMinimization_ -> computeAuxParams, MController.h:238-266), from the boundary
 * neighbor list of the last mdn_nl_build and the coordinates `coord` (NULL: the current device
 * coordinates, i.e. the last minimization's result). For filament f, load_plus[f * stride + i] is
 * Bead::loadForcesP[i] of its plus-end bead and load_minus[f * stride + i] Bead::loadForcesM[i] of
 * its minus-end bead, i < num_monomers[type] (cylinderNumMon; monomer_size[type] is monomerSize);
 * entries up to `stride` are zero. Filament types must be below n_types (<= 8). FP64. */
int mdn_boundary_load_forces(mdn_ctx* ctx, const double* coord, int32_t n_types, const double* monomer_size,
                             const int32_t* num_monomers, int32_t stride, double* load_plus, double* load_minus);

/* CGMethod::minimize (CGMethod.hpp:826-1759) driven by MController::run
 * (MController.h:108-211): minimizes from `coord` (host, in/out), writes the final forces to
 * force_out (3*n_beads, may be NULL) and the statistics to res. The neighbor list must have
 * been built (mdn_nl_build) and stays frozen during the minimization, as in the reference.
 * coord may be NULL: the minimization then starts from the context's current device
 * coordinates (the last upload, mdn_set_coords, mdn_displace or minimization result) and leaves
 * its result there, with no coordinate transfer over the host link (mdn_get_coords reads it). */
int mdn_minimize(mdn_ctx* ctx, const mdn_cg_params* params, double* coord, double* force_out,
                 mdn_min_result* res);

/* Launch statistics of the last mdn_minimize on this context: out[0] line-search launches
 * that did work, out[1] steps whose excluded-volume forces needed the numerical fallback
 * (forceN; the device state machine paused and the host resumed it), out[2] force gathers,
 * out[3] CG iterations. */
int mdn_last_min_stats(mdn_ctx* ctx, int64_t* out);

/* ---- device-resident variants (benchmarking, zero host traffic) ---------------------- */

/* Time `iters` force evaluations (mdn_forces semantics) on the context's current device
 * coordinates with HIP events on the context's stream; *ms_per_eval receives the mean
 * wall time per evaluation and *kernel_ms the mean duration of the dominant kernel. */
int mdn_bench_forces(mdn_ctx* ctx, int32_t precision, int32_t iters, double* ms_per_eval,
                     double* kernel_ms);

/* Time `iters` total-energy evaluations the same way. */
int mdn_bench_energy(mdn_ctx* ctx, int32_t precision, int32_t iters, double* ms_per_eval);

/* Time `iters` rounds of n_vec Hessian-vector products (n_vec = 1: mdn_hvp / mdn_hvp_mixed
 * semantics; 2 .. MDN_HVP_BATCH_CHUNK: one chunk of mdn_hvp_batch, its tangent terms in one
 * launch) along the tangents v (n_vec rows of n_var = 3 * n_beads host doubles, uploaded once
 * before the timed launches), then `iters` parameter gradients (mdn_param_grad /
 * mdn_param_grad_mixed), on the context's current device coordinates, with HIP events on the
 * context's stream: *ms_per_hvp receives the mean device time per product and
 * *ms_per_param_grad per gradient, without the host transfers of the per-call entry points.
 * Fails with MDN_ERR_UNSUPPORTED when an excluded-volume pair at these coordinates needs a
 * fallback (no tangent, or the Simpson energy), whose host round trips these timings would
 * leave out. */
int mdn_bench_adjoint(mdn_ctx* ctx, int32_t precision, int32_t iters, const double* v, int32_t n_var,
                      int32_t n_vec, double* ms_per_hvp, double* ms_per_param_grad);

/* Set the current device coordinates from host memory (no NL rebuild). */
int mdn_set_coords(mdn_ctx* ctx, const double* coord);
int mdn_get_coords(mdn_ctx* ctx, double* coord);

/* Displace the current device coordinates in place: x += delta, with delta a DEVICE buffer of
 * n_var = 3 * n_beads doubles on the context's device (the caller orders its writes of delta
 * before the call). The perturbation between successive minimizations (a seeded Gaussian bead
 * displacement) applied without a coordinate round trip over the host link; the neighbor lists
 * are not rebuilt (mdn_nl_build with coord NULL rebuilds them at the displaced coordinates). */
int mdn_displace(mdn_ctx* ctx, const double* delta, int32_t n_var);

/* Displace the current device coordinates by seeded Gaussian noise, x_i += sigma z_i with z_i
 * standard normal, generated on the device from (seed, stream, i) alone (SplitMix64 counters,
 * Box-Muller; picgo_amd.mech.perturbation reproduces it on the host): the seeded bead
 * perturbation between successive minimizations, with no buffer and no host transfer. The
 * neighbor lists are not rebuilt. */
int mdn_perturb(mdn_ctx* ctx, uint64_t seed, uint64_t stream, double sigma);

/* serializeDof (DofSerializer.hpp:64-92) on the device: the next system's bead coordinates,
 * n_beads of them in the order of Bead::getBeads(), gathered from the current device coordinates
 * (the last upload, displacement or minimization result) with only the new beads' coordinates
 * from the host. src[i] >= 0: bead i of the next system is bead src[i] of the current one;
 * src[i] < 0: it is a new bead, whose coordinates are the next 3 values of new_coord (new beads
 * in ascending i). The result is kept on the device until the next mdn_system_upload(ctx, desc,
 * NULL) adopts it as that system's coordinates (the current system stays usable meanwhile).
 * Beads that chemistry destroyed are the current indices that no src entry names. */
int mdn_dof_serialize(mdn_ctx* ctx, int32_t n_beads, const int32_t* src, const double* new_coord);

/* ---- binding-site pair search (HybridBindingSearchManager) ---------------------------- */

/* One filament-type pair and distance range of a binder (setbindingsearchparameter,
 * HybridBindingSearchManager.cpp:34-108, and addtoHNeighborList, :1024-1039). */
typedef struct mdn_binding_params {
    int32_t ftype1, ftype2;    /* the filament-type pair                                      */
    int32_t n_sites;           /* binding sites per cylinder (at most 1 << shift_bits)        */
    int32_t shift_bits;        /* pair entry t = (cylinder << shift_bits) | site              */
    const float* site_pos;     /* [n_sites] relative positions (binding-site monomer /
                                  cylinderNumMon, in float); the reference uses the first
                                  type's list for both cylinders (bindingsites2 = bs1)        */
    float p2_first, p2_last;   /* the second type's first and last site positions, bounds of
                                  the early rejections (minparamcyl2 / maxparamcyl2)          */
    float rmin, rmax;          /* the binder's distance range (nm), squared in float          */
    float cyl_size;            /* the larger cylinderSize of the two types (the list's range) */
    int32_t site_step;         /* > 0: only sites with index % site_step == 0 (linkers:
                                  linkerbindingskip - 1, :847-851); 0: every site             */
    const float* site_free;    /* [n_cylinders * n_sites], 1 = the site is free
                                  (speciesboundvec of the binder); NULL: every site free      */
} mdn_binding_params;

/* HybridBindingSearchManager::updateAllPossibleBindingsstencilHYBD
 * (HybridBindingSearchManager.cpp:759-966) at coord (host, 3*n_beads; NULL: the device
 * coordinates): every pair of free binding sites on two cylinders of different filaments whose
 * distance lies in [rmin, rmax], found through the hybrid neighbor list with the binder's range
 * (cylinder midpoints within cyl_size + rmax, at least max(rmin - cyl_size, 0), the
 * same-filament rule) and the reference's per-site quadratic test. Cylinders are this ABI's
 * global indices; out receives n pairs (t1, t2), t1 on the cylinder with the larger index,
 * ordered by (cylinder 1, cylinder 2, site 1, site 2). out NULL: *n_out = the count only;
 * capacity below the count: MDN_ERR_CAPACITY with *n_out = the count. */
int mdn_binding_pairs(mdn_ctx* ctx, const mdn_binding_params* p, const double* coord, uint32_t* out,
                      int64_t capacity, int64_t* n_out);

/* Mixed precision (MDN_PREC_MIXED): the bending and boundary terms and the force components of
 * the harmonic springs (stretching, linkers, motors) are computed in FP32 from bond vectors and
 * coordinate differences formed in FP64 (absolute coordinates never enter FP32 arithmetic; spring
 * lengths and extensions L - L0 stay FP64, as L0 reaches micrometres for crosslinks); excluded
 * volume stays FP64 (its closed form is a 0/0 limit for near-coplanar pairs); per-interaction
 * forces are stored in FP32; per-bead
 * force sums, energy sums, dot products and all minimizer decisions stay FP64. No reference
 * counterpart: the reference's `floatingpoint` type switches the whole program to float
 * (src/utility.h:38-43), which this replaces for the force path.
 * mdn_forces_mixed: same semantics as mdn_forces (ForceFieldManager::computeForces,
 * ForceFieldManager.cpp:242-290); mdn_energy_mixed: as mdn_energy (:209-227).
 * mdn_minimize runs in mixed precision when params->precision == MDN_PREC_MIXED (its reported
 * energies_before / energies_after stay FP64). */
int mdn_forces_mixed(mdn_ctx* ctx, const double* coord, double* force, int32_t n_var);
int mdn_energy_mixed(mdn_ctx* ctx, const double* coord, mdn_energy_report* out);

/* ---- force evaluation without allocation --------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The forces of
 * mdn_forces (precision MDN_PREC_FP64) or mdn_forces_mixed (MDN_PREC_MIXED), zero-filled then
 * every force field accumulated, bit-identical to them, with coordinates and forces in DEVICE
 * memory of the context's device (3 * n_beads doubles each). The evaluation's scratch (slot
 * vectors, bonded sums, the numerical-fallback list and partial sums) lives in a file-scope
 * device array of MDN_NOALLOC_SCRATCH_MIB MiB that each device holds until the process ends; a
 * system whose evaluation needs more is refused with MDN_ERR_CAPACITY. The call allocates
 * nothing (no hipMalloc / hipFree) and returns after the forces are written. All calls on one
 * GPU share that array: calls that overlap on one GPU (other threads, other contexts) race. */
#define MDN_NOALLOC_SCRATCH_MIB 48
int mdn_forces_noalloc(mdn_ctx* ctx, int32_t precision, const double* coord_dev, double* force_dev,
                       int32_t n_var);

/* ---- differentiation ---------------------------------------------------------------------- */

/* dE/dtheta at coord (null: the current device coordinates) in FP64, for the per-cylinder
 * parameters theta = (stretching k [n_cyl], rest length L0 [n_cyl], bending k [n_cyl]); out
 * holds 3 * n_cylinders values in that order. L0 enters the stretching terms
 * (FilamentStretchingHarmonic.cpp:185-199) and kRepScaled = krep L0_i L0_j of every excluded-
 * volume pair with the cylinder (CylinderExclVolume.cpp:73-138); a cylinder's bending k is the k
 * of the bending term whose second cylinder it is (FilamentBending.cpp:29-60; 0 when there is
 * none). No reference counterpart (the reference has no parameter derivatives). */
int mdn_param_grad(mdn_ctx* ctx, const double* coord, double* out, int64_t n_out);

/* Hessian-vector product H v = -(dF/dx) v at coord (null: the current device coordinates) in
 * FP64, by forward-mode dual arithmetic through the same force kernels (the derivative of the
 * implemented forces, including the coplanar shift of CylinderExclVolRepulsion.cpp:610-630).
 * The reference computes a dense finite-difference Hessian instead (ForceFieldManager.cpp:
 * 354-500). MDN_ERR_UNSUPPORTED when an excluded-volume pair needs the numerical force fallback. */
int mdn_hvp(mdn_ctx* ctx, const double* coord, const double* v, double* out, int32_t n_var);

/* mdn_param_grad and mdn_hvp with the mixed path's arithmetic: FP32 term arithmetic (dual numbers
 * for the tangents) for the stretching, bending and crosslink force fields, FP64 excluded volume,
 * boundary and accumulation (the precision split of mdn_forces_mixed). */
int mdn_param_grad_mixed(mdn_ctx* ctx, const double* coord, double* out, int64_t n_out);
int mdn_hvp_mixed(mdn_ctx* ctx, const double* coord, const double* v, double* out, int32_t n_var);

/* n_vec Hessian-vector products at one point: out[k n_var + i] = (H V_k)_i for the tangents
 * V_k = V + k n_var (each n_var = 3 * n_beads doubles), with the results of mdn_hvp (precision
 * MDN_PREC_FP64) or mdn_hvp_mixed (MDN_PREC_MIXED) bit for bit. Up to MDN_HVP_BATCH_CHUNK
 * products share one launch of the tangent terms (a grid dimension per product, each with its
 * own slot and piece arrays on the device, which the context keeps for later calls: about 16 MB
 * per product at 42,000 beads), and every chunk crosses the host link in one copy each way, with
 * one synchronization per call. Like mdn_hvp, the reference's counterpart is the
 * dense finite-difference Hessian (ForceFieldManager.cpp:354-500), whose columns these products
 * replace one direction at a time. */
#define MDN_HVP_BATCH_CHUNK 16
int mdn_hvp_batch(mdn_ctx* ctx, int32_t precision, const double* coord, const double* V, double* out,
                  int32_t n_var, int32_t n_vec);

/* Replace the per-cylinder parameters (any pointer may be null: unchanged): stretching k and
 * rest length L0 (the excluded-volume kRepScaled of the current neighbor list is rebuilt) per
 * cylinder, bending k per cylinder (applied to the bending term whose second cylinder it is). */
int mdn_set_params(mdn_ctx* ctx, const double* kstr, const double* eql, const double* kbend);

/* ---- batched minimization ------------------------------------------------------------- */

/* CGMethod::minimize (CGMethod.hpp:1746-1759) of n independent systems in one persistent kernel:
 * each workgroup pulls a system from a device queue and minimizes it to completion (same
 * semantics as mdn_minimize per system; only the rounding order of the block reductions
 * differs), then pulls the next. ctxs[i] must have its system uploaded and its neighbor lists
 * built, all on one device; ctxs[0] is the lead context (its stream runs the batch and it keeps
 * the tuned schedule). coords[i] (3 * n_beads of system i) is the start and receives the result
 * (coords[i] NULL: the start is the context's device coordinates, which receive the result, with
 * no coordinate transfer over the host link);
 * forces_out may be null, or forces_out[i] may be null; results[i] is filled for every system.
 * The schedule (workgroup size 128, 256, 512 or 1024 threads, given or largest-first order) is chosen on
 * the first call for a given (n, precision, total beads) by timing a complete run of each candidate
 * from the given coordinates.
 * Returns MDN_OK when every system succeeded, else the status of the first failing system (its
 * context's mdn_last_error and the lead's carry the message). No reference counterpart for the
 * batching itself: the reference minimizes one system per call. */
int mdn_minimize_batched(mdn_ctx* const* ctxs, int32_t n, const mdn_cg_params* params,
                         double* const* coords, double* const* forces_out, mdn_min_result* results);

/* The schedule of the last batched minimization on a lead context: out3 = [workgroup size,
 * largest-first order (1) or given order (0), tuned in that call (1) or reused (0)]; tune_ms (may
 * be null, MDN_BATCH_CANDIDATES values) = the tuning times in ms of the candidates 128 given,
 * 128 largest first, 256 given, 256 largest first, 512 given, 512 largest first, 1024 given,
 * 1024 largest first. */
#define MDN_BATCH_CANDIDATES 8
int mdn_last_batch_schedule(mdn_ctx* ctx, int32_t* out3, double* tune_ms);

#ifdef __cplusplus
}
#endif

#endif /* MDN_MECH_H */
