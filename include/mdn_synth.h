/*
 * mdn_synth.h -- seeded synthetic actin networks for the mechanical-equilibration path.
 *
 * Builds the mdn_system_desc + coordinates of the benchmark workloads: straight filaments of
 * 108 nm cylinders with isotropic orientations and uniform centres, a cubic (six planes at 25 nm
 * and size-25 nm, BoundaryImpl.cpp:25-53) or spherical boundary, alpha-actinin linkers and
 * myosin motor ghosts. Parameters default to MEDYAN's interactive defaults
 * (MedyanConfig.hpp:336-391): FStretchingK 100, FBendingK 672 (cosine, theta0 = 0),
 * VolumeK 1e5 with a 108 nm cutoff, BoundaryK 41 pN, screening length 2.7 nm,
 * BoundaryCutoff 300 nm. Same seed -> same bits on every host (own MT19937-64, own
 * Box-Muller), so the GPU path and the CPU oracle see identical inputs.
 */
#ifndef MDN_SYNTH_H
#define MDN_SYNTH_H

#include <stdint.h>

#include "mdn_mech.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { MDN_SYNTH_CUBE = 0, MDN_SYNTH_SPHERE = 1 };
enum {
    MDN_XL_CLOSE_SEGMENTS = 0,   /* random cylinder pairs of distinct filaments whose closest
                                    segment-segment distance is < xl_max_dist              */
    MDN_XL_NL_NEIGHBOURS = 1,    /* random midpoint-neighbour pairs (<= 108 nm), distinct filaments */
    MDN_XL_ANY_DISTINCT = 2      /* random cylinders of distinct filaments                  */
};

typedef struct mdn_synth_params {
    uint64_t seed;
    int32_t n_fil;
    int32_t n_cyl_min, n_cyl_max;   /* cylinders per filament, uniform in [min, max]          */
    int32_t boundary;               /* MDN_SYNTH_CUBE / MDN_SYNTH_SPHERE                      */
    double box;                     /* cube side (nm); for the sphere the box is 2*radius     */
    double radius;                  /* sphere radius (nm)                                      */
    double cyl_len;                 /* 108 nm                                                  */
    double k_stretch, k_bend, k_exvol, exvol_cutoff;
    double k_boundary, screen_len, boundary_cutoff, wall_margin;
    int32_t n_linkers, n_motors;
    int32_t linkers_per_fil, motors_per_fil;   /* if > 0, override n_linkers / n_motors     */
    int32_t xl_mode;                /* MDN_XL_*                                                */
    double xl_max_dist;
    double k_linker, k_motor;
    double motor_eql_min, motor_eql_max;   /* motor rest length ~ U[min,max]; min < 0: L0 = distance */
    double lognormal_sigma;         /* > 0: per-cylinder kstr, eql, kbend ~ base * exp(sigma z) */
} mdn_synth_params;

typedef struct mdn_synth mdn_synth;

/* Fill `p` with preset `config` (0: 50 x 10 filaments in a 2 um cube; 1: 2,000 x 20 in a 5 um
 * cube; 2: as 1 with per-cylinder parameters; 3: one network of the batch of 256; 4: 20,000 x 20
 * in a 10 um sphere); for the batch `index` selects one of the 256 networks (seed 3 + index) and
 * `dense` != 0 selects the 100 um^-3 density variant (the default 0.1 um^-3 is as written). */
int mdn_synth_preset(int32_t config, int32_t index, int32_t dense, mdn_synth_params* p);

int mdn_synth_create(const mdn_synth_params* p, mdn_synth** out);
void mdn_synth_destroy(mdn_synth* s);
const mdn_system_desc* mdn_synth_desc(const mdn_synth* s);
const double* mdn_synth_coord(const mdn_synth* s);
/* Add N(0, sigma) noise to every coordinate (the 420,000-bead network's per-minimization
 * perturbation), drawn from the generator's stream. */
int mdn_synth_perturb(mdn_synth* s, double* coord, double sigma);

#ifdef __cplusplus
}
#endif
#endif
