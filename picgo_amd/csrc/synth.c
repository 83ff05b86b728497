/*
 * synth.c -- seeded synthetic actin networks (see include/mdn_synth.h).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/mdn_synth.h"

/* ---- MT19937-64 (Matsumoto & Nishimura 2000) ---------------------------------------- */
typedef struct { uint64_t mt[312]; int mti; } mt64;

static void mt64_seed(mt64* r, uint64_t seed) {
    r->mt[0] = seed;
    for (int i = 1; i < 312; i++)
        r->mt[i] = 6364136223846793005ULL * (r->mt[i - 1] ^ (r->mt[i - 1] >> 62)) + (uint64_t)i;
    r->mti = 312;
}

static uint64_t mt64_next(mt64* r) {
    static const uint64_t MAG[2] = {0ULL, 0xB5026F5AA96619E9ULL};
    const uint64_t UM = 0xFFFFFFFF80000000ULL, LM = 0x7FFFFFFFULL;
    if (r->mti >= 312) {
        int i;
        for (i = 0; i < 312 - 156; i++) {
            uint64_t x = (r->mt[i] & UM) | (r->mt[i + 1] & LM);
            r->mt[i] = r->mt[i + 156] ^ (x >> 1) ^ MAG[(int)(x & 1ULL)];
        }
        for (; i < 311; i++) {
            uint64_t x = (r->mt[i] & UM) | (r->mt[i + 1] & LM);
            r->mt[i] = r->mt[i + (156 - 312)] ^ (x >> 1) ^ MAG[(int)(x & 1ULL)];
        }
        uint64_t x = (r->mt[311] & UM) | (r->mt[0] & LM);
        r->mt[311] = r->mt[155] ^ (x >> 1) ^ MAG[(int)(x & 1ULL)];
        r->mti = 0;
    }
    uint64_t x = r->mt[r->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

/* uniform in [0, 1) with 53 random bits */
static double urand(mt64* r) { return (double)(mt64_next(r) >> 11) * 0x1.0p-53; }
/* uniform integer in [lo, hi] */
static int irand(mt64* r, int lo, int hi) {
    return lo + (int)(urand(r) * (double)(hi - lo + 1));
}
/* standard normal, Box-Muller (one value per call) */
static double nrand(mt64* r) {
    double u1 = 1.0 - urand(r), u2 = urand(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ---- the network --------------------------------------------------------------------- */
struct mdn_synth {
    mdn_synth_params p;
    mt64 rng;
    mdn_system_desc desc;
    double* coord;
    int32_t *fil_bead_offset, *fil_beads, *fil_type;
    double *kstr, *eql, *kbend, *eqt, *kexv;
    mdn_crosslink *linkers, *motors;
    mdn_boundary_element bnd[6];
    int n_cyl;
    int *cyl_fil;
};

int mdn_synth_preset(int32_t config, int32_t index, int32_t dense, mdn_synth_params* p) {
    memset(p, 0, sizeof(*p));
    p->cyl_len = 108.0;
    p->k_stretch = 100.0;
    p->k_bend = 672.0;
    p->k_exvol = 1e5;
    p->exvol_cutoff = 108.0;
    p->k_boundary = 41.0;
    p->screen_len = 2.7;
    p->boundary_cutoff = 300.0;
    p->wall_margin = 25.0;
    p->k_linker = 8.0;
    p->k_motor = 2.5;
    p->motor_eql_min = 175.0;
    p->motor_eql_max = 225.0;
    p->boundary = MDN_SYNTH_CUBE;
    p->xl_max_dist = 50.0;
    switch (config) {
    case 0:
        p->seed = 0; p->n_fil = 50; p->n_cyl_min = p->n_cyl_max = 10; p->box = 2000.0;
        p->n_linkers = 20; p->xl_mode = MDN_XL_CLOSE_SEGMENTS;
        break;
    case 1:
    case 2:
        p->seed = (uint64_t)config; p->n_fil = 2000; p->n_cyl_min = p->n_cyl_max = 20; p->box = 5000.0;
        p->n_linkers = 8000; p->n_motors = 4000; p->xl_mode = MDN_XL_NL_NEIGHBOURS;
        if (config == 2) p->lognormal_sigma = 0.2;
        break;
    case 3: {
        mt64 r;
        mt64_seed(&r, 3 + (uint64_t)index);
        p->seed = 3 + (uint64_t)index;
        p->n_fil = irand(&r, 20, 400);
        int nc = irand(&r, 5, 30);
        p->n_cyl_min = p->n_cyl_max = nc;
        /* box side for a cylinder density of 0.1 um^-3 (as written) or 100 um^-3 */
        double density = dense ? 100.0 : 0.1;             /* cylinders per um^3 */
        double vol_um3 = (double)(p->n_fil * nc) / density;
        double side = cbrt(vol_um3) * 1000.0;
        double minside = 108.0 * (nc + 2) + 100.0;   /* a filament must fit diagonally */
        if (side < minside / sqrt(3.0) + 100.0) side = minside / sqrt(3.0) + 100.0;
        p->box = side;
        p->linkers_per_fil = 2; p->motors_per_fil = 1;
        p->xl_mode = MDN_XL_ANY_DISTINCT;
        p->motor_eql_min = -1.0;
        break;
    }
    case 4:
        p->seed = 4; p->n_fil = 20000; p->n_cyl_min = p->n_cyl_max = 20;
        p->boundary = MDN_SYNTH_SPHERE; p->radius = 10000.0; p->box = 20000.0;
        p->n_linkers = 80000; p->n_motors = 40000; p->xl_mode = MDN_XL_NL_NEIGHBOURS;
        break;
    default:
        return MDN_ERR_ARG;
    }
    return MDN_OK;
}

/* closest distance between segments [p0,p1] and [q0,q1] (clamped parametric solve) */
static double seg_seg_dist(const double* p0, const double* p1, const double* q0, const double* q1) {
    double d1[3], d2[3], r[3];
    for (int k = 0; k < 3; k++) {
        d1[k] = p1[k] - p0[k];
        d2[k] = q1[k] - q0[k];
        r[k] = p0[k] - q0[k];
    }
    double a = d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2];
    double e = d2[0] * d2[0] + d2[1] * d2[1] + d2[2] * d2[2];
    double f = d2[0] * r[0] + d2[1] * r[1] + d2[2] * r[2];
    double c = d1[0] * r[0] + d1[1] * r[1] + d1[2] * r[2];
    double b = d1[0] * d2[0] + d1[1] * d2[1] + d1[2] * d2[2];
    double den = a * e - b * b, s, t;
    s = den > 1e-12 * a * e ? (b * f - c * e) / den : 0.0;
    if (s < 0) s = 0; else if (s > 1) s = 1;
    t = (b * s + f) / e;
    if (t < 0) { t = 0; s = -c / a; }
    else if (t > 1) { t = 1; s = (b - c) / a; }
    if (s < 0) s = 0; else if (s > 1) s = 1;
    double dx = 0;
    for (int k = 0; k < 3; k++) {
        double v = (p0[k] + d1[k] * s) - (q0[k] + d2[k] * t);
        dx += v * v;
    }
    return sqrt(dx);
}

static int bead_inside(const mdn_synth_params* p, const double* x) {
    double m = p->wall_margin;
    if (p->boundary == MDN_SYNTH_CUBE) {
        for (int k = 0; k < 3; k++)
            if (x[k] < 25.0 + m || x[k] > p->box - 25.0 - m) return 0;
        return 1;
    }
    double c = p->box / 2, d2 = 0;
    for (int k = 0; k < 3; k++) d2 += (x[k] - c) * (x[k] - c);
    return sqrt(d2) <= p->radius - m;
}

static void interp(const double* a, const double* b, double t, double* out) {
    for (int k = 0; k < 3; k++) out[k] = a[k] * (1 - t) + t * b[k];
}

static double dist3(const double* a, const double* b) {
    double s = 0;
    for (int k = 0; k < 3; k++) s += (a[k] - b[k]) * (a[k] - b[k]);
    return sqrt(s);
}

/* midpoint-neighbour candidate pairs (c > c', distinct filaments, |m - m'| <= r) via a
 * cell list; returns a malloc'd array of 2*n ints */
static int* neighbour_pairs(struct mdn_synth* s, double r, long* n_out) {
    int nc = s->n_cyl;
    double* mid = malloc(sizeof(double) * 3 * (nc ? nc : 1));
    int c = 0;
    for (int f = 0; f < s->p.n_fil; f++) {
        int off = s->fil_bead_offset[f], nb = s->fil_bead_offset[f + 1] - off;
        for (int k = 0; k < nb - 1; k++, c++) {
            const double* x1 = &s->coord[3 * s->fil_beads[off + k]];
            const double* x2 = &s->coord[3 * s->fil_beads[off + k + 1]];
            for (int q = 0; q < 3; q++) mid[3 * c + q] = 0.5 * (x1[q] + x2[q]);
        }
    }
    int nb = (int)ceil(s->p.box / r);
    if (nb < 1) nb = 1;
    long nbins = (long)nb * nb * nb;
    int* start = calloc(nbins + 1, sizeof(int));
    int* bin = malloc(sizeof(int) * (nc ? nc : 1));
    int* mem = malloc(sizeof(int) * (nc ? nc : 1));
    for (c = 0; c < nc; c++) {
        int ix[3];
        for (int q = 0; q < 3; q++) {
            ix[q] = (int)(mid[3 * c + q] / r);
            if (ix[q] < 0) ix[q] = 0;
            if (ix[q] >= nb) ix[q] = nb - 1;
        }
        bin[c] = ix[0] + nb * (ix[1] + nb * ix[2]);
        start[bin[c] + 1]++;
    }
    for (long b = 0; b < nbins; b++) start[b + 1] += start[b];
    int* fill = malloc(sizeof(int) * nbins);
    memcpy(fill, start, sizeof(int) * nbins);
    for (c = 0; c < nc; c++) mem[fill[bin[c]]++] = c;
    free(fill);
    long cap = 1024, n = 0;
    int* out = malloc(sizeof(int) * 2 * cap);
    for (c = 0; c < nc; c++) {
        int b = bin[c], bx = b % nb, by = (b / nb) % nb, bz = b / (nb * nb);
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int x = bx + dx, y = by + dy, z = bz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nb || y >= nb || z >= nb) continue;
                    int nbn = x + nb * (y + nb * z);
                    for (int m = start[nbn]; m < start[nbn + 1]; m++) {
                        int cn = mem[m];
                        if (cn >= c) break;
                        if (s->cyl_fil[cn] == s->cyl_fil[c]) continue;
                        if (dist3(&mid[3 * c], &mid[3 * cn]) > r) continue;
                        if (n == cap) {
                            cap *= 2;
                            out = realloc(out, sizeof(int) * 2 * cap);
                        }
                        out[2 * n] = c;
                        out[2 * n + 1] = cn;
                        n++;
                    }
                }
    }
    free(mid); free(start); free(bin); free(mem);
    *n_out = n;
    return out;
}

static void cyl_beads(const struct mdn_synth* s, int c, const double** x1, const double** x2) {
    int f = s->cyl_fil[c];
    int k = c - (s->fil_bead_offset[f] - f);
    int off = s->fil_bead_offset[f];
    *x1 = &s->coord[3 * s->fil_beads[off + k]];
    *x2 = &s->coord[3 * s->fil_beads[off + k + 1]];
}

static void make_xlink(struct mdn_synth* s, int c1, int c2, double k, double eql_min,
                       double eql_max, mdn_crosslink* out) {
    const double *a1, *a2, *b1, *b2;
    out->cyl1 = c1;
    out->cyl2 = c2;
    out->pos1 = urand(&s->rng);
    out->pos2 = urand(&s->rng);
    out->kstr = k;
    if (eql_min < 0) {
        double v1[3], v2[3];
        cyl_beads(s, c1, &a1, &a2);
        cyl_beads(s, c2, &b1, &b2);
        interp(a1, a2, out->pos1, v1);
        interp(b1, b2, out->pos2, v2);
        out->eql = dist3(v1, v2);     /* rest length = distance at binding (Linker.cpp:71) */
    } else {
        out->eql = eql_min + (eql_max - eql_min) * urand(&s->rng);
    }
}

int mdn_synth_create(const mdn_synth_params* pin, mdn_synth** out) {
    struct mdn_synth* s = calloc(1, sizeof(*s));
    if (!s) return MDN_ERR_NOMEM;
    s->p = *pin;
    const mdn_synth_params* p = &s->p;
    mt64_seed(&s->rng, p->seed);
    int nf = p->n_fil;
    s->fil_bead_offset = malloc(sizeof(int32_t) * (nf + 1));
    s->fil_type = calloc(nf ? nf : 1, sizeof(int32_t));
    int* ncyl = malloc(sizeof(int) * (nf ? nf : 1));
    int tot_beads = 0;
    s->fil_bead_offset[0] = 0;
    for (int f = 0; f < nf; f++) {
        ncyl[f] = p->n_cyl_min == p->n_cyl_max ? p->n_cyl_min : irand(&s->rng, p->n_cyl_min, p->n_cyl_max);
        tot_beads += ncyl[f] + 1;
        s->fil_bead_offset[f + 1] = tot_beads;
    }
    s->n_cyl = tot_beads - nf;
    s->coord = malloc(sizeof(double) * 3 * (tot_beads ? tot_beads : 1));
    s->fil_beads = malloc(sizeof(int32_t) * (tot_beads ? tot_beads : 1));
    s->cyl_fil = malloc(sizeof(int) * (s->n_cyl ? s->n_cyl : 1));
    for (int b = 0; b < tot_beads; b++) s->fil_beads[b] = b;   /* filament-contiguous beads */
    int c = 0;
    for (int f = 0; f < nf; f++) {
        int n = ncyl[f];
        double* x = &s->coord[3 * s->fil_bead_offset[f]];
        for (int attempt = 0;; attempt++) {
            double z = 2.0 * urand(&s->rng) - 1.0, phi = 6.283185307179586 * urand(&s->rng);
            double rr = sqrt(1.0 - z * z);
            double u[3] = {rr * cos(phi), rr * sin(phi), z};
            double ctr[3];
            for (int k = 0; k < 3; k++) ctr[k] = p->box * urand(&s->rng);
            int ok = 1;
            for (int b = 0; b <= n && ok; b++) {
                double t = (b - 0.5 * n) * p->cyl_len;
                for (int k = 0; k < 3; k++) x[3 * b + k] = ctr[k] + t * u[k];
                ok = bead_inside(p, &x[3 * b]);
            }
            if (ok) break;
            if (attempt > 100000) {   /* box too small for the filament */
                free(ncyl);
                mdn_synth_destroy(s);
                return MDN_ERR_ARG;
            }
        }
        for (int k = 0; k < n; k++) s->cyl_fil[c++] = f;
    }
    free(ncyl);
    /* per-cylinder parameters (MCylinder.cpp:22-43: k scales with cylinderSize / eqLength) */
    int nc = s->n_cyl;
    s->kstr = malloc(sizeof(double) * (nc ? nc : 1));
    s->eql = malloc(sizeof(double) * (nc ? nc : 1));
    s->kbend = malloc(sizeof(double) * (nc ? nc : 1));
    s->eqt = calloc(nc ? nc : 1, sizeof(double));
    s->kexv = malloc(sizeof(double) * (nc ? nc : 1));
    for (c = 0; c < nc; c++) {
        if (p->lognormal_sigma > 0) {
            s->kstr[c] = p->k_stretch * exp(p->lognormal_sigma * nrand(&s->rng));
            s->eql[c] = p->cyl_len * exp(p->lognormal_sigma * nrand(&s->rng));
            s->kbend[c] = p->k_bend * exp(p->lognormal_sigma * nrand(&s->rng));
        } else {
            s->eql[c] = p->cyl_len;
            s->kstr[c] = p->k_stretch * (p->cyl_len / s->eql[c]);
            s->kbend[c] = p->k_bend * (p->cyl_len / s->eql[c]);
        }
        s->kexv[c] = p->k_exvol;
    }
    /* crosslinks */
    int n_l = p->linkers_per_fil > 0 ? p->linkers_per_fil * nf : p->n_linkers;
    int n_m = p->motors_per_fil > 0 ? p->motors_per_fil * nf : p->n_motors;
    s->linkers = calloc(n_l ? n_l : 1, sizeof(mdn_crosslink));
    s->motors = calloc(n_m ? n_m : 1, sizeof(mdn_crosslink));
    int got_l = 0, got_m = 0;
    if (p->xl_mode == MDN_XL_CLOSE_SEGMENTS) {
        long cap = 256, n = 0;
        int* cand = malloc(sizeof(int) * 2 * cap);
        for (int c1 = 0; c1 < nc; c1++)
            for (int c2 = 0; c2 < c1; c2++) {
                if (s->cyl_fil[c1] == s->cyl_fil[c2]) continue;
                const double *a1, *a2, *b1, *b2;
                cyl_beads(s, c1, &a1, &a2);
                cyl_beads(s, c2, &b1, &b2);
                if (seg_seg_dist(a1, a2, b1, b2) >= p->xl_max_dist) continue;
                if (n == cap) { cap *= 2; cand = realloc(cand, sizeof(int) * 2 * cap); }
                cand[2 * n] = c1; cand[2 * n + 1] = c2; n++;
            }
        /* draw without replacement (partial Fisher-Yates) */
        for (int j = 0; j < n_l && j < n; j++) {
            long pick = j + (long)(urand(&s->rng) * (double)(n - j));
            int t0 = cand[2 * j], t1 = cand[2 * j + 1];
            cand[2 * j] = cand[2 * pick]; cand[2 * j + 1] = cand[2 * pick + 1];
            cand[2 * pick] = t0; cand[2 * pick + 1] = t1;
            make_xlink(s, cand[2 * j], cand[2 * j + 1], p->k_linker, -1, -1, &s->linkers[got_l++]);
        }
        for (int j = 0; j < n_m && n > 0; j++) {
            long pick = (long)(urand(&s->rng) * (double)n);
            make_xlink(s, cand[2 * pick], cand[2 * pick + 1], p->k_motor, p->motor_eql_min,
                       p->motor_eql_max, &s->motors[got_m++]);
        }
        free(cand);
    } else if (p->xl_mode == MDN_XL_NL_NEIGHBOURS) {
        long n = 0;
        int* cand = neighbour_pairs(s, p->exvol_cutoff, &n);
        for (int j = 0; j < n_l && n > 0; j++) {
            long pick = (long)(urand(&s->rng) * (double)n);
            make_xlink(s, cand[2 * pick], cand[2 * pick + 1], p->k_linker, -1, -1, &s->linkers[got_l++]);
        }
        for (int j = 0; j < n_m && n > 0; j++) {
            long pick = (long)(urand(&s->rng) * (double)n);
            make_xlink(s, cand[2 * pick], cand[2 * pick + 1], p->k_motor, p->motor_eql_min,
                       p->motor_eql_max, &s->motors[got_m++]);
        }
        free(cand);
    } else {
        for (int which = 0; which < 2 && nf > 1; which++) {
            int want = which ? n_m : n_l;
            for (int j = 0; j < want; j++) {
                int c1, c2;
                do {
                    c1 = irand(&s->rng, 0, nc - 1);
                    c2 = irand(&s->rng, 0, nc - 1);
                } while (s->cyl_fil[c1] == s->cyl_fil[c2]);
                if (which)
                    make_xlink(s, c1, c2, p->k_motor, p->motor_eql_min, p->motor_eql_max, &s->motors[got_m++]);
                else
                    make_xlink(s, c1, c2, p->k_linker, -1, -1, &s->linkers[got_l++]);
            }
        }
    }
    /* boundary elements */
    int nbnd = 0;
    if (p->boundary == MDN_SYNTH_CUBE) {
        /* BoundaryCubic (BoundaryImpl.cpp:25-53): planes at 25 and size - 25 */
        double lo = 25.0, hi = p->box - 25.0;
        double pts[6][3] = {{lo, hi / 2, hi / 2}, {hi, hi / 2, hi / 2}, {hi / 2, lo, hi / 2},
                            {hi / 2, hi, hi / 2}, {hi / 2, hi / 2, lo}, {hi / 2, hi / 2, hi}};
        double nrm[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int e = 0; e < 6; e++) {
            mdn_boundary_element* be = &s->bnd[nbnd++];
            memset(be, 0, sizeof(*be));
            be->kind = MDN_BND_PLANE;
            be->p[0] = nrm[e][0]; be->p[1] = nrm[e][1]; be->p[2] = nrm[e][2];
            be->p[3] = -be->p[0] * pts[e][0] - be->p[1] * pts[e][1] - be->p[2] * pts[e][2];
            be->krep = p->k_boundary;
            be->slen = p->screen_len;
        }
    } else {
        /* BoundarySpherical (BoundaryImpl.cpp:364-375): centre at the box centre */
        mdn_boundary_element* be = &s->bnd[nbnd++];
        memset(be, 0, sizeof(*be));
        be->kind = MDN_BND_SPHERE;
        be->p[0] = be->p[1] = be->p[2] = p->box / 2;
        be->p[3] = p->radius;
        be->krep = p->k_boundary;
        be->slen = p->screen_len;
    }
    mdn_system_desc* d = &s->desc;
    d->n_beads = tot_beads;
    d->n_filaments = nf;
    d->fil_bead_offset = s->fil_bead_offset;
    d->fil_beads = s->fil_beads;
    d->fil_type = s->fil_type;
    d->cyl_kstr = s->kstr;
    d->cyl_eql = s->eql;
    d->cyl_kbend = s->kbend;
    d->cyl_eqtheta = s->eqt;
    d->cyl_kexvol = s->kexv;
    d->n_linkers = got_l;
    d->linkers = s->linkers;
    d->n_motors = got_m;
    d->motors = s->motors;
    d->n_boundary = nbnd;
    d->boundary = s->bnd;
    d->boundary_cutoff = p->boundary_cutoff;
    d->exvol_rmax = p->exvol_cutoff;
    d->exvol_rmin = 0.0;
    d->grid_size[0] = d->grid_size[1] = d->grid_size[2] = p->box;
    *out = s;
    return MDN_OK;
}

void mdn_synth_destroy(mdn_synth* s) {
    if (!s) return;
    free(s->coord); free(s->fil_bead_offset); free(s->fil_beads); free(s->fil_type);
    free(s->kstr); free(s->eql); free(s->kbend); free(s->eqt); free(s->kexv);
    free(s->linkers); free(s->motors); free(s->cyl_fil);
    free(s);
}

const mdn_system_desc* mdn_synth_desc(const mdn_synth* s) { return &s->desc; }
const double* mdn_synth_coord(const mdn_synth* s) { return s->coord; }

int mdn_synth_perturb(mdn_synth* s, double* coord, double sigma) {
    for (int i = 0; i < 3 * s->desc.n_beads; i++) coord[i] += sigma * nrand(&s->rng);
    return MDN_OK;
}
