/*
 * system.c -- oracle restatement of vectorizeAllForceFields, the boundary neighbor list and
 * the ForceFieldManager loops. TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define OR_ALLOC(p, n) ((p) = calloc((size_t)((n) > 0 ? (n) : 1), sizeof(*(p))))

int or_system_build(const mdn_system_desc* d, or_system* s) {
    memset(s, 0, sizeof(*s));
    if (!d || d->n_beads < 0 || d->n_filaments < 0) return MDN_ERR_ARG;
    s->n_beads = d->n_beads;
    s->n_fil = d->n_filaments;
    int n_cyl = 0;
    for (int f = 0; f < d->n_filaments; f++) {
        int nb = d->fil_bead_offset[f + 1] - d->fil_bead_offset[f];
        if (nb < 2) return MDN_ERR_ARG;
        n_cyl += nb - 1;
    }
    s->n_cyl = n_cyl;
    OR_ALLOC(s->cyl_b1, n_cyl); OR_ALLOC(s->cyl_b2, n_cyl); OR_ALLOC(s->cyl_fil, n_cyl);
    OR_ALLOC(s->cyl_pos, n_cyl); OR_ALLOC(s->cyl_type, n_cyl); OR_ALLOC(s->cyl_minus_end, n_cyl);
    OR_ALLOC(s->cyl_kexvol, n_cyl); OR_ALLOC(s->cyl_eql, n_cyl);
    /* FilamentStretching::vectorize (FilamentStretching.cpp:26-45): one term per cylinder in
     * cylinder order. */
    OR_ALLOC(s->str_beadSet, 2 * n_cyl); OR_ALLOC(s->str_k, n_cyl); OR_ALLOC(s->str_eql, n_cyl);
    int c = 0;
    for (int f = 0; f < d->n_filaments; f++) {
        int off = d->fil_bead_offset[f], nb = d->fil_bead_offset[f + 1] - off;
        for (int k = 0; k < nb - 1; k++, c++) {
            int b1 = d->fil_beads[off + k], b2 = d->fil_beads[off + k + 1];
            if (b1 < 0 || b1 >= d->n_beads || b2 < 0 || b2 >= d->n_beads) return MDN_ERR_ARG;
            s->cyl_b1[c] = b1;
            s->cyl_b2[c] = b2;
            s->cyl_fil[c] = f;
            s->cyl_pos[c] = k;
            s->cyl_type[c] = d->fil_type ? d->fil_type[f] : 0;
            s->cyl_minus_end[c] = (k == 0);
            s->cyl_kexvol[c] = d->cyl_kexvol[c];
            s->cyl_eql[c] = d->cyl_eql[c];
            s->str_beadSet[2 * c] = 3 * b1;
            s->str_beadSet[2 * c + 1] = 3 * b2;
            s->str_k[c] = d->cyl_kstr[c];
            s->str_eql[c] = d->cyl_eql[c];
        }
    }
    /* FilamentBending::vectorize (FilamentBending.cpp:29-60): consecutive cylinder pairs of
     * each filament, k and theta0 from the second cylinder. */
    int n_bend = n_cyl - d->n_filaments;
    s->n_bend = n_bend;
    OR_ALLOC(s->bend_beadSet, 3 * n_bend); OR_ALLOC(s->bend_k, n_bend); OR_ALLOC(s->bend_eqt, n_bend);
    int i = 0;
    c = 0;
    for (int f = 0; f < d->n_filaments; f++) {
        int nb = d->fil_bead_offset[f + 1] - d->fil_bead_offset[f];
        int first = c;
        for (int k = 1; k < nb - 1; k++, i++) {
            int cp = first + k - 1, cc = first + k;
            s->bend_beadSet[3 * i] = 3 * s->cyl_b1[cp];
            s->bend_beadSet[3 * i + 1] = 3 * s->cyl_b1[cc];
            s->bend_beadSet[3 * i + 2] = 3 * s->cyl_b2[cc];
            s->bend_k[i] = d->cyl_kbend[cc];
            s->bend_eqt[i] = d->cyl_eqtheta[cc];
        }
        c += nb - 1;
    }
    /* LinkerStretching::vectorize (LinkerStretching.cpp:30-57), MotorGhostStretching likewise */
    for (int which = 0; which < 2; which++) {
        int n = which ? d->n_motors : d->n_linkers;
        const mdn_crosslink* xl = which ? d->motors : d->linkers;
        int* bs; double *k, *eql, *p1, *p2, *sf;
        OR_ALLOC(bs, 4 * n); OR_ALLOC(k, n); OR_ALLOC(eql, n); OR_ALLOC(p1, n); OR_ALLOC(p2, n);
        OR_ALLOC(sf, n);
        for (int j = 0; j < n; j++) {
            int c1 = xl[j].cyl1, c2 = xl[j].cyl2;
            if (c1 < 0 || c1 >= n_cyl || c2 < 0 || c2 >= n_cyl) return MDN_ERR_ARG;
            bs[4 * j] = 3 * s->cyl_b1[c1];
            bs[4 * j + 1] = 3 * s->cyl_b2[c1];
            bs[4 * j + 2] = 3 * s->cyl_b1[c2];
            bs[4 * j + 3] = 3 * s->cyl_b2[c2];
            k[j] = xl[j].kstr;
            eql[j] = xl[j].eql;
            p1[j] = xl[j].pos1;
            p2[j] = xl[j].pos2;
        }
        if (which) {
            s->n_mot = n; s->mot_beadSet = bs; s->mot_k = k; s->mot_eql = eql; s->mot_p1 = p1;
            s->mot_p2 = p2; s->mot_sf = sf;
        } else {
            s->n_link = n; s->lnk_beadSet = bs; s->lnk_k = k; s->lnk_eql = eql; s->lnk_p1 = p1;
            s->lnk_p2 = p2; s->lnk_sf = sf;
        }
    }
    s->n_elem = d->n_boundary;
    OR_ALLOC(s->elems, d->n_boundary);
    if (d->n_boundary) memcpy(s->elems, d->boundary, sizeof(mdn_boundary_element) * d->n_boundary);
    s->bnd_cutoff = d->boundary_cutoff;
    s->nl_rmax = (float)d->exvol_rmax;
    s->nl_rmin = (float)d->exvol_rmin;
    for (int k = 0; k < 3; k++) s->grid[k] = d->grid_size[k];
    s->use_cells = 1;
    OR_ALLOC(s->bnd_nneighbors, d->n_boundary);
    return MDN_OK;
}

/* Cylinder midpoint 0.5*(x1 + x2) (Cylinder.cpp:45-49) */
static void midpoints(const or_system* s, const double* coord, double* mid) {
    for (int c = 0; c < s->n_cyl; c++) {
        const double* x1 = &coord[3 * s->cyl_b1[c]];
        const double* x2 = &coord[3 * s->cyl_b2[c]];
        for (int k = 0; k < 3; k++) mid[3 * c + k] = 0.5 * (x1[k] + x2[k]);
    }
}

int or_system_nl(or_system* s, const double* coord) {
    double* mid = malloc(sizeof(double) * 3 * (s->n_cyl > 0 ? s->n_cyl : 1));
    midpoints(s, coord, mid);
    /* excluded-volume NL */
    free(s->pairs); free(s->ev_beadSet); free(s->ev_krep); free(s->ev_eqlen);
    long needed = 0;
    long cap = 1024;
    int* pairs = malloc(sizeof(int) * 2 * cap);
    for (;;) {
        long n = s->use_cells
                     ? or_nl_cells(mid, s->cyl_fil, s->cyl_pos, s->n_cyl, s->nl_rmax, s->nl_rmin,
                                   s->grid, pairs, cap, &needed)
                     : or_nl_brute(mid, s->cyl_fil, s->cyl_pos, s->n_cyl, s->nl_rmax, s->nl_rmin,
                                   pairs, cap, &needed);
        if (n == -2) {
            free(pairs); free(mid);
            s->pairs = NULL; s->ev_beadSet = NULL; s->ev_krep = NULL; s->ev_eqlen = NULL;
            s->n_pairs = 0;
            return MDN_ERR_NL_BOUNDS;
        }
        if (n >= 0) break;
        cap = needed;
        pairs = realloc(pairs, sizeof(int) * 2 * (cap > 0 ? cap : 1));
    }
    s->n_pairs = needed;
    s->pairs = pairs;
    /* CylinderExclVolume::vectorize (CylinderExclVolume.cpp:73-138) */
    long np = s->n_pairs;
    OR_ALLOC(s->ev_beadSet, 4 * np); OR_ALLOC(s->ev_krep, np); OR_ALLOC(s->ev_eqlen, 2 * np);
    for (long p = 0; p < np; p++) {
        int ci = pairs[2 * p], cn = pairs[2 * p + 1];
        s->ev_beadSet[4 * p] = 3 * s->cyl_b1[ci];
        s->ev_beadSet[4 * p + 1] = 3 * s->cyl_b2[ci];
        s->ev_beadSet[4 * p + 2] = 3 * s->cyl_b1[cn];
        s->ev_beadSet[4 * p + 3] = 3 * s->cyl_b2[cn];
        s->ev_eqlen[2 * p] = s->cyl_eql[ci];
        s->ev_eqlen[2 * p + 1] = s->cyl_eql[cn];
        if (s->cyl_type[ci] != s->cyl_type[cn])
            s->ev_krep[p] = (s->cyl_kexvol[ci] < s->cyl_kexvol[cn]) ? s->cyl_kexvol[cn] : s->cyl_kexvol[ci];
        else
            s->ev_krep[p] = s->cyl_kexvol[ci];
    }
    /* BoundaryCylinderNL::updateNeighbors (NeighborListImpl.cpp:738-752) and
     * BoundaryCylinderRepulsion::vectorize (BoundaryCylinderRepulsion.cpp:32-87) */
    free(s->bnd_beadSet); free(s->bnd_krep); free(s->bnd_slen);
    int nint = 0;
    for (int e = 0; e < s->n_elem; e++)
        for (int c = 0; c < s->n_cyl; c++)
            if (or_boundary_distance(&s->elems[e], &mid[3 * c]) < s->bnd_cutoff)
                nint += s->cyl_minus_end[c] ? 2 : 1;
    s->n_bint = nint;
    OR_ALLOC(s->bnd_beadSet, nint); OR_ALLOC(s->bnd_krep, nint); OR_ALLOC(s->bnd_slen, nint);
    free(s->bnd_member);
    OR_ALLOC(s->bnd_member, (long)s->n_elem * s->n_cyl);
    int cum = 0;
    for (int e = 0; e < s->n_elem; e++) {
        int idx = 0;
        for (int c = 0; c < s->n_cyl; c++) {
            if (!(or_boundary_distance(&s->elems[e], &mid[3 * c]) < s->bnd_cutoff)) continue;
            s->bnd_member[(long)e * s->n_cyl + c] = 1;
            if (s->cyl_minus_end[c]) {
                s->bnd_beadSet[cum + idx] = 3 * s->cyl_b1[c];
                s->bnd_krep[cum + idx] = s->elems[e].krep;
                s->bnd_slen[cum + idx] = s->elems[e].slen;
                idx++;
            }
            s->bnd_beadSet[cum + idx] = 3 * s->cyl_b2[c];
            s->bnd_krep[cum + idx] = s->elems[e].krep;
            s->bnd_slen[cum + idx] = s->elems[e].slen;
            idx++;
        }
        s->bnd_nneighbors[e] = idx;
        cum += idx;
    }
    free(mid);
    return MDN_OK;
}

/* boundaryCylinderRepulsionLoadForce (BoundaryCylinderRepulsion.cpp:125-162) for bead bd with
 * its cylinder neighbour bo, accumulated into out[0 .. nmon) */
static void load_force_end(const mdn_boundary_element* be, const double* xd, const double* xo, double msize,
                           int nmon, double* out) {
    /* dir = normalizedVector(bd - bo): magnitude2 summed from zero, then each component / norm */
    double dir[3] = {xd[0] - xo[0], xd[1] - xo[1], xd[2] - xo[2]};
    const double norm = sqrt(dir[0] * dir[0] + dir[1] * dir[1] + dir[2] * dir[2]);
    for (int k = 0; k < 3; k++) dir[k] /= norm;
    for (int i = 0; i < nmon; i++) {
        const double step = i * msize;
        const double p[3] = {xd[0] + step * dir[0], xd[1] + step * dir[1], xd[2] + step * dir[2]};
        double n[3];
        or_boundary_normal(be, p, n);
        const double dotnd = n[0] * dir[0] + n[1] * dir[1] + n[2] * dir[2];
        const double proj = (-dotnd < 0.0) ? 0.0 : -dotnd;   /* std::max(-dot, 0.0) */
        /* BoundaryCylinderRepulsionExp::loadForces (BoundaryCylinderRepulsionExp.cpp:113-118) */
        const double R = -or_boundary_distance(be, p) / be->slen;
        const double lf = be->krep * exp(R) / be->slen;
        out[i] += proj * lf;
    }
}

int or_system_load_forces(const or_system* s, const double* coord, int n_types, const double* mon_size,
                          const int* num_mon, int stride, double* out_plus, double* out_minus) {
    for (int t = 0; t < n_types; t++)
        if (num_mon[t] < 0 || num_mon[t] > stride) return MDN_ERR_ARG;
    memset(out_plus, 0, sizeof(double) * (size_t)s->n_fil * stride);
    memset(out_minus, 0, sizeof(double) * (size_t)s->n_fil * stride);
    if (!s->bnd_member) return MDN_OK;
    /* the last cylinder of each filament is its plus end (Cylinder::isPlusEnd) */
    int* last = calloc((size_t)(s->n_fil > 0 ? s->n_fil : 1), sizeof(int));
    for (int c = 0; c < s->n_cyl; c++) last[s->cyl_fil[c]] = c;
    int st = MDN_OK;
    for (int e = 0; e < s->n_elem && st == MDN_OK; e++) {
        for (int c = 0; c < s->n_cyl; c++) {
            if (!s->bnd_member[(long)e * s->n_cyl + c]) continue;
            const int f = s->cyl_fil[c], t = s->cyl_type[c];
            if (t < 0 || t >= n_types) {
                st = MDN_ERR_ARG;
                break;
            }
            const double* x1 = &coord[3 * s->cyl_b1[c]];
            const double* x2 = &coord[3 * s->cyl_b2[c]];
            if (last[f] == c) load_force_end(&s->elems[e], x2, x1, mon_size[t], num_mon[t], &out_plus[(long)f * stride]);
            if (s->cyl_minus_end[c])
                load_force_end(&s->elems[e], x1, x2, mon_size[t], num_mon[t], &out_minus[(long)f * stride]);
        }
    }
    free(last);
    return st;
}

void or_system_free(or_system* s) {
    void* ptrs[] = {s->cyl_b1, s->cyl_b2, s->cyl_fil, s->cyl_pos, s->cyl_type, s->cyl_minus_end,
                    s->cyl_kexvol, s->cyl_eql, s->str_beadSet, s->str_k, s->str_eql,
                    s->bend_beadSet, s->bend_k, s->bend_eqt, s->lnk_beadSet, s->lnk_k,
                    s->lnk_eql, s->lnk_p1, s->lnk_p2, s->lnk_sf, s->mot_beadSet, s->mot_k,
                    s->mot_eql, s->mot_p1, s->mot_p2, s->mot_sf, s->pairs, s->ev_beadSet,
                    s->ev_krep, s->ev_eqlen, s->elems, s->bnd_beadSet, s->bnd_krep,
                    s->bnd_slen, s->bnd_nneighbors, s->bnd_member};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    memset(s, 0, sizeof(*s));
}

double or_ff_energy(const or_system* s, int ff, const double* coord) {
    switch (ff) {
    case MDN_FF_STRETCH: return or_stretch_energy(coord, s->str_beadSet, s->str_k, s->str_eql, s->n_cyl);
    case MDN_FF_BEND: return or_bend_energy(coord, s->n_bend, s->bend_beadSet, s->bend_k, s->bend_eqt);
    case MDN_FF_LINKER:
        return or_xlink_energy(coord, s->lnk_beadSet, s->lnk_k, s->lnk_eql, s->lnk_p1, s->lnk_p2, s->n_link);
    case MDN_FF_MOTOR:
        return or_xlink_energy(coord, s->mot_beadSet, s->mot_k, s->mot_eql, s->mot_p1, s->mot_p2, s->n_mot);
    case MDN_FF_EXVOL:
        return or_exvol_energy(coord, s->ev_beadSet, s->ev_krep, s->ev_eqlen, (int)s->n_pairs);
    case MDN_FF_BOUNDARY:
        return or_boundary_energy(coord, s->bnd_beadSet, s->bnd_krep, s->bnd_slen, s->bnd_nneighbors,
                                  s->elems, s->n_elem);
    }
    return 0.0;
}

void or_ff_forces(or_system* s, int ff, const double* coord, double* f) {
    switch (ff) {
    case MDN_FF_STRETCH: or_stretch_forces(coord, f, s->str_beadSet, s->str_k, s->str_eql, s->n_cyl); break;
    case MDN_FF_BEND: or_bend_forces(coord, f, s->n_bend, s->bend_beadSet, s->bend_k, s->bend_eqt); break;
    case MDN_FF_LINKER:
        or_xlink_forces(coord, f, s->lnk_beadSet, s->lnk_k, s->lnk_eql, s->lnk_p1, s->lnk_p2, s->lnk_sf, s->n_link);
        break;
    case MDN_FF_MOTOR:
        or_xlink_forces(coord, f, s->mot_beadSet, s->mot_k, s->mot_eql, s->mot_p1, s->mot_p2, s->mot_sf, s->n_mot);
        break;
    case MDN_FF_EXVOL: or_exvol_forces(coord, f, s->ev_beadSet, s->ev_krep, s->ev_eqlen, (int)s->n_pairs); break;
    case MDN_FF_BOUNDARY:
        or_boundary_forces(coord, f, s->bnd_beadSet, s->bnd_krep, s->bnd_slen, s->bnd_nneighbors, s->elems,
                           s->n_elem);
        break;
    }
}

double or_ffm_energy(const or_system* s, const double* coord, mdn_energy_report* rep) {
    /* computeEnergy (ForceFieldManager.cpp:96-207) returns +inf at the first FF energy <= -1;
     * computeIndividualEnergies (:209-227) sums every FF without the sentinel. */
    double energy = 0.0, total = 0.0;
    int failed = 0;
    for (int ff = 0; ff < MDN_NUM_FF; ff++) {
        double e = or_ff_energy(s, ff, coord);
        if (rep) rep->ff[ff] = e;
        total += e;
        if (!failed) {
            if (e <= -1) failed = 1;
            else energy += e;
        }
        if (failed && !rep) break;
    }
    if (rep) rep->total = total;
    return failed ? INFINITY : energy;
}

void or_ffm_forces(or_system* s, const double* coord, double* f) {
    /* computeForces (ForceFieldManager.cpp:242-290) */
    memset(f, 0, sizeof(double) * 3 * (size_t)s->n_beads);
    for (int ff = 0; ff < MDN_NUM_FF; ff++) or_ff_forces(s, ff, coord, f);
}

/* ---- handle API for language bindings ------------------------------------------------ */
or_system* or_system_create(const mdn_system_desc* d, int* status) {
    or_system* s = calloc(1, sizeof(or_system));
    int st = or_system_build(d, s);
    if (status) *status = st;
    if (st != MDN_OK) {
        or_system_free(s);
        free(s);
        return NULL;
    }
    return s;
}
void or_system_destroy(or_system* s) {
    if (!s) return;
    or_system_free(s);
    free(s);
}
void or_system_set_cells(or_system* s, int use_cells) { s->use_cells = use_cells; }
long or_system_npairs(const or_system* s) { return s->n_pairs; }
const int* or_system_pairs(const or_system* s) { return s->pairs; }
int or_system_nbint(const or_system* s) { return s->n_bint; }
const int* or_system_bnd_beadset(const or_system* s) { return s->bnd_beadSet; }
const int* or_system_bnd_nneighbors(const or_system* s) { return s->bnd_nneighbors; }
int or_system_ncyl(const or_system* s) { return s->n_cyl; }
int or_system_nbend(const or_system* s) { return s->n_bend; }
const double* or_system_stretchforce(const or_system* s, int motor) {
    return motor ? s->mot_sf : s->lnk_sf;
}
