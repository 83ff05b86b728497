/*
 * nl.c -- oracle restatement of the cylinder-cylinder neighbor list predicate.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * HybridCylinderCylinderNL::updateNeighborsbin (HybridNeighborListImpl.cpp:332-471) keeps, for
 * the current cylinder c and a candidate c' from a stencil bin:
 *   c.id > c'.id                                                         (:413)
 *   not (same filament and |pos - pos'| <= minCylinderDistanceSameFilament = 2) (:415-418)
 *   not (d2 < rMin^2) and not (d2 > rMax^2)                              (:429-439)
 * with d2 = distance2(mid_c, mid_c') in double (Vec.hpp:377-384) and the squared bounds
 * formed in float (HybridNeighborListImpl.h:79-80). The bin stencil only prunes candidates
 * conservatively (BinGrid.h:134-198), so the pair set is the predicate's set.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static inline int keep_pair(const double* mid, const int* fil, const int* pos, int c, int cn,
                            double rmaxsq, double rminsq) {
    if (c <= cn) return 0;
    if (fil[c] == fil[cn]) {
        int sep = abs(pos[c] - pos[cn]);
        if (sep <= 2) return 0;
    }
    const double* m = &mid[3 * c];
    const double* n = &mid[3 * cn];
    double d2 = 0.0;
    for (int k = 0; k < 3; k++) d2 += (n[k] - m[k]) * (n[k] - m[k]);
    if (d2 < rminsq || d2 > rmaxsq) return 0;
    return 1;
}

long or_nl_brute(const double* mid, const int* fil, const int* pos, int n_cyl, float rmax,
                 float rmin, int* pairs, long cap, long* needed) {
    float rmaxsq_f = rmax * rmax, rminsq_f = rmin * rmin;
    double rmaxsq = rmaxsq_f, rminsq = rminsq_f;
    long n = 0;
    for (int c = 0; c < n_cyl; c++)
        for (int cn = 0; cn < c; cn++)
            if (keep_pair(mid, fil, pos, c, cn, rmaxsq, rminsq)) {
                if (n < cap) {
                    pairs[2 * n] = c;
                    pairs[2 * n + 1] = cn;
                }
                n++;
            }
    if (needed) *needed = n;
    return n <= cap ? n : -1;
}

static int cmp_int(const void* a, const void* b) {
    int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

long or_nl_cells(const double* mid, const int* fil, const int* pos, int n_cyl, float rmax,
                 float rmin, const double* grid_size, int* pairs, long cap, long* needed) {
    float rmaxsq_f = rmax * rmax, rminsq_f = rmin * rmin;
    double rmaxsq = rmaxsq_f, rminsq = rminsq_f;
    double side = 1.125 * sqrt((double)rmaxsq_f);
    int nb[3];
    for (int k = 0; k < 3; k++) {
        nb[k] = (int)ceil(grid_size[k] / side);
        if (nb[k] < 1) nb[k] = 1;
    }
    long nbins = (long)nb[0] * nb[1] * nb[2];
    int* bin_of = (int*)malloc(sizeof(int) * (n_cyl > 0 ? n_cyl : 1));
    int* start = (int*)calloc(nbins + 1, sizeof(int));
    int* members = (int*)malloc(sizeof(int) * (n_cyl > 0 ? n_cyl : 1));
    for (int c = 0; c < n_cyl; c++) {
        int idx[3];
        for (int k = 0; k < 3; k++) {
            double x = mid[3 * c + k];
            if (!(x >= 0 && x < grid_size[k])) {  /* getBin throws outside the grid (:243-286) */
                free(bin_of); free(start); free(members);
                return -2;
            }
            idx[k] = (int)(x / side);
            if (idx[k] >= nb[k]) idx[k] = nb[k] - 1;
        }
        bin_of[c] = idx[0] + nb[0] * (idx[1] + nb[1] * idx[2]);
        start[bin_of[c] + 1]++;
    }
    for (long b = 0; b < nbins; b++) start[b + 1] += start[b];
    int* fill = (int*)malloc(sizeof(int) * nbins);
    memcpy(fill, start, sizeof(int) * nbins);
    for (int c = 0; c < n_cyl; c++) members[fill[bin_of[c]]++] = c;  /* ascending within bin */
    free(fill);

    int cand_cap = 256;
    int* cand = (int*)malloc(sizeof(int) * cand_cap);
    long n = 0;
    for (int c = 0; c < n_cyl; c++) {
        int b = bin_of[c];
        int bx = b % nb[0], by = (b / nb[0]) % nb[1], bz = b / (nb[0] * nb[1]);
        int nc = 0;
        for (int dz = -1; dz <= 1; dz++)
            for (int dy = -1; dy <= 1; dy++)
                for (int dx = -1; dx <= 1; dx++) {
                    int x = bx + dx, y = by + dy, z = bz + dz;
                    if (x < 0 || y < 0 || z < 0 || x >= nb[0] || y >= nb[1] || z >= nb[2]) continue;
                    int nbin = x + nb[0] * (y + nb[1] * z);
                    for (int m = start[nbin]; m < start[nbin + 1]; m++) {
                        int cn = members[m];
                        if (cn >= c) break;
                        if (keep_pair(mid, fil, pos, c, cn, rmaxsq, rminsq)) {
                            if (nc == cand_cap) {
                                cand_cap *= 2;
                                cand = (int*)realloc(cand, sizeof(int) * cand_cap);
                            }
                            cand[nc++] = cn;
                        }
                    }
                }
        qsort(cand, nc, sizeof(int), cmp_int);
        for (int q = 0; q < nc; q++) {
            if (n < cap) {
                pairs[2 * n] = c;
                pairs[2 * n + 1] = cand[q];
            }
            n++;
        }
    }
    free(cand); free(bin_of); free(start); free(members);
    if (needed) *needed = n;
    return n <= cap ? n : -1;
}
