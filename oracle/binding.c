/*
 * binding.c -- oracle restatement of the binding-site pair search. TEST INFRASTRUCTURE ONLY
 * (see oracle.h).
 *
 * HybridBindingSearchManager::updateAllPossibleBindingsstencilHYBD (HybridBindingSearchManager.cpp:
 * 759-966) over the binder's hybrid neighbor list (addtoHNeighborList, :1024-1039): the list holds
 * (c, c') with c > c', not (same filament and |pos - pos'| <= 2), the filament-type pair, and the
 * midpoint distance within [max(rMin - cylSize, 0), cylSize + rMax] (in float, squared in float).
 * For each listed pair on different filaments (and, as the reference tests c's type against the
 * complementary type, of equal types), each free site s1 of c (linker rule: index % step == 0)
 * gives the squared distance from x1 + s1 (x2 - x1) to x3 + t (x4 - x3) as A t^2 + B t + C; the
 * roots at rMin^2 and rMax^2 bound the accepted sites s2 of c'. Site positions and squared bounds
 * are float, the rest double, in the reference's expression order. Brute force over all pairs.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "oracle.h"

static int listed(const or_system* s, const double* mid, int c, int cn, double rmaxsq, double rminsq, int ta, int tb) {
    if (c <= cn) return 0;
    if (s->cyl_fil[c] == s->cyl_fil[cn] && abs(s->cyl_pos[c] - s->cyl_pos[cn]) <= 2) return 0;
    const int t1 = s->cyl_type[c], t2 = s->cyl_type[cn];
    const int lo = t1 < t2 ? t1 : t2, hi = t1 < t2 ? t2 : t1;
    if (lo != ta || hi != tb) return 0;
    double d2 = 0.0;
    for (int k = 0; k < 3; k++) d2 += (mid[3 * cn + k] - mid[3 * c + k]) * (mid[3 * cn + k] - mid[3 * c + k]);
    return !(d2 < rminsq || d2 > rmaxsq);
}

long or_binding_pairs(const or_system* s, const double* coord, const mdn_binding_params* p, uint32_t* out,
                      long cap, long* needed) {
    const int ta = p->ftype1 < p->ftype2 ? p->ftype1 : p->ftype2;
    const int tb = p->ftype1 < p->ftype2 ? p->ftype2 : p->ftype1;
    const float rmaxsq_b = p->rmax * p->rmax, rminsq_b = p->rmin * p->rmin;
    const float rmax_l = p->cyl_size + sqrtf(rmaxsq_b);
    float rmin_l = sqrtf(rminsq_b) - p->cyl_size;
    if (rmin_l < 0.0f) rmin_l = 0.0f;
    const float lmaxsq = rmax_l * rmax_l, lminsq = rmin_l * rmin_l;
    double* mid = malloc(sizeof(double) * 3 * (s->n_cyl > 0 ? s->n_cyl : 1));
    for (int c = 0; c < s->n_cyl; c++)
        for (int k = 0; k < 3; k++) mid[3 * c + k] = 0.5 * (coord[3 * s->cyl_b1[c] + k] + coord[3 * s->cyl_b2[c] + k]);
    long n = 0;
    for (int c = 0; c < s->n_cyl; c++) {
        const int tc = s->cyl_type[c];
        if (tc != ta && tc != tb) continue;
        const int compl_t = tc == ta ? tb : ta;
        for (int cn = 0; cn < c; cn++) {
            if (!listed(s, mid, c, cn, lmaxsq, lminsq, ta, tb)) continue;
            if (s->cyl_fil[c] == s->cyl_fil[cn]) continue;
            if (tc != compl_t) continue;
            const double* x1 = &coord[3 * s->cyl_b1[c]];
            const double* x2 = &coord[3 * s->cyl_b2[c]];
            const double* x3 = &coord[3 * s->cyl_b1[cn]];
            const double* x4 = &coord[3 * s->cyl_b2[cn]];
            const double X1X2[3] = {x2[0] - x1[0], x2[1] - x1[1], x2[2] - x1[2]};
            const double X1X3[3] = {x3[0] - x1[0], x3[1] - x1[1], x3[2] - x1[2]};
            const double X3X4[3] = {x4[0] - x3[0], x4[1] - x3[1], x4[2] - x3[2]};
            const double X1X3squared = X1X3[0] * X1X3[0] + X1X3[1] * X1X3[1] + X1X3[2] * X1X3[2];
            const double X1X2squared = X1X2[0] * X1X2[0] + X1X2[1] * X1X2[1] + X1X2[2] * X1X2[2];
            const double X1X3dotX1X2 = X1X3[0] * X1X2[0] + X1X3[1] * X1X2[1] + X1X3[2] * X1X2[2];
            const double X3X4squared = X3X4[0] * X3X4[0] + X3X4[1] * X3X4[1] + X3X4[2] * X3X4[2];
            const double X1X3dotX3X4 = X1X3[0] * X3X4[0] + X1X3[1] * X3X4[1] + X1X3[2] * X3X4[2];
            const double X3X4dotX1X2 = X3X4[0] * X1X2[0] + X3X4[1] * X1X2[1] + X3X4[2] * X1X2[2];
            for (int pos1 = 0; pos1 < p->n_sites; pos1++) {
                if (p->site_step > 0 && pos1 % p->site_step != 0) continue;
                if (p->site_free && p->site_free[(long)c * p->n_sites + pos1] != 1.0f) continue;
                const float mp1 = p->site_pos[pos1];
                const double A = X3X4squared;
                const double B = 2.0 * X1X3dotX3X4 - 2.0 * mp1 * X3X4dotX1X2;
                const double C = X1X3squared + mp1 * mp1 * X1X2squared - 2.0 * mp1 * X1X3dotX1X2;
                const double Bsq = B * B;
                const double C1 = C - rminsq_b, C2 = C - rmaxsq_b;
                const double b2m4ac1 = Bsq - 4 * A * C1, b2m4ac2 = Bsq - 4 * A * C2;
                const int status1 = b2m4ac1 < 0, status2 = b2m4ac2 < 0;
                if (status1 && status2) continue;
                double minv[2] = {0.0, 0.0}, maxv[2] = {0.0, 0.0};
                if (!status1) {
                    const double m1 = (-B + sqrt(b2m4ac1)) / (2 * A), m2 = (-B - sqrt(b2m4ac1)) / (2 * A);
                    minv[0] = m1 < m2 ? m1 : m2;
                    minv[1] = m1 < m2 ? m2 : m1;
                    if (minv[0] < p->p2_first && minv[1] > p->p2_last) continue;
                }
                if (!status2) {
                    const double m1 = (-B + sqrt(b2m4ac2)) / (2 * A), m2 = (-B - sqrt(b2m4ac2)) / (2 * A);
                    maxv[0] = m1 < m2 ? m1 : m2;
                    maxv[1] = m1 < m2 ? m2 : m1;
                    if (maxv[0] > p->p2_last || maxv[1] < p->p2_first) continue;
                }
                for (int pos2 = 0; pos2 < p->n_sites; pos2++) {
                    if (p->site_step > 0 && pos2 % p->site_step != 0) continue;
                    if (p->site_free && p->site_free[(long)cn * p->n_sites + pos2] != 1.0f) continue;
                    const float mp2 = p->site_pos[pos2];
                    if (!status2 && (mp2 < maxv[0] || mp2 > maxv[1])) continue;
                    if (!status1 && (mp2 > minv[0] && mp2 < minv[1])) continue;
                    if (n < cap) {
                        out[2 * n] = ((uint32_t)c << p->shift_bits) | (uint32_t)pos1;
                        out[2 * n + 1] = ((uint32_t)cn << p->shift_bits) | (uint32_t)pos2;
                    }
                    n++;
                }
            }
        }
    }
    free(mid);
    if (needed) *needed = n;
    return n <= cap ? n : -1;
}
