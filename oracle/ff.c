/*
 * ff.c -- oracle restatement of the six force-field interaction kernels.
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Expression trees follow the cited reference lines
 * so that IEEE evaluation without FMA contraction rounds the same way.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* mathfunc::twoPointDistance (MathFunctions.h:294-297): sqrt of ((dx*dx + dy*dy) + dz*dz)
 * with d = v2 - v1. */
static inline double two_point_distance(const double* v1, const double* v2) {
    return sqrt((v2[0] - v1[0]) * (v2[0] - v1[0]) + (v2[1] - v1[1]) * (v2[1] - v1[1]) +
                (v2[2] - v1[2]) * (v2[2] - v1[2]));
}

/* mathfunc::scalarProduct (MathFunctions.h:469-477): (v2 - v1) . (v4 - v3) */
static inline double scalar_product(const double* v1, const double* v2, const double* v3,
                                    const double* v4) {
    return ((v2[0] - v1[0]) * (v4[0] - v3[0]) + (v2[1] - v1[1]) * (v4[1] - v3[1]) +
            (v2[2] - v1[2]) * (v4[2] - v3[2]));
}

/* mathfunc::midPointCoordinate (MathFunctions.h:963-969): v = v1 * (1 - alpha) + alpha * v2 */
static inline void mid_point(double* v, const double* v1, const double* v2, double alpha) {
    double beta = 1 - alpha;
    v[0] = (v1[0] * beta + alpha * v2[0]);
    v[1] = (v1[1] * beta + alpha * v2[1]);
    v[2] = (v1[2] * beta + alpha * v2[2]);
}

/* std::max / std::min semantics: max(a, b) = (a < b) ? b : a, min(a, b) = (b < a) ? b : a */
static inline double rmax(double a, double b) { return (a < b) ? b : a; }

/* areEqual with ZERO_PREC = 1e-6 for double (src/utility.h:58-77) */
static inline int are_equal(double a, double b) { return fabs(a - b) < 1e-6; }

/* checkNaN_INF (MathFunctions.h:1376-1384) on three components */
static inline int bad3(const double* x) {
    for (int i = 0; i < 3; i++) {
        double a = fabs(x[i]);
        if (a == INFINITY || isnan(a) || isinf(a) || a > 1e15) return 1;
    }
    return 0;
}

/* ======================= FilamentStretchingHarmonic =================================== */

double or_stretch_energy(const double* coord, const int* beadSet, const double* kstr,
                         const double* eql, int nint) {
    /* FilamentStretchingHarmonic.cpp:185-207 */
    double U = 0.0;
    for (int i = 0; i < nint; i++) {
        const double* c1 = &coord[beadSet[2 * i]];
        const double* c2 = &coord[beadSet[2 * i + 1]];
        double dist = two_point_distance(c1, c2) - eql[i];
        double U_i = 0.5 * kstr[i] * dist * dist;
        U += U_i;
    }
    return U;
}

void or_stretch_forces(const double* coord, double* f, const int* beadSet, const double* kstr,
                       const double* eql, int nint) {
    /* FilamentStretchingHarmonic.cpp:209-248 */
    for (int i = 0; i < nint; i++) {
        const double* c1 = &coord[beadSet[2 * i]];
        const double* c2 = &coord[beadSet[2 * i + 1]];
        double dist = two_point_distance(c1, c2);
        double invL = 1 / dist;
        double f0 = kstr[i] * (dist - eql[i]) * invL;
        double* f1 = &f[beadSet[2 * i]];
        double* f2 = &f[beadSet[2 * i + 1]];
        double r1x = c2[0] - c1[0], r1y = c2[1] - c1[1], r1z = c2[2] - c1[2];
        double Fx = -f0 * r1x, Fy = -f0 * r1y, Fz = -f0 * r1z;
        f2[0] += Fx; f2[1] += Fy; f2[2] += Fz;
        f1[0] += -Fx; f1[1] += -Fy; f1[2] += -Fz;
    }
}

/* ======================= FilamentBendingCosine ======================================== */

double or_bend_energy(const double* coord, int nint, const int* beadSet, const double* kbend,
                      const double* eqt) {
    /* FilamentBendingCosine.cpp:199-244 */
    double U = 0.0;
    for (int i = 0; i < nint; i++) {
        const double* c1 = &coord[beadSet[3 * i]];
        const double* c2 = &coord[beadSet[3 * i + 1]];
        const double* c3 = &coord[beadSet[3 * i + 2]];
        double L1 = sqrt(scalar_product(c1, c2, c1, c2));
        double L2 = sqrt(scalar_product(c2, c3, c2, c3));
        double L1L2 = L1 * L2;
        double l1l2 = scalar_product(c1, c2, c2, c3);
        double x = l1l2 / L1L2;
        if (x < -1.0) x = -1.0;
        else if (x > 1.0) x = 1.0;
        double U_i;
        if (are_equal(eqt[i], 0.0)) {
            U_i = kbend[i] * (1 - x);
        } else {
            double cosA = x;
            double sinA = rmax(sqrt(1 - cosA * cosA), 0.0);
            double cosAminusB = cosA * cos(eqt[i]) + sinA * sin(eqt[i]);
            U_i = kbend[i] * (1 - cosAminusB);
        }
        U += U_i;
    }
    return U;
}

void or_bend_forces(const double* coord, double* f, int nint, const int* beadSet,
                    const double* kbend, const double* eqt) {
    /* FilamentBendingCosine.cpp:247-345 ("Method 2") */
    for (int i = 0; i < nint; i++) {
        const double* c1 = &coord[beadSet[3 * i]];
        const double* c2 = &coord[beadSet[3 * i + 1]];
        const double* c3 = &coord[beadSet[3 * i + 2]];
        double* f1 = &f[beadSet[3 * i]];
        double* f2 = &f[beadSet[3 * i + 1]];
        double* f3 = &f[beadSet[3 * i + 2]];
        double L1sq = scalar_product(c1, c2, c1, c2);
        double L2sq = scalar_product(c2, c3, c2, c3);
        double l1l2 = scalar_product(c1, c2, c2, c3);
        double A = 1 / sqrt(L1sq * L2sq);
        double x = l1l2 * A;
        double B = x / L1sq;
        double C = x / L2sq;
        double k;
        if (are_equal(eqt[i], 0.0)) {
            k = kbend[i];
        } else {
            if (fabs(fabs(x) - 1.0) < 0.001) x = 0.999 * x;
            if (x < -1.0) x = -1.0;
            else if (x > 1.0) x = 1.0;
            double cosp = x;
            double sinp = sqrt(rmax((1 - cosp * cosp), 0.0));
            double sinpminusq = sinp * cos(eqt[i]) - cosp * sin(eqt[i]);
            k = kbend[i] * sinpminusq / sinp;
        }
        double r1x = c2[0] - c1[0], r1y = c2[1] - c1[1], r1z = c2[2] - c1[2];
        double r2x = c3[0] - c2[0], r2y = c3[1] - c2[1], r2z = c3[2] - c2[2];
        double Fr1x = k * (r2x * A - r1x * B);
        double Fr1y = k * (r2y * A - r1y * B);
        double Fr1z = k * (r2z * A - r1z * B);
        double Fr2x = k * (r1x * A - r2x * C);
        double Fr2y = k * (r1y * A - r2y * C);
        double Fr2z = k * (r1z * A - r2z * C);
        f1[0] += -Fr1x; f1[1] += -Fr1y; f1[2] += -Fr1z;
        f2[0] += Fr1x - Fr2x; f2[1] += Fr1y - Fr2y; f2[2] += Fr1z - Fr2z;
        f3[0] += Fr2x; f3[1] += Fr2y; f3[2] += Fr2z;
    }
}

/* ======================= Linker / MotorGhost stretching =============================== */

double or_xlink_energy(const double* coord, const int* beadSet, const double* kstr,
                       const double* eql, const double* pos1, const double* pos2, int nint) {
    /* LinkerStretchingHarmonic.cpp:191-220, MotorGhostStretchingHarmonic.cpp:234-268 */
    double U = 0.0;
    double v1[3], v2[3];
    for (int i = 0; i < nint; i++) {
        const double* c1 = &coord[beadSet[4 * i]];
        const double* c2 = &coord[beadSet[4 * i + 1]];
        const double* c3 = &coord[beadSet[4 * i + 2]];
        const double* c4 = &coord[beadSet[4 * i + 3]];
        mid_point(v1, c1, c2, pos1[i]);
        mid_point(v2, c3, c4, pos2[i]);
        double dist = two_point_distance(v1, v2) - eql[i];
        double U_i = 0.5 * kstr[i] * dist * dist;
        U += U_i;
    }
    return U;
}

void or_xlink_forces(const double* coord, double* f, const int* beadSet, const double* kstr,
                     const double* eql, const double* pos1, const double* pos2,
                     double* stretchforce, int nint) {
    /* LinkerStretchingHarmonic.cpp:225-291, MotorGhostStretchingHarmonic.cpp:271-340 */
    double v1[3], v2[3];
    for (int i = 0; i < nint; i++) {
        const double* c1 = &coord[beadSet[4 * i]];
        const double* c2 = &coord[beadSet[4 * i + 1]];
        const double* c3 = &coord[beadSet[4 * i + 2]];
        const double* c4 = &coord[beadSet[4 * i + 3]];
        mid_point(v1, c1, c2, pos1[i]);
        mid_point(v2, c3, c4, pos2[i]);
        double dist = two_point_distance(v1, v2);
        double invL = 1 / dist;
        double f0 = kstr[i] * (dist - eql[i]) * invL;
        double* f1 = &f[beadSet[4 * i]];
        double* f2 = &f[beadSet[4 * i + 1]];
        double* f3 = &f[beadSet[4 * i + 2]];
        double* f4 = &f[beadSet[4 * i + 3]];
        double lx = v2[0] - v1[0], ly = v2[1] - v1[1], lz = v2[2] - v1[2];
        double Flx = -f0 * lx, Fly = -f0 * ly, Flz = -f0 * lz;
        f1[0] += -Flx * (1 - pos1[i]); f1[1] += -Fly * (1 - pos1[i]); f1[2] += -Flz * (1 - pos1[i]);
        f2[0] += -Flx * (pos1[i]); f2[1] += -Fly * (pos1[i]); f2[2] += -Flz * (pos1[i]);
        f3[0] += Flx * (1 - pos2[i]); f3[1] += Fly * (1 - pos2[i]); f3[2] += Flz * (1 - pos2[i]);
        f4[0] += Flx * (pos2[i]); f4[1] += Fly * (pos2[i]); f4[2] += Flz * (pos2[i]);
        if (stretchforce) stretchforce[i] = f0 / invL;
    }
}

/* ======================= CylinderExclVolRepulsion ===================================== */

/* normalizeVector (MathFunctions.h:178-188) */
static inline void normalize3(double* v) {
    double x2 = v[0] * v[0], y2 = v[1] * v[1], z2 = v[2] * v[2];
    double norm = sqrt(x2 + y2 + z2);
    v[0] = v[0] / norm; v[1] = v[1] / norm; v[2] = v[2] / norm;
}

/* crossProduct (MathFunctions.h:910-916) */
static inline void cross3(double* cp, const double* v1, const double* v2) {
    cp[0] = v1[1] * v2[2] - v1[2] * v2[1];
    cp[1] = v1[2] * v2[0] - v1[0] * v2[2];
    cp[2] = v1[0] * v2[1] - v1[1] * v2[0];
}

int or_are_in_plane(const double* p1, const double* p2, const double* p3, const double* p4) {
    /* MathFunctions.h:1101-1135, areEqualLT (utility.h:80-83): |a - b| < 0.1 */
    double v1[3], v2[3], v3[3], cp[3];
    for (int k = 0; k < 3; k++) {
        v1[k] = p2[k] - p1[k];
        v2[k] = p3[k] - p1[k];
        v3[k] = p4[k] - p1[k];
    }
    cross3(cp, v1, v2);
    normalize3(cp);
    normalize3(v3);
    double dot = v3[0] * cp[0] + v3[1] * cp[1] + v3[2] * cp[2];
    return fabs(dot - 0.0) < 1E-1;
}

/* movePointOutOfPlane for point i in {1,2} (MathFunctions.h:1221-1278) */
static void move_out_of_plane(const double* p1, const double* p2, const double* p3, double* newp,
                              int i, double d) {
    double v1[3], v2[3], norm[3];
    for (int k = 0; k < 3; k++) {
        v1[k] = p2[k] - p1[k];
        v2[k] = p3[k] - p1[k];
    }
    cross3(norm, v1, v2);
    normalize3(norm);
    const double* p = (i == 1) ? p1 : p2;
    for (int k = 0; k < 3; k++) newp[k] = (p[k] + norm[k] * d);
}

/* The invariants of a segment pair (CylinderExclVolRepulsion.cpp:353-397). */
typedef struct {
    double a, b, c, d, e, F, g, h, I;
    double AA, BB, CC, DD, EE, FF, GG, HH, JJ;
    double ATG1, ATG2, ATG3, ATG4;
} exv_terms;

static void exvol_terms(const double* c1, const double* c2, const double* c3, const double* c4,
                        exv_terms* t) {
    t->a = scalar_product(c1, c2, c1, c2);
    t->b = scalar_product(c3, c4, c3, c4);
    t->c = scalar_product(c3, c1, c3, c1);
    t->d = scalar_product(c1, c2, c3, c4);
    t->e = scalar_product(c1, c2, c3, c1);
    t->F = scalar_product(c3, c4, c3, c1);
    t->g = scalar_product(c1, c2, c1, c4);
    t->h = scalar_product(c3, c4, c3, c2);
    t->I = scalar_product(c3, c4, c1, c4);
    /* sqmagnitude(c1, c4), sqmagnitude(c3, c2) (MathFunctions.h:210-217) */
    double x = c1[0] - c4[0], y = c1[1] - c4[1], z = c1[2] - c4[2];
    double sqmag_D = (x * x + y * y + z * z);
    x = c3[0] - c2[0]; y = c3[1] - c2[1]; z = c3[2] - c2[2];
    double sqmag_E = (x * x + y * y + z * z);
    double vA[3], vB[3], vC[3], cp[3];
    for (int k = 0; k < 3; k++) {
        vA[k] = c1[k] - c2[k];
        vB[k] = c3[k] - c4[k];
        vC[k] = c3[k] - c1[k];
    }
    double a = t->a, b = t->b, c = t->c, d = t->d, e = t->e, F = t->F;
    double ac = sqrt(a * c);
    double bc = sqrt(b * c);
    double aD = sqrt(a * sqmag_D);
    double bE = sqrt(b * sqmag_E);
    t->AA = sqrt(rmax(0.0, (ac + e) * (ac - e)));
    t->BB = sqrt(rmax(0.0, (bc + F) * (bc - F)));
    t->CC = d * e - a * F;
    t->DD = b * e - d * F;
    t->EE = sqrt(rmax(0.0, (aD + t->g) * (aD - t->g)));
    t->FF = sqrt(rmax(0.0, (bE + t->h) * (bE - t->h)));
    t->GG = d * t->g - a * t->I;
    t->HH = t->CC + t->GG - t->DD;
    cross3(cp, vA, vB);
    /* scalarProduct(doubleprecision const*, floatingpoint const*) = plain dot (MathFunctions.h:496-499) */
    double JJ = (cp[0] * vC[0] + cp[1] * vC[1] + cp[2] * vC[2]);
    t->JJ = -JJ * JJ;
    double AA = t->AA, BB = t->BB, EE = t->EE, FF = t->FF;
    t->ATG1 = atan((a + e) / AA) - atan(e / AA);
    t->ATG2 = atan((a + e - d) / EE) - atan((e - d) / EE);
    t->ATG3 = atan((F) / BB) - atan((F - b) / BB);
    t->ATG4 = atan((d + F) / FF) - atan((d + F - b) / FF);
}

static inline int energy_bad(double U) { return fabs(U) == INFINITY || U != U || U < -1.0; }

/* getenergyintegrand / getforceintegrand (CylinderExclVolRepulsion.h:53-72) */
static inline double energy_integrand(double a, double b, double c, double d, double e, double F,
                                      double s, double t) {
    double r_2 = c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;
    return (1.0 / (r_2 * r_2));
}
static inline void force_integrand(double a, double b, double c, double d, double e, double F,
                                   double s, double t, double* out) {
    double r_2 = c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;
    double x = (1.0 / (r_2 * r_2 * r_2));
    out[0] = x; out[1] = s * x; out[2] = s * s * x; out[3] = s * t * x; out[4] = t * x;
    out[5] = t * t * x;
}

double or_exvol_energyN(const double* coord, const int* beadSet, const double* krep,
                        const double* eqLengths, int i, int movebeads) {
    /* CylinderExclVolRepulsion.cpp:884-982. The sample points are s = (i+1)/N, t = (j+1)/N with
     * the weights of the loop as written (the counter is advanced before use). */
    const double* c1 = &coord[beadSet[4 * i]];
    const double* c2 = &coord[beadSet[4 * i + 1]];
    const double* c3 = &coord[beadSet[4 * i + 2]];
    const double* c4 = &coord[beadSet[4 * i + 3]];
    double newc1[3], newc2[3];
    double kRepScaled = krep[i] * eqLengths[2 * i] * eqLengths[2 * i + 1];
    if (movebeads && or_are_in_plane(c1, c2, c3, c4)) {
        move_out_of_plane(c1, c2, c3, newc2, 2, 0.01);
        move_out_of_plane(c1, c2, c3, newc1, 1, 0.01);
        c2 = newc2;
        c1 = newc1;
    }
    double a = scalar_product(c1, c2, c1, c2), b = scalar_product(c3, c4, c3, c4);
    double c = scalar_product(c3, c1, c3, c1), d = scalar_product(c1, c2, c3, c4);
    double e = scalar_product(c1, c2, c3, c1), F = scalar_product(c3, c4, c3, c1);
    const int N1 = 400, N2 = 400;
    double deltas = 1.0 / ((double)N1), deltat = 1.0 / ((double)N2);
    double Termset1 = energy_integrand(a, b, c, d, e, F, 0.0, 0.0) +
                      energy_integrand(a, b, c, d, e, F, 0.0, 1.0) +
                      energy_integrand(a, b, c, d, e, F, 1.0, 0.0) +
                      energy_integrand(a, b, c, d, e, F, 1.0, 1.0);
    double Termset2 = 0.0, Termset3 = 0.0, Termset4 = 0.0;
    double s = 0.0, t = 0.0, factors, factort;
    for (int ii = 0; ii <= N1; ii++) {
        t = 0.0;
        s += deltas;
        factors = (ii % 2 == 0) ? 2.0 : 4.0;
        if (ii > 0 && ii < N1)
            Termset2 += factors * (energy_integrand(a, b, c, d, e, F, s, 0.0) +
                                   energy_integrand(a, b, c, d, e, F, s, 1.0));
        for (int j = 0; j <= N2; j++) {
            t += deltat;
            factort = (j % 2 == 0) ? 2.0 : 4.0;
            if (j > 0 && j < N2) {
                if (ii == 0)
                    Termset3 += factort * (energy_integrand(a, b, c, d, e, F, 0.0, t) +
                                           energy_integrand(a, b, c, d, e, F, 1.0, t));
                else if (ii < N1)
                    Termset4 += factors * factort * energy_integrand(a, b, c, d, e, F, s, t);
            }
        }
    }
    double U_i = kRepScaled * (deltas * deltat / 9.0) * (Termset1 + Termset2 + Termset3 + Termset4);
    /* The reference retries with moved beads but discards the retry's value (:975-980). */
    return U_i;
}

void or_exvol_forceN(const double* coord, double* f, const int* beadSet, const double* krep,
                     const double* eqLengths, int i, int movebeads) {
    /* CylinderExclVolRepulsion.cpp:984-1254 */
    const double* c1 = &coord[beadSet[4 * i]];
    const double* c2 = &coord[beadSet[4 * i + 1]];
    const double* c3 = &coord[beadSet[4 * i + 2]];
    const double* c4 = &coord[beadSet[4 * i + 3]];
    double newc1[3], newc2[3];
    double kRepScaled = krep[i] * eqLengths[2 * i] * eqLengths[2 * i + 1];
    if (movebeads && or_are_in_plane(c1, c2, c3, c4)) {
        move_out_of_plane(c1, c2, c3, newc2, 2, 0.01);
        move_out_of_plane(c1, c2, c3, newc1, 1, 0.01);
        c2 = newc2;
        c1 = newc1;
    }
    double* f1 = &f[beadSet[4 * i]];
    double* f2 = &f[beadSet[4 * i + 1]];
    double* f3 = &f[beadSet[4 * i + 2]];
    double* f4 = &f[beadSet[4 * i + 3]];
    double a = scalar_product(c1, c2, c1, c2), b = scalar_product(c3, c4, c3, c4);
    double c = scalar_product(c3, c1, c3, c1), d = scalar_product(c1, c2, c3, c4);
    double e = scalar_product(c1, c2, c3, c1), F = scalar_product(c3, c4, c3, c1);
    double vecA[3], vecB[3], vecC[3];
    for (int k = 0; k < 3; k++) {
        vecA[k] = c1[k] - c2[k];
        vecB[k] = c3[k] - c4[k];
        vecC[k] = c3[k] - c1[k];
    }
    const int N1 = 10000, N2 = 10000;
    double deltas = 1.0 / ((double)N1), deltat = 1.0 / ((double)N2);
    double T1[6] = {0}, T2[6] = {0}, T3[6] = {0}, T4[6] = {0}, in[6];
    const double limits[2] = {0.0, 1.0};
    for (int p = 0; p < 2; p++)
        for (int q = 0; q < 2; q++) {
            force_integrand(a, b, c, d, e, F, limits[p], limits[q], in);
            for (int k = 0; k < 6; k++) T1[k] += in[k];
        }
    double s = 0.0, t = 0.0, factors, factort;
    for (int ii = 0; ii <= N1; ii++) {
        t = 0.0;
        s += deltas;
        factors = (ii % 2 == 0) ? 2.0 : 4.0;
        if (ii > 0 && ii < N1) {
            for (int l = 0; l < 2; l++) {
                force_integrand(a, b, c, d, e, F, s, limits[l], in);
                for (int k = 0; k < 6; k++) T2[k] += factors * in[k];
            }
        }
        for (int j = 0; j <= N2; j++) {
            t += deltat;
            factort = (j % 2 == 0) ? 2.0 : 4.0;
            if (j > 0 && j < N2) {
                if (ii == 0) {
                    for (int l = 0; l < 2; l++) {
                        force_integrand(a, b, c, d, e, F, limits[l], t, in);
                        for (int k = 0; k < 6; k++) T3[k] += factort * in[k];
                    }
                } else if (ii < N1) {
                    force_integrand(a, b, c, d, e, F, s, t, in);
                    for (int k = 0; k < 6; k++) T4[k] += factors * factort * in[k];
                }
            }
        }
    }
    double I6[6];
    for (int k = 0; k < 6; k++) I6[k] = (deltas * deltat / 9.0) * (T1[k] + T2[k] + T3[k] + T4[k]);
    double f1l[3], f2l[3], f3l[3], f4l[3];
    for (int k = 0; k < 3; k++) {
        f1l[k] = 4.0 * kRepScaled *
                 (-I6[0] * vecC[k] + I6[1] * (vecC[k] - vecA[k]) + I6[2] * vecA[k] +
                  I6[4] * vecB[k] - I6[3] * vecB[k]);
        f2l[k] = 4.0 * kRepScaled * (-I6[1] * vecC[k] - I6[2] * vecA[k] + I6[3] * vecB[k]);
        f3l[k] = 4.0 * kRepScaled *
                 (I6[0] * vecC[k] + I6[1] * vecA[k] - I6[4] * (vecB[k] + vecC[k]) -
                  I6[3] * vecA[k] + I6[5] * vecB[k]);
        f4l[k] = 4.0 * kRepScaled * (I6[4] * vecC[k] + I6[3] * vecA[k] - I6[5] * vecB[k]);
    }
    if (bad3(f1l) || bad3(f2l) || bad3(f3l) || bad3(f4l)) {
        if (!movebeads) or_exvol_forceN(coord, f, beadSet, krep, eqLengths, i, 1);
        /* with moved beads still non-finite the reference exits; the oracle adds nothing */
    } else {
        for (int k = 0; k < 3; k++) {
            f1[k] += f1l[k]; f2[k] += f2l[k]; f3[k] += f3l[k]; f4[k] += f4l[k];
        }
    }
}

double or_exvol_energy(const double* coord, const int* beadSet, const double* krep,
                       const double* eqLengths, int nint) {
    /* CylinderExclVolRepulsion.cpp:330-575 */
    double U = 0.0;
    exv_terms t;
    for (int i = 0; i < nint; i++) {
        double kRepScaled = krep[i] * eqLengths[2 * i] * eqLengths[2 * i + 1];
        const double* c1 = &coord[beadSet[4 * i]];
        const double* c2 = &coord[beadSet[4 * i + 1]];
        const double* c3 = &coord[beadSet[4 * i + 2]];
        const double* c4 = &coord[beadSet[4 * i + 3]];
        exvol_terms(c1, c2, c3, c4, &t);
        double U_i = 0.5 * kRepScaled / t.JJ *
                     (t.CC / t.AA * t.ATG1 + t.GG / t.EE * t.ATG2 + t.DD / t.BB * t.ATG3 +
                      t.HH / t.FF * t.ATG4);
        if (energy_bad(U_i)) {
            U_i = or_exvol_energyN(coord, beadSet, krep, eqLengths, i, 0);
            if (energy_bad(U_i)) return -1;
        }
        U += U_i;
    }
    return U;
}

void or_exvol_forces(const double* coord, double* f, const int* beadSet, const double* krep,
                     const double* eqLengths, int nint) {
    /* CylinderExclVolRepulsion.cpp:578-882 */
    exv_terms t;
    for (int i = 0; i < nint; i++) {
        double kRepScaled = krep[i] * eqLengths[2 * i] * eqLengths[2 * i + 1];
        const double* c1 = &coord[beadSet[4 * i]];
        const double* c2 = &coord[beadSet[4 * i + 1]];
        const double* c3 = &coord[beadSet[4 * i + 2]];
        const double* c4 = &coord[beadSet[4 * i + 3]];
        double* f1 = &f[beadSet[4 * i]];
        double* f2 = &f[beadSet[4 * i + 1]];
        double* f3 = &f[beadSet[4 * i + 2]];
        double* f4 = &f[beadSet[4 * i + 3]];
        double newc2[3];
        if (or_are_in_plane(c1, c2, c3, c4)) {
            move_out_of_plane(c1, c2, c3, newc2, 2, 0.01);
            c2 = newc2;
        }
        exvol_terms(c1, c2, c3, c4, &t);
        double a = t.a, b = t.b, c = t.c, d = t.d, e = t.e, F = t.F;
        double AA = t.AA, BB = t.BB, CC = t.CC, DD = t.DD, EE = t.EE, FF = t.FF, GG = t.GG,
               HH = t.HH, JJ = t.JJ;
        double ATG1 = t.ATG1, ATG2 = t.ATG2, ATG3 = t.ATG3, ATG4 = t.ATG4;
        double invJJ = 1 / JJ;

        const double sumBlock = CC / AA * ATG1 + GG / EE * ATG2 + DD / BB * ATG3 + HH / FF * ATG4;
        const double enFactor = kRepScaled * invJJ / 2;

        const double A1 = AA * AA / (AA * AA + (a + e) * (a + e));
        const double A2 = AA * AA / (AA * AA + e * e);
        const double E1 = EE * EE / (EE * EE + (a + e - d) * (a + e - d));
        const double E2 = EE * EE / (EE * EE + (e - d) * (e - d));
        const double B1 = BB * BB / (BB * BB + F * F);
        const double B2 = BB * BB / (BB * BB + (F - b) * (F - b));
        const double F1 = FF * FF / (FF * FF + (d + F) * (d + F));
        const double F2 = FF * FF / (FF * FF + (d + F - b) * (d + F - b));

        const double blockA_C = ATG1 / AA;
        const double blockA_A = -(ATG1 * CC) / (AA * AA) + (-A1 * (a + e) + A2 * e) * CC / (AA * AA * AA);
        const double blockA_e = ((A1 - A2) * CC) / (AA * AA);
        const double blockA_a = (A1 * CC) / (AA * AA);

        const double blockE_G = ATG2 / EE;
        const double blockE_E = -(ATG2 * GG) / (EE * EE) + (-E1 * (a + e - d) + E2 * (e - d)) * GG / (EE * EE * EE);
        const double blockE_e_minus_d = ((E1 - E2) * GG) / (EE * EE);
        const double blockE_a = (E1 * GG) / (EE * EE);

        const double blockB_D = ATG3 / BB;
        const double blockB_B = -(ATG3 * DD) / (BB * BB) + (-B1 * F + B2 * (F - b)) * DD / (BB * BB * BB);
        const double blockB_f_minus_b = ((B1 - B2) * DD) / (BB * BB);
        const double blockB_b = (B1 * DD) / (BB * BB);

        const double blockF_H = ATG4 / FF;
        const double blockF_F = -(ATG4 * HH) / (FF * FF) + (-F1 * (d + F) + F2 * (d + F - b)) * HH / (FF * FF * FF);
        const double blockF_d_plus_f_minus_b = ((F1 - F2) * HH) / (FF * FF);
        const double blockF_b = (F1 * HH) / (FF * FF);

        const double block_11 = blockA_C * (-2 * F) + blockA_A * (c / AA) + blockA_a * 2 +
                                blockE_G * (2 * (F - b)) + blockE_E * ((b + c - 2 * F) / EE) + blockE_a * 2 +
                                blockF_H * (-2 * b) + blockF_F * (b / FF);
        const double block_12 = blockA_C * e + blockE_G * (2 * d - e) + blockE_E * ((e - d) / EE) +
                                blockE_e_minus_d * (-1) + blockB_D * (-F) + blockF_H * (2 * d + F) +
                                blockF_F * (-(d + F) / FF) + blockF_d_plus_f_minus_b;
        const double block_13 = blockA_C * d + blockA_A * (-e / AA) + blockA_e + blockE_G * (-d) +
                                blockE_E * (-(e - d) / EE) + blockE_e_minus_d + blockB_D * b +
                                blockF_H * (-b) + blockF_F * (b / FF);
        const double block_21 = block_12;
        const double block_22 = blockE_G * (-2 * a) + blockE_E * (a / EE) + blockB_D * (2 * e) +
                                blockB_B * (c / BB) + blockB_f_minus_b * (-2) + blockB_b * 2 +
                                blockF_H * (-2 * (a + e)) + blockF_F * ((a + c + 2 * e) / FF) +
                                blockF_d_plus_f_minus_b * (-2) + blockF_b * 2;
        const double block_23 = blockA_C * (-a) + blockE_G * a + blockE_E * (-a / EE) +
                                blockB_D * (-d) + blockB_B * (-F / BB) + blockB_f_minus_b +
                                blockF_H * d + blockF_F * (-(d + F) / FF) + blockF_d_plus_f_minus_b;
        const double block_31 = block_13;
        const double block_32 = block_23;
        const double block_33 = blockA_A * (a / AA) + blockE_E * (a / EE) + blockB_B * (b / BB) +
                                blockF_F * (b / FF);

        const double JJ_11 = 2 * (-b * c + F * F);
        const double JJ_12 = 2 * (c * d - e * F);
        const double JJ_13 = 2 * (b * e - d * F);
        const double JJ_21 = 2 * (c * d - e * F);
        const double JJ_22 = 2 * (-a * c + e * e);
        const double JJ_23 = 2 * (-d * e + a * F);
        const double JJ_31 = 2 * (b * e - d * F);
        const double JJ_32 = 2 * (-d * e + a * F);
        const double JJ_33 = 2 * (d * d - a * b);

        double deriv[4][3];
        deriv[0][0] = enFactor * (-invJJ * sumBlock * (JJ_31 - JJ_11) + (block_31 - block_11));
        deriv[0][1] = enFactor * (-invJJ * sumBlock * (JJ_32 - JJ_12) + (block_32 - block_12));
        deriv[0][2] = enFactor * (-invJJ * sumBlock * (JJ_33 - JJ_13) + (block_33 - block_13));
        deriv[1][0] = enFactor * (-invJJ * sumBlock * (JJ_11) + (block_11));
        deriv[1][1] = enFactor * (-invJJ * sumBlock * (JJ_12) + (block_12));
        deriv[1][2] = enFactor * (-invJJ * sumBlock * (JJ_13) + (block_13));
        deriv[2][0] = enFactor * (-invJJ * sumBlock * (-JJ_31 - JJ_21) + (-block_31 - block_21));
        deriv[2][1] = enFactor * (-invJJ * sumBlock * (-JJ_32 - JJ_22) + (-block_32 - block_22));
        deriv[2][2] = enFactor * (-invJJ * sumBlock * (-JJ_33 - JJ_23) + (-block_33 - block_23));
        deriv[3][0] = enFactor * (-invJJ * sumBlock * (JJ_21) + (block_21));
        deriv[3][1] = enFactor * (-invJJ * sumBlock * (JJ_22) + (block_22));
        deriv[3][2] = enFactor * (-invJJ * sumBlock * (JJ_23) + (block_23));

        double fo[4][3];
        for (int bi = 0; bi < 4; bi++)
            for (int k = 0; k < 3; k++)
                fo[bi][k] = -(deriv[bi][0] * (c2[k] - c1[k]) + deriv[bi][1] * (c4[k] - c3[k]) +
                              deriv[bi][2] * (c1[k] - c3[k]));

        if (bad3(fo[0]) || bad3(fo[1]) || bad3(fo[2]) || bad3(fo[3])) {
            or_exvol_forceN(coord, f, beadSet, krep, eqLengths, i, 0);
        } else {
            for (int k = 0; k < 3; k++) {
                f1[k] += fo[0][k]; f2[k] += fo[1][k]; f3[k] += fo[2][k]; f4[k] += fo[3][k];
            }
        }
    }
}

/* ======================= BoundaryCylinderRepulsionExp ================================= */

mdn_boundary_element or_plane_element(const double* point, const double* normal, double krep,
                                      double slen) {
    /* PlaneBoundaryElement ctor (BoundaryElementImpl.h:36-45) */
    mdn_boundary_element e;
    memset(&e, 0, sizeof(e));
    e.kind = MDN_BND_PLANE;
    e.p[0] = normal[0]; e.p[1] = normal[1]; e.p[2] = normal[2];
    e.p[3] = -e.p[0] * point[0] - e.p[1] * point[1] - e.p[2] * point[2];
    e.krep = krep;
    e.slen = slen;
    return e;
}

double or_boundary_distance(const mdn_boundary_element* e, const double* x) {
    if (e->kind == MDN_BND_PLANE) {
        /* BoundaryElementImpl.h:47-50; pow(v, 2) is exactly v*v */
        const double a = e->p[0], b = e->p[1], c = e->p[2], d = e->p[3];
        return (a * x[0] + b * x[1] + c * x[2] + d) / sqrt(a * a + b * b + c * c);
    }
    /* SphereBoundaryElement::distance = radius - twoPointDistance(center, x) (:136-142) */
    return e->p[3] - two_point_distance(e->p, x);
}

void or_boundary_normal(const mdn_boundary_element* e, const double* x, double* n) {
    if (e->kind == MDN_BND_PLANE) {
        n[0] = e->p[0]; n[1] = e->p[1]; n[2] = e->p[2];
        return;
    }
    /* twoPointDirection(x, center) (MathFunctions.h:413-421) */
    double invD = 1 / two_point_distance(x, e->p);
    n[0] = invD * (e->p[0] - x[0]);
    n[1] = invD * (e->p[1] - x[1]);
    n[2] = invD * (e->p[2] - x[2]);
}

double or_boundary_energy(const double* coord, const int* beadSet, const double* krep,
                          const double* slen, const int* nneighbors,
                          const mdn_boundary_element* elems, int nelem) {
    /* BoundaryCylinderRepulsionExp.cpp:27-60 */
    double U = 0.0;
    int cum = 0;
    for (int ib = 0; ib < nelem; ib++) {
        int nc = nneighbors[ib];
        for (int ic = 0; ic < nc; ic++) {
            const double* c1 = &coord[beadSet[cum + ic]];
            double r = or_boundary_distance(&elems[ib], c1);
            double R = -r / slen[cum + ic];
            double U_i = krep[cum + ic] * exp(R);
            U += U_i;
        }
        cum += nc;
    }
    return U;
}

void or_boundary_forces(const double* coord, double* f, const int* beadSet, const double* krep,
                        const double* slen, const int* nneighbors,
                        const mdn_boundary_element* elems, int nelem) {
    /* BoundaryCylinderRepulsionExp.cpp:63-110 */
    int cum = 0;
    double n[3];
    for (int ib = 0; ib < nelem; ib++) {
        int nc = nneighbors[ib];
        for (int ic = 0; ic < nc; ic++) {
            const double* c1 = &coord[beadSet[cum + ic]];
            double* f1 = &f[beadSet[cum + ic]];
            double r = or_boundary_distance(&elems[ib], c1);
            or_boundary_normal(&elems[ib], c1, n);
            double R = -r / slen[cum + ic];
            double f0 = krep[cum + ic] * exp(R) / slen[cum + ic];
            f1[0] += f0 * n[0];
            f1[1] += f0 * n[1];
            f1[2] += f0 * n[2];
        }
        cum += nc;
    }
}
