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
    double r14, r23, r24; /* |c4 - c1|^2, |c2 - c3|^2, |c4 - c2|^2 */
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
    t->r14 = sqmag_D;
    t->r23 = sqmag_E;
    x = c4[0] - c2[0]; y = c4[1] - c2[1]; z = c4[2] - c2[2];
    t->r24 = (x * x + y * y + z * z);
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

/* Sensitivities of one edge term X psi, psi = (atan(p / Q) - atan(q / Q)) / Q, Q^2 = P, to its
 * radicand and its two arguments: o = {psi, X dpsi/dP, X dpsi/dp, X dpsi/dq} with
 * dpsi/dp = 1 / (P + p^2), dpsi/dq = -1 / (P + q^2), dpsi/dP = -(psi + p dpsi/dp + q dpsi/dq) / (2P);
 * Pp = P + p^2 and Pq = P + q^2 are passed as products of squared lengths. */
static void edge_sens(double X, double Q, double atg, double p, double q, double Pp, double Pq,
                      double o[4]) {
    const double iQ = 1 / Q;
    o[0] = atg * iQ;
    o[2] = X / Pp;
    o[3] = -X / Pq;
    o[1] = -(X * o[0] + p * o[2] + q * o[3]) * (iQ * iQ) / 2;
}

/* Forces -dU/dc_k of the closed-form pair energy U = (K / 2) S / JJ, S = sum over the four edge
 * terms (CylinderExclVolRepulsion.cpp:353-435 defines U), by the chain rule through the six
 * scalars a = u1.u1, b = u2.u2, c = u3.u3, d = u1.u2, e = u1.u3, F = u2.u3 of u1 = c2 - c1,
 * u2 = c4 - c3, u3 = c1 - c3. The edge terms (radicand P, weight X, arguments p, q):
 *   A: P = ac - e^2, X = CC, p = a + e, q = e;  E: P = a r14 - (d - e)^2, X = GG, p = a + e - d,
 *   q = e - d;  B: P = bc - F^2, X = DD, p = F, q = F - b;  H: P = b r23 - (d + F)^2, X = HH,
 *   p = d + F, q = d + F - b;
 * with P + p^2, P + q^2 = a r23, a c (A); a r24, a r14 (E); b c, b r14 (B); b r23, b r24 (H).
 * Then dU/ds = (K / (2 JJ)) (dS/ds - (S / JJ) dJJ/ds), JJ = -(abc - aF^2 - cd^2 + 2deF - be^2),
 * and dU/du1 = 2 U_a u1 + U_d u2 + U_e u3, dU/du2 = 2 U_b u2 + U_d u1 + U_F u3,
 * dU/du3 = 2 U_c u3 + U_e u1 + U_F u2; f1 = G1 - G3, f2 = -G1, f3 = G2 + G3, f4 = -G2. */
static void exvol_pair_forces(const exv_terms* t, double K, const double* u1, const double* u2,
                              const double* u3, double fo[4][3]) {
    const double a = t->a, b = t->b, c = t->c, d = t->d, e = t->e, F = t->F;
    const double g = d - e, h = d + F;
    double A[4], E[4], B[4], H[4];
    edge_sens(t->CC, t->AA, t->ATG1, a + e, e, a * t->r23, a * c, A);
    edge_sens(t->GG, t->EE, t->ATG2, a + e - d, e - d, a * t->r24, a * t->r14, E);
    edge_sens(t->DD, t->BB, t->ATG3, F, F - b, b * c, b * t->r14, B);
    edge_sens(t->HH, t->FF, t->ATG4, d + F, d + F - b, b * t->r23, b * t->r24, H);
    const double S = t->CC * A[0] + t->GG * E[0] + t->DD * B[0] + t->HH * H[0];
    const double dSa = -F * A[0] + c * A[1] + A[2] + (F - b) * E[0] + t->r14 * E[1] + E[2] - b * H[0] + b * H[1];
    const double dSb = -a * E[0] + a * E[1] + e * B[0] + c * B[1] - B[3] - (a + e) * H[0] + t->r23 * H[1] - H[3];
    const double dSc = a * (A[1] + E[1]) + b * (B[1] + H[1]);
    const double dSd = e * A[0] + (2 * d - e) * E[0] - 2 * g * E[1] - E[2] - E[3] - F * B[0] + (2 * d + F) * H[0] -
                       2 * h * H[1] + H[2] + H[3];
    const double dSe = d * A[0] - 2 * e * A[1] + A[2] + A[3] - d * E[0] + 2 * g * E[1] + E[2] + E[3] + b * B[0] -
                       b * H[0] + 2 * b * H[1];
    const double dSF = -a * A[0] + a * E[0] - 2 * a * E[1] - d * B[0] - 2 * F * B[1] + B[2] + B[3] + d * H[0] -
                       2 * h * H[1] + H[2] + H[3];
    const double w = K / (2 * t->JJ), r = S / t->JJ;
    const double Ua = w * (dSa - r * (F * F - b * c));
    const double Ub = w * (dSb - r * (e * e - a * c));
    const double Uc = w * (dSc - r * (d * d - a * b));
    const double Ud = w * (dSd - r * (2 * (c * d - e * F)));
    const double Ue = w * (dSe - r * (2 * (b * e - d * F)));
    const double UF = w * (dSF - r * (2 * (a * F - d * e)));
    for (int k = 0; k < 3; k++) {
        const double G1 = 2 * Ua * u1[k] + Ud * u2[k] + Ue * u3[k];
        const double G2 = 2 * Ub * u2[k] + Ud * u1[k] + UF * u3[k];
        const double G3 = 2 * Uc * u3[k] + Ue * u1[k] + UF * u2[k];
        fo[0][k] = G1 - G3;
        fo[1][k] = -G1;
        fo[2][k] = G2 + G3;
        fo[3][k] = -G2;
    }
}

void or_exvol_forces(const double* coord, double* f, const int* beadSet, const double* krep,
                     const double* eqLengths, int nint) {
    /* CylinderExclVolRepulsion.cpp:578-882: the coplanar shift of bead 2 (:610-630), the closed
     * form's gradient, the Simpson forceN fallback for a non-finite or > 1e15 component (:859-872) */
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
        double u1[3], u2[3], u3[3], fo[4][3];
        for (int k = 0; k < 3; k++) {
            u1[k] = c2[k] - c1[k];
            u2[k] = c4[k] - c3[k];
            u3[k] = c1[k] - c3[k];
        }
        exvol_pair_forces(&t, kRepScaled, u1, u2, u3, fo);
        if (bad3(fo[0]) || bad3(fo[1]) || bad3(fo[2]) || bad3(fo[3])) {
            or_exvol_forceN(coord, f, beadSet, krep, eqLengths, i, 0);
        } else {
            for (int k = 0; k < 3; k++) {
                f1[k] += fo[0][k]; f2[k] += fo[1][k]; f3[k] += fo[2][k]; f4[k] += fo[3][k];
            }
        }
    }
}

/* The coplanar shift of bead i in {1, 2} by d along the unit normal of (p2 - p1, p3 - p1)
 * (movePointOutOfPlane, MathFunctions.h:1221-1278), for tests. */
void or_move_out_of_plane(const double* p1, const double* p2, const double* p3, double* newp, int i,
                          double d) {
    move_out_of_plane(p1, p2, p3, newp, i, d);
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
