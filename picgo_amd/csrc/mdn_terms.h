// mdn_terms.h -- per-interaction energy / force arithmetic of the six force fields, for
// gfx950 device code. Every expression keeps the operand order of the cited reference
// kernels so that, built with -ffp-contract=off, FP64 results round exactly like the
// reference's IEEE double arithmetic (transcendental calls aside).
//
// Citations are relative to medyan-5.4.0/ in the reference tree.
#pragma once

#include <hip/hip_runtime.h>

namespace mdn {

// std::max / std::min semantics (the reference calls std::max(a, b) = (a < b) ? b : a).
template <typename T> __device__ __host__ inline T rmax(T a, T b) { return (a < b) ? b : a; }
template <typename T> __device__ __host__ inline T rmin(T a, T b) { return (b < a) ? b : a; }

struct D3 { double x, y, z; };

__device__ inline D3 ld3(const double* __restrict__ p, int bead) {
    return D3{p[3 * bead], p[3 * bead + 1], p[3 * bead + 2]};
}
// coordinate of `bead` at x + lambda*d (CGMethod.cpp:43-46 / :350-354: coord + lambda*dir)
__device__ inline D3 ld3_at(const double* __restrict__ x, const double* __restrict__ d,
                            double lambda, int bead) {
    if (lambda == 0.0 || d == nullptr) return ld3(x, bead);
    return D3{x[3 * bead] + lambda * d[3 * bead], x[3 * bead + 1] + lambda * d[3 * bead + 1],
              x[3 * bead + 2] + lambda * d[3 * bead + 2]};
}

// mathfunc::twoPointDistance (MathFunctions.h:294-297)
__device__ inline double two_point_distance(D3 v1, D3 v2) {
    return sqrt((v2.x - v1.x) * (v2.x - v1.x) + (v2.y - v1.y) * (v2.y - v1.y) +
                (v2.z - v1.z) * (v2.z - v1.z));
}
// mathfunc::scalarProduct (MathFunctions.h:469-477): (v2 - v1).(v4 - v3)
__device__ inline double scalar_product(D3 v1, D3 v2, D3 v3, D3 v4) {
    return ((v2.x - v1.x) * (v4.x - v3.x) + (v2.y - v1.y) * (v4.y - v3.y) +
            (v2.z - v1.z) * (v4.z - v3.z));
}
// mathfunc::midPointCoordinate (MathFunctions.h:963-969)
__device__ inline D3 mid_point(D3 v1, D3 v2, double alpha) {
    double beta = 1 - alpha;
    return D3{(v1.x * beta + alpha * v2.x), (v1.y * beta + alpha * v2.y), (v1.z * beta + alpha * v2.z)};
}

// ---------------- FilamentStretchingHarmonic (FilamentStretchingHarmonic.cpp:185-248) -------
__device__ inline double stretch_energy(D3 c1, D3 c2, double k, double eql) {
    double dist = two_point_distance(c1, c2) - eql;
    return 0.5 * k * dist * dist;
}
// out[0] -> bead 1, out[1] -> bead 2
__device__ inline void stretch_forces(D3 c1, D3 c2, double k, double eql, D3* out) {
    double dist = two_point_distance(c1, c2);
    double invL = 1 / dist;
    double f0 = k * (dist - eql) * invL;
    double r1x = c2.x - c1.x, r1y = c2.y - c1.y, r1z = c2.z - c1.z;
    double Fx = -f0 * r1x, Fy = -f0 * r1y, Fz = -f0 * r1z;
    out[0] = D3{-Fx, -Fy, -Fz};
    out[1] = D3{Fx, Fy, Fz};
}

// ---------------- FilamentBendingCosine (FilamentBendingCosine.cpp:199-345) -----------------
__device__ inline bool are_equal_zero(double a) { return fabs(a - 0.0) < 1e-6; }

__device__ inline double bend_energy(D3 c1, D3 c2, D3 c3, double kbend, double eqt) {
    double L1 = sqrt(scalar_product(c1, c2, c1, c2));
    double L2 = sqrt(scalar_product(c2, c3, c2, c3));
    double L1L2 = L1 * L2;
    double l1l2 = scalar_product(c1, c2, c2, c3);
    double x = l1l2 / L1L2;
    if (x < -1.0) x = -1.0;
    else if (x > 1.0) x = 1.0;
    if (are_equal_zero(eqt)) return kbend * (1 - x);
    double cosA = x;
    double sinA = rmax(sqrt(1 - cosA * cosA), 0.0);
    double cosAminusB = cosA * cos(eqt) + sinA * sin(eqt);
    return kbend * (1 - cosAminusB);
}
__device__ inline void bend_forces(D3 c1, D3 c2, D3 c3, double kbend, double eqt, D3* out) {
    double L1sq = scalar_product(c1, c2, c1, c2);
    double L2sq = scalar_product(c2, c3, c2, c3);
    double l1l2 = scalar_product(c1, c2, c2, c3);
    double A = 1 / sqrt(L1sq * L2sq);
    double x = l1l2 * A;
    double B = x / L1sq;
    double C = x / L2sq;
    double k;
    if (are_equal_zero(eqt)) {
        k = kbend;
    } else {
        if (fabs(fabs(x) - 1.0) < 0.001) x = 0.999 * x;
        if (x < -1.0) x = -1.0;
        else if (x > 1.0) x = 1.0;
        double cosp = x;
        double sinp = sqrt(rmax((1 - cosp * cosp), 0.0));
        double sinpminusq = sinp * cos(eqt) - cosp * sin(eqt);
        k = kbend * sinpminusq / sinp;
    }
    double r1x = c2.x - c1.x, r1y = c2.y - c1.y, r1z = c2.z - c1.z;
    double r2x = c3.x - c2.x, r2y = c3.y - c2.y, r2z = c3.z - c2.z;
    double Fr1x = k * (r2x * A - r1x * B);
    double Fr1y = k * (r2y * A - r1y * B);
    double Fr1z = k * (r2z * A - r1z * B);
    double Fr2x = k * (r1x * A - r2x * C);
    double Fr2y = k * (r1y * A - r2y * C);
    double Fr2z = k * (r1z * A - r2z * C);
    out[0] = D3{-Fr1x, -Fr1y, -Fr1z};
    out[1] = D3{Fr1x - Fr2x, Fr1y - Fr2y, Fr1z - Fr2z};
    out[2] = D3{Fr2x, Fr2y, Fr2z};
}

// ---------------- Linker / MotorGhost stretching (LinkerStretchingHarmonic.cpp:191-291) -----
__device__ inline double xlink_energy(D3 c1, D3 c2, D3 c3, D3 c4, double k, double eql,
                                      double p1, double p2) {
    D3 v1 = mid_point(c1, c2, p1);
    D3 v2 = mid_point(c3, c4, p2);
    double dist = two_point_distance(v1, v2) - eql;
    return 0.5 * k * dist * dist;
}
__device__ inline double xlink_forces(D3 c1, D3 c2, D3 c3, D3 c4, double k, double eql,
                                      double p1, double p2, D3* out) {
    D3 v1 = mid_point(c1, c2, p1);
    D3 v2 = mid_point(c3, c4, p2);
    double dist = two_point_distance(v1, v2);
    double invL = 1 / dist;
    double f0 = k * (dist - eql) * invL;
    double lx = v2.x - v1.x, ly = v2.y - v1.y, lz = v2.z - v1.z;
    double Flx = -f0 * lx, Fly = -f0 * ly, Flz = -f0 * lz;
    out[0] = D3{-Flx * (1 - p1), -Fly * (1 - p1), -Flz * (1 - p1)};
    out[1] = D3{-Flx * (p1), -Fly * (p1), -Flz * (p1)};
    out[2] = D3{Flx * (1 - p2), Fly * (1 - p2), Flz * (1 - p2)};
    out[3] = D3{Flx * (p2), Fly * (p2), Flz * (p2)};
    return f0 / invL;   // stretchforce
}

// ---------------- CylinderExclVolRepulsion (CylinderExclVolRepulsion.cpp:330-882) -----------
__device__ inline void cross3(const double* v1, const double* v2, double* cp) {
    cp[0] = v1[1] * v2[2] - v1[2] * v2[1];
    cp[1] = v1[2] * v2[0] - v1[0] * v2[2];
    cp[2] = v1[0] * v2[1] - v1[1] * v2[0];
}
__device__ inline void normalize3(double* v) {
    double x2 = v[0] * v[0], y2 = v[1] * v[1], z2 = v[2] * v[2];
    double norm = sqrt(x2 + y2 + z2);
    v[0] = v[0] / norm; v[1] = v[1] / norm; v[2] = v[2] / norm;
}
// areInPlane (MathFunctions.h:1101-1135), areEqualLT threshold 0.1 (utility.h:80-83)
__device__ inline bool are_in_plane(D3 p1, D3 p2, D3 p3, D3 p4) {
    double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    double v2[3] = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    double v3[3] = {p4.x - p1.x, p4.y - p1.y, p4.z - p1.z};
    double cp[3];
    cross3(v1, v2, cp);
    normalize3(cp);
    normalize3(v3);
    double dot = v3[0] * cp[0] + v3[1] * cp[1] + v3[2] * cp[2];
    return fabs(dot - 0.0) < 1E-1;
}
// movePointOutOfPlane for bead 1 or 2 (MathFunctions.h:1221-1278)
__device__ inline D3 move_out_of_plane(D3 p1, D3 p2, D3 p3, int i, double d) {
    double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    double v2[3] = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    double n[3];
    cross3(v1, v2, n);
    normalize3(n);
    D3 p = (i == 1) ? p1 : p2;
    return D3{(p.x + n[0] * d), (p.y + n[1] * d), (p.z + n[2] * d)};
}

struct ExvTerms {
    double a, b, c, d, e, F, g, h, I;
    double AA, BB, CC, DD, EE, FF, GG, HH, JJ;
    double ATG1, ATG2, ATG3, ATG4;
};

__device__ inline void exvol_terms(D3 c1, D3 c2, D3 c3, D3 c4, ExvTerms& t) {
    t.a = scalar_product(c1, c2, c1, c2);
    t.b = scalar_product(c3, c4, c3, c4);
    t.c = scalar_product(c3, c1, c3, c1);
    t.d = scalar_product(c1, c2, c3, c4);
    t.e = scalar_product(c1, c2, c3, c1);
    t.F = scalar_product(c3, c4, c3, c1);
    t.g = scalar_product(c1, c2, c1, c4);
    t.h = scalar_product(c3, c4, c3, c2);
    t.I = scalar_product(c3, c4, c1, c4);
    double x = c1.x - c4.x, y = c1.y - c4.y, z = c1.z - c4.z;
    double sqmag_D = (x * x + y * y + z * z);
    x = c3.x - c2.x; y = c3.y - c2.y; z = c3.z - c2.z;
    double sqmag_E = (x * x + y * y + z * z);
    double vA[3] = {c1.x - c2.x, c1.y - c2.y, c1.z - c2.z};
    double vB[3] = {c3.x - c4.x, c3.y - c4.y, c3.z - c4.z};
    double vC[3] = {c3.x - c1.x, c3.y - c1.y, c3.z - c1.z};
    const double a = t.a, b = t.b, c = t.c, d = t.d, e = t.e, F = t.F;
    double ac = sqrt(a * c);
    double bc = sqrt(b * c);
    double aD = sqrt(a * sqmag_D);
    double bE = sqrt(b * sqmag_E);
    t.AA = sqrt(rmax(0.0, (ac + e) * (ac - e)));
    t.BB = sqrt(rmax(0.0, (bc + F) * (bc - F)));
    t.CC = d * e - a * F;
    t.DD = b * e - d * F;
    t.EE = sqrt(rmax(0.0, (aD + t.g) * (aD - t.g)));
    t.FF = sqrt(rmax(0.0, (bE + t.h) * (bE - t.h)));
    t.GG = d * t.g - a * t.I;
    t.HH = t.CC + t.GG - t.DD;
    double cp[3];
    cross3(vA, vB, cp);
    double JJ = (cp[0] * vC[0] + cp[1] * vC[1] + cp[2] * vC[2]);
    t.JJ = -JJ * JJ;
    const double AA = t.AA, BB = t.BB, EE = t.EE, FF = t.FF;
    t.ATG1 = atan((a + e) / AA) - atan(e / AA);
    t.ATG2 = atan((a + e - d) / EE) - atan((e - d) / EE);
    t.ATG3 = atan((F) / BB) - atan((F - b) / BB);
    t.ATG4 = atan((d + F) / FF) - atan((d + F - b) / FF);
}

__device__ inline bool energy_bad(double U) { return fabs(U) == INFINITY || U != U || U < -1.0; }

// closed-form pair energy (CylinderExclVolRepulsion.cpp:353-435)
__device__ inline double exvol_energy_closed(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled) {
    ExvTerms t;
    exvol_terms(c1, c2, c3, c4, t);
    return 0.5 * kRepScaled / t.JJ *
           (t.CC / t.AA * t.ATG1 + t.GG / t.EE * t.ATG2 + t.DD / t.BB * t.ATG3 + t.HH / t.FF * t.ATG4);
}

// energyN: 400x400 Simpson fallback with the reference's sample placement
// (CylinderExclVolRepulsion.cpp:884-982): the loop advances s (t) before use, so row i samples
// s = (i+1)/400; rows 1..399 carry the edge terms, row 0 the t-edges, rows 1..399 x cols 1..399
// the interior. Block-parallel: every thread sums a strided share, then a block tree.
__device__ inline double exvol_integrand(double a, double b, double c, double d, double e, double F,
                                         double s, double t) {
    double r_2 = c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;
    return (1.0 / (r_2 * r_2));
}
// returns this thread's share of (T2 + T3 + T4); T1 is added by the caller
__device__ inline double exvol_energyN_share(D3 c1, D3 c2, D3 c3, D3 c4, int tid, int nthreads) {
    double a = scalar_product(c1, c2, c1, c2), b = scalar_product(c3, c4, c3, c4);
    double c = scalar_product(c3, c1, c3, c1), d = scalar_product(c1, c2, c3, c4);
    double e = scalar_product(c1, c2, c3, c1), F = scalar_product(c3, c4, c3, c1);
    const int N = 400;
    const double delta = 1.0 / ((double)N);
    double acc = 0.0;
    for (int k = tid; k < (N + 1) * (N + 1); k += nthreads) {
        const int i = k / (N + 1), j = k % (N + 1);
        const double sv = (i + 1) * delta, tv = (j + 1) * delta;
        const double fs = (i % 2 == 0) ? 2.0 : 4.0, ft = (j % 2 == 0) ? 2.0 : 4.0;
        if (j == 0 && i > 0 && i < N)
            acc += fs * (exvol_integrand(a, b, c, d, e, F, sv, 0.0) + exvol_integrand(a, b, c, d, e, F, sv, 1.0));
        if (j > 0 && j < N) {
            if (i == 0)
                acc += ft * (exvol_integrand(a, b, c, d, e, F, 0.0, tv) + exvol_integrand(a, b, c, d, e, F, 1.0, tv));
            else if (i < N)
                acc += fs * ft * exvol_integrand(a, b, c, d, e, F, sv, tv);
        }
    }
    return acc;
}
__device__ inline double exvol_energyN_finish(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, double rest) {
    double a = scalar_product(c1, c2, c1, c2), b = scalar_product(c3, c4, c3, c4);
    double c = scalar_product(c3, c1, c3, c1), d = scalar_product(c1, c2, c3, c4);
    double e = scalar_product(c1, c2, c3, c1), F = scalar_product(c3, c4, c3, c1);
    const double delta = 1.0 / 400.0;
    double T1 = exvol_integrand(a, b, c, d, e, F, 0.0, 0.0) + exvol_integrand(a, b, c, d, e, F, 0.0, 1.0) +
                exvol_integrand(a, b, c, d, e, F, 1.0, 0.0) + exvol_integrand(a, b, c, d, e, F, 1.0, 1.0);
    return kRepScaled * (delta * delta / 9.0) * (T1 + rest);
}

// Closed-form pair energy; *deferred is set when it is non-finite or < -1 and the Simpson
// fallback must be evaluated (by a block-parallel pass).
__device__ inline double exvol_energy(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, int* deferred) {
    double U = exvol_energy_closed(c1, c2, c3, c4, kRepScaled);
    if (energy_bad(U)) {
        *deferred = 1;
        U = 0.0;
    }
    return U;
}

__device__ inline bool bad_component(double v) {
    double a = fabs(v);
    return a == INFINITY || isnan(a) || isinf(a) || a > 1e15;
}

// Analytic pair forces (CylinderExclVolRepulsion.cpp:610-858) on the 4 beads, from the pair
// invariants t of the (possibly shifted) geometry c1, c2, c3, c4.
__device__ inline bool exvol_forces_from(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, const ExvTerms& t, D3* out);

// Analytic pair forces (CylinderExclVolRepulsion.cpp:610-858) on the 4 beads. Returns false
// when a component is non-finite or > 1e15, in which case the reference switches to the
// 10^4 x 10^4 Simpson forceN (flagged for the fallback pass).
__device__ inline bool exvol_forces(D3 c1, D3 c2in, D3 c3, D3 c4, double kRepScaled, D3* out) {
    D3 c2 = c2in;
    if (are_in_plane(c1, c2, c3, c4)) c2 = move_out_of_plane(c1, c2, c3, 2, 0.01);
    ExvTerms t;
    exvol_terms(c1, c2, c3, c4, t);
    return exvol_forces_from(c1, c2, c3, c4, kRepScaled, t, out);
}

// Energy and forces of one pair: the invariants are shared when the coplanar shift does not
// apply (the energy never shifts, CylinderExclVolRepulsion.cpp:353-435 vs :610-630).
// *deferred: the closed-form energy needs the Simpson fallback.
__device__ inline double exvol_energy_forces(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, int* deferred,
                                             bool* fok, D3* out) {
    ExvTerms t;
    exvol_terms(c1, c2, c3, c4, t);
    double U = 0.5 * kRepScaled / t.JJ *
               (t.CC / t.AA * t.ATG1 + t.GG / t.EE * t.ATG2 + t.DD / t.BB * t.ATG3 + t.HH / t.FF * t.ATG4);
    if (are_in_plane(c1, c2, c3, c4)) {
        D3 c2s = move_out_of_plane(c1, c2, c3, 2, 0.01);
        ExvTerms ts;
        exvol_terms(c1, c2s, c3, c4, ts);
        *fok = exvol_forces_from(c1, c2s, c3, c4, kRepScaled, ts, out);
    } else {
        *fok = exvol_forces_from(c1, c2, c3, c4, kRepScaled, t, out);
    }
    if (energy_bad(U)) {
        *deferred = 1;
        U = 0.0;
    }
    return U;
}

__device__ inline bool exvol_forces_from(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, const ExvTerms& t, D3* out) {
    const double a = t.a, b = t.b, c = t.c, d = t.d, e = t.e, F = t.F;
    const double AA = t.AA, BB = t.BB, CC = t.CC, DD = t.DD, EE = t.EE, FF = t.FF, GG = t.GG,
                 HH = t.HH, JJ = t.JJ;
    const double ATG1 = t.ATG1, ATG2 = t.ATG2, ATG3 = t.ATG3, ATG4 = t.ATG4;
    const double invJJ = 1 / JJ;
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
    const double block_23 = blockA_C * (-a) + blockE_G * a + blockE_E * (-a / EE) + blockB_D * (-d) +
                            blockB_B * (-F / BB) + blockB_f_minus_b + blockF_H * d +
                            blockF_F * (-(d + F) / FF) + blockF_d_plus_f_minus_b;
    const double block_31 = block_13;
    const double block_32 = block_23;
    const double block_33 = blockA_A * (a / AA) + blockE_E * (a / EE) + blockB_B * (b / BB) + blockF_F * (b / FF);

    const double JJ_11 = 2 * (-b * c + F * F);
    const double JJ_12 = 2 * (c * d - e * F);
    const double JJ_13 = 2 * (b * e - d * F);
    const double JJ_21 = 2 * (c * d - e * F);
    const double JJ_22 = 2 * (-a * c + e * e);
    const double JJ_23 = 2 * (-d * e + a * F);
    const double JJ_31 = 2 * (b * e - d * F);
    const double JJ_32 = 2 * (-d * e + a * F);
    const double JJ_33 = 2 * (d * d - a * b);

    double dv[4][3];
    dv[0][0] = enFactor * (-invJJ * sumBlock * (JJ_31 - JJ_11) + (block_31 - block_11));
    dv[0][1] = enFactor * (-invJJ * sumBlock * (JJ_32 - JJ_12) + (block_32 - block_12));
    dv[0][2] = enFactor * (-invJJ * sumBlock * (JJ_33 - JJ_13) + (block_33 - block_13));
    dv[1][0] = enFactor * (-invJJ * sumBlock * (JJ_11) + (block_11));
    dv[1][1] = enFactor * (-invJJ * sumBlock * (JJ_12) + (block_12));
    dv[1][2] = enFactor * (-invJJ * sumBlock * (JJ_13) + (block_13));
    dv[2][0] = enFactor * (-invJJ * sumBlock * (-JJ_31 - JJ_21) + (-block_31 - block_21));
    dv[2][1] = enFactor * (-invJJ * sumBlock * (-JJ_32 - JJ_22) + (-block_32 - block_22));
    dv[2][2] = enFactor * (-invJJ * sumBlock * (-JJ_33 - JJ_23) + (-block_33 - block_23));
    dv[3][0] = enFactor * (-invJJ * sumBlock * (JJ_21) + (block_21));
    dv[3][1] = enFactor * (-invJJ * sumBlock * (JJ_22) + (block_22));
    dv[3][2] = enFactor * (-invJJ * sumBlock * (JJ_23) + (block_23));

    const double u1[3] = {c2.x - c1.x, c2.y - c1.y, c2.z - c1.z};
    const double u2[3] = {c4.x - c3.x, c4.y - c3.y, c4.z - c3.z};
    const double u3[3] = {c1.x - c3.x, c1.y - c3.y, c1.z - c3.z};
    bool ok = true;
    for (int bi = 0; bi < 4; bi++) {
        double f[3];
        for (int k = 0; k < 3; k++) {
            f[k] = -(dv[bi][0] * u1[k] + dv[bi][1] * u2[k] + dv[bi][2] * u3[k]);
            ok = ok && !bad_component(f[k]);
        }
        out[bi] = D3{f[0], f[1], f[2]};
    }
    return ok;
}

// ---------------- BoundaryCylinderRepulsionExp (BoundaryCylinderRepulsionExp.cpp:27-110) ---
struct BElem { int kind; double p[4]; };   // MDN_BND_PLANE = 0, MDN_BND_SPHERE = 1

__device__ inline double boundary_distance(const BElem& e, D3 x) {
    if (e.kind == 0) {
        const double a = e.p[0], b = e.p[1], c = e.p[2], d = e.p[3];
        return (a * x.x + b * x.y + c * x.z + d) / sqrt(a * a + b * b + c * c);
    }
    return e.p[3] - two_point_distance(D3{e.p[0], e.p[1], e.p[2]}, x);
}
__device__ inline D3 boundary_normal(const BElem& e, D3 x) {
    if (e.kind == 0) return D3{e.p[0], e.p[1], e.p[2]};
    D3 c{e.p[0], e.p[1], e.p[2]};
    double invD = 1 / two_point_distance(x, c);
    return D3{invD * (c.x - x.x), invD * (c.y - x.y), invD * (c.z - x.z)};
}
__device__ inline double boundary_energy(const BElem& e, D3 x, double krep, double slen) {
    double r = boundary_distance(e, x);
    double R = -r / slen;
    return krep * exp(R);
}
__device__ inline D3 boundary_force(const BElem& e, D3 x, double krep, double slen) {
    double r = boundary_distance(e, x);
    D3 n = boundary_normal(e, x);
    double R = -r / slen;
    double f0 = krep * exp(R) / slen;
    return D3{f0 * n.x, f0 * n.y, f0 * n.z};
}

}  // namespace mdn
