// mdn_terms.h -- per-interaction energy / force arithmetic of the six force fields, for
// gfx950 device code.
//
// Every function is a template over the arithmetic type R:
//   R = double : the FP64 path. The stretching, bending, crosslink and boundary expressions keep
//                the operand order of the cited reference kernels so that, built with
//                -ffp-contract=off, results round exactly like the reference's IEEE double
//                arithmetic (transcendental calls aside). The excluded-volume forces are this
//                file's own gradient of the closed-form energy (mdn_exvol_grad.inc).
//   R = float  : the mixed-precision path (stretching, bending, crosslinks, boundary; the kernels
//                evaluate excluded volume with R = double in both paths). Absolute coordinates
//                stay FP64; every bond vector / coordinate difference is formed in FP64 and only
//                then rounded to FP32, so the FP32 arithmetic never sees 10^3-10^4 nm absolute
//                positions. Distances to boundary planes, and the lengths and extensions of the
//                harmonic springs, are formed in FP64 for the same reason.
// Both instantiations return energies and forces widened to double; accumulation is FP64.
//
// Citations are relative to medyan-5.4.0/ in the reference tree.
#pragma once

#include <hip/hip_runtime.h>

// Host and device: the same arithmetic also compiles into host-side analysis programs.
#ifndef MDN_HD
#define MDN_HD __host__ __device__
#endif

namespace mdn {

// std::max / std::min semantics (the reference calls std::max(a, b) = (a < b) ? b : a).
template <typename T> MDN_HD inline T rmax(T a, T b) { return (a < b) ? b : a; }
template <typename T> MDN_HD inline T rmin(T a, T b) { return (b < a) ? b : a; }

struct D3 { double x, y, z; };
template <typename R> struct V3 { R x, y, z; };

MDN_HD inline D3 ld3(const double* __restrict__ p, int bead) {
    return D3{p[3 * bead], p[3 * bead + 1], p[3 * bead + 2]};
}
// a - b in FP64, then rounded to R
template <typename R> MDN_HD inline V3<R> dif(D3 a, D3 b) {
    return V3<R>{(R)(a.x - b.x), (R)(a.y - b.y), (R)(a.z - b.z)};
}
template <typename R> MDN_HD inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename R> MDN_HD inline D3 widen(V3<R> a) { return D3{(double)a.x, (double)a.y, (double)a.z}; }

// ---- forward-mode dual numbers (Hessian-vector products) -------------------------------
// Dual<T> carries a value and its directional derivative; the term functions instantiated with
// R = Dual<double> on dual positions (x, v) return the forces and their derivative along v.
// Comparisons (branches such as clamps and the coplanar test) use the value.
template <typename T> struct Dual {
    T v, d;
    MDN_HD Dual() : v(0), d(0) {}
    MDN_HD Dual(T x) : v(x), d(0) {}
    MDN_HD Dual(T x, T dx) : v(x), d(dx) {}
    template <typename U> MDN_HD explicit Dual(const Dual<U>& o) : v((T)o.v), d((T)o.d) {}
    MDN_HD friend Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
    MDN_HD friend Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
    MDN_HD friend Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
    MDN_HD friend Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
    MDN_HD friend Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
    MDN_HD friend Dual operator+(Dual a, T b) { return Dual(a.v + b, a.d); }
    MDN_HD friend Dual operator+(T a, Dual b) { return Dual(a + b.v, b.d); }
    MDN_HD friend Dual operator-(Dual a, T b) { return Dual(a.v - b, a.d); }
    MDN_HD friend Dual operator-(T a, Dual b) { return Dual(a - b.v, -b.d); }
    MDN_HD friend Dual operator*(Dual a, T b) { return Dual(a.v * b, a.d * b); }
    MDN_HD friend Dual operator*(T a, Dual b) { return Dual(a * b.v, a * b.d); }
    MDN_HD friend Dual operator/(Dual a, T b) { return Dual(a.v / b, a.d / b); }
    MDN_HD friend Dual operator/(T a, Dual b) { return Dual(a / b.v, -a * b.d / (b.v * b.v)); }
    MDN_HD friend bool operator<(Dual a, Dual b) { return a.v < b.v; }
    MDN_HD friend bool operator>(Dual a, Dual b) { return a.v > b.v; }
    MDN_HD friend bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
    MDN_HD friend bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
    MDN_HD friend bool operator==(Dual a, Dual b) { return a.v == b.v; }
    MDN_HD friend bool operator!=(Dual a, Dual b) { return a.v != b.v; }
    MDN_HD friend Dual sqrt(Dual a) {
        const T r = sqrt(a.v);
        return Dual(r, a.d / (2 * r));
    }
    MDN_HD friend Dual atan(Dual a) { return Dual(atan(a.v), a.d / (1 + a.v * a.v)); }
    MDN_HD friend Dual exp(Dual a) {
        const T e = exp(a.v);
        return Dual(e, a.d * e);
    }
    MDN_HD friend Dual fabs(Dual a) { return a.v < 0 ? -a : a; }
};
typedef Dual<double> DualD;
typedef Dual<float> DualF;   // the mixed path's tangents

// a position and its tangent
struct DualD3 { D3 v, d; };
template <typename R> MDN_HD inline V3<R> dif(DualD3 a, DualD3 b) {
    return V3<R>{R(a.v.x - b.v.x, a.d.x - b.d.x), R(a.v.y - b.v.y, a.d.y - b.d.y), R(a.v.z - b.v.z, a.d.z - b.d.z)};
}
// the FP64 companion type of an arithmetic type (spring lengths, boundary distances)
template <typename R> struct Wide { typedef double T; };
template <> struct Wide<DualD> { typedef DualD T; };
template <> struct Wide<DualF> { typedef DualD T; };

// force outputs: D3 for the value paths, DualD3 (value + derivative) for the tangent path
MDN_HD inline void put3(D3& o, double x, double y, double z) { o = D3{x, y, z}; }
MDN_HD inline void put3(D3& o, float x, float y, float z) { o = D3{(double)x, (double)y, (double)z}; }
MDN_HD inline void put3(DualD3& o, DualD x, DualD y, DualD z) { o = DualD3{D3{x.v, y.v, z.v}, D3{x.d, y.d, z.d}}; }
MDN_HD inline void put3(DualD3& o, DualF x, DualF y, DualF z) {
    o = DualD3{D3{(double)x.v, (double)y.v, (double)z.v}, D3{(double)x.d, (double)y.d, (double)z.d}};
}

// ---------------- FilamentStretchingHarmonic (FilamentStretchingHarmonic.cpp:185-248) -------
// The harmonic springs (filament stretching, linkers, motors): the length L and the extension
// L - L0 are formed in FP64 in both paths (L0 reaches micrometres for crosslinks, where an FP32
// L - L0 would lose the force to cancellation); the mixed path rounds the tension
// k (L - L0) / L to FP32 and forms the force components in FP32.
// twoPointDistance (MathFunctions.h:294-297) of r = c2 - c1
template <typename R> MDN_HD inline R stretch_energy(D3 c1, D3 c2, double k, double eql) {
    V3<double> r = dif<double>(c2, c1);
    double dist = sqrt(r.x * r.x + r.y * r.y + r.z * r.z) - eql;
    return (R)(0.5 * k * dist * dist);
}
// out[0] -> bead 1, out[1] -> bead 2
template <typename R, typename P, typename O>
MDN_HD inline void stretch_forces(P c1, P c2, double k, double eql, O* out) {
    typedef typename Wide<R>::T W;
    V3<W> r64 = dif<W>(c2, c1);
    W dist = sqrt(r64.x * r64.x + r64.y * r64.y + r64.z * r64.z);
    W invL = 1 / dist;
    const R f0 = (R)(k * (dist - eql) * invL);
    const V3<R> r{(R)r64.x, (R)r64.y, (R)r64.z};
    R Fx = -f0 * r.x, Fy = -f0 * r.y, Fz = -f0 * r.z;
    put3(out[0], -Fx, -Fy, -Fz);
    put3(out[1], Fx, Fy, Fz);
}

// ---------------- FilamentBendingCosine (FilamentBendingCosine.cpp:199-345) -----------------
MDN_HD inline bool are_equal_zero(double a) { return fabs(a - 0.0) < 1e-6; }   // utility.h:75-77

template <typename R> MDN_HD inline R bend_energy(D3 c1, D3 c2, D3 c3, double kd, double eqtd) {
    const R kbend = (R)kd;
    V3<R> r1 = dif<R>(c2, c1), r2 = dif<R>(c3, c2);
    R L1 = sqrt(dot(r1, r1));
    R L2 = sqrt(dot(r2, r2));
    if constexpr (sizeof(R) == 4) {
        // FP32 with theta0 = 0: 1 - cos(theta) as half the squared chord of the unit bond vectors,
        // |r1 / L1 - r2 / L2|^2 / 2, which does not cancel for nearly straight triples (1 - x does:
        // a relative error of eps / theta^2 there)
        if (are_equal_zero(eqtd)) {
            const R i1 = 1 / L1, i2 = 1 / L2;
            const R dx = r1.x * i1 - r2.x * i2, dy = r1.y * i1 - r2.y * i2, dz = r1.z * i1 - r2.z * i2;
            return kbend * ((dx * dx + dy * dy + dz * dz) / 2);
        }
    }
    R L1L2 = L1 * L2;
    R l1l2 = dot(r1, r2);
    R x = l1l2 / L1L2;
    if (x < (R)-1.0) x = (R)-1.0;
    else if (x > (R)1.0) x = (R)1.0;
    if (are_equal_zero(eqtd)) return kbend * (1 - x);
    const R eqt = (R)eqtd;
    R cosA = x;
    R sinA = rmax(sqrt(1 - cosA * cosA), (R)0.0);
    R cosAminusB = cosA * cos(eqt) + sinA * sin(eqt);
    return kbend * (1 - cosAminusB);
}
template <typename R, typename P, typename O>
MDN_HD inline void bend_forces(P c1, P c2, P c3, double kd, double eqtd, O* out) {
    const R kbend = (R)kd;
    V3<R> r1 = dif<R>(c2, c1), r2 = dif<R>(c3, c2);
    R L1sq = dot(r1, r1);
    R L2sq = dot(r2, r2);
    R l1l2 = dot(r1, r2);
    R A = 1 / sqrt(L1sq * L2sq);
    R x = l1l2 * A;
    R B = x / L1sq;
    R C = x / L2sq;
    R k;
    if (are_equal_zero(eqtd)) {
        k = kbend;
    } else {
        const R ce = (R)cos(eqtd), se = (R)sin(eqtd);   // constants of the term
        if (fabs(fabs(x) - (R)1.0) < (R)0.001) x = (R)0.999 * x;
        if (x < (R)-1.0) x = (R)-1.0;
        else if (x > (R)1.0) x = (R)1.0;
        R cosp = x;
        R sinp = sqrt(rmax((1 - cosp * cosp), (R)0.0));
        R sinpminusq = sinp * ce - cosp * se;
        k = kbend * sinpminusq / sinp;
    }
    R Fr1x = k * (r2.x * A - r1.x * B);
    R Fr1y = k * (r2.y * A - r1.y * B);
    R Fr1z = k * (r2.z * A - r1.z * B);
    R Fr2x = k * (r1.x * A - r2.x * C);
    R Fr2y = k * (r1.y * A - r2.y * C);
    R Fr2z = k * (r1.z * A - r2.z * C);
    put3(out[0], -Fr1x, -Fr1y, -Fr1z);
    put3(out[1], Fr1x - Fr2x, Fr1y - Fr2y, Fr1z - Fr2z);
    put3(out[2], Fr2x, Fr2y, Fr2z);
}

// ---------------- Linker / MotorGhost stretching (LinkerStretchingHarmonic.cpp:191-291) -----
// midPointCoordinate (MathFunctions.h:963-969) in FP64 (absolute positions)
MDN_HD inline D3 mid_point(D3 v1, D3 v2, double alpha) {
    double beta = 1 - alpha;
    return D3{(v1.x * beta + alpha * v2.x), (v1.y * beta + alpha * v2.y), (v1.z * beta + alpha * v2.z)};
}
MDN_HD inline DualD3 mid_point(DualD3 v1, DualD3 v2, double alpha) {
    return DualD3{mid_point(v1.v, v2.v, alpha), mid_point(v1.d, v2.d, alpha)};
}
template <typename R>
MDN_HD inline R xlink_energy(D3 c1, D3 c2, D3 c3, D3 c4, double k, double eql, double p1, double p2) {
    D3 v1 = mid_point(c1, c2, p1);
    D3 v2 = mid_point(c3, c4, p2);
    V3<double> l = dif<double>(v2, v1);
    double dist = sqrt(l.x * l.x + l.y * l.y + l.z * l.z) - eql;
    return (R)(0.5 * k * dist * dist);
}
template <typename R, typename P, typename O>
MDN_HD inline typename Wide<R>::T xlink_forces(P c1, P c2, P c3, P c4, double k, double eql, double p1d, double p2d,
                                               O* out) {
    typedef typename Wide<R>::T W;
    P v1 = mid_point(c1, c2, p1d);
    P v2 = mid_point(c3, c4, p2d);
    V3<W> l64 = dif<W>(v2, v1);
    W dist = sqrt(l64.x * l64.x + l64.y * l64.y + l64.z * l64.z);
    W invL = 1 / dist;
    W f064 = k * (dist - eql) * invL;
    const R f0 = (R)f064;
    const V3<R> l{(R)l64.x, (R)l64.y, (R)l64.z};
    R Flx = -f0 * l.x, Fly = -f0 * l.y, Flz = -f0 * l.z;
    const R p1 = (R)p1d, p2 = (R)p2d;
    put3(out[0], -Flx * (1 - p1), -Fly * (1 - p1), -Flz * (1 - p1));
    put3(out[1], -Flx * (p1), -Fly * (p1), -Flz * (p1));
    put3(out[2], Flx * (1 - p2), Fly * (1 - p2), Flz * (1 - p2));
    put3(out[3], Flx * (p2), Fly * (p2), Flz * (p2));
    return f064 / invL;   // stretchforce
}

// ---------------- CylinderExclVolRepulsion (CylinderExclVolRepulsion.cpp:330-882) -----------
MDN_HD inline void cross3(const double* v1, const double* v2, double* cp) {
    cp[0] = v1[1] * v2[2] - v1[2] * v2[1];
    cp[1] = v1[2] * v2[0] - v1[0] * v2[2];
    cp[2] = v1[0] * v2[1] - v1[1] * v2[0];
}
MDN_HD inline void normalize3(double* v) {
    double x2 = v[0] * v[0], y2 = v[1] * v[1], z2 = v[2] * v[2];
    double norm = sqrt(x2 + y2 + z2);
    v[0] = v[0] / norm; v[1] = v[1] / norm; v[2] = v[2] / norm;
}
// areInPlane (MathFunctions.h:1101-1135), areEqualLT threshold 0.1 (utility.h:80-83); FP64
MDN_HD inline bool are_in_plane(D3 p1, D3 p2, D3 p3, D3 p4) {
    double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    double v2[3] = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    double v3[3] = {p4.x - p1.x, p4.y - p1.y, p4.z - p1.z};
    double cp[3];
    cross3(v1, v2, cp);
    normalize3(cp);
    normalize3(v3);
    double dot3 = v3[0] * cp[0] + v3[1] * cp[1] + v3[2] * cp[2];
    return fabs(dot3 - 0.0) < 1E-1;
}
// movePointOutOfPlane for bead 1 or 2 (MathFunctions.h:1221-1278); FP64
MDN_HD inline D3 move_out_of_plane(D3 p1, D3 p2, D3 p3, int i, double d) {
    double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    double v2[3] = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    double n[3];
    cross3(v1, v2, n);
    normalize3(n);
    D3 p = (i == 1) ? p1 : p2;
    return D3{(p.x + n[0] * d), (p.y + n[1] * d), (p.z + n[2] * d)};
}

// the coplanar test on the values; the shift of bead 2 with its derivative
MDN_HD inline bool are_in_plane(DualD3 p1, DualD3 p2, DualD3 p3, DualD3 p4) {
    return are_in_plane(p1.v, p2.v, p3.v, p4.v);
}
MDN_HD inline DualD3 move_out_of_plane(DualD3 p1, DualD3 p2, DualD3 p3, int i, double d) {
    const V3<DualD> v1 = dif<DualD>(p2, p1), v2 = dif<DualD>(p3, p1);
    V3<DualD> n{v1.y * v2.z - v1.z * v2.y, v1.z * v2.x - v1.x * v2.z, v1.x * v2.y - v1.y * v2.x};
    const DualD norm = sqrt(n.x * n.x + n.y * n.y + n.z * n.z);
    n = V3<DualD>{n.x / norm, n.y / norm, n.z / norm};
    const DualD3 p = (i == 1) ? p1 : p2;
    const DualD px = DualD(p.v.x, p.d.x) + n.x * d, py = DualD(p.v.y, p.d.y) + n.y * d,
                pz = DualD(p.v.z, p.d.z) + n.z * d;
    return DualD3{D3{px.v, py.v, pz.v}, D3{px.d, py.d, pz.d}};
}

template <typename R> struct ExvTerms {
    R a, b, c, d, e, F, g, h, I;
    R AA, BB, CC, DD, EE, FF, GG, HH, JJ;
    R ATG1, ATG2, ATG3, ATG4;
    R r14, r23, r24;    // |c4 - c1|^2, |c2 - c3|^2, |c4 - c2|^2 (the force sensitivities' denominators)
    V3<R> u1, u2, u3;   // c2 - c1, c4 - c3, c1 - c3 (the scalars' gradient directions)
};

// The pair invariants (CylinderExclVolRepulsion.cpp:353-397) from five FP64 difference vectors:
// scalarProduct(p1,p2,p3,p4) = (p2 - p1).(p4 - p3), c1 - c2 = -(c2 - c1) exactly, etc.
template <typename R, typename P, bool FAST = false> MDN_HD inline void exvol_terms(P c1, P c2, P c3, P c4, ExvTerms<R>& t) {
    const V3<R> u1 = dif<R>(c2, c1), u2 = dif<R>(c4, c3), u3 = dif<R>(c1, c3);
    const V3<R> w14 = dif<R>(c4, c1), w32 = dif<R>(c2, c3);
    t.u1 = u1; t.u2 = u2; t.u3 = u3;
    t.a = dot(u1, u1);
    t.b = dot(u2, u2);
    t.c = dot(u3, u3);
    t.d = dot(u1, u2);
    t.e = dot(u1, u3);
    t.F = dot(u2, u3);
    t.g = dot(u1, w14);
    t.h = dot(u2, w32);
    t.I = dot(u2, w14);
    const R sqmag_D = (-w14.x * -w14.x + -w14.y * -w14.y + -w14.z * -w14.z);   // |c1 - c4|^2
    const R sqmag_E = (-w32.x * -w32.x + -w32.y * -w32.y + -w32.z * -w32.z);   // |c3 - c2|^2
    const V3<R> w24 = dif<R>(c4, c2);
    t.r14 = sqmag_D;
    t.r23 = sqmag_E;
    t.r24 = dot(w24, w24);
    const R a = t.a, b = t.b, c = t.c, d = t.d, e = t.e, F = t.F;
    R ac = sqrt(a * c);
    R bc = sqrt(b * c);
    R aD = sqrt(a * sqmag_D);
    R bE = sqrt(b * sqmag_E);
    t.AA = sqrt(rmax((R)0.0, (ac + e) * (ac - e)));
    t.BB = sqrt(rmax((R)0.0, (bc + F) * (bc - F)));
    t.CC = d * e - a * F;
    t.DD = b * e - d * F;
    t.EE = sqrt(rmax((R)0.0, (aD + t.g) * (aD - t.g)));
    t.FF = sqrt(rmax((R)0.0, (bE + t.h) * (bE - t.h)));
    t.GG = d * t.g - a * t.I;
    t.HH = t.CC + t.GG - t.DD;
    // vA = c1 - c2 = -u1, vB = c3 - c4 = -u2, vC = c3 - c1 = -u3; JJ = -((vA x vB) . vC)^2
    const R vA0 = -u1.x, vA1 = -u1.y, vA2 = -u1.z, vB0 = -u2.x, vB1 = -u2.y, vB2 = -u2.z;
    const R cp0 = vA1 * vB2 - vA2 * vB1, cp1 = vA2 * vB0 - vA0 * vB2, cp2 = vA0 * vB1 - vA1 * vB0;
    R JJ = (cp0 * -u3.x + cp1 * -u3.y + cp2 * -u3.z);
    t.JJ = -JJ * JJ;
    const R AA = t.AA, BB = t.BB, EE = t.EE, FF = t.FF;
    if constexpr (FAST) {   // the mixed path's form (below)
        t.ATG1 = atan2(a * AA, AA * AA + (a + e) * e);
        t.ATG2 = atan2(a * EE, EE * EE + (a + e - d) * (e - d));
        t.ATG3 = atan2(b * BB, BB * BB + F * (F - b));
        t.ATG4 = atan2(b * FF, FF * FF + (d + F) * (d + F - b));
    } else {
        t.ATG1 = atan((a + e) / AA) - atan(e / AA);
        t.ATG2 = atan((a + e - d) / EE) - atan((e - d) / EE);
        t.ATG3 = atan((F) / BB) - atan((F - b) / BB);
        t.ATG4 = atan((d + F) / FF) - atan((d + F - b) / FF);
    }
}

template <typename R> MDN_HD inline bool energy_bad(R U) {
    return fabs(U) == (R)INFINITY || U != U || U < (R)-1.0;
}

template <typename R> MDN_HD inline R exvol_energy_of(const ExvTerms<R>& t, R kRepScaled) {
    return (R)0.5 * kRepScaled / t.JJ *
           (t.CC / t.AA * t.ATG1 + t.GG / t.EE * t.ATG2 + t.DD / t.BB * t.ATG3 + t.HH / t.FF * t.ATG4);
}

// energyN: 400x400 Simpson fallback with the reference's sample placement
// (CylinderExclVolRepulsion.cpp:884-982): the loop advances s (t) before use, so row i samples
// s = (i+1)/400; rows 1..399 carry the edge terms, row 0 the t-edges, rows 1..399 x cols 1..399
// the interior. Block-parallel (FP64 in both paths): every thread sums a strided share, then a
// block tree.
MDN_HD inline double exvol_integrand(double a, double b, double c, double d, double e, double F,
                                         double s, double t) {
    double r_2 = c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;
    return (1.0 / (r_2 * r_2));
}
struct SixInv { double a, b, c, d, e, F; };
MDN_HD inline SixInv six_inv(D3 c1, D3 c2, D3 c3, D3 c4) {
    const V3<double> u1 = dif<double>(c2, c1), u2 = dif<double>(c4, c3), u3 = dif<double>(c1, c3);
    return SixInv{dot(u1, u1), dot(u2, u2), dot(u3, u3), dot(u1, u2), dot(u1, u3), dot(u2, u3)};
}
// returns this thread's share of (T2 + T3 + T4); T1 is added by exvol_energyN_finish
MDN_HD inline double exvol_energyN_share(D3 c1, D3 c2, D3 c3, D3 c4, int tid, int nthreads) {
    const SixInv v = six_inv(c1, c2, c3, c4);
    const int N = 400;
    const double delta = 1.0 / ((double)N);
    double acc = 0.0;
    for (int k = tid; k < (N + 1) * (N + 1); k += nthreads) {
        const int i = k / (N + 1), j = k % (N + 1);
        const double sv = (i + 1) * delta, tv = (j + 1) * delta;
        const double fs = (i % 2 == 0) ? 2.0 : 4.0, ft = (j % 2 == 0) ? 2.0 : 4.0;
        if (j == 0 && i > 0 && i < N)
            acc += fs * (exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, sv, 0.0) +
                         exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, sv, 1.0));
        if (j > 0 && j < N) {
            if (i == 0)
                acc += ft * (exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, 0.0, tv) +
                             exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, 1.0, tv));
            else if (i < N)
                acc += fs * ft * exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, sv, tv);
        }
    }
    return acc;
}
MDN_HD inline double exvol_energyN_finish(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, double rest) {
    const SixInv v = six_inv(c1, c2, c3, c4);
    const double delta = 1.0 / 400.0;
    double T1 = exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, 0.0, 0.0) + exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, 0.0, 1.0) +
                exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, 1.0, 0.0) + exvol_integrand(v.a, v.b, v.c, v.d, v.e, v.F, 1.0, 1.0);
    return kRepScaled * (delta * delta / 9.0) * (T1 + rest);
}

// Closed-form pair energy; *deferred is set when it is non-finite or < -1 and the Simpson
// fallback must be evaluated (by a block-parallel pass).
template <typename R>
MDN_HD inline double exvol_energy(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, int* deferred) {
    ExvTerms<R> t;
    exvol_terms<R>(c1, c2, c3, c4, t);
    R U = exvol_energy_of<R>(t, (R)kRepScaled);
    if (energy_bad<R>(U)) {
        *deferred = 1;
        return 0.0;
    }
    return (double)U;
}

template <typename R> MDN_HD inline bool bad_component(R v) {
    R a = fabs(v);
    return a != a || a > (R)1e15;   // NaN, +-inf or beyond 1e15 (checkNaN_INF, CylinderExclVolRepulsion.cpp:862-864)
}

// Pair forces -dU/dc_k on the 4 beads from the invariants of the (possibly shifted) geometry,
// by the chain rule through the six scalars (derivation in mdn_exvol_grad.inc). Returns false
// when a component is non-finite or > 1e15, in which case the reference switches to the
// 10^4 x 10^4 Simpson forceN (CylinderExclVolRepulsion.cpp:859-872).
template <typename R, typename O>
MDN_HD inline bool exvol_forces_from(const ExvTerms<R>& t, R K, O* out)
#include "mdn_exvol_grad.inc"

// Forces only: the coplanar shift of c2 (CylinderExclVolRepulsion.cpp:610-630) applies.
template <typename R, typename P, typename O>
MDN_HD inline bool exvol_forces(P c1, P c2in, P c3, P c4, double kRepScaled, O* out) {
    P c2 = c2in;
    if (are_in_plane(c1, c2, c3, c4)) c2 = move_out_of_plane(c1, c2, c3, 2, 0.01);
    ExvTerms<R> t;
    exvol_terms<R>(c1, c2, c3, c4, t);
    return exvol_forces_from<R>(t, (R)kRepScaled, out);
}

// Energy and forces of one pair: the invariants are shared when the coplanar shift does not
// apply (the energy never shifts, CylinderExclVolRepulsion.cpp:353-435 vs :610-630).
// *deferred: the closed-form energy needs the Simpson fallback.
template <typename R>
MDN_HD inline double exvol_energy_forces(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, int* deferred,
                                             bool* fok, D3* out) {
    ExvTerms<R> t;
    exvol_terms<R>(c1, c2, c3, c4, t);
    R U = exvol_energy_of<R>(t, (R)kRepScaled);
    // a coplanar pair's invariants are replaced by the shifted geometry's, so a wave holding
    // both kinds of pairs runs the force assembly once (two calls would run it under both masks)
    if (are_in_plane(c1, c2, c3, c4)) {
        D3 c2s = move_out_of_plane(c1, c2, c3, 2, 0.01);
        exvol_terms<R>(c1, c2s, c3, c4, t);
    }
    *fok = exvol_forces_from<R>(t, (R)kRepScaled, out);
    if (energy_bad<R>(U)) {
        *deferred = 1;
        return 0.0;
    }
    return (double)U;
}

// ---- the mixed path's excluded volume: the same closed form in fewer FP64 operations -------
// The energy form above spends most of its issue slots on eight atan calls. With AA > 0 both
// angles of atan(p / AA) - atan(q / AA) lie in (-pi/2, pi/2), so
//     atan(p / AA) - atan(q / AA) = atan2((p - q) AA, AA^2 + p q)
// (tangent of a difference; the cosine of the difference has the sign of AA^2 + p q): four
// atan2 instead of eight atan, with p - q (= a or b) entering exactly instead of through the
// cancellation of two angles. The force gradient is the same as the FP64 path's with FMA
// contraction. In exact arithmetic every quantity equals the FP64 path's; the rounding differs.
// The invariants a ... JJ are the reference's, bit for bit, so the degenerate geometries (AA,
// BB, EE, FF or JJ exactly zero) give non-finite values exactly where the reference's do and the
// Simpson fallbacks still trigger there.
// The coplanar test without its two normalizations: |n.v3| < 0.1 for the unit normal n of
// (p2 - p1, p3 - p1) and the unit v3 = (p4 - p1) / |p4 - p1| is (cp.w)^2 < 0.01 |cp|^2 |w|^2 for
// the unnormalized cp and w (a degenerate cp or w, where the reference's test reads NaN and
// fails, gives 0 < 0 and fails too). Pairs within rounding of the threshold may be classified
// differently from the reference's form.
MDN_HD inline bool are_in_plane_fast(D3 p1, D3 p2, D3 p3, D3 p4) {
    const double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    const double v2[3] = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    const double w[3] = {p4.x - p1.x, p4.y - p1.y, p4.z - p1.z};
    double cp[3];
    cross3(v1, v2, cp);
    const double cw = cp[0] * w[0] + cp[1] * w[1] + cp[2] * w[2];
    const double cc = cp[0] * cp[0] + cp[1] * cp[1] + cp[2] * cp[2];
    const double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    return cw * cw < 0.01 * (cc * ww);
}
// bead 2 moved by d along the unit normal of (p2 - p1, p3 - p1), with one reciprocal square root
MDN_HD inline D3 move_out_of_plane_fast(D3 p1, D3 p2, D3 p3, double d) {
    const double v1[3] = {p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
    const double v2[3] = {p3.x - p1.x, p3.y - p1.y, p3.z - p1.z};
    double n[3];
    cross3(v1, v2, n);
    const double sc = d / sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    return D3{p2.x + n[0] * sc, p2.y + n[1] * sc, p2.z + n[2] * sc};
}

MDN_HD inline double exvol_energy_of_fast(const ExvTerms<double>& t, double kRepScaled) {
    return exvol_energy_of<double>(t, kRepScaled);   // the atan2 differences are in t
}

// the same gradient with FMA contraction (the mixed path)
template <typename O> MDN_HD inline bool exvol_forces_from_fast(const ExvTerms<double>& t, double K, O* out) {
#pragma clang fp contract(fast)
    typedef double R;
#include "mdn_exvol_grad.inc"
}

// forces only (the coplanar shift applies), energy only, energy and forces: as the reference
// forms above, on the atan2 terms
MDN_HD inline bool exvol_forces_fast(D3 c1, D3 c2in, D3 c3, D3 c4, double kRepScaled, D3* out) {
    D3 c2 = c2in;
    if (are_in_plane_fast(c1, c2, c3, c4)) c2 = move_out_of_plane_fast(c1, c2, c3, 0.01);
    ExvTerms<double> t;
    exvol_terms<double, D3, true>(c1, c2, c3, c4, t);
    return exvol_forces_from_fast(t, kRepScaled, out);
}
MDN_HD inline double exvol_energy_fast(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, int* deferred) {
    ExvTerms<double> t;
    exvol_terms<double, D3, true>(c1, c2, c3, c4, t);
    const double U = exvol_energy_of_fast(t, kRepScaled);
    if (energy_bad<double>(U)) {
        *deferred = 1;
        return 0.0;
    }
    return U;
}
MDN_HD inline double exvol_energy_forces_fast(D3 c1, D3 c2, D3 c3, D3 c4, double kRepScaled, int* deferred,
                                              bool* fok, D3* out) {
    ExvTerms<double> t;
    exvol_terms<double, D3, true>(c1, c2, c3, c4, t);
    const double U = exvol_energy_of_fast(t, kRepScaled);
    if (are_in_plane_fast(c1, c2, c3, c4)) {
        const D3 c2s = move_out_of_plane_fast(c1, c2, c3, 0.01);
        exvol_terms<double, D3, true>(c1, c2s, c3, c4, t);
    }
    *fok = exvol_forces_from_fast(t, kRepScaled, out);
    if (energy_bad<double>(U)) {
        *deferred = 1;
        return 0.0;
    }
    return U;
}

// ---------------- BoundaryCylinderRepulsionExp (BoundaryCylinderRepulsionExp.cpp:27-110) ---
struct BElem { int kind; double p[4]; };   // MDN_BND_PLANE = 0, MDN_BND_SPHERE = 1

// distance in FP64 (PlaneBoundaryElement::distance / SphereBoundaryElement::distance,
// BoundaryElementImpl.h:47-50, :136-142)
MDN_HD inline double boundary_distance(const BElem& e, D3 x) {
    if (e.kind == 0) {
        const double a = e.p[0], b = e.p[1], c = e.p[2], d = e.p[3];
        return (a * x.x + b * x.y + c * x.z + d) / sqrt(a * a + b * b + c * c);
    }
    const double dx = x.x - e.p[0], dy = x.y - e.p[1], dz = x.z - e.p[2];
    return e.p[3] - sqrt(dx * dx + dy * dy + dz * dz);
}
template <typename R> MDN_HD inline V3<R> boundary_normal(const BElem& e, D3 x) {
    if (e.kind == 0) return V3<R>{(R)e.p[0], (R)e.p[1], (R)e.p[2]};
    // twoPointDirection(x, centre) (MathFunctions.h:413-421)
    D3 c{e.p[0], e.p[1], e.p[2]};
    V3<R> v = dif<R>(c, x);
    R invD = 1 / sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
    return V3<R>{invD * v.x, invD * v.y, invD * v.z};
}
template <typename R> MDN_HD inline R boundary_energy(const BElem& e, D3 x, double krep, double slen) {
    R r = (R)boundary_distance(e, x);
    R Rr = -r / (R)slen;
    return (R)krep * exp(Rr);
}
template <typename R> MDN_HD inline D3 boundary_force(const BElem& e, D3 x, double krep, double slen) {
    R r = (R)boundary_distance(e, x);
    V3<R> n = boundary_normal<R>(e, x);
    R Rr = -r / (R)slen;
    R f0 = (R)krep * exp(Rr) / (R)slen;
    return D3{(double)(f0 * n.x), (double)(f0 * n.y), (double)(f0 * n.z)};
}

// the tangent path: distance, normal and force with their derivatives along the tangent
MDN_HD inline DualD boundary_distance(const BElem& e, DualD3 x) {
    if (e.kind == 0) {
        const double a = e.p[0], b = e.p[1], c = e.p[2], d = e.p[3];
        const double nrm = sqrt(a * a + b * b + c * c);
        return DualD((a * x.v.x + b * x.v.y + c * x.v.z + d) / nrm, (a * x.d.x + b * x.d.y + c * x.d.z) / nrm);
    }
    const V3<DualD> r = dif<DualD>(x, DualD3{D3{e.p[0], e.p[1], e.p[2]}, D3{0.0, 0.0, 0.0}});
    return e.p[3] - sqrt(r.x * r.x + r.y * r.y + r.z * r.z);
}
MDN_HD inline DualD3 boundary_force_dual(const BElem& e, DualD3 x, double krep, double slen) {
    const DualD r = boundary_distance(e, x);
    V3<DualD> n;
    if (e.kind == 0) {
        n = V3<DualD>{DualD(e.p[0]), DualD(e.p[1]), DualD(e.p[2])};
    } else {
        const V3<DualD> v = dif<DualD>(DualD3{D3{e.p[0], e.p[1], e.p[2]}, D3{0.0, 0.0, 0.0}}, x);
        const DualD invD = 1 / sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
        n = V3<DualD>{invD * v.x, invD * v.y, invD * v.z};
    }
    const DualD Rr = -r / slen;
    const DualD f0 = krep * exp(Rr) / slen;
    DualD3 o;
    put3(o, f0 * n.x, f0 * n.y, f0 * n.z);
    return o;
}

}  // namespace mdn
