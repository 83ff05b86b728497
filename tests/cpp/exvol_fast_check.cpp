// Host-side checks of the excluded-volume pair arithmetic (picgo_amd/csrc/mdn_terms.h, gradient
// in mdn_exvol_grad.inc) on seeded random pairs of 108 nm cylinders. Built host-only.
//
// 1. The gradient formula itself: the forces of the extended-precision (x87, 64-bit significand)
//    instantiation against Richardson-extrapolated central differences of the closed-form energy
//    (CylinderExclVolRepulsion.cpp:353-435) in the same precision, on pairs outside the coplanar
//    test (where the force is -dE/dx). Prints "gradient_fd <cases> <max relative error>".
// 2. Rounding: the FP64 path's form (atan differences, no contraction) and the mixed path's form
//    (atan2 differences, FMA contraction) against the extended-precision evaluation of the same
//    closed form and gradient. Prints per class (by the normalized triple product tau, the closed
//    form's conditioning ~1/tau^2, and the coplanar-shifted pairs)
//     "<class> <cases> <energy: fast-vs-fp64 max> <force: fast-vs-fp64 max>
//      <energy err 99th pct / max, fp64> <same, fast> <force err 99th pct / max, fp64> <same, fast>"
//    (errors relative to |E| and to the pair's max |F|), then "flags <cases> <mismatches>" (the
//    deferred / non-finite flags and the entry points' agreement) and "degenerate <cases>
//    <mismatches>" on exactly degenerate integer geometries (parallel, an end on the other line,
//    crossing, identical), where both forms must flag the same pairs for the Simpson fallbacks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

// the extended-precision instantiation must find the long double overloads of the functions the
// term code calls unqualified (otherwise the double ones of the global namespace are used)
using std::atan;
using std::atan2;
using std::fabs;
using std::sqrt;

#include "../../picgo_amd/csrc/mdn_terms.h"

namespace mdn {
// the extended-precision instantiation's force output
inline void put3(D3& o, long double x, long double y, long double z) { o = D3{(double)x, (double)y, (double)z}; }
}
using namespace mdn;

static std::mt19937_64 rng(777);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static D3 rnd_dir() {
    std::normal_distribution<double> N(0.0, 1.0);
    double x = N(rng), y = N(rng), z = N(rng), n = std::sqrt(x * x + y * y + z * z);
    return D3{x / n, y / n, z / n};
}
static D3 add(D3 a, D3 b, double s) { return D3{a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }

struct Cls {
    const char* name;
    int n = 0;
    double e_err = 0.0, f_err = 0.0;   // fast vs the FP64 path's form
    std::vector<double> er_ref, er_fast, fr_ref, fr_fast;   // each vs extended precision
    void add(double e_rel, double f_rel, double e_ref, double e_fast, double f_ref, double f_fast) {
        n++;
        e_err = std::fmax(e_err, e_rel);
        f_err = std::fmax(f_err, f_rel);
        er_ref.push_back(e_ref); er_fast.push_back(e_fast); fr_ref.push_back(f_ref); fr_fast.push_back(f_fast);
    }
    static double q(std::vector<double> v, double p) {
        if (v.empty()) return 0.0;
        std::sort(v.begin(), v.end());
        return v[(size_t)(p * (v.size() - 1))];
    }
    void print() const {
        std::printf("%s %d %.3e %.3e %.3e %.3e %.3e %.3e %.3e %.3e %.3e %.3e\n", name, n, e_err, f_err,
                    q(er_ref, 0.99), q(er_ref, 1.0), q(er_fast, 0.99), q(er_fast, 1.0),
                    q(fr_ref, 0.99), q(fr_ref, 1.0), q(fr_fast, 0.99), q(fr_fast, 1.0));
    }
};

static double tau_of(const D3* x) {
    V3<double> u1 = dif<double>(x[1], x[0]), u2 = dif<double>(x[3], x[2]), u3 = dif<double>(x[0], x[2]);
    V3<double> cr{u1.y * u2.z - u1.z * u2.y, u1.z * u2.x - u1.x * u2.z, u1.x * u2.y - u1.y * u2.x};
    return std::fabs(dot(cr, u3)) / std::sqrt(dot(u1, u1) * dot(u2, u2) * dot(u3, u3));
}

// both forms on one pair: energy (with the deferred flag) and forces of the energy-and-forces
// entry, plus the forces-only and energy-only entries, which must agree with it; the errors of
// both against the extended-precision reference form (energy at the pair, forces at the
// possibly shifted geometry, as the entry points evaluate them)
struct Res { double e_rel, f_rel, e_ref, e_fast, f_ref, f_fast; };
static bool compare(const D3* x, double ks, Res* r, bool* flag_mismatch) {
    int dr = 0, df = 0;
    bool okr = true, okf = true;
    D3 fr[4], ff[4], fr2[4], ff2[4];
    const double Ur = exvol_energy_forces<double>(x[0], x[1], x[2], x[3], ks, &dr, &okr, fr);
    const double Uf = exvol_energy_forces_fast(x[0], x[1], x[2], x[3], ks, &df, &okf, ff);
    const bool okr2 = exvol_forces<double>(x[0], x[1], x[2], x[3], ks, fr2);
    const bool okf2 = exvol_forces_fast(x[0], x[1], x[2], x[3], ks, ff2);
    int de = 0;
    const double Ue = exvol_energy_fast(x[0], x[1], x[2], x[3], ks, &de);
    *flag_mismatch = (dr != df) || (okr != okf) || (okr2 != okf2) || (de != df);
    if (dr || df || !okr || !okf) return false;
    if (Ue != Uf) *flag_mismatch = true;   // the energy-only entry is the same arithmetic
    // extended precision
    ExvTerms<long double> tl;
    exvol_terms<long double>(x[0], x[1], x[2], x[3], tl);
    const long double Ul = exvol_energy_of<long double>(tl, (long double)ks);
    D3 c2 = x[1];
    if (are_in_plane(x[0], x[1], x[2], x[3])) c2 = move_out_of_plane(x[0], x[1], x[2], 2, 0.01);
    exvol_terms<long double>(x[0], c2, x[2], x[3], tl);
    D3 fl[4];
    (void)exvol_forces_from<long double>(tl, (long double)ks, fl);
    r->e_rel = std::fabs(Uf - Ur) / std::fabs(Ur);
    r->e_ref = (double)(std::fabs((long double)Ur - Ul) / std::fabs(Ul));
    r->e_fast = (double)(std::fabs((long double)Uf - Ul) / std::fabs(Ul));
    double sc = 0.0, e = 0.0, el = 0.0, efl = 0.0;
    for (int k = 0; k < 4; k++) {
        const double R[3] = {fr[k].x, fr[k].y, fr[k].z}, F[3] = {ff[k].x, ff[k].y, ff[k].z};
        const double F2[3] = {ff2[k].x, ff2[k].y, ff2[k].z}, L[3] = {fl[k].x, fl[k].y, fl[k].z};
        for (int q = 0; q < 3; q++) {
            sc = std::fmax(sc, std::fabs(L[q]));
            e = std::fmax(e, std::fabs(F[q] - R[q]));
            el = std::fmax(el, std::fabs(R[q] - L[q]));
            efl = std::fmax(efl, std::fabs(F[q] - L[q]));
            if (F2[q] != F[q]) *flag_mismatch = true;
        }
    }
    r->f_rel = e / sc;
    r->f_ref = el / sc;
    r->f_fast = efl / sc;
    return true;
}

// closed-form energy of a pair in extended precision
static long double energy_ext(const D3* x, double ks) {
    ExvTerms<long double> t;
    exvol_terms<long double>(x[0], x[1], x[2], x[3], t);
    return exvol_energy_of<long double>(t, (long double)ks);
}

// forces of the extended-precision gradient vs -dE/dx by Richardson-extrapolated central
// differences (step 2^-10 nm and 2^-11 nm, exact in FP64 at these coordinates); relative to the
// pair's max |F|
static double gradient_fd_error(const D3* x, double ks) {
    ExvTerms<long double> t;
    exvol_terms<long double>(x[0], x[1], x[2], x[3], t);
    D3 f[4];
    (void)exvol_forces_from<long double>(t, (long double)ks, f);
    double sc = 0.0, err = 0.0;
    for (int k = 0; k < 4; k++) {
        const double fk[3] = {f[k].x, f[k].y, f[k].z};
        for (int q = 0; q < 3; q++) {
            long double dd[2];
            for (int r = 0; r < 2; r++) {
                const double h = r ? 0x1p-11 : 0x1p-10;
                D3 xp[4] = {x[0], x[1], x[2], x[3]}, xm[4] = {x[0], x[1], x[2], x[3]};
                double* cp = q == 0 ? &xp[k].x : q == 1 ? &xp[k].y : &xp[k].z;
                double* cm = q == 0 ? &xm[k].x : q == 1 ? &xm[k].y : &xm[k].z;
                *cp += h;
                *cm -= h;
                dd[r] = (energy_ext(xp, ks) - energy_ext(xm, ks)) / (2.0L * h);
            }
            const long double g = (4.0L * dd[1] - dd[0]) / 3.0L;
            sc = std::fmax(sc, std::fabs(fk[q]));
            err = std::fmax(err, (double)std::fabs((long double)fk[q] + g));
        }
    }
    return err / sc;
}

int main() {
    const double ks = 1e5 * 108.0 * 108.0;
    Cls well{"tau_ge_0.2"}, mid{"tau_0.01_0.2"}, low{"tau_1e-3_0.01"}, shifted{"coplanar_shifted"};
    int n_flag_mismatch = 0, n_cases = 0;
    D3 x[4];
    for (int i = 0; i < 200000 && (well.n < 4000 || mid.n < 4000 || low.n < 1000 || shifted.n < 2000); i++) {
        x[0] = D3{U(-20, 20), U(-20, 20), U(-20, 20)};
        const D3 a = rnd_dir();
        x[1] = add(x[0], a, 108.0);
        const D3 mid_pt = add(x[0], a, 54.0);
        const D3 m2 = add(mid_pt, rnd_dir(), U(15, 100));
        const D3 b = rnd_dir();
        x[2] = add(m2, b, -54.0);
        x[3] = add(m2, b, 54.0);
        const double tau = tau_of(x);
        const bool inplane = are_in_plane(x[0], x[1], x[2], x[3]);
        Res r;
        bool mm;
        if (!compare(x, ks, &r, &mm)) { n_flag_mismatch += mm; continue; }
        n_flag_mismatch += mm;
        Cls* c = inplane ? &shifted : tau >= 0.2 ? &well : tau >= 0.01 ? &mid : tau >= 1e-3 ? &low : nullptr;
        if (c) c->add(r.e_rel, r.f_rel, r.e_ref, r.e_fast, r.f_ref, r.f_fast);
    }
    // exactly degenerate geometries on integer coordinates along the axes (every invariant exact):
    // parallel in a plane (JJ = 0), c3 on the line of c1 c2 (AA = 0), crossing in a plane,
    // identical cylinders
    int n_deg = 0, n_deg_mm = 0;
    const D3 axes[6] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int i = 0; i < 600; i++) {
        const D3 a = axes[i % 6], perp = axes[(i % 6 + 2 + (i / 6) % 4) % 6];
        const D3 o{(double)(int)U(-20, 20), (double)(int)U(-20, 20), (double)(int)U(-20, 20)};
        x[0] = o;
        x[1] = add(o, a, 108.0);
        switch ((i / 24) % 4) {
        case 0: x[2] = add(o, perp, (double)(int)U(10, 90)); x[3] = add(x[2], a, 108.0); break;
        case 1: x[2] = add(o, a, (double)(int)U(120, 200)); x[3] = add(x[2], perp, 108.0); break;
        case 2: x[2] = add(add(o, a, 54.0), perp, -54.0); x[3] = add(x[2], perp, 108.0); break;
        default: x[2] = x[0]; x[3] = x[1]; break;
        }
        Res r;
        bool mm = false;
        (void)compare(x, ks, &r, &mm);
        n_deg++;
        n_deg_mm += mm;
    }
    // the gradient formula against the energy's finite differences (tau >= 0.01, outside the
    // coplanar test)
    int n_fd = 0;
    double fd_err = 0.0;
    while (n_fd < 2000) {
        x[0] = D3{(double)(int)U(-20, 20), (double)(int)U(-20, 20), (double)(int)U(-20, 20)};
        const D3 a = rnd_dir();
        x[1] = add(x[0], a, 108.0);
        const D3 m2 = add(add(x[0], a, 54.0), rnd_dir(), U(15, 100));
        const D3 b = rnd_dir();
        x[2] = add(m2, b, -54.0);
        x[3] = add(m2, b, 54.0);
        if (tau_of(x) < 0.01 || are_in_plane(x[0], x[1], x[2], x[3])) continue;
        fd_err = std::fmax(fd_err, gradient_fd_error(x, ks));
        n_fd++;
    }
    n_cases = well.n + mid.n + low.n + shifted.n;
    for (const Cls* c : {&well, &mid, &low, &shifted}) c->print();
    std::printf("flags %d %d\n", n_cases, n_flag_mismatch);
    std::printf("degenerate %d %d\n", n_deg, n_deg_mm);
    std::printf("gradient_fd %d %.3e\n", n_fd, fd_err);
    return 0;
}
