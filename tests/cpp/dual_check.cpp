// Host-side check of the forward-mode tangents of the term forces (picgo_amd/csrc/mdn_terms.h):
// the Dual<double> instantiation on dual positions (x, v) against central differences of the
// FP64 value instantiation, (F(x + h v) - F(x - h v)) / 2h, on seeded random geometries of every
// force field (well-conditioned excluded-volume pairs: normalized triple product >= 0.2, plus
// pairs inside the coplanar test away from its threshold, whose shifted geometry the tangent
// differentiates too; their small triple products make the central difference itself noisy at
// the 1e-5 level). Prints one line per force field: "<name> <cases> <max relative error>".
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <random>

#include "../../picgo_amd/csrc/mdn_terms.h"

using namespace mdn;

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static D3 rnd_dir() {
    std::normal_distribution<double> N(0.0, 1.0);
    double x = N(rng), y = N(rng), z = N(rng), n = std::sqrt(x * x + y * y + z * z);
    return D3{x / n, y / n, z / n};
}
static D3 add(D3 a, D3 b, double s) { return D3{a.x + s * b.x, a.y + s * b.y, a.z + s * b.z}; }

struct Err {
    double worst = 0.0;
    int n = 0;
    // tangent t[k] vs finite difference fd[k] over nb beads, relative to the tangent's scale
    void add_case(const D3* t, const D3* fd, int nb) {
        double sc = 0.0, e = 0.0;
        for (int k = 0; k < nb; k++) {
            sc = std::fmax(sc, std::fmax(std::fabs(t[k].x), std::fmax(std::fabs(t[k].y), std::fabs(t[k].z))));
            e = std::fmax(e, std::fmax(std::fabs(t[k].x - fd[k].x),
                                       std::fmax(std::fabs(t[k].y - fd[k].y), std::fabs(t[k].z - fd[k].z))));
        }
        if (sc > 0) worst = std::fmax(worst, e / sc);
        n++;
    }
};

// central difference of a force function f(positions, out) along directions v
template <typename F>
static void central(F f, const D3* x, const D3* v, int nb, double h, D3* fd) {
    D3 xp[4], xm[4], op[4], om[4];
    for (int k = 0; k < nb; k++) { xp[k] = add(x[k], v[k], h); xm[k] = add(x[k], v[k], -h); }
    f(xp, op);
    f(xm, om);
    for (int k = 0; k < nb; k++)
        fd[k] = D3{(op[k].x - om[k].x) / (2 * h), (op[k].y - om[k].y) / (2 * h), (op[k].z - om[k].z) / (2 * h)};
}
static void duals(const D3* x, const D3* v, int nb, DualD3* p) {
    for (int k = 0; k < nb; k++) p[k] = DualD3{x[k], v[k]};
}
static void tangents(const DualD3* o, int nb, D3* t) {
    for (int k = 0; k < nb; k++) t[k] = o[k].d;
}

int main() {
    const double h = 1e-3;   // truncation error <= 1e-8 here; smaller h only adds rounding noise
    const int N = 2000;
    Err es, eb, ex, ev, evp, ep;   // evp: pairs inside the coplanar test
    Err fs, fb, fb0, fx;           // FP32 tangents (the mixed path's) against the FP64 ones
                                   // (fb0: bending with theta0 != 0)
    DualD3 o32[4];
    D3 t32[4];
    D3 x[4], v[4], fd[4], t[4];
    DualD3 p[4], o[4];
    for (int i = 0; i < N; i++) {   // stretching
        x[0] = D3{U(-50, 50), U(-50, 50), U(-50, 50)};
        x[1] = add(x[0], rnd_dir(), U(80, 140));
        for (int k = 0; k < 2; k++) v[k] = rnd_dir();
        const double kk = U(50, 150), l0 = U(90, 130);
        central([&](const D3* y, D3* out) { stretch_forces<double>(y[0], y[1], kk, l0, out); }, x, v, 2, h, fd);
        duals(x, v, 2, p);
        stretch_forces<DualD>(p[0], p[1], kk, l0, o);
        tangents(o, 2, t);
        es.add_case(t, fd, 2);
        stretch_forces<DualF>(p[0], p[1], kk, l0, o32);
        tangents(o32, 2, t32);
        fs.add_case(t, t32, 2);
    }
    for (int i = 0; i < N; i++) {   // bending, theta0 = 0 and theta0 != 0
        x[0] = D3{U(-50, 50), U(-50, 50), U(-50, 50)};
        D3 a = rnd_dir(), b = rnd_dir();
        x[1] = add(x[0], a, U(90, 120));
        x[2] = add(x[1], b, U(90, 120));
        for (int k = 0; k < 3; k++) v[k] = rnd_dir();
        const double kb = U(300, 900), eqt = (i & 1) ? U(0.1, 0.6) : 0.0;
        central([&](const D3* y, D3* out) { bend_forces<double>(y[0], y[1], y[2], kb, eqt, out); }, x, v, 3, h, fd);
        duals(x, v, 3, p);
        bend_forces<DualD>(p[0], p[1], p[2], kb, eqt, o);
        tangents(o, 3, t);
        eb.add_case(t, fd, 3);
        bend_forces<DualF>(p[0], p[1], p[2], kb, eqt, o32);
        tangents(o32, 3, t32);
        (eqt == 0.0 ? fb : fb0).add_case(t, t32, 3);
    }
    for (int i = 0; i < N; i++) {   // linkers / motors
        for (int k = 0; k < 4; k += 2) {
            x[k] = D3{U(-60, 60), U(-60, 60), U(-60, 60)};
            x[k + 1] = add(x[k], rnd_dir(), 108.0);
        }
        for (int k = 0; k < 4; k++) v[k] = rnd_dir();
        const double kk = U(2, 10), l0 = U(20, 200), p1 = U(0, 1), p2 = U(0, 1);
        central([&](const D3* y, D3* out) { (void)xlink_forces<double>(y[0], y[1], y[2], y[3], kk, l0, p1, p2, out); },
                x, v, 4, h, fd);
        duals(x, v, 4, p);
        (void)xlink_forces<DualD>(p[0], p[1], p[2], p[3], kk, l0, p1, p2, o);
        tangents(o, 4, t);
        ex.add_case(t, fd, 4);
        (void)xlink_forces<DualF>(p[0], p[1], p[2], p[3], kk, l0, p1, p2, o32);
        tangents(o32, 4, t32);
        fx.add_case(t, t32, 4);
    }
    int n_inplane = 0;
    for (int i = 0; i < N;) {   // excluded volume
        x[0] = D3{U(-20, 20), U(-20, 20), U(-20, 20)};
        const D3 a = rnd_dir();
        x[1] = add(x[0], a, 108.0);
        const D3 mid = add(x[0], a, 54.0);
        const D3 off = rnd_dir();
        const D3 m2 = add(mid, off, U(15, 100));
        D3 b = rnd_dir();
        x[2] = add(m2, b, -54.0);
        x[3] = add(m2, b, 54.0);
        // conditioning of the closed form: normalized triple product, and the coplanar test value
        V3<double> u1 = dif<double>(x[1], x[0]), u2 = dif<double>(x[3], x[2]), u3 = dif<double>(x[0], x[2]);
        V3<double> cr{u1.y * u2.z - u1.z * u2.y, u1.z * u2.x - u1.x * u2.z, u1.x * u2.y - u1.y * u2.x};
        const double tau = std::fabs(dot(cr, u3)) / std::sqrt(dot(u1, u1) * dot(u2, u2) * dot(u3, u3));
        const bool inplane = are_in_plane(x[0], x[1], x[2], x[3]);
        double dd[3] = {x[1].x - x[0].x, x[1].y - x[0].y, x[1].z - x[0].z};
        double ee[3] = {x[2].x - x[0].x, x[2].y - x[0].y, x[2].z - x[0].z};
        double ff[3] = {x[3].x - x[0].x, x[3].y - x[0].y, x[3].z - x[0].z};
        double nn[3];
        cross3(dd, ee, nn);
        normalize3(nn);
        normalize3(ff);
        const double ip = std::fabs(nn[0] * ff[0] + nn[1] * ff[1] + nn[2] * ff[2]);
        if (!inplane && tau < 0.2) continue;                                       // ill conditioned
        if (inplane && (ip < 0.02 || ip > 0.08)) continue;                          // near the threshold
        if (inplane) n_inplane++;
        i++;
        for (int k = 0; k < 4; k++) v[k] = rnd_dir();
        const double ks = 1e5 * 108.0 * 108.0;
        bool ok = true;
        central([&](const D3* y, D3* out) { ok = exvol_forces<double>(y[0], y[1], y[2], y[3], ks, out) && ok; },
                x, v, 4, h, fd);
        duals(x, v, 4, p);
        ok = exvol_forces<DualD>(p[0], p[1], p[2], p[3], ks, o) && ok;
        if (!ok) continue;
        tangents(o, 4, t);
        (inplane ? evp : ev).add_case(t, fd, 4);
    }
    for (int i = 0; i < N; i++) {   // boundary: plane and sphere
        BElem e;
        if (i & 1) {
            e.kind = 0;
            const D3 n = rnd_dir();
            e.p[0] = n.x; e.p[1] = n.y; e.p[2] = n.z; e.p[3] = U(-10, 10);
            x[0] = add(D3{0, 0, 0}, n, U(2, 40) - e.p[3]);
        } else {
            e.kind = 1;
            e.p[0] = U(-5, 5); e.p[1] = U(-5, 5); e.p[2] = U(-5, 5); e.p[3] = 1000.0;
            x[0] = add(D3{e.p[0], e.p[1], e.p[2]}, rnd_dir(), 1000.0 - U(2, 40));
        }
        v[0] = rnd_dir();
        central([&](const D3* y, D3* out) { out[0] = boundary_force<double>(e, y[0], 41.0, 2.7); }, x, v, 1, h, fd);
        const DualD3 q = boundary_force_dual(e, DualD3{x[0], v[0]}, 41.0, 2.7);
        t[0] = q.d;
        ep.add_case(t, fd, 1);
    }
    std::printf("stretch %d %.3e\n", es.n, es.worst);
    std::printf("bend %d %.3e\n", eb.n, eb.worst);
    std::printf("xlink %d %.3e\n", ex.n, ex.worst);
    std::printf("exvol %d %.3e\n", ev.n, ev.worst);
    std::printf("exvol_inplane %d %.3e\n", evp.n, evp.worst);
    std::printf("boundary %d %.3e\n", ep.n, ep.worst);
    std::printf("stretch_fp32 %d %.3e\n", fs.n, fs.worst);
    std::printf("bend_fp32 %d %.3e\n", fb.n, fb.worst);
    std::printf("bend_theta0_fp32 %d %.3e\n", fb0.n, fb0.worst);
    std::printf("xlink_fp32 %d %.3e\n", fx.n, fx.worst);
    return 0;
}
