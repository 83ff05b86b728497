"""The oracle against the reference's own unit-test fixtures.

Fixture data below is transcribed from the reference's Catch2 tests (inputs and pass criteria):
  src/TESTS/Mechanics/ForceField/Volume/TestCylinderExclVolume.cpp:13-243
  src/TESTS/Mechanics/ForceField/Boundary/TestBoundaryBeadRepulsion.cpp:13-96
  src/TESTS/Mechanics/Minimizer/TestCGMethod.cpp:9-93
  src/TESTS/Mechanics/ForceField/TestFFCommon.hpp:81-134 (equalRelEps, energy-force consistency)
"""
import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi


def equal_rel_eps(a, b, rel):
    # TestFFCommon.hpp:81-83
    return abs(a - b) <= min(abs(a), abs(b)) * rel


def energy_force_consistency(c0, energy, force, move_mag, rel_eps, rng):
    # TestFFCommon.hpp:142-170: central difference of E against -2 dc.F
    dc = rng.normal(0.0, move_mag, size=c0.shape)
    e1 = energy(c0 - dc)
    e2 = energy(c0 + dc)
    f = force(c0)
    de_actual = e2 - e1
    de_expected = -2 * float(np.dot(dc, f))
    return equal_rel_eps(de_actual, de_expected, rel_eps), de_actual, de_expected


# ---------------------------------------------------------------- excluded volume ------

SCALING_COORD = np.array([
    -60.0, 0.0, 0.0, 0.0, 0.0, 0.0, 60.0, 0.0, 0.0,
    -40.0, -40.0, 10.0, -28.0, -24.0, 10.0, -10.0, 0.0, 10.0, 20.0, 40.0, 10.0])


def test_exvol_scaling_additivity():
    """TestCylinderExclVolume.cpp:13-47: one 2x3-segment pair == sum of the 6 sub-pairs with
    rescaled rest lengths (Catch Approx: relative 100 * float epsilon)."""
    e0 = of.exvol_energy(SCALING_COORD, [0, 6, 9, 18], [1e-2], [100.0, 100.0])
    bs1 = [0, 3, 9, 12, 0, 3, 12, 15, 0, 3, 15, 18, 3, 6, 9, 12, 3, 6, 12, 15, 3, 6, 15, 18]
    eq1 = [50.0, 20.0, 50.0, 30.0, 50.0, 50.0, 50.0, 20.0, 50.0, 30.0, 50.0, 50.0]
    e1 = of.exvol_energy(SCALING_COORD, bs1, [1e-2] * 6, eq1)
    assert e0 > 0
    assert abs(e0 - e1) <= 100 * np.finfo(np.float32).eps * max(abs(e0), abs(e1))


def monomer_energy(coord, c11, c12, m_lo1, m_hi1, int1, eq1, c21, c22, m_lo2, m_hi2, int2, eq2, kvol):
    """CylinderVolumeMon.hpp:31-58: pairwise 1/d^4 between monomer centres."""
    def ceildiv(a, b):   # MathFunctions.h:1479-1481 with C++ truncating division
        q = int(a / b)
        r = a - b * q
        return q + ((0 if ((a > 0) ^ (b > 0)) else 1) if r != 0 else 0)
    x11, x12 = coord[c11:c11 + 3], coord[c12:c12 + 3]
    x21, x22 = coord[c21:c21 + 3], coord[c22:c22 + 3]
    n1, n2 = m_hi1 - m_lo1, m_hi2 - m_lo2
    s1 = ceildiv(m_lo1, int1) * int1
    s2 = ceildiv(m_lo2, int2) * int2
    coeff = kvol * (int1 * eq1) * (int2 * eq2)
    en = 0.0
    for m1 in range(s1, m_hi1, int1):
        l1 = (m1 - m_lo1 + 0.5) / n1
        p1 = x11 + l1 * (x12 - x11)
        for m2 in range(s2, m_hi2, int2):
            l2 = (m2 - m_lo2 + 0.5) / n2
            p2 = x21 + l2 * (x22 - x21)
            d2 = float(np.sum((p2 - p1) ** 2))
            en += coeff / (d2 * d2)
    return en


def test_exvol_vs_monomer_sum():
    """TestCylinderExclVolume.cpp:49-100: closed form vs 40-monomer sum within 1e-2."""
    coord = np.array([10, 0, 0, -10, 0, 0, 0, 10, 5, 0, -10, 5], float)
    eq, kvol = 20.0, 0.25
    e_int = of.exvol_energy(coord, [0, 3, 6, 9], [kvol], [eq, eq])
    e_mon = monomer_energy(coord, 0, 3, 0, 40, 1, eq / 40, 6, 9, -9, -9 + 40, 4, eq / 40, kvol)
    assert equal_rel_eps(e_int, e_mon, 1e-2), (e_int, e_mon)


EXVOL_INPUTS = [
    ("standard-symmetric", [0, 3, 6, 9], [0.25], [20.0, 20.0],
     [10, 0, 0, -10, 0, 0, 0, 10, 5, 0, -10, 5]),
    ("general", [0, 3, 6, 9, 0, 3, 12, 15], [100.0, 2e3], [20.0, 20.0, 20.0, 100.0],
     [10, 0, 0, -10, 0, 0, 0, 10, 5, 3, -10, 5, 3, 70, 30, 2, -20, 0]),
]


@pytest.mark.parametrize("name,bs,k,eq,coord", EXVOL_INPUTS, ids=[x[0] for x in EXVOL_INPUTS])
def test_exvol_energy_force_consistency(name, bs, k, eq, coord):
    """TestCylinderExclVolume.cpp:103-243: moveMag 1e-5, relEps 5e-4, >= 9 of 10 pass."""
    c0 = np.array(coord, float)
    rng = np.random.default_rng(0)
    passed = 0
    for _ in range(10):
        ok, _, _ = energy_force_consistency(
            c0, lambda c: of.exvol_energy(c, bs, k, eq), lambda c: of.exvol_forces(c, bs, k, eq), 1e-5, 5e-4, rng)
        passed += ok
    assert passed >= 9


# ---------------------------------------------------------------- boundary --------------

def test_boundary_bead_repulsion_consistency():
    """TestBoundaryBeadRepulsion.cpp:13-96: plane through (1,2,3) with normal z, beads on,
    above and below the plane."""
    elem = of.plane_element([1.0, 2.0, 3.0], [0.0, 0.0, 1.0], 10, 2.7)
    assert elem.p[3] == -3.0
    coords = np.array([5.0, -100.0, 3.0, 0.0, 0.0, 13.0, -1.0, -1e8, -7.0])
    kreps, slens, bs, nn = [1e5, 1e4, 1e3], [1.0, 1.5, 2.0], [0, 3, 6], [3]
    rng = np.random.default_rng(1)
    passed = 0
    for _ in range(10):
        ok, _, _ = energy_force_consistency(
            coords, lambda c: of.boundary_energy(c, bs, kreps, slens, nn, [elem]),
            lambda c: of.boundary_forces(c, bs, kreps, slens, nn, [elem]), 1e-5, 5e-4, rng)
        passed += ok
    assert passed >= 9
    # closed form: U = k exp(-d / lambda), d = z - 3
    e = of.boundary_energy(coords, bs, kreps, slens, nn, [elem])
    assert e == pytest.approx(1e5 * np.exp(0.0) + 1e4 * np.exp(-10 / 1.5) + 1e3 * np.exp(10 / 2.0), rel=1e-15)


# ---------------------------------------------------------------- CG minimizer ---------

OPT = np.array([-500.0, -400.0, -300.0, -200.0, -100.0, 0.0, 100.0, 200.0, 300.0, 400.0])
W = np.array([3, 1, 0.1, 0.5, 1, 10, 20, 50, 100, 0.1])


@pytest.mark.parametrize("ls", [abi.MDN_LS_BACKTRACKING, abi.MDN_LS_QUADRATIC], ids=["backtracking", "quadratic"])
@pytest.mark.parametrize("method", [abi.MDN_CG_STEEPEST, abi.MDN_CG_FLETCHER_REEVES, abi.MDN_CG_POLAK_RIBIERE],
                         ids=["steepest", "fletcher_rieves", "polak_ribiere"])
@pytest.mark.parametrize("init", ["input_0", "input_1_already_minimized", "input_2"])
def test_cg_harmonic10(method, init, ls):
    """TestCGMethod.cpp:9-93: 10-D separable harmonic, gradTol 0.1, maxDist 10, lambdaMax 1,
    recovery off; success() and every |F| <= gradTol. input_0 is the reference's
    vector<floatingpoint>(800, n): 800 entries equal to 10, of which the first 10 are DOFs. The
    reference's fixture runs the default (backtracking) line search; the quadratic one must meet
    the same criteria."""
    n = 10
    if init == "input_0":
        x0, nvar = np.full(800, 10.0), 800
    elif init == "input_1_already_minimized":
        x0, nvar = OPT.copy(), n
    else:
        x0, nvar = np.array([1000, 900, 800, 700, 600, 500, 400, 300, 200, 100], float), n

    def energy(c):
        return float(np.sum(0.5 * W * (c[:n] - OPT) * (c[:n] - OPT)))

    def force(c, f):
        f[:n] = -W * (c[:n] - OPT)

    p = abi.default_cg_params(method=method, grad_tol=0.1, max_dist=10, lambda_max=1, try_recover=0, line_search=ls)
    st, x, f, res = of.cg_minimize(p, x0, n, energy, force, nvar=nvar)
    assert st == 0
    assert res.success()
    assert np.count_nonzero(np.abs(f) > 0.1) == 0
    np.testing.assert_allclose(x[:n], OPT, atol=0.1 / W.min() + 1e-9)


@pytest.mark.parametrize("ratio,step,trials", [(5.0, 1.0, 5), (3.0, 1.0, 4), (8.0, 1.0, 4), (2.2, 1.1, 2)])
def test_quadratic_line_search_lands_on_the_line_minimizer(ratio, step, trials):
    """On a quadratic energy the parabola through (0, E0), the accepted backtracking trial and the
    first trial is the energy along the line itself, so the quadratic search (CGMethod.hpp:522-764,
    trials at x + lambda d) must step exactly to the line minimizer lambda* = F.F / F.H.F whenever
    its parabola phase runs. With lambda0 = ratio * lambda* the Armijo rule (slope 0.4) rejects
    every trial above 1.2 lambda*:
      ratio 5: 5, 2.5, 1.25 rejected, 0.625 accepted; the vertex is 60 % away, so one evaluation
               there (within 1e-3 of the accepted energy) returns it: 5 trial evaluations;
      ratio 3: 3, 1.5 rejected, 0.75 accepted, one evaluation at the vertex: 4;
      ratio 8: 8, 4, 2 rejected, lambda* accepted; the vertex is within 1e-3 of it and is
               returned without an evaluation: 4;
      ratio 2.2: one rejection, 1.1 accepted; the parabola phase needs an entry after the first
               with more energy than the accepted trial, which a single rejection never leaves,
               so the backtracking step stays: 2."""
    rng = np.random.default_rng(31)
    n = 6
    h = rng.uniform(1.0, 10.0, n)
    c = rng.normal(0.0, 5.0, n)
    x0 = c + rng.normal(0.0, 20.0, n)

    def energy(x):
        return float(1e6 + 0.5 * np.sum(h * (x - c) ** 2))

    f0 = -h * (x0 - c)
    lam_star = (f0 @ f0) / (f0 @ (h * f0))
    calls, forces_at = [], []

    def energy_rec(x):
        calls.append(x.copy())
        return energy(x)

    def force(x, f):
        forces_at.append((x.copy(), len(calls)))
        f[:] = -h * (x - c)

    lam0 = ratio * lam_star
    p = abi.default_cg_params(method=abi.MDN_CG_STEEPEST, grad_tol=1e-12, max_dist=1e9, lambda_max=lam0,
                              try_recover=0, line_search=abi.MDN_LS_QUADRATIC, energy_call_limit=40)
    st, x, f, res = of.cg_minimize(p, x0, n, energy_rec, force)
    assert st == 0
    x1, ncalls = forces_at[1]             # the step's forces: at x0 + lambda d
    assert ncalls - 2 == trials           # less the start report and the search's current energy
    np.testing.assert_allclose(x1, x0 + step * lam_star * f0, rtol=1e-12)
    # a quadratic energy with the search's exact steps converges (steepest descent)
    assert res.num_iter >= 2


def test_quadratic_line_search_refits_near_the_line_minimizer():
    """On a quadratic energy with a quartic perturbation the parabola through three points is a
    model, so the search re-fits it around its best point (CGMethod.hpp:660-764) until the vertex
    or the energy moves by less than 1e-3 relative. The far first trial stays a bracket point, so
    the re-fits creep from the backtracking step toward the true line minimizer (found here by
    bisection on the directional derivative) and stop short of it: the step must lie between the
    two, at lower energy than the backtracking step (a wrong bracket update moves away or past)."""
    rng = np.random.default_rng(32)
    n = 5
    h = rng.uniform(1.0, 3.0, n)
    c = rng.normal(0.0, 2.0, n)
    x0 = c + rng.normal(0.0, 3.0, n)
    eps = 0.005

    def grad(x):
        return h * (x - c) + 4 * eps * h * (x - c) ** 3

    def energy(x):
        return float(np.sum(0.5 * h * (x - c) ** 2 + eps * h * (x - c) ** 4))

    def force(x, f):
        forces.append(x.copy())
        f[:] = -grad(x)

    f0 = -grad(x0)
    lo, hi = 0.0, 1.0
    while -grad(x0 + hi * f0) @ f0 > 0:
        hi *= 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        lo, hi = (mid, hi) if -grad(x0 + mid * f0) @ f0 > 0 else (lo, mid)
    lam_true = 0.5 * (lo + hi)
    for ratio, near in ((3.0, 0.15), (5.0, 0.2)):
        results = {}
        for ls in (abi.MDN_LS_BACKTRACKING, abi.MDN_LS_QUADRATIC):
            forces = []
            p = abi.default_cg_params(method=abi.MDN_CG_STEEPEST, grad_tol=1e-12, max_dist=1e9,
                                      lambda_max=ratio * lam_true, try_recover=0, line_search=ls,
                                      energy_call_limit=60)
            st, _, _, _ = of.cg_minimize(p, x0, n, energy, force)
            assert st == 0
            results[ls] = ((forces[1] - x0) @ f0) / (f0 @ f0)
        lam_q, lam_bt = results[abi.MDN_LS_QUADRATIC], results[abi.MDN_LS_BACKTRACKING]
        assert lam_bt < lam_q <= lam_true * (1 + 1e-3), (ratio, lam_bt / lam_true, lam_q / lam_true)
        assert lam_true - lam_q <= near * lam_true
        assert energy(x0 + lam_q * f0) < energy(x0 + lam_bt * f0)


def test_database_restatement_reproduces_the_reference_fixture():
    """TestDatabase.cpp:100-154 ("vectorized data"): the element indices that serializeDof orders
    beads by (DatabaseBase: append, swap-with-last removal) and the stable indices with their
    rearrange (Database<T, true>), at every point the reference's test checks."""
    from oracle.database import ElementList, StableIndex
    el, si = ElementList(), StableIndex()

    def make(v):
        el.create(v)
        si.create(v)

    def drop(v):
        el.destroy(v)
        si.destroy(v)

    make(0)
    assert el.index(0) == 0 and si.where[0] == 0
    for v in (1, 2, 3, 4):
        make(v)
    drop(3)
    make(5)
    assert len(si.stable) == 5 and el.index(5) == 4 and si.where[5] == 3 and si.stable[3] == 5
    make(6)
    drop(5)
    drop(1)
    drop(2)
    assert len(el.elems) == 3 and len(si.stable) == 6
    assert el.index(4) == 2 and si.where[4] == 4 and si.stable[4] == 4
    assert el.index(6) == 1 and si.where[6] == 5 and si.stable[5] == 6
    si.rearrange()
    assert len(el.elems) == 3 and len(si.stable) == 3
    assert el.index(4) == 2 and si.stable[1] == 4 and si.where[4] == 1
    assert el.index(6) == 1 and si.stable[2] == 6 and si.where[6] == 2


def test_database_restatement_invariants():
    """Random create / destroy sequences: the element order holds every live element once; after
    rearrange the stable indices are 0 .. n-1, every live element keeps its stable index when it
    was below the final size, and serialize_map names each surviving element's previous index."""
    from oracle.database import ElementList, StableIndex, serialize_map
    rng = np.random.default_rng(5)
    el, si = ElementList(), StableIndex()
    nxt = 0
    for _ in range(40):
        el.create(nxt)
        si.create(nxt)
        nxt += 1
    for _ in range(20):
        before = list(el.elems)
        live = set(el.elems)
        for _ in range(int(rng.integers(1, 15))):
            if rng.random() < 0.5 and len(el.elems) > 1:
                v = el.elems[int(rng.integers(len(el.elems)))]
                el.destroy(v)
                si.destroy(v)
                live.discard(v)
            else:
                el.create(nxt)
                si.create(nxt)
                live.add(nxt)
                nxt += 1
        assert sorted(el.elems) == sorted(live)
        keep = {v: i for v, i in si.where.items()}
        final = len(si.stable) - len(si.deleted)
        si.rearrange()
        assert len(si.stable) == len(live) and sorted(si.stable) == sorted(live)
        for v, i in keep.items():
            if i < final:
                assert si.where[v] == i
        m = serialize_map(before, el.elems)
        for i, v in enumerate(el.elems):
            assert m[i] == (before.index(v) if v in before else -1)
