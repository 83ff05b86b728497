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


@pytest.mark.parametrize("method", [abi.MDN_CG_STEEPEST, abi.MDN_CG_FLETCHER_REEVES, abi.MDN_CG_POLAK_RIBIERE],
                         ids=["steepest", "fletcher_rieves", "polak_ribiere"])
@pytest.mark.parametrize("init", ["input_0", "input_1_already_minimized", "input_2"])
def test_cg_harmonic10(method, init):
    """TestCGMethod.cpp:9-93: 10-D separable harmonic, gradTol 0.1, maxDist 10, lambdaMax 1,
    recovery off; success() and every |F| <= gradTol. input_0 is the reference's
    vector<floatingpoint>(800, n): 800 entries equal to 10, of which the first 10 are DOFs."""
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

    p = abi.default_cg_params(method=method, grad_tol=0.1, max_dist=10, lambda_max=1, try_recover=0)
    st, x, f, res = of.cg_minimize(p, x0, n, energy, force, nvar=nvar)
    assert st == 0
    assert res.success()
    assert np.count_nonzero(np.abs(f) > 0.1) == 0
    np.testing.assert_allclose(x[:n], OPT, atol=0.1 / W.min() + 1e-9)
