"""Boundary load forces on the filament ends (BoundaryCylinderRepulsion::computeLoadForces,
BoundaryCylinderRepulsion.cpp:122-187): the oracle's restatement pinned on hand geometries whose
values follow from the physics alone. The plus-end bead of a filament advanced by i monomers
along the filament direction sits at p_i = bd + i m dir; the load on it is the component of the
boundary's exponential repulsion against the growth direction, max(-n.dir, 0) k/lambda
exp(-d(p_i)/lambda), summed over the boundary elements whose neighbor list holds the end
cylinder. A filament pointing at a plane along its normal feels k/lambda exp(-(d0 - i m)/lambda);
at an angle theta, cos(theta) k/lambda exp(-(d0 - i m cos(theta))/lambda); pointing away, zero;
beyond the neighbor-list cutoff, zero."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi
from tests.hand_systems import HandSystem

K, LAM, MON, NMON = 41.0, 2.7, 2.7, 40


def plane(point, normal):
    return of.plane_element(np.array(point, float), np.array(normal, float), K, LAM)


def sphere(center, radius):
    e = abi.BoundaryElement()
    e.kind = 1
    e.p[0], e.p[1], e.p[2], e.p[3] = center[0], center[1], center[2], radius
    e.krep, e.slen = K, LAM
    return e


def chain(end, direction, n_beads, spacing=108.0):
    """Beads of a straight filament ending (plus end) at `end`, growing along `direction`."""
    u = np.asarray(direction, float) / np.linalg.norm(direction)
    return [np.asarray(end, float) - (n_beads - 1 - j) * spacing * u for j in range(n_beads)]


def hand_systems():
    i = np.arange(NMON)
    th = 0.6
    top = plane((0, 0, 1000.0), (0, 0, -1.0))   # plane z = 1000, normal into the box
    fils = [
        chain((500, 500, 916.0), (0, 0, 1), 3),                                  # at the plane, d0 = 84
        chain((1500, 500, 1000 - 60.0), (np.sin(th), 0, np.cos(th)), 2),         # one cylinder, oblique
        [np.array([500, 1500, 916.0 - 108 * j]) for j in range(3)],              # grows away; minus end at 84
        chain((1500, 1500, 300.0), (0, 0, 1), 3),                                # beyond the cutoff
    ]
    expect_p = np.zeros((4, NMON))
    expect_m = np.zeros((4, NMON))
    expect_p[0] = K / LAM * np.exp(-(84.0 - i * MON) / LAM)
    expect_p[1] = np.cos(th) * K / LAM * np.exp(-(60.0 - i * MON * np.cos(th)) / LAM)
    expect_m[2] = K / LAM * np.exp(-(84.0 - i * MON) / LAM)
    cube = HandSystem(fils, [top])
    ball = HandSystem([chain((1500 + 950.0, 1500, 1500), (1, 0, 0), 2)], [sphere((1500, 1500, 1500), 1000.0)])
    expect_ball = K / LAM * np.exp(-(50.0 - i * MON) / LAM)
    return cube, expect_p, expect_m, ball, expect_ball


def oracle_loads(hs, coord=None):
    o = of.OracleSystem(hs.desc)
    o.build_nl(hs.coord)
    return o.load_forces(hs.coord if coord is None else coord, [MON], [NMON], NMON)


def test_oracle_load_forces_closed_forms():
    cube, ep, em, ball, eb = hand_systems()
    lp, lm = oracle_loads(cube)
    np.testing.assert_allclose(lp, ep, rtol=1e-13, atol=0)
    np.testing.assert_allclose(lm, em, rtol=1e-13, atol=0)
    # the one-cylinder filament is both ends: its minus end points away from the plane
    assert (lm[1] == 0).all()
    bp, bm = oracle_loads(ball)
    np.testing.assert_allclose(bp[0], eb, rtol=1e-13, atol=0)
    assert (bm == 0).all()


def test_oracle_load_forces_invariances():
    """Two planes add in element order; a rigid rotation of coordinates and planes changes
    nothing beyond rounding; the neighbor list is the one of the last build."""
    th = 0.6
    fil = chain((500, 500, 1000 - 60.0), (np.sin(th), 0, np.cos(th)), 2)
    one = HandSystem([fil], [plane((0, 0, 1000.0), (0, 0, -1.0))])
    two = HandSystem([fil], [plane((0, 0, 1000.0), (0, 0, -1.0)), plane((620.0, 0, 0), (-1.0, 0, 0))])
    p1, _ = oracle_loads(one)
    p2, _ = oracle_loads(two)
    i = np.arange(NMON)
    dx = 620.0 - fil[-1][0]
    side = np.sin(th) * K / LAM * np.exp(-(dx - i * MON * np.sin(th)) / LAM)
    np.testing.assert_allclose(p2[0], p1[0] + side, rtol=1e-13)
    # rotation about z by 30 degrees (the z plane is invariant; the x plane rotates with it)
    a = np.pi / 6
    R = np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1]])
    shift = np.array([1000.0, 1000.0, 0.0])
    rfil = [R @ p + shift for p in fil]
    n2 = R @ np.array([-1.0, 0, 0])
    rot = HandSystem([rfil], [plane((0, 0, 1000.0), (0, 0, -1.0)), plane(R @ np.array([620.0, 0, 0]) + shift, n2)])
    pr, _ = oracle_loads(rot)
    np.testing.assert_allclose(pr[0], p2[0], rtol=1e-11)
    # the neighbor list is frozen at the build: moving the filament out of the cutoff afterwards
    # still counts the element (and the distance term decays accordingly)
    o = of.OracleSystem(one.desc)
    o.build_nl(one.coord)
    far = one.coord.copy().reshape(-1, 3)
    far[:, 2] -= 300.0   # the end cylinder's midpoint is now 405 nm from the plane
    lp, _ = o.load_forces(far.ravel(), [MON], [NMON], NMON)
    expect = np.cos(th) * K / LAM * np.exp(-(360.0 - i * MON * np.cos(th)) / LAM)
    np.testing.assert_allclose(lp[0], expect, rtol=1e-12)


def test_oracle_load_forces_rejects_unknown_type():
    fil = chain((500, 500, 916.0), (0, 0, 1), 2)
    hs = HandSystem([fil], [plane((0, 0, 1000.0), (0, 0, -1.0))], types=[1])
    o = of.OracleSystem(hs.desc)
    o.build_nl(hs.coord)
    with pytest.raises(ValueError):
        o.load_forces(hs.coord, [MON], [NMON], NMON)
    lp, _ = o.load_forces(hs.coord, [MON, MON], [NMON, 10], NMON)
    assert (lp[0, 10:] == 0).all() and lp[0, 9] > 0
