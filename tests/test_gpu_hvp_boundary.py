"""GPU: the tangent of the boundary repulsion (BoundaryCylinderRepulsionExp.cpp:27-110, plane and
sphere elements) inside mdn_hvp, on hand systems where it is the only term.

Filaments sit 3-12 nm inside a sphere (radius 1 um) or below a plane, some beads oblique to the
wall, with the filament stretching, bending and excluded-volume stiffnesses set to zero, so H v is
the boundary's alone: exponential repulsion k exp(-d / lambda) with lambda = 2.7 nm. Central
differences of the device forces with h = 1e-3 nm (truncation about (h / lambda)^2 / 6 = 2e-8
relative, rounding about 1e-12) pin it to 1e-7 normwise on 8 seeded unit vectors (measured 4e-9
on both shapes); the mixed path computes the boundary in FP64 too, so its products equal the FP64
ones; mdn_hvp_batch equals the single products bit for bit.
"""
import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi
from tests.hand_systems import HandSystem

pytestmark = pytest.mark.gpu

K, LAM = 41.0, 2.7
C0 = np.array([1500.0, 1500.0, 1500.0])
R = 1000.0


def _sphere():
    e = abi.BoundaryElement()
    e.kind = 1
    e.p[0], e.p[1], e.p[2], e.p[3] = C0[0], C0[1], C0[2], R
    e.krep, e.slen = K, LAM
    return e


def _unit(v):
    v = np.asarray(v, float)
    return v / np.linalg.norm(v)


def _systems():
    # inside the sphere: beads at 3-12 nm from the wall, one filament along the wall, one oblique
    along = [C0 + (R - d) * _unit((np.cos(t), np.sin(t), 0.1)) for d, t in ((4.0, 0.0), (7.0, 0.1), (11.0, 0.2))]
    oblique = [C0 + (R - d) * _unit((1.0, -0.5 + 0.3 * j, 0.8)) for j, d in enumerate((3.0, 6.0, 12.0))]
    ball = HandSystem([along, oblique], [_sphere()], grid=(3000.0, 3000.0, 3000.0))
    # below the plane z = 1000 (normal into the box), beads 3-10 nm from it
    top = of.plane_element(np.array([0.0, 0.0, 1000.0]), np.array([0.0, 0.0, -1.0]), K, LAM)
    flat = HandSystem([[np.array([500.0 + 100 * j, 500.0, 1000.0 - d]) for j, d in enumerate((3.0, 5.0, 10.0))],
                       [np.array([1500.0, 500.0 + 90 * j, 1000.0 - d]) for j, d in enumerate((9.0, 4.0, 6.0))]], [top])
    for hs in (ball, flat):
        hs._keep[3][:] = 0.0   # stretching k
        hs._keep[5][:] = 0.0   # bending k
        hs._keep[7][:] = 0.0   # excluded-volume k
    return {"sphere": ball, "plane": flat}


@pytest.mark.parametrize("shape", ["sphere", "plane"])
def test_hvp_of_the_boundary_repulsion(gpu, shape):
    hs = _systems()[shape]
    g = gpu()
    g.upload(hs.desc, hs.coord)
    g.reset_neighbor_lists(hs.coord)
    x = hs.coord
    o = of.OracleSystem(hs.desc)
    o.build_nl(x)
    assert len(o.boundary_list()) > 0
    f = g.compute_forces(x)
    fb = g.ff_compute_forces(abi.MDN_FF_BOUNDARY, x)
    np.testing.assert_array_equal(f, fb)   # the boundary is the only force
    assert float(np.abs(f).max()) > 1.0
    rng = np.random.default_rng(31)
    h = 1e-3
    V = np.stack([_unit(rng.standard_normal(x.size)) for _ in range(8)])
    worst = 0.0
    for v in V:
        hv = g.hvp(v, x)
        fd = -(g.compute_forces(x + h * v) - g.compute_forces(x - h * v)) / (2 * h)
        err = float(np.linalg.norm(hv - fd) / np.linalg.norm(hv))
        worst = max(worst, err)
        assert err <= 1e-7
        np.testing.assert_array_equal(g.hvp(v, x, precision=abi.MDN_PREC_MIXED), hv)
    print(f"{shape}: max|F| {float(np.abs(f).max()):.4f} pN, worst |Hv - FD| / |Hv| {worst:.3e} over 8 vectors",
          flush=True)
    single = np.stack([g.hvp(v, x) for v in V])
    np.testing.assert_array_equal(g.hvp_batch(V, x), single)
