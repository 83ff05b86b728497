"""GPU: the failure paths of the device minimizer, each forced from the host and checked against
the oracle's CGMethod::minimize where the reference defines the outcome.

* Line-search failure with the recovery disabled (CGMethod.hpp:1337-1367, MController.h:142-155):
  on a dense network whose first search fails, the device sets the line-search error flag, stops
  after the same calls as the oracle and reports MDN_ERR_LINESEARCH; the coordinates are the
  oracle's.
* Call limit with the minimum-energy backup (CGMethod.hpp:794-814, 1564-1566, 1579-1590): a small
  compressed system stopped by its force-call limit restores the backup coordinates, as the
  oracle does, and reports the FP64 energies there.
* A non-finite start (maxF not finite, CGMethod.hpp:202-212): MDN_ERR_NONFINITE, and the context
  minimizes the next valid input exactly as a fresh context does.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi, synth
from tests.hand_systems import HandSystem

pytestmark = pytest.mark.gpu

P = np.array


def _raw_minimize(g, p, coord):
    """mdn_minimize through the C ABI, returning (status, coordinates, result) also on errors."""
    x = np.array(coord, dtype=np.float64)
    f = np.zeros_like(x)
    res = abi.MinResult()
    dp = C.POINTER(C.c_double)
    st = g._L.mdn_minimize(g._h, C.byref(p), x.ctypes.data_as(dp), f.ctypes.data_as(dp), C.byref(res))
    return st, x, res


def test_line_search_failure_without_recovery(gpu):
    net = synth.Network.config(3, index=4, dense=True)
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    o = of.OracleSystem(net.desc)
    o.build_nl(net.coord)
    p = abi.default_cg_params(try_recover=0)
    st, x, r = _raw_minimize(g, p, net.coord)
    sto, xo, fo, ro = o.minimize(p, net.coord)
    print("gpu", st, r.error_flags, r.num_iter, r.num_energy_call, r.num_force_call, "oracle", sto, ro.error_flags,
          ro.num_iter, ro.num_energy_call, ro.num_force_call, flush=True)
    assert sto == 5 and ro.error_flags & 1                        # the reference's errorLineSearch
    assert st == 5 and r.error_flags == ro.error_flags            # MDN_ERR_LINESEARCH
    assert (r.num_iter, r.num_energy_call, r.num_force_call, r.num_safe_mode) == \
        (ro.num_iter, ro.num_energy_call, ro.num_force_call, ro.num_safe_mode)
    np.testing.assert_allclose(x, xo, rtol=1e-8, atol=0)


@pytest.mark.parametrize("limit", [20, 30])
def test_call_limit_restores_minimum_energy(gpu, limit):
    # two 3-bead filaments compressed to 63 / 60 nm bonds (rest length 108 nm), far apart
    hs = HandSystem([[P([500., 500, 500 + 63 * j]) for j in range(3)], [P([900., 500, 500 + 60 * j]) for j in range(3)]], [])
    g = gpu()
    g.upload(hs.desc, hs.coord)
    g.reset_neighbor_lists(hs.coord)
    o = of.OracleSystem(hs.desc)
    o.build_nl(hs.coord)
    p = abi.default_cg_params(grad_tol=300.0, force_call_limit=limit)
    st, x, r = _raw_minimize(g, p, hs.coord)
    sto, xo, fo, ro = o.minimize(p, hs.coord)
    print(limit, "gpu", st, r.num_iter, r.num_force_call, r.call_limit_exceeded, r.restored_min_energy,
          r.energies_after.total, "oracle", sto, ro.num_iter, ro.num_force_call, ro.call_limit_exceeded,
          ro.restored_min_energy, ro.energies_after.total, flush=True)
    assert sto == 0 and ro.call_limit_exceeded and ro.restored_min_energy
    assert st == 0 and r.call_limit_exceeded and r.restored_min_energy and r.error_flags == 0
    assert (r.num_iter, r.num_energy_call, r.num_force_call) == (ro.num_iter, ro.num_energy_call, ro.num_force_call)
    np.testing.assert_allclose(x, xo, rtol=1e-12, atol=0)
    assert r.energies_after.total == pytest.approx(ro.energies_after.total, rel=1e-12)
    assert r.energies_after.total == pytest.approx(o.energy(x)[0], rel=1e-12)


def test_nonfinite_start_then_recovers(gpu):
    net = synth.Network.config(0)
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    bad = net.coord.copy()
    bad[3 * 7 + 1] = np.nan
    st, _, _ = _raw_minimize(g, abi.default_cg_params(), bad)
    print("non-finite start:", st, g._L.mdn_last_error(g._h).decode(), flush=True)
    assert st == 4   # MDN_ERR_NONFINITE
    p = abi.default_cg_params(energy_call_limit=40)
    x1, f1, r1 = g.minimize(p, net.coord)
    g2 = gpu()
    g2.upload(net.desc, net.coord)
    g2.reset_neighbor_lists(net.coord)
    x2, f2, r2 = g2.minimize(p, net.coord)
    assert (r1.num_iter, r1.num_energy_call) == (r2.num_iter, r2.num_energy_call)
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(f1, f2)
