"""GPU: the order in which threads take the linkers and motors (sorted by the first bead of their
first cylinder at upload; MDN_XL_ORDER=0 keeps the interaction order) changes no force: every
crosslink's forces land in its own slots, which the gather reads in the reference's order.

Bars on the 42,000-bead network (8,000 linkers and 4,000 motors in random order): forces, per
force field forces, crosslink stretch forces and Hessian-vector products bit-identical between the
two orders in both precisions; energies within 1e-13 relative (block partial sums hold other
crosslinks); a minimization prefix of 60 energy calls with identical counts and coordinates within
1e-8."""
import os

import numpy as np
import pytest

from picgo_amd import abi, synth

pytestmark = pytest.mark.gpu


def _ctx(gpu, net, order):
    if order is not None:
        os.environ["MDN_XL_ORDER"] = order
    try:
        g = gpu()
        g.upload(net.desc, net.coord)
    finally:
        os.environ.pop("MDN_XL_ORDER", None)
    g.reset_neighbor_lists(net.coord)
    return g


def test_crosslink_order_changes_no_force(gpu):
    net = synth.Network.config(1)
    a, b = _ctx(gpu, net, "0"), _ctx(gpu, net, None)
    x = net.coord + np.random.default_rng(41).normal(0.0, 2.0, net.coord.size)
    for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
        np.testing.assert_array_equal(a.compute_forces(x, precision=prec), b.compute_forces(x, precision=prec))
        la, ma = a.stretch_forces(net.desc.n_linkers, net.desc.n_motors)
        lb, mb = b.stretch_forces(net.desc.n_linkers, net.desc.n_motors)
        np.testing.assert_array_equal(la, lb)
        np.testing.assert_array_equal(ma, mb)
    for ff in (abi.MDN_FF_LINKER, abi.MDN_FF_MOTOR):
        np.testing.assert_array_equal(a.ff_compute_forces(ff, x), b.ff_compute_forces(ff, x))
    for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
        ea = a.compute_individual_energies(x, precision=prec)
        eb = b.compute_individual_energies(x, precision=prec)
        for k in range(abi.MDN_NUM_FF):
            assert eb.ff[k] == pytest.approx(ea.ff[k], rel=1e-13, abs=1e-300)
    v = np.random.default_rng(42).normal(size=x.size)
    for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
        np.testing.assert_array_equal(a.hvp(v, x, precision=prec), b.hvp(v, x, precision=prec))
    p = abi.default_cg_params(energy_call_limit=60)
    xa, _, ra = a.minimize(p, net.coord)
    xb, _, rb = b.minimize(p, net.coord)
    assert (ra.num_iter, ra.num_energy_call, ra.num_force_call) == (rb.num_iter, rb.num_energy_call, rb.num_force_call)
    np.testing.assert_allclose(xb, xa, rtol=1e-8, atol=0)
