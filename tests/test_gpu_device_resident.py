"""GPU: the device-resident entry points that the successive-minimization workload uses.

* mdn_perturb: the device's seeded Gaussian displacement equals the host generator
  (picgo_amd.mech.perturbation) to a few ulps of the noise (the device's log / cos / sqrt), so the
  displaced coordinates agree to 1e-9 nm.
* mdn_minimize with no host buffer: starts from the device coordinates and leaves its result
  there; bit-identical (counts, coordinates, forces) to the same minimization from a host buffer.
* mdn_displace: x += delta from a device buffer, bit-identical to the host sum.
* mdn_nl_build with no host buffer: the lists of the displaced device coordinates, identical to
  those built from the same coordinates passed from the host.
"""
import numpy as np
import pytest
import torch

from picgo_amd import abi, mech, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def net1(gpu):
    net = synth.Network.config(1)
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    return net, g


@pytest.mark.parametrize("stream", [0, 7])
def test_perturb_matches_host_generator(net1, stream):
    net, g = net1
    g.set_coords(net.coord)
    g.perturb(4, stream, 5.0)
    x = g.get_coords()
    ref = net.coord + mech.perturbation(4, stream, 5.0, net.coord.size)
    err = float(np.abs(x - ref).max())
    print(f"stream {stream}: max |device - host| {err:.3e} nm", flush=True)
    assert err <= 1e-9
    assert float(np.abs(x - net.coord).std()) == pytest.approx(5.0 * np.sqrt(1 - 2 / np.pi), rel=0.01)


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_minimize_from_device_coordinates(net1, prec):
    net, g = net1
    p = abi.default_cg_params(precision=prec, energy_call_limit=120)
    g.set_coords(net.coord)
    g.perturb(4, 1, 5.0)
    x0 = g.get_coords()
    g.reset_neighbor_lists(x0)
    xh, fh, rh = g.minimize(p, x0)
    g.set_coords(x0)
    _, fd, rd = g.minimize(p, None)
    xd = g.get_coords()
    assert (rh.num_iter, rh.num_energy_call, rh.num_force_call) == (rd.num_iter, rd.num_energy_call, rd.num_force_call)
    np.testing.assert_array_equal(xd, xh)
    np.testing.assert_array_equal(fd, fh)
    assert rd.energies_after.total == rh.energies_after.total
    # a second minimization continues from the first one's device result
    _, _, r2 = g.minimize(p, None)
    xh2, _, rh2 = g.minimize(p, xh)
    np.testing.assert_array_equal(g.get_coords(), xh2)
    assert r2.num_iter == rh2.num_iter


def test_displace_and_rebuild_from_device_coordinates(net1):
    net, g = net1
    delta = np.random.default_rng(3).normal(0.0, 5.0, net.coord.size)
    d_dev = torch.from_numpy(delta).to("cuda:0")
    torch.cuda.synchronize()
    g.set_coords(net.coord)
    g.displace(d_dev.data_ptr())
    x = g.get_coords()
    np.testing.assert_array_equal(x, net.coord + delta)
    n_dev = g.reset_neighbor_lists(None)
    pairs_dev = g.neighbor_pairs()
    n_host = g.reset_neighbor_lists(net.coord + delta)
    assert n_dev == n_host
    np.testing.assert_array_equal(pairs_dev, g.neighbor_pairs())
    g.reset_neighbor_lists(net.coord)


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_direct_slot_launches_equal_graph_replay(gpu, prec):
    """MDN_CG_GRAPH=0 (a context that launches the minimizer's slots directly instead of replaying
    the captured graph): the same kernels and decisions, so the same minimization bit for bit. The
    graph is captured when a layout is minimized a second time, so the default context minimizes
    twice and its second (graph-replayed) minimization is compared, and its first too."""
    import os
    net = synth.Network.config(1)
    os.environ["MDN_CG_GRAPH"] = "0"
    try:
        d = gpu()
    finally:
        os.environ.pop("MDN_CG_GRAPH", None)
    g = gpu()
    for c in (d, g):
        c.upload(net.desc, net.coord)
        c.reset_neighbor_lists(net.coord)
    p = abi.default_cg_params(precision=prec, energy_call_limit=150)
    xd, fd, rd = d.minimize(p, net.coord)
    for _ in range(2):   # direct launches, then the captured graph
        xg, fg, rg = g.minimize(p, net.coord)
        assert (rd.num_iter, rd.num_energy_call, rd.num_force_call) == (rg.num_iter, rg.num_energy_call, rg.num_force_call)
        np.testing.assert_array_equal(xd, xg)
        np.testing.assert_array_equal(fd, fg)
