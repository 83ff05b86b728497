"""The oracle reproduces its committed golden fixture bit for bit (tests/golden/make_golden.py)."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import compute

HERE = os.path.dirname(os.path.abspath(__file__))


def test_oracle_matches_golden_fixture():
    g = np.load(os.path.join(HERE, "golden", "config0_oracle.npz"))
    now = compute()
    assert set(g.files) == set(now)
    for k in g.files:
        np.testing.assert_array_equal(now[k], g[k], err_msg=k)


@pytest.mark.gpu
def test_device_matches_golden_fixture(gpu):
    """The device path against the committed fixture (no oracle at run time): the 50-filament
    network's neighbor pairs and boundary beads identical, the stretching / bending / linker / motor
    forces bit-identical, excluded volume and boundary within 1e-12 of max|F|, energies 1e-12."""
    from picgo_amd import abi, synth
    g_ = np.load(os.path.join(HERE, "golden", "config0_oracle.npz"))
    net = synth.Network.config(0)
    np.testing.assert_array_equal(net.coord, g_["coord"])
    dev = gpu()
    dev.upload(net.desc, net.coord)
    dev.reset_neighbor_lists(net.coord)
    np.testing.assert_array_equal(dev.neighbor_pairs(), g_["pairs"])
    beads, _ = dev.boundary_interactions(net.desc.n_boundary)
    np.testing.assert_array_equal(beads, g_["boundary_beads"])
    exact = (abi.MDN_FF_STRETCH, abi.MDN_FF_BEND, abi.MDN_FF_LINKER, abi.MDN_FF_MOTOR)
    for ff in range(abi.MDN_NUM_FF):
        fd, fg = dev.ff_compute_forces(ff, net.coord), g_[f"force_ff{ff}"]
        if ff in exact:
            np.testing.assert_array_equal(fd, fg, err_msg=abi.FF_NAMES[ff])
        else:
            assert float(np.abs(fd - fg).max()) <= 1e-12 * max(1.0, float(np.abs(fg).max())), abi.FF_NAMES[ff]
    rep = dev.compute_individual_energies(net.coord)
    for ff in range(abi.MDN_NUM_FF):
        assert rep.ff[ff] == pytest.approx(float(g_["energy_ff"][ff]), rel=1e-12, abs=1e-300)
