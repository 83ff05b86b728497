"""GPU: mdn_forces_noalloc, the force evaluation whose scratch lives in a file-scope device array
(no allocation per call), against mdn_forces / mdn_forces_mixed through the C ABI.

Bars: bit-identical forces in both precisions (the same kernels in the same order), on calls in a
row on systems of decreasing size that share the array: the 42,000-bead network, then two
cylinders on one line (whose pair needs the numerical force fallback, forceN), then the 550-bead
network (no fallback: the previous call's fallback count must not leak into it), then the
42,000-bead network again. A system whose evaluation needs more than the array holds is refused
with MDN_ERR_CAPACITY, leaving the force buffer untouched and the context usable.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from picgo_amd import abi, mech, synth
from tests.hand_systems import HandSystem

pytestmark = pytest.mark.gpu


def _noalloc(g, prec, x):
    L = mech.lib()
    L.mdn_forces_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]
    xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to("cuda:0")
    fd = torch.full_like(xd, float("nan"))
    torch.cuda.synchronize()
    st = L.mdn_forces_noalloc(g._h, prec, C.c_void_p(xd.data_ptr()), C.c_void_p(fd.data_ptr()), x.size)
    torch.cuda.synchronize()
    return st, fd.cpu().numpy()


def _ctx(gpu, desc, coord):
    g = gpu()
    g.upload(desc, coord)
    g.reset_neighbor_lists(coord)
    return g


@pytest.fixture(scope="module")
def systems(gpu):
    n1, n0 = synth.Network.config(1), synth.Network.config(0)
    line = HandSystem([[np.array([500., 500, 500]), np.array([560., 500, 500])],
                       [np.array([580., 500, 500]), np.array([640., 500, 500])]], [])
    rng = np.random.default_rng(11)
    x1 = n1.coord + rng.normal(0.0, 2.0, n1.coord.size)
    x0 = n0.coord + rng.normal(0.0, 2.0, n0.coord.size)
    return [(_ctx(gpu, n1.desc, n1.coord), x1, "42,000 beads"),
            (_ctx(gpu, line.desc, line.coord), line.coord.copy(), "collinear pair (forceN)"),
            (_ctx(gpu, n0.desc, n0.coord), x0, "550 beads"),
            (None, None, "42,000 beads again")], (n1, n0, line)


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_noalloc_equals_forces(systems, prec):
    seq, _ = systems
    seq = seq[:3] + [(seq[0][0], seq[0][1], seq[3][2])]
    for g, x, name in seq:
        st, f = _noalloc(g, prec, x)
        assert st == abi.MDN_OK, (name, g._L.mdn_last_error(g._h).decode())
        ref = g.compute_forces(x, precision=prec)
        assert np.all(np.isfinite(f)), name
        np.testing.assert_array_equal(f, ref, err_msg=name)
        print(f"precision {prec}, {name}: identical to mdn_forces (max|F| {np.abs(ref).max():.3e})", flush=True)


def test_noalloc_refuses_beyond_capacity(gpu):
    big = synth.Network.config(4, n_fil=60000)   # about 1.26 M beads at three times the density
    g = _ctx(gpu, big.desc, big.coord)
    st, f = _noalloc(g, abi.MDN_PREC_FP64, big.coord)
    msg = g._L.mdn_last_error(g._h).decode()
    print(f"{big.n_beads} beads: status {st}: {msg}", flush=True)
    assert st == abi.MDN_ERR_CAPACITY and "scratch" in msg
    assert np.all(np.isnan(f))   # the force buffer is untouched
    n0 = synth.Network.config(0)
    h = _ctx(gpu, n0.desc, n0.coord)
    st, f = _noalloc(h, abi.MDN_PREC_FP64, n0.coord)
    assert st == abi.MDN_OK
    np.testing.assert_array_equal(f, h.compute_forces(n0.coord))
