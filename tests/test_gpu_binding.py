"""GPU: the binding-site pair search (mdn_binding_pairs), HybridBindingSearchManager.cpp:759-966.

Bars: the device's pairs equal the oracle's (oracle/binding.c, itself pinned by site-to-site
distances in tests/test_oracle_math.py) element for element and in order, on the 550-bead network
for linker (30-40 nm) and motor (175-225 nm) ranges with and without the linker site step and
random free-site masks, and on the 42,000-bead network for the linker range (124,036 pairs); on
the 420,000-bead network a sample of the pairs lies in range by direct distance. Bad parameters
and a short output buffer are refused."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi, mech, synth

pytestmark = pytest.mark.gpu

SITES = np.arange(4, 40, 4) / 40.0


def _setup(gpu, cfg):
    net = synth.Network.config(cfg)
    g = gpu()
    g.upload(net.desc, net.coord)
    return net, g


@pytest.mark.parametrize("rmin,rmax,step,with_free", [(30.0, 40.0, 0, False), (30.0, 40.0, 1, True),
                                                     (175.0, 225.0, 0, False), (175.0, 225.0, 2, True)])
def test_binding_pairs_equal_oracle_550(gpu, rmin, rmax, step, with_free):
    net, g = _setup(gpu, 0)
    x = net.coord + np.random.default_rng(int(rmax) + step).normal(0.0, 4.0, net.coord.size)
    free = None
    if with_free:
        free = (np.random.default_rng(5).random((net.n_cylinders, len(SITES))) < 0.7).astype(np.float32).reshape(-1)
    bp = abi.binding_params(SITES, rmin, rmax, 108.0, site_step=step, site_free=free)
    got = g.binding_pairs(bp, x)
    ref = of.OracleSystem(net.desc).binding_pairs(bp, x)
    print(rmin, rmax, step, with_free, "pairs", len(got), len(ref), flush=True)
    assert len(ref) > 0
    np.testing.assert_array_equal(got, ref)


def test_binding_pairs_equal_oracle_42k(gpu):
    net, g = _setup(gpu, 1)
    bp = abi.binding_params(SITES, 30.0, 40.0, 108.0)
    got = g.binding_pairs(bp, net.coord)
    ref = of.OracleSystem(net.desc).binding_pairs(bp, net.coord)
    print("pairs", len(got), len(ref), flush=True)
    np.testing.assert_array_equal(got, ref)
    # from the device coordinates (the uploaded ones) as well
    np.testing.assert_array_equal(g.binding_pairs(bp), ref)


def test_binding_pairs_in_range_420k(gpu):
    net, g = _setup(gpu, 4)
    bp = abi.binding_params(SITES, 30.0, 40.0, 108.0)
    got = g.binding_pairs(bp)
    assert len(got) > 1000
    assert np.all(got[:, 0] >> 4 > got[:, 1] >> 4)
    rng = np.random.default_rng(9)
    nf = net.desc.n_filaments
    off = np.ctypeslib.as_array(net.desc.fil_bead_offset, shape=(nf + 1,))
    fb = np.ctypeslib.as_array(net.desc.fil_beads, shape=(int(off[-1]),))
    cyl_b = np.array([(fb[k], fb[k + 1]) for f in range(nf) for k in range(off[f], off[f + 1] - 1)])
    x = net.coord.reshape(-1, 3)
    s = SITES.astype(np.float32).astype(np.float64)
    for a, b in got[rng.choice(len(got), 2000, replace=False)]:
        (c1, p1), (c2, p2) = (a >> 4, a & 15), (b >> 4, b & 15)
        q1 = x[cyl_b[c1, 0]] + s[p1] * (x[cyl_b[c1, 1]] - x[cyl_b[c1, 0]])
        q2 = x[cyl_b[c2, 0]] + s[p2] * (x[cyl_b[c2, 1]] - x[cyl_b[c2, 0]])
        d = np.linalg.norm(q1 - q2)
        assert 30.0 * (1 - 1e-6) <= d <= 40.0 * (1 + 1e-6)


def test_binding_pairs_reject_bad_input(gpu):
    net, g = _setup(gpu, 0)
    bad = abi.binding_params(SITES, 30.0, 40.0, 108.0, shift_bits=3)    # 9 sites do not fit 3 bits
    with pytest.raises(mech.MdnError):
        g.binding_pairs(bad)
    bp = abi.binding_params(SITES, 175.0, 225.0, 108.0)
    n = C.c_int64()
    out = np.zeros(2, np.uint32)
    st = g._L.mdn_binding_pairs(g._h, C.byref(bp), None, out.ctypes.data_as(C.POINTER(C.c_uint32)), 1, C.byref(n))
    assert st == abi.MDN_ERR_CAPACITY and n.value > 1
