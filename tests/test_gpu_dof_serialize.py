"""GPU: serializeDof on the device (mdn_dof_serialize), DofSerializer.hpp:64-92.

Chemistry between two minimizations destroys and creates beads; the next system's coordinate
vector lists the beads in Bead::getBeads() order, which the element database keeps by appending
new elements and moving the last element into a destroyed one's index (oracle/database.py, pinned
by the reference's TestDatabase fixture). Here beads of the 550-bead network are destroyed and
re-created as new elements (their coordinates passed as new beads), so the next system is the same
network under the database's new bead order. Bars: the gathered device coordinates equal the
host-side reordering bit for bit; a minimization from them equals the same minimization from a
host buffer bit for bit; bad maps and a coordinate-less upload without a gather are refused."""
import ctypes as C

import numpy as np
import pytest

from oracle.database import ElementList, serialize_map
from picgo_amd import abi, mech, synth

pytestmark = pytest.mark.gpu


def relabelled(desc, new_index):
    """The system `desc` with bead b renamed new_index[b] (cylinders keep their filament-major
    numbering, so the crosslink and boundary tables carry over)."""
    nf = desc.n_filaments
    off = np.ctypeslib.as_array(desc.fil_bead_offset, shape=(nf + 1,)).copy()
    fb = np.ctypeslib.as_array(desc.fil_beads, shape=(int(off[-1]),))
    beads = np.ascontiguousarray(new_index[fb], dtype=np.int32)
    keep = [off.astype(np.int32), beads]
    d = abi.SystemDesc()
    C.pointer(d)[0] = desc
    d.fil_bead_offset = keep[0].ctypes.data_as(C.POINTER(C.c_int32))
    d.fil_beads = keep[1].ctypes.data_as(C.POINTER(C.c_int32))
    return d, keep


def test_device_serialize_equals_host_order(gpu):
    net = synth.Network.config(0)
    n = net.n_beads
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    x1, _, _ = g.minimize(abi.default_cg_params(energy_call_limit=30), net.coord)   # device holds x1 too
    # chemistry: beads destroyed and re-created as new elements, interleaved
    rng = np.random.default_rng(8)
    el = ElementList()
    for b in range(n):
        el.create(("old", b))
    for b in rng.choice(n, 60, replace=False):
        el.destroy(("old", int(b)))
        if rng.random() < 0.7:
            el.create(("new", int(b)))
    for b in range(n):   # every bead exists again
        if ("old", b) not in el.elems and ("new", b) not in el.elems:
            el.create(("new", b))
    order = [b for _, b in el.elems]              # the physical bead at each new index
    assert sorted(order) == list(range(n))
    src = serialize_map([("old", b) for b in range(n)], el.elems)
    new_coord = np.concatenate([x1[3 * b:3 * b + 3] for (kind, b) in el.elems if kind == "new"])
    assert sum(1 for s in src if s < 0) == len(new_coord) // 3 > 0
    g.dof_serialize(src, new_coord)
    new_index = np.empty(n, np.int64)
    new_index[order] = np.arange(n)
    desc2, keep = relabelled(net.desc, new_index)
    x_next = x1.reshape(-1, 3)[order].reshape(-1)
    g.upload(desc2, None)
    np.testing.assert_array_equal(g.get_coords(), x_next)
    h = gpu()
    h.upload(desc2, x_next)
    p = abi.default_cg_params(energy_call_limit=80)
    assert g.reset_neighbor_lists(None) == h.reset_neighbor_lists(x_next)
    _, fg, rg = g.minimize(p, None)
    xh, fh, rh = h.minimize(p, x_next)
    assert (rg.num_iter, rg.num_energy_call, rg.num_force_call) == (rh.num_iter, rh.num_energy_call, rh.num_force_call)
    np.testing.assert_array_equal(g.get_coords(), xh)
    np.testing.assert_array_equal(fg, fh)
    del keep


def test_device_serialize_rejects_bad_maps(gpu):
    net = synth.Network.config(0)
    g = gpu()
    g.upload(net.desc, net.coord)
    n = net.n_beads
    with pytest.raises(mech.MdnError):
        g.dof_serialize([0, 0] + list(range(2, n)))            # a current bead named twice
    with pytest.raises(mech.MdnError):
        g.dof_serialize(list(range(n - 1)) + [n])              # beyond the current beads
    with pytest.raises(mech.MdnError):
        g.dof_serialize(list(range(n - 1)) + [-1])             # a new bead without coordinates
    h = gpu()
    with pytest.raises(mech.MdnError):
        h.upload(net.desc, None)                               # nothing gathered
    g.dof_serialize(list(range(n)))
    with pytest.raises(mech.MdnError):
        small, keep = relabelled(net.desc, np.arange(n))
        small.n_beads = n - 1                                  # gathered n beads, not n - 1
        g.upload(small, None)
