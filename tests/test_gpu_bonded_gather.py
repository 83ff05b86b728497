"""GPU: the two ways the filament stretching and bending forces reach the per-bead sums.

Either the terms (or line-search) kernel writes each bonded term's forces into bead-major piece
arrays and the gather reads them, or bead blocks of that kernel recompute each bead's pieces from
the coordinates (the same term functions on the same inputs) and store their sum, which the
gather starts from, while the filament terms' own blocks leave their forces out. By default
force evaluations and the minimizer use the sums (from 100,000 beads the line search's bead
blocks take the filament terms' energies as well); MDN_BONDED_GATHER=0 / 1 at context creation
fixes one way for everything. The two must agree bit for bit: forces and per-force-field
forces in both precisions (per-force-field evaluations carry the caller's base forces through the
sums), whole minimization trajectories, the Hessian-vector product (whose tangent pieces always
go through the piece arrays) and the batched minimizer (which always uses them). The recomputed
FP64 bonded forces are also checked against the oracle bit for bit.
"""
import os

import numpy as np
import pytest

import ctypes as C

from oracle import oracle_ffi as of
from picgo_amd import abi, mech, synth
from tests.hand_systems import HandSystem

pytestmark = pytest.mark.gpu


def _ctx(gpu, net, mode):
    """A context with MDN_BONDED_GATHER=mode at creation (None: the defaults)."""
    if mode is not None:
        os.environ["MDN_BONDED_GATHER"] = str(mode)
    try:
        g = gpu()
    finally:
        os.environ.pop("MDN_BONDED_GATHER", None)
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    return g


@pytest.mark.parametrize("cfg", [0, 1])
def test_recomputed_bonded_forces_equal_pieces(gpu, cfg):
    net = synth.Network.config(cfg)
    a, b = _ctx(gpu, net, 0), _ctx(gpu, net, 1)
    x = net.coord
    for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
        fa, fb = a.compute_forces(x, precision=prec), b.compute_forces(x, precision=prec)
        print(f"config {cfg} precision {prec}: max |dF| {float(np.abs(fa - fb).max()):.3e}", flush=True)
        np.testing.assert_array_equal(fa, fb)
    base = np.random.default_rng(10 + cfg).normal(scale=50.0, size=x.size)
    for ff in range(6):
        np.testing.assert_array_equal(a.ff_compute_forces(ff, x), b.ff_compute_forces(ff, x))
        # accumulating into given forces, as ForceField::computeForces does
        np.testing.assert_array_equal(a.ff_compute_forces(ff, x, force=base.copy()),
                                      b.ff_compute_forces(ff, x, force=base.copy()))
    o = of.OracleSystem(net.desc)
    o.build_nl(x)
    for ff in (abi.MDN_FF_STRETCH, abi.MDN_FF_BEND):
        np.testing.assert_array_equal(b.ff_compute_forces(ff, x), o.ff_forces(ff, x))
    v = np.random.default_rng(cfg).normal(size=x.size)

    def hv(g):   # an unrelaxed network may hold a pair without an analytic tangent: same error
        try:
            return g.hvp(v, x)
        except mech.MdnError as e:
            return e.code
    ha, hb = hv(a), hv(b)
    print(f"config {cfg}: hvp", "error code" if np.isscalar(ha) else "vectors", flush=True)
    np.testing.assert_array_equal(ha, hb)


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_recomputed_bonded_minimization_equals_pieces(gpu, prec):
    net = synth.Network.config(1)
    a, b = _ctx(gpu, net, 0), _ctx(gpu, net, 1)
    p = abi.default_cg_params(precision=prec, energy_call_limit=200)
    xa, fa, ra = a.minimize(p, net.coord)
    xb, fb, rb = b.minimize(p, net.coord)
    print("precision", prec, "iterations", ra.num_iter, rb.num_iter, "energy calls", ra.num_energy_call,
          rb.num_energy_call, flush=True)
    assert (ra.num_iter, ra.num_energy_call, ra.num_force_call, ra.num_safe_mode) == \
        (rb.num_iter, rb.num_energy_call, rb.num_force_call, rb.num_safe_mode)
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(fa, fb)
    assert ra.energies_after.total == rb.energies_after.total


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_large_network_minimization_equals_pieces(gpu, prec):
    """From 100,000 beads the line search's bead blocks also take the filament terms' energies
    (their own blocks leave the grid): the 420,000-bead network's trajectory with the defaults
    must be the piece-array trajectory, bit for bit (the line-search energies are summed in
    another order, which moves no decision here)."""
    net = synth.Network.config(4)
    a, b = _ctx(gpu, net, 0), _ctx(gpu, net, None)
    p = abi.default_cg_params(precision=prec, energy_call_limit=40)
    xa, fa, ra = a.minimize(p, net.coord)
    xb, fb, rb = b.minimize(p, net.coord)
    print("precision", prec, "iterations", ra.num_iter, rb.num_iter, flush=True)
    assert (ra.num_iter, ra.num_energy_call, ra.num_force_call) == (rb.num_iter, rb.num_energy_call, rb.num_force_call)
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(fa, fb)
    np.testing.assert_array_equal(a.compute_forces(net.coord, precision=prec), b.compute_forces(net.coord, precision=prec))


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_bead_neighbours_from_indices_equal_tables(gpu, prec):
    """With the beads numbered filament after filament (the generator's and serializeDof's order)
    the bead blocks take a bead's neighbours from the adjacent indices and its bending term from
    its stretching term; MDN_BEAD_TABLES=1 makes them read the neighbour tables instead. On the
    420,000-bead network (where the line search's bead blocks also take the filament energies)
    the forces, energies and a 40-call trajectory must be bit-identical."""
    net = synth.Network.config(4)
    os.environ["MDN_BEAD_TABLES"] = "1"
    try:
        a = gpu()
        a.upload(net.desc, net.coord)
    finally:
        os.environ.pop("MDN_BEAD_TABLES", None)
    a.reset_neighbor_lists(net.coord)
    b = _ctx(gpu, net, None)
    x = net.coord + np.random.default_rng(12).normal(0.0, 3.0, net.coord.size)
    np.testing.assert_array_equal(a.compute_forces(x, precision=prec), b.compute_forces(x, precision=prec))
    ea, eb = a.compute_individual_energies(x, precision=prec), b.compute_individual_energies(x, precision=prec)
    assert list(ea.ff) == list(eb.ff)
    p = abi.default_cg_params(precision=prec, energy_call_limit=40)
    xa, fa, ra = a.minimize(p, net.coord)
    xb, fb, rb = b.minimize(p, net.coord)
    assert (ra.num_iter, ra.num_energy_call, ra.num_force_call) == (rb.num_iter, rb.num_energy_call, rb.num_force_call)
    np.testing.assert_array_equal(xa, xb)
    np.testing.assert_array_equal(fa, fb)


def test_batched_with_recomputing_contexts(gpu):
    nets = [synth.Network.config(3, index=i) for i in range(3)]
    sa = [_ctx(gpu, n, 0) for n in nets]
    sb = [_ctx(gpu, n, 1) for n in nets]
    p = abi.default_cg_params(energy_call_limit=80)
    xa, fa, ra = mech.minimize_batched(sa, p, [n.coord for n in nets])
    xb, fb, rb = mech.minimize_batched(sb, p, [n.coord for n in nets])
    for i in range(len(nets)):
        assert (ra[i].num_iter, ra[i].num_energy_call) == (rb[i].num_iter, rb[i].num_energy_call)
        np.testing.assert_array_equal(xa[i], xb[i])
        np.testing.assert_array_equal(fa[i], fb[i])


def _relabelled(desc, coord, seed):
    """The same system with its beads renumbered by a random permutation: bead b becomes perm[b]
    (coordinates moved accordingly, filament bead lists mapped), so a filament's beads are no
    longer consecutive in the bead array. Cylinders keep their numbering."""
    n = desc.n_beads
    perm = np.random.default_rng(seed).permutation(n).astype(np.int32)
    m = desc.fil_bead_offset[desc.n_filaments]
    fb = np.array([perm[desc.fil_beads[i]] for i in range(m)], dtype=np.int32)
    x = np.empty_like(coord)
    x.reshape(-1, 3)[perm] = coord.reshape(-1, 3)
    d = abi.SystemDesc.from_buffer_copy(desc)
    d.fil_beads = fb.ctypes.data_as(C.POINTER(C.c_int32))
    d._keep = fb
    return d, x, perm


def _short_filaments():
    """Filaments of 2, 3, 4 and 5 beads (1 to 4 cylinders), bent, far from each other and from
    the boundary."""
    P = np.array
    fils = []
    for j, nb in enumerate((2, 3, 4, 5, 2, 3)):
        base = P([400.0 + 350.0 * j, 600.0, 700.0])
        fils.append([base + P([100.0 * k, 15.0 * k * k, 5.0 * k]) for k in range(nb)])
    return HandSystem(fils, [])


@pytest.mark.parametrize("case", ["network", "short"])
def test_relabelled_beads(gpu, case):
    """Bonded sums use per-bead neighbour tables built from the filament bead lists: with the beads
    renumbered, every force and energy must be the original one at the renumbered bead, bit for
    bit, in both layouts and precisions (and the stretching and bending forces the oracle's)."""
    if case == "network":
        net = synth.Network.config(0)
        desc, coord = net.desc, net.coord
    else:
        hs = _short_filaments()
        desc, coord = hs.desc, hs.coord
    d2, x2, perm = _relabelled(desc, coord, 5)
    for mode in (0, 1):
        a, b = _ctx_desc(gpu, desc, coord, mode), _ctx_desc(gpu, d2, x2, mode)
        for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
            fa = a.compute_forces(coord, precision=prec).reshape(-1, 3)
            fb = b.compute_forces(x2, precision=prec).reshape(-1, 3)
            np.testing.assert_array_equal(fb[perm], fa)
            ea, eb = a.compute_individual_energies(coord, precision=prec), b.compute_individual_energies(x2, precision=prec)
            assert list(ea.ff) == list(eb.ff)
        o = of.OracleSystem(d2)
        o.build_nl(x2)
        for ff in (abi.MDN_FF_STRETCH, abi.MDN_FF_BEND):
            np.testing.assert_array_equal(b.ff_compute_forces(ff, x2), o.ff_forces(ff, x2))
        print(f"{case} mode {mode}: forces and energies identical after renumbering", flush=True)


def _ctx_desc(gpu, desc, coord, mode):
    os.environ["MDN_BONDED_GATHER"] = str(mode)
    try:
        g = gpu()
    finally:
        os.environ.pop("MDN_BONDED_GATHER", None)
    g.upload(desc, coord)
    g.reset_neighbor_lists(coord)
    return g
