"""GPU: the persistent batched minimizer (mdn_minimize_batched) on many independent networks.

Bars:
  trajectory prefix: with an energy-call limit, every system of the batch follows the oracle's
    CGMethod::minimize trajectory (same iteration / energy-call / force-call counts, coordinates
    within rtol 1e-8: only the rounding order of the reductions differs), through the
    safe-mode entries and recovery moves these dense networks start with;
  full batch (all 256 networks): every system stops by the reference criterion without an error; the reported
    max|F| is the FP64 max|F| at the returned coordinates (FP64 batch: exactly; mixed batch:
    within 1e-3 pN), the reported final energy is the FP64 energy there, and it does not exceed
    the start energy;
  dense networks (excluded volume, safe mode, recovery moves): the same checks;
  queue: with the persistent grid capped below the number of systems (every workgroup pulls
    several systems in turn, reusing its shared state), results are bit-identical to the
    uncapped batch and still follow the oracle's trajectory prefixes.
"""
import os

import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi, mech, synth

pytestmark = pytest.mark.gpu


def _batch(gpu, n, dense=False, first=0):
    nets = [synth.Network.config(3, index=first + i, dense=dense) for i in range(n)]
    ffms = []
    for net in nets:
        g = gpu()
        g.upload(net.desc, net.coord)
        g.reset_neighbor_lists(net.coord)
        ffms.append(g)
    return nets, ffms


@pytest.mark.parametrize("ncalls", [10, 22])
def test_batched_trajectory_prefix_vs_oracle(gpu, ncalls):
    nets, ffms = _batch(gpu, 6, dense=True)
    p = abi.default_cg_params(energy_call_limit=ncalls)
    xs, fs, res = mech.minimize_batched(ffms, p, [n.coord for n in nets])
    for i, net in enumerate(nets):
        o = of.OracleSystem(net.desc)
        o.build_nl(net.coord)
        st, xo, fo, ro = o.minimize(p, net.coord)
        rb = res[i]
        print(f"system {i}: beads {net.n_beads} batched {rb.num_iter}/{rb.num_energy_call}/{rb.num_force_call} "
              f"oracle {ro.num_iter}/{ro.num_energy_call}/{ro.num_force_call} "
              f"max|dx| {float(np.abs(xs[i] - xo).max()):.3e}", flush=True)
        assert (rb.num_iter, rb.num_energy_call, rb.num_force_call) == (ro.num_iter, ro.num_energy_call, ro.num_force_call)
        np.testing.assert_allclose(xs[i], xo, rtol=1e-8, atol=0)


def test_batched_relabelled_beads_prefix_vs_oracle(gpu):
    """Dense networks with their beads renumbered by a random permutation, so a filament's beads
    are not consecutive: the batched workgroups then take each bead's filament neighbours from the
    neighbour tables. The trajectory prefix equals the oracle's on the renumbered systems."""
    from tests.test_gpu_bonded_gather import _relabelled
    nets = [synth.Network.config(3, index=i, dense=True) for i in range(4)]
    systems, ffms = [], []
    for k, net in enumerate(nets):
        d2, x2, _ = _relabelled(net.desc, net.coord, 11 + k)
        g = gpu()
        g.upload(d2, x2)
        g.reset_neighbor_lists(x2)
        systems.append((d2, x2))
        ffms.append(g)
    p = abi.default_cg_params(energy_call_limit=22)
    xs, _, res = mech.minimize_batched(ffms, p, [x for _, x in systems])
    for i, (d2, x2) in enumerate(systems):
        o = of.OracleSystem(d2)
        o.build_nl(x2)
        st, xo, fo, ro = o.minimize(p, x2)
        rb = res[i]
        print(f"relabelled system {i}: batched {rb.num_iter}/{rb.num_energy_call} oracle {ro.num_iter}/{ro.num_energy_call} "
              f"max|dx| {float(np.abs(xs[i] - xo).max()):.3e}", flush=True)
        assert (rb.num_iter, rb.num_energy_call, rb.num_force_call) == (ro.num_iter, ro.num_energy_call, ro.num_force_call)
        np.testing.assert_allclose(xs[i], xo, rtol=1e-8, atol=0)


def test_batched_short_filaments_vs_oracle(gpu):
    """Filaments of 2 to 5 beads (every absent-neighbour case of the bead-by-bead bonded sums), as
    given and with the beads renumbered, minimized in one batch beside a network: each follows the
    oracle's whole minimization (counts identical, coordinates 1e-8)."""
    from tests.test_gpu_bonded_gather import _relabelled, _short_filaments
    hs = _short_filaments()
    d2, x2, _ = _relabelled(hs.desc, hs.coord, 7)
    net = synth.Network.config(3, index=5)
    systems = [(hs.desc, hs.coord), (d2, x2), (net.desc, net.coord)]
    ffms = []
    for d, x in systems:
        g = gpu()
        g.upload(d, x)
        g.reset_neighbor_lists(x)
        ffms.append(g)
    p = abi.default_cg_params()
    xs, _, res = mech.minimize_batched(ffms, p, [x for _, x in systems])
    for i, (d, x) in enumerate(systems[:2]):
        o = of.OracleSystem(d)
        o.build_nl(x)
        st, xo, fo, ro = o.minimize(p, x)
        rb = res[i]
        print(f"short filaments {i}: batched {rb.num_iter}/{rb.num_energy_call} oracle {ro.num_iter}/{ro.num_energy_call} "
              f"max|dx| {float(np.abs(xs[i] - xo).max()):.3e}", flush=True)
        assert (rb.num_iter, rb.num_energy_call, rb.num_force_call) == (ro.num_iter, ro.num_energy_call, ro.num_force_call)
        np.testing.assert_allclose(xs[i], xo, rtol=1e-8, atol=1e-9)
    assert res[2].converged_grad or res[2].converged_energy


def test_batched_mixed_reaches_the_force_tolerance(gpu):
    """The mixed batched minimizer with the energy criterion off: every network of a batch (32 of
    the 256-network set, after seeded 5 nm bead noise so that each holds real forces) reaches the
    reference's force tolerance max|F| < 1 pN, and FP64 re-evaluation at each result confirms it:
    the mixed forces there are within 1e-5 relative (normwise) of FP64."""
    rng = np.random.default_rng(9)
    nets = [synth.Network.config(3, index=8 * i) for i in range(32)]
    ffms, starts = [], []
    for net in nets:
        g = gpu()
        x = net.coord + 5.0 * rng.standard_normal(net.coord.size)
        g.upload(net.desc, x)
        g.reset_neighbor_lists(x)
        ffms.append(g)
        starts.append(x)
    p = abi.default_cg_params(precision=abi.MDN_PREC_MIXED, energy_rel_tol=0.0)
    xs, fs, res = mech.minimize_batched(ffms, p, starts)
    worst_rel, worst_f64 = 0.0, 0.0
    for i, (g, r) in enumerate(zip(ffms, res)):
        assert r.converged_grad and not r.call_limit_exceeded and r.error_flags == 0, i
        assert r.grad_inf_norm < 1.0, i
        f64 = g.compute_forces(xs[i])
        fm = g.compute_forces(xs[i], precision=abi.MDN_PREC_MIXED)
        np.testing.assert_array_equal(fm, fs[i])   # the returned forces are the mixed path's at the result
        rel = float(np.abs(fm - f64).max() / np.abs(f64).max())
        worst_rel, worst_f64 = max(worst_rel, rel), max(worst_f64, float(np.abs(f64).max()))
        assert rel <= 1e-5, (i, rel)
        assert float(np.abs(f64).max()) < 1.0 + 1e-5, i
    print(f"32 networks, force-only mixed batch: worst FP64 max|F| {worst_f64:.6f} pN, worst relative "
          f"force error {worst_rel:.3e}", flush=True)


def test_batched_quadratic_line_search_prefix_vs_oracle(gpu):
    """The quadratic line search in the persistent kernel (the same ls_control as the
    single-system minimizer): networks 0-3 of the batch (on three of them the oracle's parabola
    phase runs within 40 energy calls) follow the oracle's trajectories."""
    nets, ffms = _batch(gpu, 4)
    kw = dict(energy_call_limit=40, lambda_max=0.1, max_dist=5.0)
    p = abi.default_cg_params(line_search=abi.MDN_LS_QUADRATIC, **kw)
    xs, fs, res = mech.minimize_batched(ffms, p, [n.coord for n in nets])
    changed = 0
    for i, net in enumerate(nets):
        o = of.OracleSystem(net.desc)
        o.build_nl(net.coord)
        st, xo, fo, ro = o.minimize(p, net.coord)
        _, xb, _, _ = o.minimize(abi.default_cg_params(**kw), net.coord)
        changed += not np.array_equal(xo, xb)
        rb = res[i]
        print(f"system {i}: batched {rb.num_iter}/{rb.num_energy_call}/{rb.num_force_call} "
              f"oracle {ro.num_iter}/{ro.num_energy_call}/{ro.num_force_call} "
              f"max|dx| {float(np.abs(xs[i] - xo).max()):.3e}", flush=True)
        assert (rb.num_iter, rb.num_energy_call, rb.num_force_call) == (ro.num_iter, ro.num_energy_call, ro.num_force_call)
        np.testing.assert_allclose(xs[i], xo, rtol=1e-8, atol=0)
    assert changed >= 3


@pytest.fixture(scope="module")
def batch256(gpu):
    """The full batch: 256 networks of 20-400 filaments x 5-30 cylinders (about 1.03 M beads)."""
    return _batch(gpu, 256)


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_batched_config3_converges(batch256, prec):
    nets, ffms = batch256
    p = abi.default_cg_params(precision=prec)
    xs, fs, res = mech.minimize_batched(ffms, p, [n.coord for n in nets])
    lead_sched = mech.last_batch_schedule(ffms[0])
    print(f"precision {prec}: schedule {lead_sched}, iterations max {max(r.num_iter for r in res)}", flush=True)
    bar = 0.0 if prec == abi.MDN_PREC_FP64 else 1e-3
    for i, (g, x, f, r) in enumerate(zip(ffms, xs, fs, res)):
        assert r.error_flags == 0 and not r.call_limit_exceeded, i
        assert r.converged_grad or r.converged_energy, i
        fx = g.compute_forces(x)
        assert abs(float(np.abs(fx).max()) - r.grad_inf_norm) <= bar, i
        assert float(np.abs(f - fx).max()) <= max(bar, 1e-9 * float(np.abs(fx).max())), i
        assert r.energies_after.total == pytest.approx(g.compute_energy(x), rel=1e-12), i
        assert r.energies_after.total <= r.energies_before.total * (1 + 1e-12), i


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_batched_dense_converges(gpu, prec):
    """Dense networks (excluded-volume pairs, safe-mode entries and recovery moves at the start):
    every system converges by the reference criterion and its result is consistent under FP64
    re-evaluation."""
    nets, ffms = _batch(gpu, 16, dense=True)
    p = abi.default_cg_params(precision=prec)
    xs, fs, res = mech.minimize_batched(ffms, p, [n.coord for n in nets])
    bar = 0.0 if prec == abi.MDN_PREC_FP64 else 1e-3
    print(f"precision {prec}: iterations {[r.num_iter for r in res]}, safe modes {sum(r.num_safe_mode for r in res)}",
          flush=True)
    for i, (g, x, r) in enumerate(zip(ffms, xs, res)):
        assert r.error_flags == 0 and not r.call_limit_exceeded, i
        assert r.converged_grad or r.converged_energy, i
        fx = g.compute_forces(x)
        assert abs(float(np.abs(fx).max()) - r.grad_inf_norm) <= bar + 1e-12 * float(np.abs(fx).max()), i
        assert r.energies_after.total == pytest.approx(g.compute_energy(x), rel=1e-12), i
        assert r.energies_after.total < r.energies_before.total, i


def test_batched_rejects_bad_input(gpu):
    nets, ffms = _batch(gpu, 2)
    p = abi.default_cg_params()
    with pytest.raises(mech.MdnError):
        mech.minimize_batched([ffms[0], ffms[0]], p, [nets[0].coord, nets[0].coord])   # repeated context
    g = gpu()
    g.upload(nets[1].desc, nets[1].coord)   # no neighbor lists yet
    with pytest.raises(mech.MdnError):
        mech.minimize_batched([ffms[0], g], p, [nets[0].coord, nets[1].coord])


def _capped(fn, groups):
    os.environ["MDN_BATCH_MAX_WORKGROUPS"] = str(groups)
    try:
        return fn()
    finally:
        os.environ.pop("MDN_BATCH_MAX_WORKGROUPS", None)


@pytest.mark.parametrize("prec", [abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED])
def test_batched_queue_multiple_pulls(gpu, prec):
    """40 systems on 3 workgroups: each workgroup minimizes about 13 systems one after another.
    Every system's result is the one the uncapped grid (one system per workgroup) gives, bit for
    bit, and the dense systems among them follow the oracle's trajectory prefix."""
    nets, ffms = _batch(gpu, 34)
    dn, dffms = _batch(gpu, 6, dense=True)
    nets, ffms = nets + dn, ffms + dffms
    coords = [n.coord for n in nets]
    p = abi.default_cg_params(precision=prec, energy_call_limit=60)
    xa, fa, ra = mech.minimize_batched(ffms, p, coords)
    xb, fb, rb = _capped(lambda: mech.minimize_batched(ffms, p, coords), 3)
    for i in range(len(nets)):
        assert (ra[i].num_iter, ra[i].num_energy_call, ra[i].num_force_call) == \
            (rb[i].num_iter, rb[i].num_energy_call, rb[i].num_force_call), i
        np.testing.assert_array_equal(xa[i], xb[i])
        np.testing.assert_array_equal(fa[i], fb[i])
        assert ra[i].energies_after.total == rb[i].energies_after.total
    print(f"precision {prec}: 40 systems on 3 workgroups identical to one system per workgroup", flush=True)
    if prec != abi.MDN_PREC_FP64:
        return
    q = abi.default_cg_params(energy_call_limit=22)
    xc, _, rc = _capped(lambda: mech.minimize_batched(dffms, q, [n.coord for n in dn]), 2)
    for i, net in enumerate(dn):
        o = of.OracleSystem(net.desc)
        o.build_nl(net.coord)
        st, xo, fo, ro = o.minimize(q, net.coord)
        assert (rc[i].num_iter, rc[i].num_energy_call, rc[i].num_force_call) == \
            (ro.num_iter, ro.num_energy_call, ro.num_force_call)
        np.testing.assert_allclose(xc[i], xo, rtol=1e-8, atol=0)


def test_batched_from_device_coordinates(gpu):
    """Systems whose coordinates entry is None start from their contexts' device coordinates and
    keep the result there: the same minimizations, bit for bit, as from host buffers."""
    nets, ffms = _batch(gpu, 12, dense=True)
    p = abi.default_cg_params(energy_call_limit=40)
    xh, fh, rh = mech.minimize_batched(ffms, p, [n.coord for n in nets])
    for g, n in zip(ffms, nets):
        g.set_coords(n.coord)
    coords = [None if i % 2 else n.coord for i, n in enumerate(nets)]   # half device-resident
    xd, fd, rd = mech.minimize_batched(ffms, p, coords)
    for i, g in enumerate(ffms):
        x = g.get_coords() if coords[i] is None else xd[i]
        if coords[i] is None:
            assert xd[i] is None
        assert (rh[i].num_iter, rh[i].num_energy_call) == (rd[i].num_iter, rd[i].num_energy_call), i
        np.testing.assert_array_equal(x, xh[i])
        np.testing.assert_array_equal(fd[i], fh[i])
