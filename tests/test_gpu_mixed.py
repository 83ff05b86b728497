"""GPU: the mixed-precision path (MDN_PREC_MIXED) against the FP64 path on the same inputs.

The mixed path computes the stretching, bending, crosslink and boundary terms in FP32 from bond
vectors formed in FP64, keeps excluded volume in FP64 (the same closed form with four atan2 and
reciprocal force blocks instead of the reference's eight atan and fifty divisions; its accuracy
against extended precision is checked on the host by test_exvol_fast.py) and accumulates in FP64.
Bars (the FP64 path itself is pinned to the oracle by test_gpu_parity.py):
  unrelaxed networks (forces up to 1e14 pN from overlapping cylinders): normwise relative force
    error max|dF| / max|F| <= 1e-6 (the target is 1e-5; measured 2e-8 .. 7e-8);
  total energy: relative 1e-6;
  relaxed network (the 42,000-bead network after an FP64 minimization, max|F| ~ 50 pN, per-term forces much
    larger): absolute force error <= 1e-2 pN, i.e. 1 % of the 1 pN force tolerance, outside the
    beads of excluded-volume pairs with normalized triple product below 1e-3 (measured 1e-4 pN;
    FP32 rounding of the large cancelling bonded terms);
  minimization: stops by the reference criterion with a state whose FP64 forces agree with the
    mixed path's reported max|F| and forces to that absolute bar (outside the beads of
    excluded-volume pairs with normalized triple product below 1e-3, where both FP64 forms of the
    closed form are inaccurate), and lowers the FP64 energy.
"""
import numpy as np
import pytest

from picgo_amd import abi, synth
from tests.exvol_conditioning import well_conditioned

pytestmark = pytest.mark.gpu


def _sys(gpu, cfg):
    net = synth.Network.config(cfg)
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    return net, g


@pytest.mark.parametrize("cfg", [0, 1, 2])
def test_mixed_force_error_unrelaxed(gpu, cfg):
    net, g = _sys(gpu, cfg)
    x = net.coord
    f64 = g.compute_forces(x)
    f32 = g.compute_forces(x, precision=abi.MDN_PREC_MIXED)
    err = float(np.abs(f32 - f64).max() / np.abs(f64).max())
    print(f"config {cfg}: mixed normwise force error {err:.3e} (max|F| {np.abs(f64).max():.3e})", flush=True)
    assert err <= 1e-6
    # the FP32 path really ran: the results are not bit-identical to FP64
    assert not np.array_equal(f32, f64)
    e64 = g.compute_individual_energies(x)
    e32 = g.compute_individual_energies(x, precision=abi.MDN_PREC_MIXED)
    assert abs(e32.total - e64.total) <= 1e-6 * abs(e64.total)
    # excluded volume is FP64 in both paths (the mixed path's reduced-operation form of the same
    # closed form): energies agree to rounding on these unrelaxed networks
    ev32, ev64 = e32.ff[abi.MDN_FF_EXVOL], e64.ff[abi.MDN_FF_EXVOL]
    print(f"config {cfg}: excluded-volume energy rel diff {abs(ev32 - ev64) / abs(ev64):.3e}", flush=True)
    assert abs(ev32 - ev64) <= 1e-12 * abs(ev64)


def test_mixed_force_error_batch_networks(gpu):
    """Config 3's networks (20-400 filaments x 5-30 cylinders), every eighth of the 256: the mixed
    path's force error against FP64. At the generated start states many networks hold almost no
    force (filaments and crosslinks at their rest lengths, no excluded-volume pair or boundary
    contact: max|F| can be 1e-10 pN); where max|F| < 1 pN the error is bounded in pN against the
    1 pN force tolerance, elsewhere normwise. After seeded 5 nm bead noise every network holds
    real forces. In every network the FP32 path really ran."""
    g = gpu()
    rng = np.random.default_rng(3)
    worst_abs, worst_rel, n_rel = 0.0, 0.0, 0
    for i in range(0, 256, 8):
        net = synth.Network.config(3, index=i)
        g.upload(net.desc, net.coord)
        for x in (net.coord, net.coord + 5.0 * rng.standard_normal(net.coord.size)):
            g.reset_neighbor_lists(x)
            f64 = g.compute_forces(x)
            f32 = g.compute_forces(x, precision=abi.MDN_PREC_MIXED)
            d, fmax = float(np.abs(f32 - f64).max()), float(np.abs(f64).max())
            if fmax >= 1.0:
                worst_rel = max(worst_rel, d / fmax)
                n_rel += 1
                assert d <= 1e-6 * fmax, (i, d, fmax)
            else:
                worst_abs = max(worst_abs, d)
                assert d <= 1e-6, (i, d)                    # pN
        assert not np.array_equal(f32, f64), i
    assert n_rel >= 32   # every perturbed state, at least
    print(f"config 3 (32 networks, start and perturbed): worst normwise mixed force error {worst_rel:.3e} "
          f"over {n_rel} states with max|F| >= 1 pN, worst {worst_abs:.3e} pN elsewhere", flush=True)


def test_mixed_force_error_relaxed(gpu):
    net, g = _sys(gpu, 1)
    x, _, res = g.minimize(abi.default_cg_params(), net.coord)
    assert res.converged_grad or res.converged_energy
    f64 = g.compute_forces(x)
    f32 = g.compute_forces(x, precision=abi.MDN_PREC_MIXED)
    # outside the beads of excluded-volume pairs with tau < 1e-3, where the two FP64 forms of the
    # closed form carry rounding errors of their own (test_exvol_fast.py)
    m = well_conditioned(g, net, x)
    dF = np.abs(f32 - f64)
    err = float(dF[m].max())
    print(f"relaxed 42,000-bead network: max|F| {np.abs(f64).max():.3f} pN, mixed abs error {err:.3e} pN outside "
          f"{int((~m).sum())} coordinates on pairs with tau < 1e-3 (max there {float(dF.max()):.3e})", flush=True)
    assert err <= 1e-2 and np.all(np.isfinite(f32))


@pytest.mark.parametrize("cfg", [0, 1])
def test_mixed_minimize(gpu, cfg):
    net, g = _sys(gpu, cfg)
    p = abi.default_cg_params(precision=abi.MDN_PREC_MIXED)
    x, f, res = g.minimize(p, net.coord)
    print(f"config {cfg}: mixed minimization {res.num_iter} iterations, max|F| {res.grad_inf_norm:.4f}, "
          f"grad {res.converged_grad} energy {res.converged_energy}, safe mode {res.num_safe_mode}", flush=True)
    assert res.error_flags == 0
    assert res.converged_grad or res.converged_energy
    assert not res.call_limit_exceeded
    fx = g.compute_forces(x)
    assert abs(float(np.abs(fx).max()) - res.grad_inf_norm) <= 1e-2
    # forces returned are the mixed path's forces at the result: the FP64 path's within 1e-2 pN
    # outside the beads of excluded-volume pairs with tau < 1e-3. On those pairs the closed form
    # loses most of its digits in FP64 whatever the operation order: against extended precision
    # the reference form and the mixed path's form err alike (test_exvol_fast.py; a relative
    # error of order 1 on the worst shifted pairs), so the two only agree to that error there.
    m = well_conditioned(g, net, x)
    dF = np.abs(f - fx)
    print(f"config {cfg}: {int((~m).sum())} coordinates on pairs with tau < 1e-3 (max |dF| "
          f"{float(dF[~m].max()) if (~m).any() else 0.0:.3e} pN), max |dF| elsewhere "
          f"{float(dF[m].max()):.3e} pN", flush=True)
    assert np.all(np.isfinite(f))
    assert float(dF[m].max()) <= 1e-2
    # reported energies are FP64 and consistent with an independent FP64 evaluation
    assert res.energies_after.total == pytest.approx(g.compute_energy(x), rel=1e-12)
    assert res.energies_after.total < res.energies_before.total
    if cfg == 0:   # small network: the force tolerance itself is reached
        assert res.converged_grad and res.grad_inf_norm < 1.0


def test_mixed_minimize_force_tolerance(gpu):
    """Same force tolerance as the FP64 path: with the energy criterion off, the mixed
    minimizer reaches max|F| < 1 pN and FP64 re-evaluation confirms it within the error bar."""
    net, g = _sys(gpu, 1)
    p = abi.default_cg_params(precision=abi.MDN_PREC_MIXED, energy_rel_tol=0.0)
    x, f, res = g.minimize(p, net.coord)
    fx = g.compute_forces(x)
    print(f"42,000-bead network, force-only: {res.num_iter} iterations, mixed max|F| {res.grad_inf_norm:.4f}, "
          f"FP64 max|F| {np.abs(fx).max():.4f}", flush=True)
    assert res.converged_grad and not res.call_limit_exceeded
    assert res.grad_inf_norm < 1.0
    assert float(np.abs(fx).max()) < 1.0 + 1e-2
