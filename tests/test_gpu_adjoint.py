"""GPU: parameter gradients dE/dtheta (mdn_param_grad) and Hessian-vector products (mdn_hvp) on
the 42,000-bead network with per-cylinder parameters (lognormal per-cylinder stretching k, rest
length and bending k), near the relaxed state of an FP64 minimization (the unrelaxed network has overlapping
cylinders with forces up to 1e13 pN, where finite differences say nothing).

Central finite differences with h = 1e-4 relative to the perturbed quantity (1e-4 x the mean rest
length for coordinates) pin:
  -dE/dx against the forces per force field (the forces equal the oracle's, the reference
  restatement: bit-identical for the bonded and crosslink terms, to the closed form's
  conditioning for excluded volume, whose force is -dE/dx only outside the coplanar shift);
  H v against -(F(x + h v) - F(x - h v)) / 2h for 64 seeded random unit vectors;
  dE/dtheta against (E(theta + h) - E(theta - h)) / 2h for sampled cylinders and every kind.
Exact identities (no finite-difference error) pin the tangent arithmetic further: H v is linear in
v, and u.Hv = v.Hu up to the coplanar-shift pairs, whose force is not -dE/dx by construction
(CylinderExclVolRepulsion.cpp:610-630). The mixed path's HVP and dE/dtheta (FP32 term arithmetic)
are held to 1e-5 of the FP64 ones (normwise).
"""
import numpy as np
import pytest

from oracle import oracle_ffi as of
from picgo_amd import abi, mech, synth
from tests.exvol_conditioning import exvol_bead_bars as _exvol_bead_bars, well_conditioned as _well_conditioned

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def relaxed(gpu):
    net = synth.Network.config(2)
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    x, _, res = g.minimize(abi.default_cg_params(), net.coord)
    assert res.converged_grad or res.converged_energy
    # the minimizer can stop on the kink of the reference's coplanar test (|n.v| = 0.1, where the
    # excluded-volume force jumps by the 0.01 nm shift): 0.01 nm of seeded noise moves every pair
    # off that surface, so the finite differences below see a smooth force
    x = x + np.random.default_rng(22).normal(0.0, 0.01, x.size)
    g.set_coords(x)
    return net, g, x


def _unit(rng, n):
    v = rng.standard_normal(n)
    return v / np.linalg.norm(v)


def _h(net):
    d = net.desc
    nc = d.fil_bead_offset[d.n_filaments] - d.n_filaments
    return 1e-4 * float(np.mean([d.cyl_eql[i] for i in range(nc)]))


def test_forces_are_minus_energy_gradient(relaxed):
    net, g, x = relaxed
    o = of.OracleSystem(net.desc)
    o.build_nl(net.coord)
    f = g.compute_forces(x)
    fo = o.forces(x)
    # bonded, crosslink and boundary forces are bit-identical; excluded volume agrees to the
    # conditioning of its closed form (nearly coplanar pairs at a relaxed state)
    bars = _exvol_bead_bars(g, net, x, 1e-5 * float(np.abs(fo).max()))
    for ff in range(abi.MDN_NUM_FF):
        fg_ff, fo_ff = g.ff_compute_forces(ff, x), o.ff_forces(ff, x)
        if ff == abi.MDN_FF_EXVOL:
            assert np.all(np.abs(fg_ff - fo_ff) <= bars)
        elif ff != abi.MDN_FF_BOUNDARY:
            np.testing.assert_array_equal(fg_ff, fo_ff)
        else:
            assert float(np.abs(fg_ff - fo_ff).max()) <= 1e-12 * float(np.abs(fo).max())
    assert np.all(np.abs(f - fo) <= bars + 1e-12 * float(np.abs(fo).max()))
    # per force field: -F_ff.v against the central difference of E_ff along v. Excluded volume
    # only to a few per cent: its force is taken at the coplanar-shifted geometry for the pairs
    # within the reference's in-plane test, so it is not -dE/dx there by construction.
    rng = np.random.default_rng(20)
    h = _h(net)
    fff = [g.ff_compute_forces(ff, x) for ff in range(abi.MDN_NUM_FF)]
    for _ in range(4):
        v = _unit(rng, x.size)
        ep = g.compute_individual_energies(x + h * v).ff
        em = g.compute_individual_energies(x - h * v).ff
        for ff in range(abi.MDN_NUM_FF):
            fd = (ep[ff] - em[ff]) / (2 * h)
            an = -float(fff[ff] @ v)
            scale = max(abs(an), 1e-6 * float(np.linalg.norm(fff[ff])))
            print(f"{abi.FF_NAMES[ff]:26s} -F.v {an:+.9e}  FD dE/dv {fd:+.9e}", flush=True)
            bar = 5e-2 if ff == abi.MDN_FF_EXVOL else 1e-5
            assert abs(fd - an) <= bar * scale + 1e-12, abi.FF_NAMES[ff]


def test_hvp_64_vectors_vs_finite_differences(relaxed):
    """The full system: the finite differences are limited by FP64 rounding noise of the
    excluded-volume closed form in near-coplanar pairs (~1e-4 pN on a force at tau ~ 1e-3,
    growing like 1/tau^2, divided by 2h), so the bar is 1e-3 on the coordinates outside the
    pairs with tau < 1e-3 (a few hundred of 42,000 beads); the tangents of every term are pinned
    to 1e-6 .. 1e-5 on well-conditioned geometries by tests/test_dual_tangents.py, and the
    variant below checks the other force fields of this network at 1e-6."""
    net, g, x = relaxed
    rng = np.random.default_rng(2)
    h = _h(net)
    m = _well_conditioned(g, net, x)
    print(f"coordinates outside ill-conditioned pairs: {int(m.sum())} of {m.size}", flush=True)
    worst = 0.0
    for k in range(64):
        v = _unit(rng, x.size)
        hv = g.hvp(v, x)
        fd = -(g.compute_forces(x + h * v) - g.compute_forces(x - h * v)) / (2 * h)
        err = float(np.linalg.norm((hv - fd)[m]) / np.linalg.norm(hv[m]))
        worst = max(worst, err)
        assert err <= 1e-3, k
    print(f"64 HVPs: worst normwise |Hv - FD| / |Hv| = {worst:.3e}", flush=True)


def test_hvp_without_excluded_volume(gpu, relaxed):
    """The same network and state with the excluded-volume stiffness set to zero: stretching,
    bending, linkers, motors and boundary, whose finite differences are clean."""
    net0, _, x = relaxed
    net = synth.Network.config(2)
    d = net.desc
    nc = d.fil_bead_offset[d.n_filaments] - d.n_filaments
    for i in range(nc):
        d.cyl_kexvol[i] = 0.0
    g = gpu()
    g.upload(d, net.coord)
    g.reset_neighbor_lists(net.coord)
    rng = np.random.default_rng(5)
    h = _h(net)
    worst = 0.0
    for k in range(16):
        v = _unit(rng, x.size)
        hv = g.hvp(v, x)
        fd = -(g.compute_forces(x + h * v) - g.compute_forces(x - h * v)) / (2 * h)
        err = float(np.linalg.norm(hv - fd) / np.linalg.norm(hv))
        worst = max(worst, err)
        assert err <= 1e-6, k
    print(f"16 HVPs without excluded volume: worst {worst:.3e}", flush=True)


def test_hvp_linear_and_symmetric(relaxed):
    net, g, x = relaxed
    rng = np.random.default_rng(7)
    u, v = _unit(rng, x.size), _unit(rng, x.size)
    hu, hv = g.hvp(u, x), g.hvp(v, x)
    hw = g.hvp(2.0 * u - 3.0 * v, x)
    # linear up to rounding: 1e-8 of max|Hw| on every coordinate outside the excluded-volume pairs
    # with tau < 1e-3, where the rounding of the closed form's tangent is amplified by a power of
    # 1/tau (measured 3e-5 .. 1.4e-4 absolute on |Hw| ~ 12 in a few dozen of 126,000 components,
    # depending on the relaxed state); their other terms run the same code as everywhere else
    m = _well_conditioned(g, net, x)
    scale = float(np.abs(hw).max())
    lin = hw - (2.0 * hu - 3.0 * hv)
    assert float(np.abs(lin[m]).max()) <= 1e-8 * scale
    print(f"linearity: max |dev| {float(np.abs(lin[m]).max()):.3e} on {int(m.sum())} coordinates, "
          f"{float(np.abs(lin[~m]).max()) if (~m).any() else 0.0:.3e} on the {int((~m).sum())} others", flush=True)
    a, b = float(u @ hv), float(v @ hu)
    print(f"u.Hv {a:.12e} v.Hu {b:.12e}", flush=True)
    assert abs(a - b) <= 1e-3 * max(abs(a), abs(b))


def test_param_grad_vs_finite_differences(relaxed):
    net, g, x = relaxed
    d = net.desc
    nc = d.fil_bead_offset[d.n_filaments] - d.n_filaments
    kstr = np.array([d.cyl_kstr[i] for i in range(nc)])
    eql = np.array([d.cyl_eql[i] for i in range(nc)])
    kb = np.array([d.cyl_kbend[i] for i in range(nc)])
    grad = g.param_grad(x)
    assert grad.shape == (3, nc)
    rng = np.random.default_rng(3)
    # sampled cylinders, plus ones with excluded-volume partners (L0 scales kRepScaled)
    pairs = g.neighbor_pairs()
    sample = list(rng.choice(nc, 8, replace=False)) + [int(pairs[0, 0]), int(pairs[len(pairs) // 2, 1])]
    base = [kstr, eql, kb]
    try:
        for kind in range(3):
            for c in sample:
                th = base[kind][c]
                h = 1e-4 * th
                e = []
                for sgn in (1.0, -1.0):
                    p = [a.copy() for a in base]
                    p[kind][c] = th + sgn * h
                    g.set_params(*p)
                    e.append(g.compute_energy(x))
                fd = (e[0] - e[1]) / (2 * h)
                an = grad[kind, c]
                print(f"kind {kind} cylinder {c}: dE/dtheta {an:.9e} FD {fd:.9e}", flush=True)
                assert abs(fd - an) <= 1e-5 * abs(an) + 1e-6, (kind, c)
    finally:
        g.set_params(kstr, eql, kb)
    # cylinder 0 (the first of filament 0) is no bending term's second cylinder
    assert grad[2, 0] == 0.0


def test_mixed_hvp_and_param_grad(relaxed):
    """The mixed path's Hessian-vector products (FP32 dual arithmetic for the stretching, bending
    and crosslink terms, FP64 excluded volume, boundary and accumulation) and parameter gradients
    against the FP64 ones, on the same 64 seeded unit vectors: normwise max|dHv| / max|Hv| <= 1e-5
    (the FP32 rounding of the term arithmetic), dE/dtheta 1e-5 relative per component (above a
    floor of 1e-6 of the kind's largest value)."""
    net, g, x = relaxed
    rng = np.random.default_rng(64)
    worst = 0.0
    for _ in range(64):
        v = _unit(rng, x.size)
        h64 = g.hvp(v, x)
        h32 = g.hvp(v, x, precision=abi.MDN_PREC_MIXED)
        worst = max(worst, float(np.abs(h32 - h64).max()) / float(np.abs(h64).max()))
    g64 = g.param_grad(x)
    g32 = g.param_grad(x, precision=abi.MDN_PREC_MIXED)
    norm = [float(np.abs(g32[k] - g64[k]).max() / np.abs(g64[k]).max()) for k in range(3)]
    comp = [float((np.abs(g32[k] - g64[k]) / np.abs(g64[k]).clip(1e-300)).max()) for k in (0, 2)]
    print(f"mixed HVP: max normwise relative error {worst:.3e} over 64 vectors; mixed dE/dtheta normwise "
          f"relative error (k, L0, kbend) {norm}, per component (k, kbend) {comp}", flush=True)
    assert 0.0 < worst <= 1e-5
    assert max(norm) <= 1e-5
    # per component: dE/dk = (L - L0)^2 / 2 keeps FP32 relative accuracy; dE/dkbend = 1 - cos(theta)
    # (half the squared chord of the unit bond vectors in FP32, no cancellation) carries the FP32
    # rounding of the bond directions, a relative error of eps / theta (measured 9e-4 on the
    # straightest triples of the relaxed network, where 1 - x in FP32 gave 0.12); dE/dL0 =
    # -k (L - L0) + sum U / L0 can cancel between its two parts, so it is held normwise only
    assert comp[0] <= 1e-6 and comp[1] <= 1e-2


def test_bench_adjoint_times_the_entry_points_work(relaxed):
    """mdn_bench_adjoint times, on the device, the launches that mdn_hvp and mdn_param_grad make
    (tangent terms and gather; pair energies and the per-cylinder kernel): after it, the
    context's results are those of the per-call entry points at the same coordinates."""
    net, g, x = relaxed
    rng = np.random.default_rng(9)
    v = _unit(rng, x.size)
    g.set_coords(x)
    for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
        ms_hvp, ms_grad = g.bench_adjoint(4, v, precision=prec)
        print(f"precision {prec}: {1e3 * ms_hvp:.1f} us per HVP, {1e3 * ms_grad:.1f} us per dE/dtheta", flush=True)
        assert 0.0 < ms_hvp < 100.0 and 0.0 < ms_grad < 100.0
        np.testing.assert_array_equal(g.hvp(v, precision=prec), g.hvp(v, x, precision=prec))
        np.testing.assert_array_equal(g.param_grad(precision=prec), g.param_grad(x, precision=prec))
    with pytest.raises(mech.MdnError) as ei:   # more products than one launch takes
        g.bench_adjoint(1, np.tile(v, (17, 1)))
    assert ei.value.code == abi.MDN_ERR_ARG


def test_hvp_batch_equals_single_products(relaxed):
    """mdn_hvp_batch: 20 products (a chunk of 16 in one tangent launch, then 4) equal mdn_hvp's
    and mdn_hvp_mixed's bit for bit, at given coordinates and at the device's current ones; the
    batched launches as timed by mdn_bench_adjoint leave the same products on the device."""
    net, g, x = relaxed
    rng = np.random.default_rng(12)
    V = np.stack([_unit(rng, x.size) for _ in range(20)])
    g.set_coords(x)
    for prec in (abi.MDN_PREC_FP64, abi.MDN_PREC_MIXED):
        single = np.stack([g.hvp(v, x, precision=prec) for v in V])
        np.testing.assert_array_equal(g.hvp_batch(V, x, precision=prec), single)
        np.testing.assert_array_equal(g.hvp_batch(V[:3], precision=prec), single[:3])
        ms_one, _ = g.bench_adjoint(2, V[0], precision=prec)
        ms_16, _ = g.bench_adjoint(2, V[:16], precision=prec)
        print(f"precision {prec}: {1e3 * ms_one:.1f} us per product alone, {1e3 * ms_16:.1f} us per product "
              f"in a launch of 16", flush=True)
        np.testing.assert_array_equal(g.hvp_batch(V[:1], precision=prec), single[:1])
    with pytest.raises(mech.MdnError) as ei:
        g.hvp_batch(V[:0])
    assert ei.value.code == abi.MDN_ERR_ARG
