#!/usr/bin/env python3
"""Benchmark of the mechanical-equilibration hot path on one MI355X.

Default workload (the stress case): successive minimizations of the 420,000-bead network (20,000
filaments x 20 cylinders = 400,000 cylinders inside a 10 um-radius sphere, 80,000 linkers, 40,000
motors). Before each minimization every bead gets seeded N(0, 5 nm) noise and the neighbor lists
are rebuilt; then CGMethod::minimize runs to the reference criterion (max|F| < 1 pN or
|dE/E| < 1e-9). A "step" is one such minimization: --steps K times K of them (default 100, the
workload's count) after --warmup W untimed ones (the sequence then restarts from the seeded start
state). The coordinates stay in HBM for the whole sequence: each step displaces the device
coordinates by the seeded noise, generated in place (mdn_perturb), rebuilds the lists from them
(mdn_nl_build) and minimizes them in place (mdn_minimize with no host buffer); only the result
statistics cross the host link.

`value` = the mixed-precision path's minimizations/s (FP32 term arithmetic, FP64 excluded volume
and FP64 reductions); the FP64 path runs the same sequence beside it. The line also carries:
  force_only       the first few minimizations of the sequence with the energy criterion off, so
                   both paths stop at the same force tolerance (1 pN);
  roofline         the force evaluation of the 420,000-bead network (the terms kernel and the
                   gather, back to back, timed with HIP events on the library's stream), with the
                   HBM traffic of both kernels from the committed rocprofv3 counter summary;
  roofline_cg_iteration   the algorithmic bytes of a CG iteration over its measured time;
  cpu_baseline     the oracle (built with the reference's release flags) on one host core: a
                   bounded sample of the first minimization, scaled to minimizations/s;
  network_42k      the 42,000-bead network (5 um cube): force evaluations/s and whole
                   minimizations in both precisions;
  network_42k_adjoint  the same network with lognormal per-cylinder parameters near its relaxed
                   state: Hessian-vector products and dE/dtheta, device time per product (HIP
                   events; one per launch and 16 per launch), 64 HVPs through the host-buffer
                   entry points (one call each, and one mdn_hvp_batch call), mixed error against FP64;
  batch_256        the 256-network batch in the persistent batched kernel.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
       python bench.py --workload single|adjoint|batched|minimize ...   (one part alone, for profiling)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from picgo_amd import abi, mech, synth  # noqa: E402

METRIC = "force-evals/sec & CG minimizations/sec, mixed vs FP64 GPU, same force tol"
HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
STRESS = ("successive minimizations: 20000 filaments x 20 cylinders (420000 beads) in a 10 um-radius sphere, "
          "80000 linkers, 40000 motors; seeded 5 nm bead noise and a neighbor-list rebuild before each")
SINGLE = "single network: 2000 filaments x 20 cylinders (42000 beads), 8000 linkers, 4000 motors, 5 um cube"
BATCH = "batch: 256 networks, 20-400 filaments x 5-30 cylinders, 2 linkers + 1 motor per filament"
DTYPE = "f32 term math + f64 excluded volume / reductions"

TABLES = {"FilamentStretching": (2, 2), "FilamentBending": (3, 2), "LinkerStretching": (4, 4),
          "MotorGhostStretching": (4, 4), "CylinderExclVolume": (4, 3), "BoundaryCylinderRepulsion": (1, 2)}


def table_bytes(counts, idx=4, sp=8):
    """Each interaction table read once: (bead indices, parameters) per interaction."""
    return sum(counts[name] * (idx * ni + sp * npar) for name, (ni, npar) in TABLES.items())


def force_eval_bytes(counts, n_beads):
    """Algorithmic bytes of a force evaluation: coordinates in, forces out, the tables."""
    return 3 * n_beads * (8 + 8) + table_bytes(counts)


def energy_eval_bytes(counts, n_beads):
    return 3 * n_beads * 8 + table_bytes(counts)


def cg_iteration_bytes(counts, n_beads, energies_per_iter):
    """A CG iteration: one force evaluation, its energy evaluations, and the vector passes
    (read x, d, F, F_prev; write x, d, F_prev)."""
    return (force_eval_bytes(counts, n_beads) + energies_per_iter * energy_eval_bytes(counts, n_beads)
            + 7 * 3 * n_beads * 8)


def load_traffic(unit, net_name):
    p = os.path.join(ROOT, "profiles", f"traffic_{unit}_{net_name}.json")
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    return None


def criterion(res):
    return "grad" if res.converged_grad else ("energy" if res.converged_energy else "limit")


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------------------------
# the stress sequence (default)
# ---------------------------------------------------------------------------------------------

NOISE_SEED, NOISE_SIGMA = 4, 5.0


def run_sequence(ffm, net, steps, params):
    """The first `steps` minimizations of the sequence from the seeded start state, with the
    coordinates resident in HBM: minimization k starts from the previous result displaced by the
    seeded N(0, 5 nm) noise of stream k (generated in place on the device, mdn_perturb), with the
    neighbor lists rebuilt there. Returns per-minimization statistics and the wall times."""
    ffm.set_coords(net.coord)
    stats, t_nl = [], 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        ffm.perturb(NOISE_SEED, k, NOISE_SIGMA)
        t1 = time.perf_counter()
        n_pairs = ffm.reset_neighbor_lists(None)
        t_nl += time.perf_counter() - t1
        _, _, res = ffm.minimize(params, None, want_forces=False)
        if res.error_flags or not (res.converged_grad or res.converged_energy or res.call_limit_exceeded):
            raise RuntimeError(f"minimization {k} failed: flags {res.error_flags}")
        stats.append((res.num_iter, res.num_energy_call, res.num_force_call, criterion(res), n_pairs,
                      res.grad_inf_norm, res.energies_after.total))
    total = time.perf_counter() - t0
    return stats, total, t_nl


def summarize(stats, total, t_nl, steps):
    iters = [s[0] for s in stats]
    crit = {"grad": 0, "energy": 0, "limit": 0}
    for s in stats:
        crit[s[3]] += 1
    return {"minimizations_per_s": steps / total, "total_s": total, "ms_per_minimization": 1e3 * total / steps,
            "nl_rebuild_ms_mean": 1e3 * t_nl / steps, "cg_iterations_total": int(sum(iters)),
            "cg_iterations_mean": float(np.mean(iters)), "cg_iterations_median": float(np.median(iters)),
            "energy_calls_total": int(sum(s[1] for s in stats)), "force_calls_total": int(sum(s[2] for s in stats)),
            "us_per_cg_iteration": 1e6 * (total - t_nl) / max(1, sum(iters)), "criteria": crit,
            "exvol_pairs_last": int(stats[-1][4]), "final_max_abs_force_pN": float(stats[-1][5]),
            "final_energy_fp64": float(stats[-1][6])}


def precisions(args):
    both = (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED))
    return [b for b in both if args.stress_prec in ("both", b[0])]


def stress_part(args, net, ffm):
    out = {}
    for name, prec in precisions(args):
        p = abi.default_cg_params(precision=prec)
        if args.warmup > 0:   # untimed: the first W minimizations of the sequence
            run_sequence(ffm, net, args.warmup, p)
        stats, total, t_nl = run_sequence(ffm, net, args.steps, p)
        out[name] = summarize(stats, total, t_nl, args.steps)
        log(f"stress {name}: {args.steps} minimizations in {total:.3f} s ({out[name]['minimizations_per_s']:.2f}/s), "
            f"{out[name]['cg_iterations_total']} CG iterations, criteria {out[name]['criteria']}")
    fo = {}
    if args.force_only > 0:
        for name, prec in precisions(args):
            p = abi.default_cg_params(precision=prec, energy_rel_tol=0.0)
            stats, total, t_nl = run_sequence(ffm, net, args.force_only, p)
            fo[name] = summarize(stats, total, t_nl, args.force_only)
            log(f"force-only {name}: {args.force_only} minimizations in {total:.3f} s, "
                f"{fo[name]['cg_iterations_total']} CG iterations, criteria {fo[name]['criteria']}")
        if len(fo) == 2:
            fo["mixed_vs_fp64_minimizations"] = fo["mixed"]["minimizations_per_s"] / fo["fp64"]["minimizations_per_s"]
        fo["minimizations"] = args.force_only
    return out, fo


def force_evals(ffm, steps, warmup):
    ev = {}
    for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
        ffm.bench_forces(max(1, warmup), prec)
        ms_eval, ms_kernel = ffm.bench_forces(steps, prec)
        ev[name] = {"force_evals_per_s": 1000.0 / ms_eval, "us_per_force_eval": 1e3 * ms_eval,
                    "us_terms_kernel": 1e3 * ms_kernel}
    ev["mixed_vs_fp64_force_evals"] = ev["mixed"]["force_evals_per_s"] / ev["fp64"]["force_evals_per_s"]
    return ev


def cpu_baseline_stress(net, noise_host0, fp64_iters_mean, target_s):
    """The oracle (liboracle_fast.so, the reference's release flags) on one core: the first
    minimization of the sequence, bounded by an energy-call limit sized to about target_s, then
    scaled to whole minimizations with the FP64 GPU path's mean iteration count (the same
    algorithm to the same criterion)."""
    from oracle import oracle_ffi as of
    o = of.OracleSystem(net.desc, fast=True)
    x = net.coord + noise_host0
    o.build_nl(x)
    t0 = time.perf_counter()
    o.forces(x)
    t_force = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.energy(x)
    t_energy = time.perf_counter() - t0
    limit = max(10, int(target_s / (t_force + 2.0 * t_energy) * 2.0))
    t0 = time.perf_counter()
    st, _, _, res = o.minimize(abi.default_cg_params(energy_call_limit=limit), x)
    el = time.perf_counter() - t0
    s_per_iter = el / max(1, res.num_iter)
    return {"value": 1.0 / (s_per_iter * fp64_iters_mean), "unit": "minimizations/s", "cores": 1, "kind": "port",
            "sample": (f"first minimization of the 420,000-bead sequence stopped after {res.num_energy_call} energy "
                       f"calls ({res.num_iter} CG iterations, {el:.1f} s, single thread, oracle built -O2 "
                       f"-funroll-loops -flto -ffast-math): {1e3 * s_per_iter:.1f} ms per CG iteration, scaled by "
                       f"the {fp64_iters_mean:.0f} iterations per minimization of the FP64 GPU path"),
            "force_evals_per_s": 1.0 / t_force, "energy_evals_per_s": 1.0 / t_energy,
            "host_nproc": os.cpu_count()}


# ---------------------------------------------------------------------------------------------
# the 42,000-bead network and the 256-network batch
# ---------------------------------------------------------------------------------------------

def single_network(args):
    net = synth.Network.config(1)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    n_pairs = ffm.reset_neighbor_lists(net.coord)
    counts = ffm.counts()
    ev = force_evals(ffm, 200, 20)
    f64 = ffm.compute_forces(net.coord)
    f32 = ffm.compute_forces(net.coord, precision=abi.MDN_PREC_MIXED)
    out = {"workload": SINGLE, "n_beads": net.n_beads, "exvol_pairs": n_pairs, "force_eval": ev,
           "mixed_max_rel_force_error": float(np.abs(f32 - f64).max() / np.abs(f64).max())}
    B = force_eval_bytes(counts, net.n_beads)
    out["force_eval_roofline_frac"] = B / (ev["mixed"]["us_per_force_eval"] * 1e-6) / 1e9 / HBM_PEAK_GBS
    mins = {}
    for crit, ert in (("reference", 1e-9), ("force_only", 0.0)):
        for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
            ffm.minimize(abi.default_cg_params(precision=prec, energy_call_limit=50), net.coord)   # warm
            t1 = time.perf_counter()
            x, f, res = ffm.minimize(abi.default_cg_params(precision=prec, energy_rel_tol=ert), net.coord)
            tmin = time.perf_counter() - t1
            mins[f"{name}_{crit}"] = {
                "minimizations_per_s": 1.0 / tmin, "us_per_cg_iteration": 1e6 * tmin / max(1, res.num_iter),
                "cg_iterations": res.num_iter, "energy_calls": res.num_energy_call, "criterion": criterion(res),
                "final_max_abs_force_pN": res.grad_inf_norm,
                "fp64_max_abs_force_at_result_pN": float(np.abs(ffm.compute_forces(x)).max())}
    out["minimization"] = mins
    out["mixed_vs_fp64_minimizations"] = (mins["mixed_reference"]["minimizations_per_s"] /
                                          mins["fp64_reference"]["minimizations_per_s"])
    out["mixed_vs_fp64_minimizations_force_only"] = (mins["mixed_force_only"]["minimizations_per_s"] /
                                                     mins["fp64_force_only"]["minimizations_per_s"])
    log(f"42k network: force eval {ev['mixed']['us_per_force_eval']:.1f} us mixed / "
        f"{ev['fp64']['us_per_force_eval']:.1f} us FP64")
    ffm.close()
    return out


def adjoint_42k(args):
    """dE/dtheta and Hessian-vector products on the 42,000-bead network with lognormal
    per-cylinder parameters, near its relaxed state (an FP64 minimization, then 0.01 nm of seeded
    noise, which moves every pair off the coplanar test's kink): 64 seeded unit vectors."""
    net = synth.Network.config(2)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    ffm.reset_neighbor_lists(net.coord)
    x, _, res = ffm.minimize(abi.default_cg_params(), net.coord, want_forces=False)
    x = x + np.random.default_rng(22).normal(0.0, 0.01, x.size)
    ffm.set_coords(x)
    rng = np.random.default_rng(64)
    vs = rng.standard_normal((64, x.size))
    vs /= np.linalg.norm(vs, axis=1, keepdims=True)
    counts = ffm.counts()
    B_hvp = force_eval_bytes(counts, net.n_beads) + 24 * net.n_beads   # + the tangent vector in
    out = {"workload": "adjoint: " + SINGLE + "; lognormal per-cylinder parameters; 64 seeded unit vectors",
           "n_beads": net.n_beads, "relax_cg_iterations": res.num_iter}
    hv = {}
    for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
        ffm.bench_adjoint(8, vs[0], prec)
        ms_hvp, ms_grad = ffm.bench_adjoint(64, vs[0], prec)
        ffm.bench_adjoint(2, vs[:16], prec)
        ms_hvp16, _ = ffm.bench_adjoint(4, vs[:16], prec)   # 4 launches of 16: the 64 vectors' device work
        ffm.hvp(vs[0], precision=prec)
        t0 = time.perf_counter()
        hv[name] = np.stack([ffm.hvp(v, precision=prec) for v in vs])
        t_host = time.perf_counter() - t0
        ffm.hvp_batch(vs[:16], precision=prec)
        t0 = time.perf_counter()
        hb = ffm.hvp_batch(vs, precision=prec)
        t_batch = time.perf_counter() - t0
        batch_same = bool(np.array_equal(hb, hv[name]))
        ffm.param_grad(precision=prec)
        t0 = time.perf_counter()
        for _ in range(16):
            ffm.param_grad(precision=prec)
        t_grad_host = (time.perf_counter() - t0) / 16
        out[name] = {"us_per_hvp_device": 1e3 * ms_hvp, "us_per_param_grad_device": 1e3 * ms_grad,
                     "ms_64_hvps_device": 64 * ms_hvp,
                     "us_per_hvp_device_16_per_launch": 1e3 * ms_hvp16,
                     "ms_64_hvps_device_16_per_launch": 64 * ms_hvp16,
                     "ms_64_hvps_host_buffers": 1e3 * t_host,
                     "ms_64_hvps_host_buffers_one_call": 1e3 * t_batch,
                     "hvp_batch_equals_single_calls": batch_same,
                     "us_per_param_grad_host_buffers": 1e6 * t_grad_host,
                     "hvp_roofline_frac": B_hvp / (ms_hvp * 1e-3) / 1e9 / HBM_PEAK_GBS}
    err = np.abs(hv["mixed"] - hv["fp64"]).max(axis=1) / np.abs(hv["fp64"]).max(axis=1)
    g64, g32 = ffm.param_grad(), ffm.param_grad(precision=abi.MDN_PREC_MIXED)
    out["mixed_max_rel_hvp_error"] = float(err.max())
    out["mixed_param_grad_rel_error"] = [float(np.abs(g32[k] - g64[k]).max() / np.abs(g64[k]).max())
                                         for k in range(3)]
    out["hvp_algorithmic_bytes"] = B_hvp
    log(f"42k adjoint: HVP {out['mixed']['us_per_hvp_device']:.1f} us mixed / "
        f"{out['fp64']['us_per_hvp_device']:.1f} us FP64 ({out['fp64']['us_per_hvp_device_16_per_launch']:.1f} us "
        f"16 per launch), dE/dtheta {out['fp64']['us_per_param_grad_device']:.1f} us")
    ffm.close()
    return out


def batch_256(args):
    nets = [synth.Network.config(3, index=i) for i in range(args.batch_count)]
    ffms = []
    for n in nets:
        g = mech.ForceFieldManager(0)
        g.upload(n.desc, n.coord)
        g.reset_neighbor_lists(n.coord)
        ffms.append(g)
    coords = [n.coord for n in nets]
    out = {"workload": BATCH, "systems": len(nets), "n_beads_total": int(sum(n.n_beads for n in nets))}
    # the mixed path's force error against FP64 at every network's start state: the worst in pN,
    # and the worst normwise one over the networks that hold forces of 1 pN or more (many start at
    # rest, with forces of 1e-10 pN, where a relative error says nothing)
    err_abs, err_rel, n_rel = 0.0, 0.0, 0
    for g, c in zip(ffms, coords):
        f64 = g.compute_forces(c)
        f32 = g.compute_forces(c, precision=abi.MDN_PREC_MIXED)
        d, fmax = float(np.abs(f32 - f64).max()), float(np.abs(f64).max())
        if fmax >= 1.0:
            err_rel = max(err_rel, d / fmax)
            n_rel += 1
        else:
            err_abs = max(err_abs, d)
    out["mixed_max_rel_force_error"] = {"value": err_rel, "networks": n_rel,
                                        "over": "networks whose start state holds max|F| >= 1 pN"}
    out["mixed_max_force_error_pN_at_rest"] = {"value": err_abs, "networks": len(nets) - n_rel,
                                               "over": "networks whose start state holds max|F| < 1 pN"}
    for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
        p = abi.default_cg_params(precision=prec)
        mech.minimize_batched(ffms, p, coords)   # schedule tuning + warm-up
        for g, c in zip(ffms, coords):   # the start coordinates resident in HBM before the timed call
            g.set_coords(c)
        t0 = time.perf_counter()
        _, _, res = mech.minimize_batched(ffms, p, None, want_forces=False)
        dt = time.perf_counter() - t0
        nt, srt, _, tune = mech.last_batch_schedule(ffms[0])
        # the network with the most work (CG iterations x beads) alone, after its own schedule tuning
        heavy = int(np.argmax([r.num_iter * n.n_beads for r, n in zip(res, nets)]))
        mech.minimize_batched([ffms[heavy]], p, [coords[heavy]], want_forces=False)
        ffms[heavy].set_coords(coords[heavy])
        t1 = time.perf_counter()
        mech.minimize_batched([ffms[heavy]], p, None, want_forces=False)
        solo = time.perf_counter() - t1
        out[name] = {"minimizations_per_s": len(nets) / dt, "batch_ms": 1e3 * dt,
                     "heaviest_network": {"beads": nets[heavy].n_beads, "cg_iterations": res[heavy].num_iter},
                     "heaviest_network_alone_ms": 1e3 * solo, "batch_over_heaviest_alone": dt / solo,
                     "converged": sum(1 for r in res if r.converged_grad or r.converged_energy),
                     "cg_iterations_max": max(r.num_iter for r in res),
                     "schedule": {"workgroup": nt, "largest_first": srt, "tuning_ms": tune}}
        log(f"batch {name}: {dt * 1e3:.1f} ms (the heaviest network alone {solo * 1e3:.1f} ms)")
    p = abi.default_cg_params()
    t0 = time.perf_counter()
    for g, c in zip(ffms, coords):
        g.minimize(p, c, want_forces=False)
    dseq = time.perf_counter() - t0
    out["sequential_fp64_minimizations_per_s"] = len(nets) / dseq
    out["batched_mixed_vs_sequential_fp64"] = out["mixed"]["minimizations_per_s"] / out["sequential_fp64_minimizations_per_s"]
    for g in ffms:
        g.close()
    return out


def base_line(value, unit, steps, warmup, ms_per_step, config):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": DTYPE, "data": "synthetic (seeded generator)", "config": config}


def run_stress(args):
    net = synth.Network.config(4)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    n_pairs0 = ffm.reset_neighbor_lists(net.coord)
    counts = ffm.counts()
    log(f"420k network: {n_pairs0} excluded-volume pairs at the start; counts {counts}")
    # the force evaluation at the start state (its roofline)
    if args.stress_prec == "both":
        ev = force_evals(ffm, 200, 20)
        log(f"force eval {ev['mixed']['us_per_force_eval']:.1f} us mixed / {ev['fp64']['us_per_force_eval']:.1f} us FP64")
    seq, fo = stress_part(args, net, ffm)
    if args.stress_prec != "both":   # one precision alone (profiling): its statistics only
        print(json.dumps({"workload": STRESS, "precision": args.stress_prec, **seq[args.stress_prec]}), flush=True)
        return
    mixed, fp64 = seq["mixed"], seq["fp64"]

    B_F = force_eval_bytes(counts, net.n_beads)
    achieved = B_F / (ev["mixed"]["us_per_force_eval"] * 1e-6) / 1e9
    n_e = mixed["energy_calls_total"] / max(1, mixed["cg_iterations_total"])
    B_it = cg_iteration_bytes(counts, net.n_beads, n_e)
    ach_it = B_it / (mixed["us_per_cg_iteration"] * 1e-6) / 1e9
    line = base_line(mixed["minimizations_per_s"], "minimizations/s", args.steps, args.warmup,
                     mixed["ms_per_minimization"],
                     {"workload": f"{STRESS}; {args.steps} minimizations timed", "n_beads": net.n_beads,
                      "n_cylinders": net.n_cylinders, "minimizations": args.steps, "precision_path": "mixed"})
    line["data"] = "synthetic (seeded generator; seeded N(0, 5 nm) bead noise generated in place in HBM)"
    line["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": load_traffic("force_eval_mixed", "420k_bead_network"),
                        "kernel": "force evaluation of the 420,000-bead network: k_forces_terms<float> + "
                                  "k_gather_plain<float, 128, true>",
                        "algorithmic_bytes": B_F}
    line["roofline_cg_iteration"] = {"bound": "hbm", "achieved": ach_it, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                     "frac": ach_it / HBM_PEAK_GBS, "algorithmic_bytes": B_it,
                                     "energy_evals_per_iteration": n_e,
                                     "us_per_cg_iteration": mixed["us_per_cg_iteration"]}
    line["mixed"], line["fp64"] = mixed, fp64
    line["mixed_vs_fp64_minimizations"] = mixed["minimizations_per_s"] / fp64["minimizations_per_s"]
    line["mixed_vs_fp64_us_per_cg_iteration"] = fp64["us_per_cg_iteration"] / mixed["us_per_cg_iteration"]
    if fo:
        line["force_only"] = fo
    line["force_eval"] = ev
    f64 = ffm.compute_forces(net.coord)
    f32 = ffm.compute_forces(net.coord, precision=abi.MDN_PREC_MIXED)
    line["mixed_max_rel_force_error"] = float(np.abs(f32 - f64).max() / np.abs(f64).max())
    ffm.close()
    if not args.no_extras:
        line["network_42k"] = single_network(args)
        line["network_42k_adjoint"] = adjoint_42k(args)
        line["batch_256"] = batch_256(args)
    if not args.no_cpu:
        first = mech.perturbation(NOISE_SEED, 0, NOISE_SIGMA, 3 * net.n_beads)
        line["cpu_baseline"] = cpu_baseline_stress(net, first, fp64["cg_iterations_mean"], args.cpu_seconds)
    print(json.dumps(line), flush=True)


# ---------------------------------------------------------------------------------------------
# single parts (profiling)
# ---------------------------------------------------------------------------------------------

def run_single(args):
    out = single_network(args)
    line = base_line(out["force_eval"]["mixed"]["force_evals_per_s"], "force-evals/s", 200, 20,
                     out["force_eval"]["mixed"]["us_per_force_eval"] / 1e3, {"workload": SINGLE})
    line.update(out)
    print(json.dumps(line), flush=True)


def run_batched(args):
    out = batch_256(args)
    line = base_line(out["mixed"]["minimizations_per_s"], "minimizations/s", 1, 1, out["mixed"]["batch_ms"],
                     {"workload": BATCH})
    line.update(out)
    print(json.dumps(line), flush=True)


def run_adjoint(args):
    out = adjoint_42k(args)
    line = base_line(1e6 / out["mixed"]["us_per_hvp_device"], "HVPs/s", 64, 8,
                     out["mixed"]["us_per_hvp_device"] / 1e3, {"workload": out["workload"]})
    line.update(out)
    print(json.dumps(line), flush=True)


def run_minimize(args):
    """One full minimization (reference criterion) of the 42,000- or 420,000-bead network from the
    seeded start, after a 50-energy-call warm-up: the unit the CG kernel profiles are taken on."""
    net = synth.Network.config(args.net)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    ffm.reset_neighbor_lists(net.coord)
    prec = abi.MDN_PREC_MIXED if args.prec == "mixed" else abi.MDN_PREC_FP64
    ffm.minimize(abi.default_cg_params(precision=prec, energy_call_limit=50), net.coord)
    t0 = time.perf_counter()
    x, f, res = ffm.minimize(abi.default_cg_params(precision=prec), net.coord, want_forces=False)
    dt = time.perf_counter() - t0
    st = ffm.last_min_stats()
    line = base_line(1.0 / dt, "minimizations/s", 1, 1, 1e3 * dt,
                     {"workload": f"one minimization of the {net.n_beads}-bead network", "n_beads": net.n_beads})
    line.update({"cg_iterations": res.num_iter, "us_per_cg_iteration": 1e6 * dt / max(1, res.num_iter),
                 "line_search_launches": st["ls_launches"], "criterion": criterion(res)})
    print(json.dumps(line), flush=True)


def run_force_evals(args):
    """Force evaluations of the --net network in --prec precision (the unit of the force-kernel
    profiles): 20 warm-up evaluations, then --steps timed ones."""
    net = synth.Network.config(args.net)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    ffm.reset_neighbor_lists(net.coord)
    prec = abi.MDN_PREC_MIXED if args.prec == "mixed" else abi.MDN_PREC_FP64
    ffm.bench_forces(20, prec)
    ms, ms_k = ffm.bench_forces(args.steps, prec)
    print(json.dumps({"workload": f"force evaluations of the {net.n_beads}-bead network", "precision": args.prec,
                      "us_per_force_eval": 1e3 * ms, "us_terms_kernel": 1e3 * ms_k, "evaluations": args.steps}),
          flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="minimizations of the stress sequence timed")
    ap.add_argument("--warmup", type=int, default=2, help="untimed minimizations before the timed sequence")
    ap.add_argument("--force-only", type=int, default=3,
                    help="minimizations of the sequence also run with the energy criterion off")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="skip the 42,000-bead network and the 256-network batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["stress", "single", "adjoint", "batched", "minimize", "forces"], default="stress")
    ap.add_argument("--net", type=int, choices=[1, 4], default=4)
    ap.add_argument("--prec", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--stress-prec", choices=["both", "fp64", "mixed"], default="both",
                    help="run the stress sequence in one precision only and print its statistics (profiling)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run 20 device copies of 64 MiB (FETCH_SIZE / WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.calib_copy:
        import torch
        a = torch.ones(8 << 20, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(20):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calib_copy_bytes": a.numel() * 8, "launches": 20}))
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of this path is a design note, not a measured configuration.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be at least 1")
    {"stress": run_stress, "single": run_single, "adjoint": run_adjoint, "batched": run_batched, "minimize": run_minimize,
     "forces": run_force_evals}[args.workload](args)


if __name__ == "__main__":
    main()
