#!/usr/bin/env python3
"""Benchmark of the mechanical-equilibration hot path on one MI355X.

Workload: the single large network (2,000 filaments x 20 cylinders = 40,000 cylinders /
42,000 beads in a 5 um cube, 8,000 linkers, 4,000 motors, seeded synthetic coordinates),
neighbor list built once from the start coordinates, as the reference freezes it per
minimization.

A "step" is one force evaluation (ForceFieldManager::computeForces: six force fields, the
gather into the 3 x 42,000 force vector) with the coordinates resident in HBM. `value` is the
mixed-precision path's force-evaluations per second over the K timed steps (FP32 term math for
the bonded, crosslink and boundary force fields, FP64 excluded volume, FP64 accumulation); the
FP64 path is timed beside it, and the max force error of the mixed path against FP64 is
reported. The same JSON line carries the CG minimizations/s of full minimizations from the
seeded start in both precisions, with the reference criterion (max|F| < 1 pN or |dE/E| < 1e-9)
and with the force tolerance alone, the roofline of the force evaluation (its two dependent
launches, the terms kernel and the gather, timed with HIP events on the library's stream) and the
single-core CPU baseline (the oracle built with the reference's release flags, timed on a
bounded sample of the same workload).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-min]
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

HBM_PEAK_GBS = 8000.0   # MI355X HBM3E
WORKLOAD = "single network: 2000 filaments x 20 cylinders (42000 beads), 8000 linkers, 4000 motors, 5 um cube"


def algorithmic_bytes(counts, n_beads, sx=8, sf=8, sp=8, idx=4):
    """Bytes a force evaluation must move at minimum: coordinates in, forces out, each
    interaction table read once: (bead indices, parameters) = stretch (2,2), bend (3,2),
    linker / motor (4,4), excluded volume (4,3), boundary (1,2)."""
    per = {"FilamentStretching": (2, 2), "FilamentBending": (3, 2), "LinkerStretching": (4, 4),
           "MotorGhostStretching": (4, 4), "CylinderExclVolume": (4, 3), "BoundaryCylinderRepulsion": (1, 2)}
    b = 3 * n_beads * (sx + sf)
    for name, (ni, npar) in per.items():
        b += counts[name] * (idx * ni + sp * npar)
    return b


def cpu_baseline(net, target_s=15.0):
    """Oracle (liboracle_fast.so: the reference's release flags) on one core: force
    evaluations of the same 42,000-bead network for about target_s seconds."""
    from oracle import oracle_ffi as of
    o = of.OracleSystem(net.desc, fast=True)
    o.build_nl(net.coord)
    o.forces(net.coord)
    n, t0 = 0, time.perf_counter()
    while True:
        o.forces(net.coord)
        n += 1
        el = time.perf_counter() - t0
        if el >= target_s:
            break
    return {"value": n / el, "unit": "force-evals/s", "cores": 1, "kind": "port",
            "sample": f"{n} force evaluations of the 42,000-bead network ({el:.1f} s, oracle built "
                      f"-O2 -funroll-loops -flto -ffast-math, single thread)"}


def load_traffic(unit):
    p = os.path.join(ROOT, "profiles", f"traffic_{unit}_42k_bead_network.json")
    if os.path.exists(p):
        with open(p) as fh:
            return json.load(fh).get("hbm_bytes_per_launch")
    return None


STRESS = "successive minimizations: 20000 filaments x 20 cylinders (420000 beads), 80000 linkers, 40000 motors, 10 um sphere"
BATCH = "batch: 256 networks, 20-400 filaments x 5-30 cylinders, 2 linkers + 1 motor per filament"


def run_stress(args):
    """Successive minimizations of the 420,000-bead network: before each one the beads get seeded
    N(0, 5 nm) noise and the neighbor lists are rebuilt; the same noise sequence for both
    precisions, drawn before the timed region (it is the workload's input). Everything between
    the first perturbation and the last result is timed: applying the noise, the neighbor-list
    rebuild, the minimization with its start and end energy reports, the coordinates back."""
    net = synth.Network.config(4)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    ffm.reset_neighbor_lists(net.coord)
    rng = np.random.default_rng(4)
    noise = [rng.normal(0.0, 5.0, net.coord.size) for _ in range(args.stress_count)]
    out = {}
    for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
        p = abi.default_cg_params(precision=prec)
        ffm.minimize(abi.default_cg_params(precision=prec, energy_call_limit=20), net.coord)   # warm
        x = net.coord.copy()
        iters, crit, t_nl, pairs = [], {"grad": 0, "energy": 0, "limit": 0}, 0.0, []
        t0 = time.perf_counter()
        for k in range(args.stress_count):
            x = x + noise[k]
            t1 = time.perf_counter()
            pairs.append(ffm.reset_neighbor_lists(x))
            t_nl += time.perf_counter() - t1
            x, f, res = ffm.minimize(p, x, want_forces=False)
            iters.append(res.num_iter)
            crit["grad" if res.converged_grad else ("energy" if res.converged_energy else "limit")] += 1
            if k % 10 == 0:
                print(f"# stress {name}: minimization {k} {res.num_iter} iterations, "
                      f"{time.perf_counter() - t0:.2f} s", file=sys.stderr, flush=True)
        total = time.perf_counter() - t0
        out[name] = {"minimizations_per_s": args.stress_count / total, "total_s": total,
                     "nl_rebuild_ms_mean": 1e3 * t_nl / args.stress_count, "cg_iterations_total": int(sum(iters)),
                     "cg_iterations_first": iters[0], "cg_iterations_median": float(np.median(iters)),
                     "us_per_cg_iteration": 1e6 * (total - t_nl) / max(1, sum(iters)), "criteria": crit,
                     "exvol_pairs_last": int(pairs[-1]), "final_energy_fp64": res.energies_after.total}
    line = {
        "metric": "force-evals/sec & CG minimizations/sec, mixed vs FP64 GPU, same force tol",
        "value": out["mixed"]["minimizations_per_s"], "unit": "minimizations/s",
        "n_gpus": 1, "steps": args.stress_count, "warmup": 1,
        "ms_per_step": 1e3 * out["mixed"]["total_s"] / args.stress_count,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 term math + f64 excluded volume / accumulation",
        "data": "synthetic (seeded generator; seeded N(0, 5 nm) perturbations between minimizations)",
        "config": {"workload": STRESS, "n_beads": net.n_beads, "minimizations": args.stress_count},
        "mixed": out["mixed"], "fp64": out["fp64"],
        "mixed_vs_fp64_minimizations": out["mixed"]["minimizations_per_s"] / out["fp64"]["minimizations_per_s"],
    }
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
    line = {
        "metric": "force-evals/sec & CG minimizations/sec, mixed vs FP64 GPU, same force tol",
        "value": 1.0 / dt, "unit": "minimizations/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1e3 * dt, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == abi.MDN_PREC_FP64 else "f32 term math + f64 excluded volume / accumulation",
        "data": "synthetic (seeded generator)",
        "config": {"workload": f"one minimization of the {net.n_beads}-bead network", "n_beads": net.n_beads},
        "cg_iterations": res.num_iter, "us_per_cg_iteration": 1e6 * dt / max(1, res.num_iter),
        "line_search_launches": st["ls_launches"],
        "criterion": "grad" if res.converged_grad else ("energy" if res.converged_energy else "limit"),
    }
    print(json.dumps(line), flush=True)


def run_batched(args):
    """The 256-network batch in one persistent kernel (mixed and FP64), against the same systems
    minimized one after another by the single-system minimizer (FP64)."""
    nets = [synth.Network.config(3, index=i) for i in range(args.batch_count)]
    ffms = []
    for n in nets:
        g = mech.ForceFieldManager(0)
        g.upload(n.desc, n.coord)
        g.reset_neighbor_lists(n.coord)
        ffms.append(g)
    coords = [n.coord for n in nets]
    out = {}
    for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
        p = abi.default_cg_params(precision=prec)
        mech.minimize_batched(ffms, p, coords)   # schedule tuning + warm-up
        t0 = time.perf_counter()
        xs, fs, res = mech.minimize_batched(ffms, p, coords)
        dt = time.perf_counter() - t0
        nt, srt, _, tune = mech.last_batch_schedule(ffms[0])
        out[name] = {"minimizations_per_s": len(nets) / dt, "batch_s": dt,
                     "converged": sum(1 for r in res if r.converged_grad or r.converged_energy),
                     "cg_iterations_max": max(r.num_iter for r in res),
                     "cg_iterations_total": int(sum(r.num_iter for r in res)),
                     "schedule": {"workgroup": nt, "largest_first": srt, "tuning_ms": tune}}
        print(f"# batched {name}: {dt * 1e3:.1f} ms", file=sys.stderr, flush=True)
    p = abi.default_cg_params()
    t0 = time.perf_counter()
    for g, c in zip(ffms, coords):
        g.minimize(p, c)
    dseq = time.perf_counter() - t0
    out["sequential_fp64"] = {"minimizations_per_s": len(nets) / dseq, "total_s": dseq}
    line = {
        "metric": "force-evals/sec & CG minimizations/sec, mixed vs FP64 GPU, same force tol",
        "value": out["mixed"]["minimizations_per_s"], "unit": "minimizations/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * out["mixed"]["batch_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 term math + f64 excluded volume / accumulation",
        "data": "synthetic (seeded generator)",
        "config": {"workload": BATCH, "systems": len(nets), "n_beads_total": int(sum(n.n_beads for n in nets))},
        "batched_mixed": out["mixed"], "batched_fp64": out["fp64"], "sequential_fp64": out["sequential_fp64"],
        "batched_mixed_vs_sequential_fp64": out["mixed"]["minimizations_per_s"] / out["sequential_fp64"]["minimizations_per_s"],
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-min", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--workload", choices=["single", "stress", "batched", "minimize"], default="single",
                    help="single: the metric's single network (default); stress: 20,000 x 20 filaments in a "
                         "10 um sphere, successive minimizations with 5 nm noise and a neighbor-list rebuild "
                         "each; batched: 256 uneven networks in one persistent kernel; minimize: one "
                         "minimization of the --net network in --prec precision (for profiling)")
    ap.add_argument("--net", type=int, choices=[1, 4], default=1)
    ap.add_argument("--prec", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--stress-count", type=int, default=100)
    ap.add_argument("--batch-count", type=int, default=256)
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
    if args.workload == "stress":
        return run_stress(args)
    if args.workload == "batched":
        return run_batched(args)
    if args.workload == "minimize":
        return run_minimize(args)

    net = synth.Network.config(1)
    ffm = mech.ForceFieldManager(0)
    ffm.upload(net.desc, net.coord)
    n_pairs = ffm.reset_neighbor_lists(net.coord)
    counts = ffm.counts()
    print(f"# {WORKLOAD}; {n_pairs} excluded-volume pairs; counts {counts}", file=sys.stderr, flush=True)
    bytes_alg = algorithmic_bytes(counts, net.n_beads)

    # force evaluations: warm-up, then exactly K timed evaluations per precision (the library
    # times them with HIP events on its stream and synchronizes at the end)
    evals = {}
    for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
        ffm.bench_forces(max(1, args.warmup), prec)
        t0 = time.perf_counter()
        ms_eval, ms_kernel = ffm.bench_forces(args.steps, prec)
        wall = time.perf_counter() - t0
        evals[name] = {"force_evals_per_s": 1000.0 / ms_eval, "us_per_force_eval": 1e3 * ms_eval,
                       "us_terms_kernel": 1e3 * ms_kernel, "wall_s": wall}
        print(f"# {name}: force eval {ms_eval * 1e3:.1f} us (terms kernel {ms_kernel * 1e3:.1f} us), "
              f"{args.steps} steps in {wall:.3f} s wall", file=sys.stderr, flush=True)
    # max force error of the mixed path against FP64 at the same coordinates
    f64 = ffm.compute_forces(net.coord)
    f32 = ffm.compute_forces(net.coord, precision=abi.MDN_PREC_MIXED)
    max_rel_err = float(np.abs(f32 - f64).max() / np.abs(f64).max())

    mins = {}
    if not args.no_min:
        # the reference criterion (max|F| < 1 pN or |dE/E| < 1e-9), and the force tolerance alone
        for crit, ert in (("reference", 1e-9), ("force_only", 0.0)):
            for name, prec in (("fp64", abi.MDN_PREC_FP64), ("mixed", abi.MDN_PREC_MIXED)):
                p = abi.default_cg_params(precision=prec, energy_rel_tol=ert)
                ffm.minimize(abi.default_cg_params(precision=prec, energy_call_limit=50), net.coord)   # warm
                t1 = time.perf_counter()
                x, f, res = ffm.minimize(p, net.coord)
                tmin = time.perf_counter() - t1
                fx = ffm.compute_forces(x)   # FP64 forces at the result
                mins[f"{name}_{crit}"] = {
                    "minimizations_per_s": 1.0 / tmin, "minimization_s": tmin,
                    "us_per_cg_iteration": 1e6 * tmin / max(1, res.num_iter),
                    "cg_iterations": res.num_iter, "energy_calls": res.num_energy_call,
                    "force_calls": res.num_force_call, "safe_mode": res.num_safe_mode,
                    "criterion": "grad" if res.converged_grad else ("energy" if res.converged_energy else "limit"),
                    "final_max_abs_force_pN": res.grad_inf_norm,
                    "fp64_max_abs_force_at_result_pN": float(np.abs(fx).max()),
                    "final_energy_fp64": res.energies_after.total}
                print(f"# minimization {name} ({crit}): {tmin:.3f} s, {res.num_iter} iterations, "
                      f"criterion {mins[f'{name}_{crit}']['criterion']}", file=sys.stderr, flush=True)

    mixed, fp64 = evals["mixed"], evals["fp64"]
    # the roofline unit is one force evaluation (the algorithmic bytes are per evaluation): the
    # terms kernel (slots of the crosslinks, excluded volume and boundary) and the gather (which
    # recomputes the filament terms' forces and sums every bead's forces), back to back
    achieved = bytes_alg / (mixed["us_per_force_eval"] * 1e-6) / 1e9
    line = {
        "metric": "force-evals/sec & CG minimizations/sec, mixed vs FP64 GPU, same force tol",
        "value": mixed["force_evals_per_s"],
        "unit": "force-evals/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": mixed["us_per_force_eval"] / 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 term math + f64 excluded volume / accumulation",
        "data": "synthetic (seeded generator)",
        "config": {"workload": WORKLOAD, "n_beads": net.n_beads, "n_cylinders": net.n_cylinders,
                   "exvol_pairs": n_pairs, "precision_path": "mixed"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": load_traffic("force_eval_mixed"),
                     "kernel": "force evaluation: k_forces_terms<float> + k_gather_plain<float, 64, true>",
                     "algorithmic_bytes": bytes_alg},
        "mixed": mixed, "fp64": fp64,
        "mixed_vs_fp64_force_evals": mixed["force_evals_per_s"] / fp64["force_evals_per_s"],
        "mixed_max_rel_force_error": max_rel_err,
    }
    if mins:
        line["minimization"] = mins
        line["minimizations_per_s"] = mins["mixed_reference"]["minimizations_per_s"]
        line["mixed_vs_fp64_minimizations"] = (mins["mixed_reference"]["minimizations_per_s"] /
                                               mins["fp64_reference"]["minimizations_per_s"])
        line["mixed_vs_fp64_minimizations_force_only"] = (mins["mixed_force_only"]["minimizations_per_s"] /
                                                          mins["fp64_force_only"]["minimizations_per_s"])
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(net, args.cpu_seconds)
        line["cpu_baseline"]["host_nproc"] = os.cpu_count()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
