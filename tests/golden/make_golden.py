"""Regenerate tests/golden/config0_oracle.npz: the CPU oracle's outputs on the seeded
50-filament network (neighbor pairs, boundary interactions, per-force-field forces and energies,
and the state after 60 energy evaluations of the minimizer). The fixture pins the oracle against
regressions; the GPU parity tests compare the device against the live oracle.

Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import oracle_ffi as of  # noqa: E402
from picgo_amd import abi, synth  # noqa: E402


def compute():
    net = synth.Network.config(0)
    o = of.OracleSystem(net.desc)
    o.build_nl(net.coord)
    out = {"coord": net.coord, "pairs": o.pairs(), "boundary_beads": o.boundary_list()}
    for ff in range(abi.MDN_NUM_FF):
        out[f"force_ff{ff}"] = o.ff_forces(ff, net.coord)
    e, rep = o.energy(net.coord)
    out["energy_ff"] = np.array(list(rep.ff))
    out["energy_total"] = np.array([e])
    st, x, f, res = o.minimize(abi.default_cg_params(energy_call_limit=60), net.coord)
    out["min60_coord"] = x
    out["min60_counts"] = np.array([res.num_iter, res.num_energy_call, res.num_force_call])
    return out


if __name__ == "__main__":
    np.savez_compressed(os.path.join(HERE, "config0_oracle.npz"), **compute())
    print("wrote", os.path.join(HERE, "config0_oracle.npz"))
