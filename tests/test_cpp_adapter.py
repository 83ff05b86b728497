"""The header-only C++ adapters (ForceFieldManager, the six ForceField adapters, CGMethod) compile
against the C ABI and drive the path end to end."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "adapter_demo")


def build_demo():
    pk = os.path.join(ROOT, "picgo_amd")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", os.path.join(ROOT, "include"), "-o", EXE,
                    os.path.join(ROOT, "tests", "cpp", "adapter_demo.cpp"), f"-L{pk}", "-lmdn_mech", "-lmdn_synth",
                    f"-Wl,-rpath,{pk}"], check=True)
    return EXE


def test_adapter_compiles_and_links():
    assert os.path.exists(build_demo())


@pytest.mark.gpu
def test_adapter_runs_minimization():
    exe = EXE if os.path.exists(EXE) else build_demo()
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    pairs, e0, e1, iters, mx, conv, nsf, loaded, ff_ok, dev_ok = out.stdout.split()
    assert int(pairs) > 0 and int(conv) == 1
    assert float(e1) < float(e0) and float(mx) < 1.0 and int(iters) > 10
    assert int(nsf) == 20 and int(loaded) > 0   # the 550-bead network: 20 linkers; filament ends near the walls
    assert int(ff_ok) == 1   # the six ForceField adapters add up to the manager's forces and energy
    assert int(dev_ok) == 1  # device-resident minimization = host-buffer one; device noise of unit variance
