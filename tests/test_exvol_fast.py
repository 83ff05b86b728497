"""The mixed path's excluded-volume form (atan2 differences, reciprocal force blocks; see
exvol_*_fast in picgo_amd/csrc/mdn_terms.h) checked on the host (no GPU) against the FP64 path's
reference-order form of the same closed form (CylinderExclVolRepulsion.cpp:330-858), and both
against that form evaluated in x87 extended precision (tests/cpp/exvol_fast_check.cpp): the
mixed form must be as accurate as the reference form in every conditioning class, agree with it
to rounding on well-conditioned pairs, and flag exactly the same degenerate pairs for the
Simpson fallbacks."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "exvol_fast_check.cpp")
EXE = os.path.join(ROOT, "tests", "cpp", "exvol_fast_check")


def test_exvol_fast_form_matches_reference_form():
    hipcc = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    subprocess.run([hipcc, "--offload-host-only", "-O2", "-std=c++17", "-ffp-contract=off", "-o", EXE, SRC],
                   check=True, capture_output=True)
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=120, check=True).stdout
    rows = {}
    for line in out.splitlines():
        f = line.split()
        rows[f[0]] = [int(f[1])] + [float(v) for v in f[2:]]
    print(out)
    assert rows["flags"][1] == 0 and rows["flags"][0] > 100000
    assert rows["degenerate"] == [600, 0]
    need = {"tau_ge_0.2": 4000, "tau_0.01_0.2": 4000, "tau_1e-3_0.01": 500, "coplanar_shifted": 2000}
    for name, n_min in need.items():
        n, e_fr, f_fr, e99_r, emax_r, e99_f, emax_f, f99_r, fmax_r, f99_f, fmax_f = rows[name]
        assert n >= n_min, name
        # as accurate as the reference form against extended precision: the 99th percentile
        # within 1.5x, the worst case (a few hundred samples in the smallest class) within 3x
        assert e99_f <= 1.5 * e99_r + 1e-15 and emax_f <= 3 * emax_r + 1e-15, (name, "energy")
        assert f99_f <= 1.5 * f99_r + 1e-15 and fmax_f <= 3 * fmax_r + 1e-15, (name, "forces")
    # well-conditioned pairs: the two FP64 forms agree to rounding
    assert rows["tau_ge_0.2"][1] <= 1e-13 and rows["tau_ge_0.2"][2] <= 1e-12
