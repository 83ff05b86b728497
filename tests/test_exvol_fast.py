"""The excluded-volume pair arithmetic checked on the host (no GPU) by
tests/cpp/exvol_fast_check.cpp:

* the gradient formula (picgo_amd/csrc/mdn_exvol_grad.inc): forces of its extended-precision
  instantiation against Richardson-extrapolated central differences of the closed-form energy
  (CylinderExclVolRepulsion.cpp:353-435) in the same precision, on pairs outside the coplanar test;
* rounding: the FP64 path's form (atan differences, no contraction) and the mixed path's form
  (atan2 differences, FMA contraction) against the extended-precision evaluation; the mixed form
  must be as accurate as the FP64 form in every conditioning class, agree with it to rounding on
  well-conditioned pairs, and flag exactly the same degenerate pairs for the Simpson fallbacks."""
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
    # the gradient is the energy's: a dropped term, a wrong sign or a swapped scalar gives errors
    # of order 1e-3 .. 1 (measured 7e-12 at tau >= 0.01)
    n_fd, fd_err = rows["gradient_fd"]
    assert n_fd == 2000 and fd_err <= 1e-9
