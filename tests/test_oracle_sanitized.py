"""The CPU oracle (test infrastructure) and the seeded network generator built with
AddressSanitizer and UndefinedBehaviorSanitizer and run over their main paths
(tests/cpp/oracle_sanitized.c): neighbor lists (brute force and cells), energies and forces per
force field, a minimization with safe mode and recovery moves, boundary load forces and the
Simpson fallbacks. Any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = [os.path.join(ROOT, "tests", "cpp", "oracle_sanitized.c")] + \
    [os.path.join(ROOT, "oracle", f) for f in ("ff.c", "nl.c", "system.c", "cg.c")] + \
    [os.path.join(ROOT, "picgo_amd", "csrc", "synth.c")]
EXE = os.path.join(ROOT, "tests", "cpp", "oracle_sanitized")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_under_sanitizers():
    flags = ["-O1", "-g", "-std=c11", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
             "-fno-sanitize-recover=undefined", "-ffp-contract=off"]
    subprocess.run(["gcc"] + flags + ["-o", EXE] + SRC + ["-lm"], check=True, capture_output=True)
    # verify_asan_link_order=0: the sanitizer runtime need not be the first library loaded
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=env)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert r.stdout.strip().startswith("ok 3")
