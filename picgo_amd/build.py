"""Build the in-tree libraries (run on the CPU host; hipcc cross-compiles for gfx950).

    libmdn_mech.so   the gfx950 HIP kernels + C ABI (include/mdn_mech.h)
    libmdn_synth.so  the seeded synthetic-network generator (include/mdn_synth.h)
    _torch_mech.so   libtorch autograd functions over the C ABI (csrc/torch_mech.cpp)

Usage: python -m picgo_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

# -ffp-contract=off: the FP64 path keeps the reference's rounding (no FMA contraction).
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-Wno-unused-command-line-argument",
]


def _hipcc():
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build_synth(verbose=False):
    out = os.path.join(HERE, "libmdn_synth.so")
    _run(["gcc", "-O2", "-std=c11", "-fPIC", "-shared", "-Wall", "-o", out,
          os.path.join(CSRC, "synth.c"), "-lm"], verbose)
    return out


def build_mech(verbose=False):
    out = os.path.join(HERE, "libmdn_mech.so")
    tmp = out + ".tmp"
    _run([_hipcc()] + HIP_FLAGS + ["-o", tmp, os.path.join(CSRC, "mdn_mech.hip")], verbose)
    os.replace(tmp, out)
    return out


def build_torch_ext(verbose=False):
    """The libtorch extension: host C++ only (it drives the GPU through libmdn_mech.so)."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    out = os.path.join(HERE, "_torch_mech.so")
    tmp = out + ".tmp"
    tlib = os.path.join(os.path.dirname(torch.__file__), "lib")
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-DTORCH_EXTENSION_NAME=_torch_mech", "-DTORCH_API_INCLUDE_EXTENSION_H"]
    for inc in ce.include_paths() + [sysconfig.get_paths()["include"]]:
        cmd += ["-I", inc]
    cmd += ["-o", tmp, os.path.join(CSRC, "torch_mech.cpp"),
            f"-L{tlib}", "-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python", f"-Wl,-rpath,{tlib}",
            f"-L{HERE}", "-lmdn_mech", "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build_all(verbose=False):
    return [build_synth(verbose), build_mech(verbose), build_torch_ext(verbose)]


if __name__ == "__main__":
    for p in build_all(verbose=True):
        print("built", p)
