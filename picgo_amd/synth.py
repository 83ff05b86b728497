"""Seeded synthetic actin networks (binding of libmdn_synth.so, include/mdn_synth.h)."""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libmdn_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.mdn_synth_preset.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.POINTER(abi.SynthParams)]
        L.mdn_synth_create.argtypes = [C.POINTER(abi.SynthParams), C.POINTER(C.c_void_p)]
        L.mdn_synth_destroy.argtypes = [C.c_void_p]
        L.mdn_synth_desc.argtypes = [C.c_void_p]
        L.mdn_synth_desc.restype = C.POINTER(abi.SystemDesc)
        L.mdn_synth_coord.argtypes = [C.c_void_p]
        L.mdn_synth_coord.restype = C.POINTER(C.c_double)
        L.mdn_synth_perturb.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_double]
        _lib = L
    return _lib


def preset(config, index=0, dense=False, **overrides):
    p = abi.SynthParams()
    st = lib().mdn_synth_preset(config, index, 1 if dense else 0, C.byref(p))
    if st != 0:
        raise ValueError(f"unknown config {config}")
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


class Network:
    """A generated system: ``desc`` (abi.SystemDesc pointer, valid while the object lives)
    and ``coord`` (numpy float64, 3*n_beads, a private copy)."""

    def __init__(self, params):
        self._h = C.c_void_p()
        st = lib().mdn_synth_create(C.byref(params), C.byref(self._h))
        if st != 0:
            raise RuntimeError(f"mdn_synth_create failed: {abi.status_name(st)}")
        self.params = params
        self.desc_ptr = lib().mdn_synth_desc(self._h)
        self.desc = self.desc_ptr.contents
        n = self.desc.n_beads
        self.coord = np.ctypeslib.as_array(lib().mdn_synth_coord(self._h), shape=(3 * n,)).copy()

    @classmethod
    def config(cls, config, index=0, dense=False, **overrides):
        return cls(preset(config, index, dense, **overrides))

    @property
    def n_beads(self):
        return self.desc.n_beads

    @property
    def n_cylinders(self):
        d = self.desc
        return d.fil_bead_offset[d.n_filaments] - d.n_filaments

    def perturb(self, coord, sigma):
        """Add N(0, sigma) noise to ``coord`` in place from the generator's stream."""
        assert coord.dtype == np.float64 and coord.flags.c_contiguous
        lib().mdn_synth_perturb(self._h, coord.ctypes.data_as(C.POINTER(C.c_double)), sigma)
        return coord

    def summary(self):
        d = self.desc
        return {"n_beads": d.n_beads, "n_filaments": d.n_filaments,
                "n_cylinders": self.n_cylinders, "n_linkers": d.n_linkers,
                "n_motors": d.n_motors, "n_boundary_elements": d.n_boundary}

    def __del__(self):
        try:
            if self._h:
                lib().mdn_synth_destroy(self._h)
                self._h = C.c_void_p()
        except Exception:
            pass
