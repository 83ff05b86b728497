"""Hand-built mdn_system_desc instances for tests: explicit filaments (lists of bead coordinates)
and boundary elements, with the per-cylinder constants of the synthetic networks (stretching k
100, rest length 108 nm, bending k 672, excluded-volume k 1e5)."""
import ctypes as C

import numpy as np

from picgo_amd import abi


class HandSystem:
    """A custom mdn_system_desc from explicit filaments (lists of bead coordinates); keeps its
    arrays alive."""

    def __init__(self, filaments, elements, cutoff=300.0, grid=(3000.0, 3000.0, 3000.0), types=None):
        coords, offs, beads = [], [0], []
        for fil in filaments:
            for p in fil:
                beads.append(len(coords))
                coords.append(p)
            offs.append(len(beads))
        self.coord = np.array(coords, dtype=np.float64).ravel()
        n_cyl = offs[-1] - len(filaments)
        self._keep = [
            np.array(offs, np.int32), np.array(beads, np.int32),
            np.array(types if types is not None else [0] * len(filaments), np.int32),
            np.full(n_cyl, 100.0), np.full(n_cyl, 108.0), np.full(n_cyl, 672.0), np.zeros(n_cyl),
            np.full(n_cyl, 1e5),
        ]
        self.elems = (abi.BoundaryElement * max(1, len(elements)))(*elements)
        d = abi.SystemDesc()
        d.n_beads = len(coords)
        d.n_filaments = len(filaments)
        i32 = lambda a: a.ctypes.data_as(C.POINTER(C.c_int32))
        f64 = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        d.fil_bead_offset, d.fil_beads, d.fil_type = i32(self._keep[0]), i32(self._keep[1]), i32(self._keep[2])
        d.cyl_kstr, d.cyl_eql, d.cyl_kbend = f64(self._keep[3]), f64(self._keep[4]), f64(self._keep[5])
        d.cyl_eqtheta, d.cyl_kexvol = f64(self._keep[6]), f64(self._keep[7])
        d.n_linkers = d.n_motors = 0
        d.n_boundary = len(elements)
        d.boundary = self.elems
        d.boundary_cutoff = cutoff
        d.exvol_rmax, d.exvol_rmin = 108.0, 0.0
        d.grid_size[:] = grid
        self.desc = d
