"""Python binding of the CPU oracle (liboracle.so / liboracle_fast.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the oracle and the CPU baseline. The product (picgo_amd) never imports it.
"""
import ctypes as C
import os

import numpy as np

from picgo_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_libs = {}

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def _arr(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a


def _p(a):
    if a.dtype == np.float64:
        return a.ctypes.data_as(dp)
    return a.ctypes.data_as(ip)


def lib(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    if name not in _libs:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run `make -C oracle`")
        L = C.CDLL(path)
        L.or_stretch_energy.restype = C.c_double
        L.or_stretch_energy.argtypes = [dp, ip, dp, dp, C.c_int]
        L.or_stretch_forces.argtypes = [dp, dp, ip, dp, dp, C.c_int]
        L.or_bend_energy.restype = C.c_double
        L.or_bend_energy.argtypes = [dp, C.c_int, ip, dp, dp]
        L.or_bend_forces.argtypes = [dp, dp, C.c_int, ip, dp, dp]
        L.or_xlink_energy.restype = C.c_double
        L.or_xlink_energy.argtypes = [dp, ip, dp, dp, dp, dp, C.c_int]
        L.or_xlink_forces.argtypes = [dp, dp, ip, dp, dp, dp, dp, dp, C.c_int]
        L.or_exvol_energy.restype = C.c_double
        L.or_exvol_energy.argtypes = [dp, ip, dp, dp, C.c_int]
        L.or_exvol_forces.argtypes = [dp, dp, ip, dp, dp, C.c_int]
        L.or_exvol_energyN.restype = C.c_double
        L.or_exvol_energyN.argtypes = [dp, ip, dp, dp, C.c_int, C.c_int]
        L.or_exvol_forceN.argtypes = [dp, dp, ip, dp, dp, C.c_int, C.c_int]
        L.or_are_in_plane.restype = C.c_int
        L.or_are_in_plane.argtypes = [dp, dp, dp, dp]
        L.or_move_out_of_plane.argtypes = [dp, dp, dp, dp, C.c_int, C.c_double]
        bep = C.POINTER(abi.BoundaryElement)
        L.or_boundary_energy.restype = C.c_double
        L.or_boundary_energy.argtypes = [dp, ip, dp, dp, ip, bep, C.c_int]
        L.or_boundary_forces.argtypes = [dp, dp, ip, dp, dp, ip, bep, C.c_int]
        L.or_boundary_distance.restype = C.c_double
        L.or_boundary_distance.argtypes = [bep, dp]
        L.or_plane_element.restype = abi.BoundaryElement
        L.or_plane_element.argtypes = [dp, dp, C.c_double, C.c_double]
        L.or_nl_brute.restype = C.c_long
        L.or_nl_brute.argtypes = [dp, ip, ip, C.c_int, C.c_float, C.c_float, ip, C.c_long,
                                  C.POINTER(C.c_long)]
        L.or_nl_cells.restype = C.c_long
        L.or_nl_cells.argtypes = [dp, ip, ip, C.c_int, C.c_float, C.c_float, dp, ip, C.c_long,
                                  C.POINTER(C.c_long)]
        L.or_binding_pairs.restype = C.c_long
        L.or_binding_pairs.argtypes = [C.c_void_p, dp, C.POINTER(abi.BindingParams), C.POINTER(C.c_uint32), C.c_long,
                                       C.POINTER(C.c_long)]
        L.or_system_create.restype = C.c_void_p
        L.or_system_create.argtypes = [C.POINTER(abi.SystemDesc), C.POINTER(C.c_int)]
        L.or_system_destroy.argtypes = [C.c_void_p]
        L.or_system_set_cells.argtypes = [C.c_void_p, C.c_int]
        L.or_system_nl.restype = C.c_int
        L.or_system_nl.argtypes = [C.c_void_p, dp]
        L.or_system_npairs.restype = C.c_long
        L.or_system_npairs.argtypes = [C.c_void_p]
        L.or_system_pairs.restype = ip
        L.or_system_pairs.argtypes = [C.c_void_p]
        L.or_system_nbint.restype = C.c_int
        L.or_system_nbint.argtypes = [C.c_void_p]
        L.or_system_bnd_beadset.restype = ip
        L.or_system_bnd_beadset.argtypes = [C.c_void_p]
        L.or_system_bnd_nneighbors.restype = ip
        L.or_system_bnd_nneighbors.argtypes = [C.c_void_p]
        L.or_system_ncyl.restype = C.c_int
        L.or_system_ncyl.argtypes = [C.c_void_p]
        L.or_system_nbend.restype = C.c_int
        L.or_system_nbend.argtypes = [C.c_void_p]
        L.or_system_stretchforce.restype = dp
        L.or_system_stretchforce.argtypes = [C.c_void_p, C.c_int]
        L.or_system_load_forces.restype = C.c_int
        L.or_system_load_forces.argtypes = [C.c_void_p, dp, C.c_int, dp, ip, C.c_int, dp, dp]
        L.or_ffm_energy.restype = C.c_double
        L.or_ffm_energy.argtypes = [C.c_void_p, dp, C.POINTER(abi.EnergyReport)]
        L.or_ffm_forces.argtypes = [C.c_void_p, dp, dp]
        L.or_ff_forces.argtypes = [C.c_void_p, C.c_int, dp, dp]
        L.or_ff_energy.restype = C.c_double
        L.or_ff_energy.argtypes = [C.c_void_p, C.c_int, dp]
        L.or_system_minimize.restype = C.c_int
        L.or_system_minimize.argtypes = [C.c_void_p, C.POINTER(abi.CgParams), dp, dp,
                                         C.POINTER(abi.MinResult)]
        L.or_cg_minimize.restype = C.c_int
        L.or_cg_minimize.argtypes = [C.POINTER(abi.CgParams), dp, C.c_int, C.c_int,
                                     C.c_void_p, dp, C.POINTER(abi.MinResult)]
        _libs[name] = L
    return _libs[name]


# ---- interaction kernels (reference signatures, beadSet = flat offsets 3*bead) --------

def stretch_energy(coord, beadSet, k, eql):
    c, b, k, e = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(k, np.float64), _arr(eql, np.float64)
    return lib().or_stretch_energy(_p(c), _p(b), _p(k), _p(e), len(k))


def stretch_forces(coord, beadSet, k, eql, f=None):
    c, b, k, e = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(k, np.float64), _arr(eql, np.float64)
    f = np.zeros_like(c) if f is None else f
    lib().or_stretch_forces(_p(c), _p(f), _p(b), _p(k), _p(e), len(k))
    return f


def bend_energy(coord, beadSet, k, eqt):
    c, b, k, t = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(k, np.float64), _arr(eqt, np.float64)
    return lib().or_bend_energy(_p(c), len(k), _p(b), _p(k), _p(t))


def bend_forces(coord, beadSet, k, eqt, f=None):
    c, b, k, t = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(k, np.float64), _arr(eqt, np.float64)
    f = np.zeros_like(c) if f is None else f
    lib().or_bend_forces(_p(c), _p(f), len(k), _p(b), _p(k), _p(t))
    return f


def xlink_energy(coord, beadSet, k, eql, p1, p2):
    a = [_arr(coord, np.float64), _arr(beadSet, np.int32)] + [_arr(x, np.float64) for x in (k, eql, p1, p2)]
    return lib().or_xlink_energy(*[_p(x) for x in a], len(a[2]))


def xlink_forces(coord, beadSet, k, eql, p1, p2, f=None):
    a = [_arr(coord, np.float64), _arr(beadSet, np.int32)] + [_arr(x, np.float64) for x in (k, eql, p1, p2)]
    f = np.zeros_like(a[0]) if f is None else f
    sf = np.zeros(len(a[2]))
    lib().or_xlink_forces(_p(a[0]), _p(f), _p(a[1]), _p(a[2]), _p(a[3]), _p(a[4]), _p(a[5]), _p(sf), len(a[2]))
    return f, sf


def exvol_energy(coord, beadSet, krep, eqlen):
    c, b, k, e = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(krep, np.float64), _arr(eqlen, np.float64)
    return lib().or_exvol_energy(_p(c), _p(b), _p(k), _p(e), len(k))


def exvol_forces(coord, beadSet, krep, eqlen, f=None):
    c, b, k, e = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(krep, np.float64), _arr(eqlen, np.float64)
    f = np.zeros_like(c) if f is None else f
    lib().or_exvol_forces(_p(c), _p(f), _p(b), _p(k), _p(e), len(k))
    return f


def exvol_energyN(coord, beadSet, krep, eqlen, i, movebeads=0):
    c, b, k, e = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(krep, np.float64), _arr(eqlen, np.float64)
    return lib().or_exvol_energyN(_p(c), _p(b), _p(k), _p(e), i, movebeads)


def exvol_forceN(coord, beadSet, krep, eqlen, i, movebeads=0):
    """The 10^4 x 10^4 Simpson force rule of pair i, accumulated into a zero force vector."""
    c, b, k, e = _arr(coord, np.float64), _arr(beadSet, np.int32), _arr(krep, np.float64), _arr(eqlen, np.float64)
    f = np.zeros_like(c)
    lib().or_exvol_forceN(_p(c), _p(f), _p(b), _p(k), _p(e), i, movebeads)
    return f


def are_in_plane(p1, p2, p3, p4):
    a = [_arr(x, np.float64) for x in (p1, p2, p3, p4)]
    return bool(lib().or_are_in_plane(*[_p(x) for x in a]))


def move_out_of_plane(p1, p2, p3, i, d):
    """movePointOutOfPlane: point i (1 or 2) moved by d along the unit normal of (p2 - p1, p3 - p1)."""
    a = [_arr(x, np.float64) for x in (p1, p2, p3)]
    out = np.zeros(3)
    lib().or_move_out_of_plane(*[_p(x) for x in a], _p(out), int(i), float(d))
    return out


def plane_element(point, normal, krep, slen):
    return lib().or_plane_element(_p(_arr(point, np.float64)), _p(_arr(normal, np.float64)), krep, slen)


def _elems(elems):
    arr = (abi.BoundaryElement * max(1, len(elems)))(*elems)
    return arr


def boundary_energy(coord, beadSet, krep, slen, nneighbors, elems):
    c, b = _arr(coord, np.float64), _arr(beadSet, np.int32)
    k, s, n = _arr(krep, np.float64), _arr(slen, np.float64), _arr(nneighbors, np.int32)
    ea = _elems(elems)
    return lib().or_boundary_energy(_p(c), _p(b), _p(k), _p(s), _p(n), ea, len(elems))


def boundary_forces(coord, beadSet, krep, slen, nneighbors, elems, f=None):
    c, b = _arr(coord, np.float64), _arr(beadSet, np.int32)
    k, s, n = _arr(krep, np.float64), _arr(slen, np.float64), _arr(nneighbors, np.int32)
    f = np.zeros_like(c) if f is None else f
    ea = _elems(elems)
    lib().or_boundary_forces(_p(c), _p(f), _p(b), _p(k), _p(s), _p(n), ea, len(elems))
    return f


def nl_pairs(mid, fil, pos, rmax, rmin=0.0, grid=None):
    """Pair array (n, 2) from the brute-force predicate, or the cell list when grid is given."""
    m, fi, po = _arr(mid, np.float64), _arr(fil, np.int32), _arr(pos, np.int32)
    n_cyl = len(fi)
    cap = 1024
    while True:
        out = np.zeros(2 * cap, np.int32)
        need = C.c_long()
        if grid is None:
            n = lib().or_nl_brute(_p(m), _p(fi), _p(po), n_cyl, rmax, rmin, _p(out), cap, C.byref(need))
        else:
            g = _arr(grid, np.float64)
            n = lib().or_nl_cells(_p(m), _p(fi), _p(po), n_cyl, rmax, rmin, _p(g), _p(out), cap, C.byref(need))
        if n == -2:
            raise ValueError("midpoint outside the neighbor-list grid")
        if n >= 0:
            return out[: 2 * n].reshape(-1, 2)
        cap = need.value


# ---- CG with Python callbacks --------------------------------------------------------

ENERGY_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, dp)
FORCE_FN = C.CFUNCTYPE(None, C.c_void_p, dp, dp, C.c_int)
ENERGY_IND_FN = C.CFUNCTYPE(None, C.c_void_p, dp, C.POINTER(abi.EnergyReport))
RECOVERY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, dp, dp, C.c_int)


class _Callbacks(C.Structure):
    _fields_ = [("energy", ENERGY_FN), ("force", FORCE_FN), ("energy_individual", ENERGY_IND_FN),
                ("recovery", RECOVERY_FN), ("user", C.c_void_p), ("n_beads", C.c_int)]


def cg_minimize(params, coord, ndof, energy, force, nvar=None, recovery=None, n_beads=0):
    """or_cg_minimize with Python energy(x)->float and force(x, f) callbacks on numpy views."""
    nvar = nvar or ndof
    x = np.ascontiguousarray(coord, dtype=np.float64).copy()

    def e_cb(_u, cp):
        return float(energy(np.ctypeslib.as_array(cp, shape=(nvar,))))

    def f_cb(_u, cp, fp, n):
        force(np.ctypeslib.as_array(cp, shape=(nvar,)), np.ctypeslib.as_array(fp, shape=(nvar,)))

    def ei_cb(_u, cp, rep):
        rep.contents.total = float(energy(np.ctypeslib.as_array(cp, shape=(nvar,))))

    def r_cb(_u, cp, dp_, n):
        if recovery is None:
            return 0
        return int(recovery(np.ctypeslib.as_array(cp, shape=(nvar,)), np.ctypeslib.as_array(dp_, shape=(n,))))

    cbs = _Callbacks(ENERGY_FN(e_cb), FORCE_FN(f_cb), ENERGY_IND_FN(ei_cb), RECOVERY_FN(r_cb), None, n_beads)
    res = abi.MinResult()
    fout = np.zeros(nvar)
    st = lib().or_cg_minimize(C.byref(params), _p(x), ndof, nvar, C.cast(C.pointer(cbs), C.c_void_p),
                              _p(fout), C.byref(res))
    return st, x, fout, res


# ---- the system ---------------------------------------------------------------------

class OracleSystem:
    """The reference path on the CPU for one mdn_system_desc (vectorize + FFM + CG)."""

    def __init__(self, desc, fast=False, use_cells=True):
        self._L = lib(fast)
        st = C.c_int()
        ptr = desc if isinstance(desc, C._Pointer) else C.pointer(desc)
        self._h = self._L.or_system_create(ptr, C.byref(st))
        if not self._h:
            raise ValueError(f"or_system_create: {abi.status_name(st.value)}")
        self._L.or_system_set_cells(self._h, 1 if use_cells else 0)
        self.n_beads = desc.contents.n_beads if isinstance(desc, C._Pointer) else desc.n_beads
        self.n_filaments = desc.contents.n_filaments if isinstance(desc, C._Pointer) else desc.n_filaments

    def build_nl(self, coord):
        c = _arr(coord, np.float64)
        st = self._L.or_system_nl(self._h, _p(c))
        if st != 0:
            raise ValueError(f"or_system_nl: {abi.status_name(st)}")
        return self.n_pairs

    @property
    def n_pairs(self):
        return self._L.or_system_npairs(self._h)

    def binding_pairs(self, params, coord):
        """or_binding_pairs: (n, 2) uint32 (cylinder << shift_bits | site) pairs by brute force."""
        c = _arr(coord, np.float64)
        need = C.c_long()
        self._L.or_binding_pairs(self._h, _p(c), C.byref(params), None, 0, C.byref(need))
        out = np.zeros(2 * max(1, need.value), np.uint32)
        n = self._L.or_binding_pairs(self._h, _p(c), C.byref(params), out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                     need.value, C.byref(need))
        assert n == need.value
        return out[: 2 * n].reshape(-1, 2)

    def pairs(self):
        n = self.n_pairs
        if n == 0:
            return np.zeros((0, 2), np.int32)
        return np.ctypeslib.as_array(self._L.or_system_pairs(self._h), shape=(2 * n,)).reshape(-1, 2).copy()

    def boundary_list(self):
        n = self._L.or_system_nbint(self._h)
        beads = (np.ctypeslib.as_array(self._L.or_system_bnd_beadset(self._h), shape=(n,)).copy() // 3
                 if n else np.zeros(0, np.int32))
        return beads

    def boundary_nneighbors(self, n_elem):
        return np.ctypeslib.as_array(self._L.or_system_bnd_nneighbors(self._h), shape=(n_elem,)).copy()

    def energy(self, coord):
        """(computeEnergy total with the +inf sentinel, EnergyReport of computeIndividualEnergies)"""
        rep = abi.EnergyReport()
        c = _arr(coord, np.float64)
        e = self._L.or_ffm_energy(self._h, _p(c), C.byref(rep))
        return e, rep

    def ff_energy(self, ff, coord):
        return self._L.or_ff_energy(self._h, ff, _p(_arr(coord, np.float64)))

    def forces(self, coord):
        c = _arr(coord, np.float64)
        f = np.zeros_like(c)
        self._L.or_ffm_forces(self._h, _p(c), _p(f))
        return f

    def ff_forces(self, ff, coord, f=None):
        c = _arr(coord, np.float64)
        f = np.zeros_like(c) if f is None else f
        self._L.or_ff_forces(self._h, ff, _p(c), _p(f))
        return f

    def stretch_forces_of(self, n, motor=False):
        """The linker (or motor) stretchForce values k (L - L0) of the last force evaluation
        (LinkerStretchingHarmonic.cpp:291), n of them."""
        if n == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(self._L.or_system_stretchforce(self._h, 1 if motor else 0), shape=(n,)).copy()

    def load_forces(self, coord, mon_size, num_mon, stride):
        """Boundary load forces on the filament ends: (plus, minus), each [n_filaments, stride]
        (or_system_load_forces)."""
        ms = _arr(mon_size, np.float64)
        nm = _arr(num_mon, np.int32)
        c = _arr(coord, np.float64)
        nf = self.n_filaments
        outp, outm = np.zeros(nf * stride), np.zeros(nf * stride)
        st = self._L.or_system_load_forces(self._h, _p(c), len(ms), _p(ms), nm.ctypes.data_as(ip), stride,
                                           _p(outp), _p(outm))
        if st != 0:
            raise ValueError(f"or_system_load_forces: {abi.status_name(st)}")
        return outp.reshape(nf, stride), outm.reshape(nf, stride)

    def minimize(self, params, coord):
        x = _arr(coord, np.float64).copy()
        f = np.zeros_like(x)
        res = abi.MinResult()
        st = self._L.or_system_minimize(self._h, C.byref(params), _p(x), _p(f), C.byref(res))
        return st, x, f, res

    def __del__(self):
        try:
            if self._h:
                self._L.or_system_destroy(self._h)
                self._h = None
        except Exception:
            pass
