"""Python host mirror of the reference's mechanical-equilibration interfaces over the C ABI.

``ForceFieldManager`` mirrors medyan::ForceFieldManager (ForceFieldManager.h:64-123):
``vectorize_all_force_fields`` / ``compute_energy`` / ``compute_individual_energies`` /
``compute_forces`` (zero-fill then accumulate), plus per-force-field accumulation
(ForceField::computeForces). ``CGMethod.minimize`` mirrors CGMethod::minimize
(CGMethod.hpp:1746-1759) as driven by MController::run (MController.h:108-211).
Everything runs in libmdn_mech.so on the GPU; there is no CPU fallback: a missing library or
device raises.

Citations are relative to medyan-5.4.0/ in the reference tree.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None

dp = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)

EXPORTED = [
    "mdn_abi_version", "mdn_ctx_create", "mdn_ctx_destroy", "mdn_last_error", "mdn_system_upload",
    "mdn_system_counts", "mdn_nl_build", "mdn_nl_pairs", "mdn_boundary_list", "mdn_energy",
    "mdn_forces", "mdn_ff_forces", "mdn_crosslink_stretch_forces", "mdn_minimize",
    "mdn_bench_forces", "mdn_bench_energy", "mdn_set_coords", "mdn_get_coords", "mdn_forces_mixed",
    "mdn_last_min_stats", "mdn_energy_mixed", "mdn_minimize_batched", "mdn_last_batch_schedule",
    "mdn_param_grad", "mdn_hvp", "mdn_set_params", "mdn_boundary_load_forces", "mdn_crosslinks_update",
    "mdn_forces_noalloc", "mdn_displace", "mdn_perturb", "mdn_param_grad_mixed", "mdn_hvp_mixed",
    "mdn_dof_serialize", "mdn_binding_pairs", "mdn_bench_adjoint", "mdn_hvp_batch",
]


class MdnError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{abi.status_name(code)}: {msg}")
        self.code = code


def lib_path():
    return os.path.join(_HERE, "libmdn_mech.so")


def lib():
    """Load libmdn_mech.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        L = C.CDLL(path)
        L.mdn_abi_version.restype = C.c_int
        L.mdn_ctx_create.argtypes = [C.c_int, C.POINTER(C.c_void_p)]
        L.mdn_ctx_destroy.argtypes = [C.c_void_p]
        L.mdn_last_error.argtypes = [C.c_void_p]
        L.mdn_last_error.restype = C.c_char_p
        L.mdn_system_upload.argtypes = [C.c_void_p, C.POINTER(abi.SystemDesc), dp]
        L.mdn_system_counts.argtypes = [C.c_void_p, i64p]
        L.mdn_nl_build.argtypes = [C.c_void_p, dp, i64p]
        L.mdn_nl_pairs.argtypes = [C.c_void_p, i32p, C.c_int64, i64p]
        L.mdn_boundary_list.argtypes = [C.c_void_p, i32p, C.c_int64, i32p, i64p]
        L.mdn_energy.argtypes = [C.c_void_p, dp, C.POINTER(abi.EnergyReport)]
        L.mdn_forces.argtypes = [C.c_void_p, dp, dp, C.c_int32]
        L.mdn_ff_forces.argtypes = [C.c_void_p, C.c_int32, dp, dp, C.c_int32]
        L.mdn_crosslink_stretch_forces.argtypes = [C.c_void_p, dp, dp]
        L.mdn_boundary_load_forces.argtypes = [C.c_void_p, dp, C.c_int32, dp, i32p, C.c_int32, dp, dp]
        xlp = C.POINTER(abi.Crosslink)
        L.mdn_crosslinks_update.argtypes = [C.c_void_p, xlp, C.c_int32, xlp, C.c_int32]
        L.mdn_minimize.argtypes = [C.c_void_p, C.POINTER(abi.CgParams), dp, dp, C.POINTER(abi.MinResult)]
        L.mdn_bench_forces.argtypes = [C.c_void_p, C.c_int32, C.c_int32, dp, dp]
        L.mdn_bench_energy.argtypes = [C.c_void_p, C.c_int32, C.c_int32, dp]
        L.mdn_bench_adjoint.argtypes = [C.c_void_p, C.c_int32, C.c_int32, dp, C.c_int32, C.c_int32, dp, dp]
        L.mdn_set_coords.argtypes = [C.c_void_p, dp]
        L.mdn_get_coords.argtypes = [C.c_void_p, dp]
        L.mdn_forces_mixed.argtypes = [C.c_void_p, dp, dp, C.c_int32]
        L.mdn_energy_mixed.argtypes = [C.c_void_p, dp, C.POINTER(abi.EnergyReport)]
        L.mdn_minimize_batched.argtypes = [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(abi.CgParams),
                                           C.POINTER(dp), C.POINTER(dp), C.POINTER(abi.MinResult)]
        L.mdn_last_batch_schedule.argtypes = [C.c_void_p, i32p, dp]
        L.mdn_param_grad.argtypes = [C.c_void_p, dp, dp, C.c_int64]
        L.mdn_hvp.argtypes = [C.c_void_p, dp, dp, dp, C.c_int32]
        L.mdn_hvp_mixed.argtypes = [C.c_void_p, dp, dp, dp, C.c_int32]
        L.mdn_hvp_batch.argtypes = [C.c_void_p, C.c_int32, dp, dp, dp, C.c_int32, C.c_int32]
        L.mdn_param_grad_mixed.argtypes = [C.c_void_p, dp, dp, C.c_int64]
        L.mdn_set_params.argtypes = [C.c_void_p, dp, dp, dp]
        L.mdn_last_min_stats.argtypes = [C.c_void_p, i64p]
        L.mdn_forces_noalloc.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32]
        L.mdn_displace.argtypes = [C.c_void_p, C.c_void_p, C.c_int32]
        L.mdn_perturb.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_double]
        L.mdn_dof_serialize.argtypes = [C.c_void_p, C.c_int32, i32p, dp]
        L.mdn_binding_pairs.argtypes = [C.c_void_p, C.POINTER(abi.BindingParams), dp, C.POINTER(C.c_uint32),
                                        C.c_int64, i64p]
        _lib = L
    return _lib


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class ForceFieldManager:
    """A device context holding one vectorized system (one HIP stream, not thread-safe)."""

    def __init__(self, device=0):
        self._L = lib()
        self._h = C.c_void_p()
        st = self._L.mdn_ctx_create(device, C.byref(self._h))
        if st != 0:
            raise MdnError(st, f"mdn_ctx_create(device={device}) failed (is a GPU visible?)")
        self.n_beads = 0

    def _chk(self, st):
        if st != 0:
            raise MdnError(st, self._L.mdn_last_error(self._h).decode())

    # ForceFieldManager::vectorizeAllForceFields (bonded + crosslink tables); coord None: the
    # coordinates gathered on the device by dof_serialize
    def vectorize_all_force_fields(self, desc, coord):
        ptr = desc if isinstance(desc, C._Pointer) else C.pointer(desc)
        c = None if coord is None else _f64(coord)
        self._chk(self._L.mdn_system_upload(self._h, ptr, None if c is None else c.ctypes.data_as(dp)))
        self.n_beads = ptr.contents.n_beads
        self.n_filaments = ptr.contents.n_filaments
        return self

    upload = vectorize_all_force_fields

    def counts(self):
        out = (C.c_int64 * 8)()
        self._chk(self._L.mdn_system_counts(self._h, out))
        keys = ["n_beads", "n_cylinders"] + list(abi.FF_NAMES)
        return dict(zip(keys, list(out)))

    # SubSystem::resetNeighborLists -> HybridCylinderCylinderNL::reset + BoundaryCylinderNL::reset
    def reset_neighbor_lists(self, coord=None):
        n = C.c_int64()
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        self._chk(self._L.mdn_nl_build(self._h, cp, C.byref(n)))
        return n.value

    def neighbor_pairs(self):
        n = C.c_int64()
        self._L.mdn_nl_pairs(self._h, None, 0, C.byref(n))
        out = np.zeros(2 * max(1, n.value), np.int32)
        self._chk(self._L.mdn_nl_pairs(self._h, out.ctypes.data_as(i32p), n.value, C.byref(n)))
        return out[: 2 * n.value].reshape(-1, 2)

    def boundary_interactions(self, n_elem):
        n = C.c_int64()
        self._L.mdn_boundary_list(self._h, None, 0, None, C.byref(n))
        beads = np.zeros(max(1, n.value), np.int32)
        nn = np.zeros(max(1, n_elem), np.int32)
        self._chk(self._L.mdn_boundary_list(self._h, beads.ctypes.data_as(i32p), n.value,
                                            nn.ctypes.data_as(i32p), C.byref(n)))
        return beads[: n.value], nn[:n_elem]

    def compute_individual_energies(self, coord=None, precision=abi.MDN_PREC_FP64):
        """EnergyReport per force field (ForceFieldManager::computeIndividualEnergies)."""
        rep = abi.EnergyReport()
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        fn = self._L.mdn_energy_mixed if precision == abi.MDN_PREC_MIXED else self._L.mdn_energy
        self._chk(fn(self._h, cp, C.byref(rep)))
        return rep

    def compute_energy(self, coord=None, precision=abi.MDN_PREC_FP64):
        """ForceFieldManager::computeEnergy: +inf when a force field reports <= -1."""
        rep = self.compute_individual_energies(coord, precision)
        tot = 0.0
        for e in rep.ff:
            if e <= -1:
                return float("inf")
            tot += e
        return tot

    def compute_forces(self, coord=None, precision=abi.MDN_PREC_FP64):
        """ForceFieldManager::computeForces: zero-filled, every force field accumulated."""
        f = np.zeros(3 * self.n_beads)
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        fn = self._L.mdn_forces_mixed if precision == abi.MDN_PREC_MIXED else self._L.mdn_forces
        self._chk(fn(self._h, cp, f.ctypes.data_as(dp), f.size))
        return f

    def ff_compute_forces(self, ff, coord=None, force=None):
        """ForceField::computeForces of force field ``ff``: accumulates into ``force``."""
        f = np.zeros(3 * self.n_beads) if force is None else force
        assert f.dtype == np.float64 and f.flags.c_contiguous
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        self._chk(self._L.mdn_ff_forces(self._h, ff, cp, f.ctypes.data_as(dp), f.size))
        return f

    def stretch_forces(self, n_linkers, n_motors):
        lo, mo = np.zeros(max(1, n_linkers)), np.zeros(max(1, n_motors))
        self._chk(self._L.mdn_crosslink_stretch_forces(self._h, lo.ctypes.data_as(dp), mo.ctypes.data_as(dp)))
        return lo[:n_linkers], mo[:n_motors]

    def set_coords(self, coord):
        self._chk(self._L.mdn_set_coords(self._h, _f64(coord).ctypes.data_as(dp)))

    def get_coords(self):
        x = np.zeros(3 * self.n_beads)
        self._chk(self._L.mdn_get_coords(self._h, x.ctypes.data_as(dp)))
        return x

    def minimize(self, params, coord, want_forces=True):
        """CGMethod::minimize from ``coord`` with the neighbor lists frozen.
        Returns (coord_out, force_out or None, MinResult). ``coord`` None: start from the
        context's device coordinates and leave the result there (coord_out is None)."""
        x = None if coord is None else _f64(coord).copy()
        f = np.zeros(3 * self.n_beads) if want_forces else None
        res = abi.MinResult()
        self._chk(self._L.mdn_minimize(self._h, C.byref(params), None if x is None else x.ctypes.data_as(dp),
                                       f.ctypes.data_as(dp) if want_forces else None, C.byref(res)))
        return x, f, res

    def binding_pairs(self, params, coord=None):
        """HybridBindingSearchManager's site pairs (mdn_binding_pairs): an (n, 2) uint32 array of
        (cylinder << shift_bits | site) entries, the first on the larger cylinder index."""
        cp = None if coord is None else _f64(coord).ctypes.data_as(dp)
        n = C.c_int64()
        self._chk(self._L.mdn_binding_pairs(self._h, C.byref(params), cp, None, 0, C.byref(n)))
        out = np.zeros(2 * max(1, n.value), np.uint32)
        self._chk(self._L.mdn_binding_pairs(self._h, C.byref(params), cp, out.ctypes.data_as(C.POINTER(C.c_uint32)),
                                            n.value, C.byref(n)))
        return out[: 2 * n.value].reshape(-1, 2)

    def dof_serialize(self, src, new_coord=None):
        """serializeDof on the device (mdn_dof_serialize): bead i of the next system is current
        bead src[i], or (src[i] < 0) the next new bead of ``new_coord`` (3 values each). The next
        ``upload(desc, None)`` adopts the result."""
        sv = np.ascontiguousarray(src, dtype=np.int32)
        nc = None if new_coord is None else _f64(new_coord)
        self._chk(self._L.mdn_dof_serialize(self._h, int(sv.size), sv.ctypes.data_as(i32p),
                                            None if nc is None else nc.ctypes.data_as(dp)))

    def perturb(self, seed, stream, sigma):
        """x_i += sigma z_i with seeded standard normals z generated on the device (mdn_perturb;
        ``perturbation`` gives the same values on the host)."""
        self._chk(self._L.mdn_perturb(self._h, int(seed), int(stream), float(sigma)))

    def displace(self, delta_dev_ptr):
        """x += delta on the device coordinates; ``delta_dev_ptr``: address of 3 * n_beads doubles
        in device memory (mdn_displace)."""
        self._chk(self._L.mdn_displace(self._h, C.c_void_p(delta_dev_ptr), 3 * self.n_beads))

    def update_crosslinks(self, linkers, motors):
        """LinkerStretching / MotorGhostStretching::vectorize alone (mdn_crosslinks_update): new
        linker and motor-ghost tables (sequences of abi.Crosslink) for the uploaded system; the
        filament tables and the neighbor lists are kept."""
        arrs = []
        for xl in (linkers, motors):
            n = len(xl)
            a = (abi.Crosslink * max(1, n))(*xl)
            arrs.append((a, n))
        (la, ln), (ma, mn) = arrs
        self._chk(self._L.mdn_crosslinks_update(self._h, la, ln, ma, mn))

    def boundary_load_forces(self, monomer_size, num_monomers, coord=None, stride=None):
        """Boundary load forces on the filament ends (mdn_boundary_load_forces): (plus, minus),
        each [n_filaments, stride] (stride defaults to max(num_monomers)). ``monomer_size`` and
        ``num_monomers`` are per filament type."""
        ms = np.ascontiguousarray(np.atleast_1d(monomer_size), dtype=np.float64)
        nm = np.ascontiguousarray(np.atleast_1d(num_monomers), dtype=np.int32)
        stride = int(nm.max()) if stride is None else int(stride)
        nf = self.n_filaments
        lp, lm = np.zeros(nf * stride), np.zeros(nf * stride)
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        self._chk(self._L.mdn_boundary_load_forces(self._h, cp, len(ms), ms.ctypes.data_as(dp),
                                                   nm.ctypes.data_as(i32p), stride, lp.ctypes.data_as(dp),
                                                   lm.ctypes.data_as(dp)))
        return lp.reshape(nf, stride), lm.reshape(nf, stride)

    def last_min_stats(self):
        out = (C.c_int64 * 4)()
        self._chk(self._L.mdn_last_min_stats(self._h, out))
        return {"ls_launches": out[0], "forceN_steps": out[1], "gathers": out[2], "iterations": out[3]}

    def param_grad(self, coord=None, precision=abi.MDN_PREC_FP64):
        """dE/dtheta for theta = (stretching k, rest length L0, bending k) per cylinder; returns
        an array of shape (3, n_cylinders) (mdn_param_grad, mdn_param_grad_mixed)."""
        nc = self.counts()["n_cylinders"]
        out = np.zeros(3 * nc)
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        fn = self._L.mdn_param_grad_mixed if precision == abi.MDN_PREC_MIXED else self._L.mdn_param_grad
        self._chk(fn(self._h, cp, out.ctypes.data_as(dp), out.size))
        return out.reshape(3, nc)

    def hvp(self, v, coord=None, precision=abi.MDN_PREC_FP64):
        """Hessian-vector product H v = -(dF/dx) v (mdn_hvp, mdn_hvp_mixed)."""
        vv = _f64(v)
        out = np.zeros(3 * self.n_beads)
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        fn = self._L.mdn_hvp_mixed if precision == abi.MDN_PREC_MIXED else self._L.mdn_hvp
        self._chk(fn(self._h, cp, vv.ctypes.data_as(dp), out.ctypes.data_as(dp), out.size))
        return out

    def hvp_batch(self, V, coord=None, precision=abi.MDN_PREC_FP64):
        """Hessian-vector products of the rows of V (n_vec x 3 n_beads) at one point, as rows
        (mdn_hvp_batch)."""
        VV = _f64(V).reshape(-1, 3 * self.n_beads)
        out = np.zeros_like(VV)
        cp = _f64(coord).ctypes.data_as(dp) if coord is not None else None
        self._chk(self._L.mdn_hvp_batch(self._h, precision, cp, VV.ctypes.data_as(dp), out.ctypes.data_as(dp),
                                        VV.shape[1], VV.shape[0]))
        return out

    def set_params(self, kstr=None, eql=None, kbend=None):
        """Replace per-cylinder stretching k, rest length and bending k (mdn_set_params)."""
        arrs = [None if a is None else _f64(a) for a in (kstr, eql, kbend)]
        ptrs = [None if a is None else a.ctypes.data_as(dp) for a in arrs]
        self._chk(self._L.mdn_set_params(self._h, *ptrs))

    def bench_forces(self, iters, precision=abi.MDN_PREC_FP64):
        a, b = C.c_double(), C.c_double()
        self._chk(self._L.mdn_bench_forces(self._h, precision, iters, C.byref(a), C.byref(b)))
        return a.value, b.value

    def bench_energy(self, iters, precision=abi.MDN_PREC_FP64):
        a = C.c_double()
        self._chk(self._L.mdn_bench_energy(self._h, precision, iters, C.byref(a)))
        return a.value

    def bench_adjoint(self, iters, v, precision=abi.MDN_PREC_FP64):
        """Mean device ms per Hessian-vector product along the rows of v (one row: mdn_hvp's
        launches; several: one chunk of mdn_hvp_batch) and per dE/dtheta at the current device
        coordinates (mdn_bench_adjoint)."""
        vv = _f64(v).reshape(-1, 3 * self.n_beads)
        a, b = C.c_double(), C.c_double()
        self._chk(self._L.mdn_bench_adjoint(self._h, precision, iters, vv.ctypes.data_as(dp), vv.shape[1],
                                            vv.shape[0], C.byref(a), C.byref(b)))
        return a.value, b.value

    def close(self):
        if getattr(self, "_h", None):
            self._L.mdn_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def perturbation(seed, stream, sigma, n):
    """The seeded Gaussian displacement mdn_perturb adds to n coordinates, on the host: SplitMix64
    counters (key = mix(seed + G (stream + 1)), r(j) = mix(key + G (j + 1))) give u1 = ((r(2i) >> 11)
    + 1) 2^-53 and u2 = (r(2i + 1) >> 11) 2^-53, and z_i = sqrt(-2 ln u1) cos(2 pi u2). The device
    evaluates log, sqrt and cos with its own math library, so values agree to a few ulps."""
    G = np.uint64(0x9E3779B97F4A7C15)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        key = mix(np.uint64(seed) + G * np.uint64(stream + 1))
        i = np.arange(n, dtype=np.uint64)
        r1 = mix(key + G * (np.uint64(2) * i + np.uint64(1)))
        r2 = mix(key + G * (np.uint64(2) * i + np.uint64(2)))
    u1 = ((r1 >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * 2.0 ** -53
    u2 = (r2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return sigma * (np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2))


def minimize_batched(systems, params, coords, want_forces=True):
    """CGMethod::minimize of many independent systems in one persistent kernel
    (mdn_minimize_batched). ``systems``: ForceFieldManager objects with uploaded systems and
    built neighbor lists (the first one is the batch's lead context); ``coords``: one start
    vector per system (an entry None, or coords None: that system starts from its device
    coordinates and keeps the result there). Returns (list of coordinates or None, list of forces
    or None, list of MinResult)."""
    L = lib()
    n = len(systems)
    xs = [None if c is None else _f64(c).copy() for c in (coords if coords is not None else [None] * n)]
    fs = [np.zeros(3 * s.n_beads) for s in systems] if want_forces else None
    ctxs = (C.c_void_p * n)(*[s._h.value for s in systems])
    xp = (dp * n)(*[None if x is None else x.ctypes.data_as(dp) for x in xs])
    fp = (dp * n)(*[f.ctypes.data_as(dp) for f in fs]) if want_forces else None
    res = (abi.MinResult * n)()
    st = L.mdn_minimize_batched(ctxs, n, C.byref(params), xp, fp, res)
    if st != 0:
        raise MdnError(st, L.mdn_last_error(systems[0]._h).decode())
    return xs, fs, list(res)


def last_batch_schedule(lead):
    """(workgroup size, largest-first order, tuned in the last call, tuning times in ms)."""
    out = (C.c_int32 * 3)()
    ms = (C.c_double * abi.MDN_BATCH_CANDIDATES)()
    st = lib().mdn_last_batch_schedule(lead._h, out, ms)
    if st != 0:
        raise MdnError(st, "mdn_last_batch_schedule")
    return out[0], out[1], out[2], list(ms)
