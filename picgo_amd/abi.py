"""ctypes mirror of include/mdn_mech.h and include/mdn_synth.h.

The structures here are byte-compatible with the C declarations; every binding (the product
host API in ``picgo_amd.mech``, the synthetic-network generator in ``picgo_amd.synth``) uses
them.
"""
import ctypes as C

MDN_OK = 0
STATUS_NAMES = {
    0: "MDN_OK", 1: "MDN_ERR_ARG", 2: "MDN_ERR_HIP", 3: "MDN_ERR_STATE",
    4: "MDN_ERR_NONFINITE", 5: "MDN_ERR_LINESEARCH", 6: "MDN_ERR_NL_BOUNDS",
    7: "MDN_ERR_RECOVERY", 8: "MDN_ERR_NOMEM", 9: "MDN_ERR_UNSUPPORTED", 10: "MDN_ERR_CAPACITY",
}
MDN_ERR_ARG = 1
MDN_ERR_HIP = 2
MDN_ERR_STATE = 3
MDN_ERR_NONFINITE = 4
MDN_ERR_LINESEARCH = 5
MDN_ERR_NL_BOUNDS = 6
MDN_ERR_RECOVERY = 7
MDN_ERR_NOMEM = 8
MDN_ERR_UNSUPPORTED = 9
MDN_ERR_CAPACITY = 10

FF_NAMES = ("FilamentStretching", "FilamentBending", "LinkerStretching",
            "MotorGhostStretching", "CylinderExclVolume", "BoundaryCylinderRepulsion")
MDN_NUM_FF = 6
MDN_FF_STRETCH, MDN_FF_BEND, MDN_FF_LINKER, MDN_FF_MOTOR, MDN_FF_EXVOL, MDN_FF_BOUNDARY = range(6)

MDN_BND_PLANE, MDN_BND_SPHERE = 0, 1
MDN_CG_STEEPEST, MDN_CG_FLETCHER_REEVES, MDN_CG_POLAK_RIBIERE = 0, 1, 2
MDN_PREC_FP64, MDN_PREC_MIXED = 0, 1
MDN_LS_BACKTRACKING, MDN_LS_QUADRATIC = 0, 1
MDN_BATCH_CANDIDATES = 8   # schedules the batched minimizer's tuner times (include/mdn_mech.h)
MDN_MIN_ERR_LINESEARCH = 1

c_int32_p = C.POINTER(C.c_int32)
c_double_p = C.POINTER(C.c_double)


class Crosslink(C.Structure):
    _fields_ = [("cyl1", C.c_int32), ("cyl2", C.c_int32), ("pos1", C.c_double),
                ("pos2", C.c_double), ("kstr", C.c_double), ("eql", C.c_double)]


class BoundaryElement(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("p", C.c_double * 4),
                ("krep", C.c_double), ("slen", C.c_double)]


class SystemDesc(C.Structure):
    _fields_ = [
        ("n_beads", C.c_int32), ("n_filaments", C.c_int32),
        ("fil_bead_offset", c_int32_p), ("fil_beads", c_int32_p), ("fil_type", c_int32_p),
        ("cyl_kstr", c_double_p), ("cyl_eql", c_double_p), ("cyl_kbend", c_double_p),
        ("cyl_eqtheta", c_double_p), ("cyl_kexvol", c_double_p),
        ("n_linkers", C.c_int32), ("linkers", C.POINTER(Crosslink)),
        ("n_motors", C.c_int32), ("motors", C.POINTER(Crosslink)),
        ("n_boundary", C.c_int32), ("boundary", C.POINTER(BoundaryElement)),
        ("boundary_cutoff", C.c_double), ("exvol_rmax", C.c_double),
        ("exvol_rmin", C.c_double), ("grid_size", C.c_double * 3),
    ]


class BindingParams(C.Structure):
    _fields_ = [("ftype1", C.c_int32), ("ftype2", C.c_int32), ("n_sites", C.c_int32), ("shift_bits", C.c_int32),
                ("site_pos", C.POINTER(C.c_float)), ("p2_first", C.c_float), ("p2_last", C.c_float),
                ("rmin", C.c_float), ("rmax", C.c_float), ("cyl_size", C.c_float), ("site_step", C.c_int32),
                ("site_free", C.POINTER(C.c_float))]


def binding_params(site_pos, rmin, rmax, cyl_size, site_step=0, site_free=None, ftype=(0, 0), shift_bits=4):
    """A BindingParams with its arrays kept alive on the returned object (``.keep``)."""
    import numpy as np
    sp = np.ascontiguousarray(site_pos, dtype=np.float32)
    sf = None if site_free is None else np.ascontiguousarray(site_free, dtype=np.float32)
    p = BindingParams(ftype1=ftype[0], ftype2=ftype[1], n_sites=sp.size, shift_bits=shift_bits,
                      site_pos=sp.ctypes.data_as(C.POINTER(C.c_float)), p2_first=float(sp[0]),
                      p2_last=float(sp[-1]), rmin=rmin, rmax=rmax, cyl_size=cyl_size, site_step=site_step,
                      site_free=None if sf is None else sf.ctypes.data_as(C.POINTER(C.c_float)))
    p.keep = (sp, sf)
    return p


class EnergyReport(C.Structure):
    _fields_ = [("total", C.c_double), ("ff", C.c_double * MDN_NUM_FF)]

    def as_dict(self):
        d = {name: self.ff[i] for i, name in enumerate(FF_NAMES)}
        d["total"] = self.total
        return d


class CgParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("try_recover", C.c_int32), ("max_dist", C.c_double),
                ("lambda_max", C.c_double), ("grad_tol", C.c_double),
                ("energy_rel_tol", C.c_double), ("force_call_limit", C.c_int32),
                ("energy_call_limit", C.c_int32), ("report_on_call_limit", C.c_int32),
                ("precision", C.c_int32), ("line_search", C.c_int32), ("reserved", C.c_int32)]


def default_cg_params(**kw):
    """MController's ConjugateGradientParams with MEDYAN's defaults (MController.cpp:34-63,
    MedyanConfig.hpp:355-362): Polak-Ribiere, maxDist 1.0, lambdaMax 0.01, gradTol 1 pN,
    energyRelTol 1e-9, forceCallLimit 10000, recovery on."""
    p = CgParams(method=MDN_CG_POLAK_RIBIERE, try_recover=1, max_dist=1.0, lambda_max=0.01,
                 grad_tol=1.0, energy_rel_tol=1e-9, force_call_limit=10000,
                 energy_call_limit=0, report_on_call_limit=1, precision=MDN_PREC_FP64)
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class MinResult(C.Structure):
    _fields_ = [
        ("energies_before", EnergyReport), ("energies_after", EnergyReport),
        ("error_flags", C.c_uint32), ("num_energy_call", C.c_int32),
        ("num_force_call", C.c_int32), ("force_call_limit", C.c_int32),
        ("energy_call_limit", C.c_int32), ("num_iter", C.c_int32),
        ("num_safe_mode", C.c_int32), ("num_recovery", C.c_int32),
        ("grad_inf_norm", C.c_double), ("grad_tol", C.c_double),
        ("energy_rel_change", C.c_double), ("energy_rel_tol", C.c_double),
        ("converged_grad", C.c_int32), ("converged_energy", C.c_int32),
        ("call_limit_exceeded", C.c_int32), ("restored_min_energy", C.c_int32),
    ]

    def success(self):
        return self.error_flags == 0

    def converged(self):
        return bool(self.converged_grad or self.converged_energy)

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = v.as_dict() if isinstance(v, EnergyReport) else v
        return out


class SynthParams(C.Structure):
    _fields_ = [
        ("seed", C.c_uint64), ("n_fil", C.c_int32), ("n_cyl_min", C.c_int32),
        ("n_cyl_max", C.c_int32), ("boundary", C.c_int32), ("box", C.c_double),
        ("radius", C.c_double), ("cyl_len", C.c_double), ("k_stretch", C.c_double),
        ("k_bend", C.c_double), ("k_exvol", C.c_double), ("exvol_cutoff", C.c_double),
        ("k_boundary", C.c_double), ("screen_len", C.c_double),
        ("boundary_cutoff", C.c_double), ("wall_margin", C.c_double),
        ("n_linkers", C.c_int32), ("n_motors", C.c_int32), ("linkers_per_fil", C.c_int32),
        ("motors_per_fil", C.c_int32), ("xl_mode", C.c_int32), ("xl_max_dist", C.c_double),
        ("k_linker", C.c_double), ("k_motor", C.c_double), ("motor_eql_min", C.c_double),
        ("motor_eql_max", C.c_double), ("lognormal_sigma", C.c_double),
    ]


def status_name(code):
    return STATUS_NAMES.get(code, f"status {code}")
