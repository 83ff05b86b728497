"""libtorch autograd functions of the mechanical energy (the in-tree extension _torch_mech.so,
built from csrc/torch_mech.cpp by __graft_entry__.build()).

    E = energy(ffm, x, theta)
        the energy of the network held by ForceFieldManager ``ffm`` at coordinates x
        [3 n_beads] with per-cylinder parameters theta [3, n_cylinders] = (stretching k,
        rest length L0, bending k); differentiable in x (dE/dx = -F) and theta (dE/dtheta), and
        twice in x: a double backward yields Hessian-vector products (mdn_hvp).
    theta0(desc) -> the parameters of a system description as a (3, n_cylinders) tensor.

The forward pass installs theta in the context (mdn_set_params).
"""
import numpy as np
import torch

from . import mech


def _ext():
    try:
        from . import _torch_mech
    except ImportError as e:   # the product path has no fallback
        raise RuntimeError("picgo_amd/_torch_mech.so is not built; run __graft_entry__.build()") from e
    return _torch_mech


def energy(ffm, x, theta):
    return _ext().mech_energy(x, theta, ffm._h.value)


def grad_x(ffm, x):
    """dE/dx = -F(x) as a differentiable function of x (its backward is H v)."""
    return _ext().mech_grad_x(x, ffm._h.value)


def theta0(desc):
    nc = desc.fil_bead_offset[desc.n_filaments] - desc.n_filaments
    th = np.array([[desc.cyl_kstr[i] for i in range(nc)], [desc.cyl_eql[i] for i in range(nc)],
                   [desc.cyl_kbend[i] for i in range(nc)]], dtype=np.float64)
    return torch.from_numpy(th)


__all__ = ["energy", "grad_x", "theta0", "mech"]
