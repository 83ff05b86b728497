"""Conditioning of the excluded-volume closed form (CylinderExclVolRepulsion.cpp:330-858) at a
given state, for the GPU tests' per-coordinate bars: the closed form divides a bracket that
vanishes like tau^2 by JJ ~ -tau^2 (tau: the pair's normalized triple product), so rounding
differences of any origin (the device's vs the host's atan, or two operation orders of the same
formula) are amplified like 1/tau^2 on the beads of nearly coplanar pairs."""
import numpy as np


def pair_tau(g, net, x):
    """(bead indices [n_pairs, 4], normalized triple product tau) of the excluded-volume pairs."""
    d = net.desc
    cyl = np.array([(b, b + 1) for f in range(d.n_filaments)
                    for b in range(d.fil_bead_offset[f], d.fil_bead_offset[f + 1] - 1)])
    pairs = g.neighbor_pairs()
    X = x.reshape(-1, 3)
    b4 = np.concatenate([cyl[pairs[:, 0]], cyl[pairs[:, 1]]], axis=1)
    u1 = X[b4[:, 1]] - X[b4[:, 0]]
    u2 = X[b4[:, 3]] - X[b4[:, 2]]
    u3 = X[b4[:, 0]] - X[b4[:, 2]]
    tau = np.abs(np.einsum("ij,ij->i", np.cross(u1, u2), u3)) / np.sqrt(
        np.einsum("ij,ij->i", u1, u1) * np.einsum("ij,ij->i", u2, u2) * np.einsum("ij,ij->i", u3, u3))
    return b4, tau


def exvol_bead_bars(g, net, x, base):
    """Per-coordinate bar for excluded-volume forces from two evaluations of the closed form:
    `base`, grown like 1/tau^2 on the beads of pairs whose normalized triple product tau is below
    1e-3 (a pair with tau = 7e-5 moves by 0.08 pN under random 1-ulp atan changes on the host)."""
    b4, tau = pair_tau(g, net, x)
    grow = np.ones(x.size // 3)
    for k in range(4):
        np.maximum.at(grow, b4[:, k], np.maximum(1.0, (1e-3 / np.maximum(tau, 1e-12)) ** 2))
    return base * np.repeat(grow, 3)


def well_conditioned(g, net, x, tau_min=1e-3):
    """Coordinate mask: False on the beads of excluded-volume pairs with tau < tau_min, whose
    forces carry rounding noise of order eps / tau^2 (which a finite difference divides by 2h)."""
    b4, tau = pair_tau(g, net, x)
    ok = np.ones(x.size // 3, bool)
    ok[b4[tau < tau_min].ravel()] = False
    return np.repeat(ok, 3)
