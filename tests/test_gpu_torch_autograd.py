"""GPU: the libtorch autograd functions (picgo_amd/_torch_mech.so, csrc/torch_mech.cpp) on the
42,000-bead network with per-cylinder parameters near its relaxed state: torch.autograd gives dE/dx = -F (mdn_forces), dE/dtheta
(mdn_param_grad), and a double backward gives H v (mdn_hvp); the parameter gradient through
torch matches central differences of the energy (h = 1e-4 relative)."""
import numpy as np
import pytest
import torch

from picgo_amd import abi, autograd, synth

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sysx(gpu):
    net = synth.Network.config(2)
    g = gpu()
    g.upload(net.desc, net.coord)
    g.reset_neighbor_lists(net.coord)
    x, _, _ = g.minimize(abi.default_cg_params(), net.coord)
    x = x + np.random.default_rng(22).normal(0.0, 0.01, x.size)
    return net, g, x


def test_energy_and_first_derivatives(sysx):
    net, g, x0 = sysx
    th0 = autograd.theta0(net.desc)
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    th = th0.clone().requires_grad_(True)
    E = autograd.energy(g, x, th)
    assert float(E.detach()) == pytest.approx(g.compute_energy(x0), rel=1e-14)
    gx, gt = torch.autograd.grad(E, (x, th))
    np.testing.assert_array_equal(gx.numpy(), -g.compute_forces(x0))
    np.testing.assert_array_equal(gt.numpy(), g.param_grad(x0))


def test_double_backward_is_hvp(sysx):
    net, g, x0 = sysx
    th0 = autograd.theta0(net.desc)
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    v = torch.from_numpy(np.random.default_rng(4).standard_normal(x0.size))
    E = autograd.energy(g, x, th0)
    (gx,) = torch.autograd.grad(E, x, create_graph=True)
    (hv,) = torch.autograd.grad(gx, x, grad_outputs=v)
    np.testing.assert_allclose(hv.numpy(), g.hvp(v.numpy(), x0), rtol=0, atol=1e-12 * float(hv.abs().max()))


def test_theta_gradient_vs_finite_differences(sysx):
    net, g, x0 = sysx
    th0 = autograd.theta0(net.desc)
    x = torch.tensor(x0, dtype=torch.float64)
    th = th0.clone().requires_grad_(True)
    (gt,) = torch.autograd.grad(autograd.energy(g, x, th), th)
    rng = np.random.default_rng(9)
    for kind in range(3):
        for c in rng.choice(th0.shape[1], 4, replace=False):
            h = 1e-4 * float(th0[kind, c])
            tp, tm = th0.clone(), th0.clone()
            tp[kind, c] += h
            tm[kind, c] -= h
            fd = (float(autograd.energy(g, x, tp)) - float(autograd.energy(g, x, tm))) / (2 * h)
            assert abs(fd - float(gt[kind, c])) <= 1e-5 * abs(float(gt[kind, c])) + 1e-6, (kind, c)
    autograd.energy(g, x, th0)   # restore the parameters in the context


def test_two_parameter_sets_share_one_backward(sysx):
    """Two energies of one context at different theta, summed, then one backward pass: each term's
    dE/dx, dE/dtheta and Hessian-vector product are taken at its own theta (every backward call
    reinstalls the parameters its forward ran with), identical to separate backward passes."""
    net, g, x0 = sysx
    th0 = autograd.theta0(net.desc)
    th1 = th0 * torch.exp(0.1 * torch.from_numpy(np.random.default_rng(5).standard_normal(tuple(th0.shape))))
    v = torch.from_numpy(np.random.default_rng(6).standard_normal(x0.size))

    def separate(th):
        x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
        t = th.clone().requires_grad_(True)
        gx, gt = torch.autograd.grad(autograd.energy(g, x, t), (x, t), create_graph=True)
        (hv,) = torch.autograd.grad(gx, x, grad_outputs=v)
        return gx.detach(), gt.detach(), hv

    gx0, gt0, hv0 = separate(th0)
    gx1, gt1, hv1 = separate(th1)
    x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
    ta, tb = th0.clone().requires_grad_(True), th1.clone().requires_grad_(True)
    E = autograd.energy(g, x, ta) + autograd.energy(g, x, tb)
    gx, gta, gtb = torch.autograd.grad(E, (x, ta, tb), create_graph=True)
    (hv,) = torch.autograd.grad(gx, x, grad_outputs=v)
    np.testing.assert_array_equal(gta.detach().numpy(), gt0.numpy())
    np.testing.assert_array_equal(gtb.detach().numpy(), gt1.numpy())
    np.testing.assert_array_equal(gx.detach().numpy(), (gx0 + gx1).numpy())
    np.testing.assert_array_equal(hv.numpy(), (hv0 + hv1).numpy())
    assert float(np.abs(gt1.numpy() - gt0.numpy()).max()) > 0   # the two sets differ
    autograd.energy(g, torch.tensor(x0), th0)   # restore the parameters in the context
