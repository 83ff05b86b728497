// torch_mech.cpp -- libtorch autograd functions over the C ABI of include/mdn_mech.h.
//
// mech_energy(x, theta, ctx) is the mechanical energy of the network held by the context
// (ForceFieldManager::computeEnergy, ForceFieldManager.cpp:96-207) as a differentiable function
// of the bead coordinates x [3 n_beads] and of the per-cylinder parameters
// theta [3, n_cylinders] = (stretching k, rest length L0, bending k):
//   dE/dx     = -F(x)                 (mdn_forces; MechGradX, itself differentiable)
//   dE/dtheta = mdn_param_grad
//   d/dx (g . dE/dx) = H g            (mdn_hvp: the backward of MechGradX, so a double backward
//                                      through the energy yields Hessian-vector products)
// Every call into the context (forward and backward) first installs the theta it is evaluated
// at (mdn_set_params), so several energies of one context at different theta can share one
// backward pass. Mixed second derivatives d2E/dx dtheta are not provided (a backward that needs
// them raises). Tensors may live on any device; the C ABI takes FP64 host buffers, so inputs are
// staged through FP64 CPU copies.
#include <torch/extension.h>

#include "../../include/mdn_mech.h"

using torch::autograd::AutogradContext;
using torch::autograd::tensor_list;

namespace {

mdn_ctx* ctx_of(int64_t h) { return reinterpret_cast<mdn_ctx*>(h); }

void check(mdn_ctx* c, int st, const char* what) {
    TORCH_CHECK(st == MDN_OK, what, " failed: ", mdn_last_error(c), " (status ", st, ")");
}

torch::Tensor host64(const torch::Tensor& t) { return t.detach().to(torch::kCPU, torch::kFloat64).contiguous(); }

int64_t n_var(mdn_ctx* c) {
    int64_t counts[8];
    check(c, mdn_system_counts(c, counts), "mdn_system_counts");
    return 3 * counts[0];
}

// theta as an FP64 host copy of shape (3, n_cylinders), or undefined (keep the context's)
torch::Tensor theta_host(mdn_ctx* c, const torch::Tensor& theta) {
    if (!theta.defined()) return torch::Tensor();
    torch::Tensor th = host64(theta);
    int64_t counts[8];
    check(c, mdn_system_counts(c, counts), "mdn_system_counts");
    TORCH_CHECK(th.dim() == 2 && th.size(0) == 3 && th.size(1) == counts[1],
                "theta must have shape (3, n_cylinders = ", counts[1], ")");
    return th;
}

void install(mdn_ctx* c, const torch::Tensor& th) {
    if (!th.defined()) return;
    check(c, mdn_set_params(c, th[0].contiguous().data_ptr<double>(), th[1].contiguous().data_ptr<double>(),
                            th[2].contiguous().data_ptr<double>()),
          "mdn_set_params");
}

// dE/dx = -F(x) at theta (undefined: the context's parameters); backward: the Hessian-vector
// product with the incoming gradient.
struct MechGradX : public torch::autograd::Function<MechGradX> {
    static torch::Tensor forward(AutogradContext* actx, torch::Tensor x, torch::Tensor theta, int64_t h) {
        mdn_ctx* c = ctx_of(h);
        const int64_t n = n_var(c);
        TORCH_CHECK(x.numel() == n, "x must hold 3 * n_beads = ", n, " values");
        torch::Tensor th = theta_host(c, theta);
        install(c, th);
        actx->save_for_backward({x});
        actx->saved_data["h"] = h;
        if (th.defined()) actx->saved_data["theta"] = th;
        torch::Tensor xc = host64(x);
        torch::Tensor f = torch::empty_like(xc);
        check(c, mdn_forces(c, xc.data_ptr<double>(), f.data_ptr<double>(), (int32_t)n), "mdn_forces");
        return (-f).reshape(x.sizes()).to(x.device(), x.scalar_type());
    }
    static tensor_list backward(AutogradContext* actx, tensor_list grad) {
        TORCH_CHECK(!actx->needs_input_grad(1),
                    "mixed second derivatives d2E/dx dtheta are not provided by the mechanical energy");
        mdn_ctx* c = ctx_of(actx->saved_data["h"].toInt());
        if (actx->saved_data.count("theta")) install(c, actx->saved_data["theta"].toTensor());
        torch::Tensor x = actx->get_saved_variables()[0];
        torch::Tensor xc = host64(x), g = host64(grad[0]);
        torch::Tensor hv = torch::empty_like(xc);
        check(c, mdn_hvp(c, xc.data_ptr<double>(), g.data_ptr<double>(), hv.data_ptr<double>(), (int32_t)xc.numel()),
              "mdn_hvp");
        return {hv.reshape(x.sizes()).to(x.device(), x.scalar_type()), torch::Tensor(), torch::Tensor()};
    }
};

struct MechEnergy : public torch::autograd::Function<MechEnergy> {
    static torch::Tensor forward(AutogradContext* actx, torch::Tensor x, torch::Tensor theta, int64_t h) {
        mdn_ctx* c = ctx_of(h);
        const int64_t n = n_var(c);
        TORCH_CHECK(x.numel() == n, "x must hold 3 * n_beads = ", n, " values");
        torch::Tensor th = theta_host(c, theta);
        TORCH_CHECK(th.defined(), "theta is required");
        const int64_t nc = th.size(1);
        install(c, th);
        torch::Tensor xc = host64(x);
        mdn_energy_report rep;
        check(c, mdn_energy(c, xc.data_ptr<double>(), &rep), "mdn_energy");
        double e = 0.0;   // computeEnergy: a force field at or below -1 makes the total +inf
        for (int ff = 0; ff < MDN_NUM_FF; ff++) {
            if (rep.ff[ff] <= -1) {
                e = INFINITY;
                break;
            }
            e += rep.ff[ff];
        }
        actx->save_for_backward({x, theta});
        actx->saved_data["h"] = h;
        actx->saved_data["nc"] = nc;
        actx->saved_data["theta"] = th;
        return torch::tensor(e, torch::dtype(torch::kFloat64)).to(x.device());
    }
    static tensor_list backward(AutogradContext* actx, tensor_list grad) {
        const int64_t h = actx->saved_data["h"].toInt();
        mdn_ctx* c = ctx_of(h);
        auto saved = actx->get_saved_variables();
        torch::Tensor x = saved[0], theta = saved[1];
        torch::Tensor go = grad[0];
        // the forward's parameters, whatever the context holds now
        const torch::Tensor th = actx->saved_data["theta"].toTensor();
        torch::Tensor gx = MechGradX::apply(x, th, h) * go;   // differentiable: H g on a second pass
        install(c, th);
        const int64_t nc = actx->saved_data["nc"].toInt();
        torch::Tensor xc = host64(x);
        torch::Tensor gt = torch::empty({3, nc}, torch::dtype(torch::kFloat64));
        check(c, mdn_param_grad(c, xc.data_ptr<double>(), gt.data_ptr<double>(), 3 * nc), "mdn_param_grad");
        gt = gt.to(theta.device(), theta.scalar_type()) * go.to(theta.device());
        return {gx, gt, torch::Tensor()};
    }
};

torch::Tensor mech_energy(torch::Tensor x, torch::Tensor theta, int64_t ctx) { return MechEnergy::apply(x, theta, ctx); }
torch::Tensor mech_grad_x(torch::Tensor x, int64_t ctx) { return MechGradX::apply(x, torch::Tensor(), ctx); }

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "libtorch autograd functions over the mechanical-equilibration C ABI";
    m.def("mech_energy", &mech_energy, "E(x, theta) of the network of a context (handle as int)");
    m.def("mech_grad_x", &mech_grad_x, "dE/dx = -F(x), differentiable (backward: H v)");
}
