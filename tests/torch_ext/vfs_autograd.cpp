// Test-harness wrapper of the differentiable substeps as a libtorch
// torch::autograd::Function: loss(x0, v0, viscosity; target) with gradients
// from vfs_adjoint_backward.  Built against libvfs.so by __graft_entry__.build().
#include <torch/cuda.h>
#include <torch/extension.h>

#include <stdexcept>
#include <string>

#include "vfs_c.h"

namespace {

void check(int rc, const char* what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + vfs_last_error());
}

class WcsphDensityLoss : public torch::autograd::Function<WcsphDensityLoss> {
public:
    // x0, v0: (n, 3) float32 on the GPU; viscosity: float64 scalar on the CPU;
    // target: grid values, float32 on the GPU.  sim/adj: vfs_sim* and
    // vfs_adjoint* as integers.
    static torch::Tensor forward(torch::autograd::AutogradContext* ctx, torch::Tensor x0, torch::Tensor v0,
                                 torch::Tensor viscosity, torch::Tensor target, int64_t sim, int64_t adj) {
        TORCH_CHECK(x0.is_cuda() && v0.is_cuda() && target.is_cuda(), "state and target must be on the GPU");
        TORCH_CHECK(x0.scalar_type() == torch::kFloat32 && v0.scalar_type() == torch::kFloat32,
                    "state must be float32");
        auto* s = reinterpret_cast<vfs_sim*>(sim);
        auto* a = reinterpret_cast<vfs_adjoint*>(adj);
        vfs_params p;
        check(vfs_sim_get_params(s, &p), "vfs_sim_get_params");
        p.viscosity_strenght = (float)viscosity.item<double>();
        check(vfs_sim_set_params(s, &p), "vfs_sim_set_params");
        auto x = x0.contiguous(), v = v0.contiguous(), t = target.to(torch::kFloat32).contiguous();
        torch::cuda::synchronize();  // torch's work on the inputs precedes the simulation stream
        double loss = 0.0;
        check(vfs_adjoint_forward(a, x.data_ptr<float>(), v.data_ptr<float>(), t.data_ptr<float>(), nullptr, &loss),
              "vfs_adjoint_forward");
        ctx->saved_data["adj"] = adj;
        ctx->saved_data["n"] = x0.size(0);
        return torch::tensor(loss, torch::dtype(torch::kFloat64));
    }

    static torch::autograd::tensor_list backward(torch::autograd::AutogradContext* ctx,
                                                 torch::autograd::tensor_list grad_out) {
        auto* a = reinterpret_cast<vfs_adjoint*>(ctx->saved_data["adj"].toInt());
        const int64_t n = ctx->saved_data["n"].toInt();
        auto opts = torch::dtype(torch::kFloat32).device(torch::kCUDA);
        auto gx = torch::empty({n, 3}, opts), gv = torch::empty({n, 3}, opts);
        double gnu = 0.0;
        torch::cuda::synchronize();
        check(vfs_adjoint_backward(a, gx.data_ptr<float>(), gv.data_ptr<float>(), &gnu), "vfs_adjoint_backward");
        const double scale = grad_out[0].item<double>();
        return {gx * scale, gv * scale, torch::tensor(gnu * scale, torch::dtype(torch::kFloat64)), torch::Tensor(),
                torch::Tensor(), torch::Tensor()};
    }
};

torch::Tensor wcsph_density_loss(torch::Tensor x0, torch::Tensor v0, torch::Tensor viscosity, torch::Tensor target,
                                 int64_t sim, int64_t adj) {
    return WcsphDensityLoss::apply(x0, v0, viscosity, target, sim, adj);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("wcsph_density_loss", &wcsph_density_loss,
          "density-grid loss of n substeps from (x0, v0) with viscosity; differentiable in all three");
}
