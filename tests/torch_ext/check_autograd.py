"""Child process of test_torch_autograd_function_wrapper: torch first (its HIP
runtime), then libvfs.so, then the autograd wrapper against the C ABI."""
import sys
from pathlib import Path

import torch

torch.zeros(1, device="cuda")  # initialise torch's HIP runtime before libvfs.so is loaded

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE))

import numpy as np  # noqa: E402

import build as torch_ext_build  # noqa: E402
import test_gpu_adjoint as T  # noqa: E402


def main() -> int:
    mod = torch_ext_build.load()
    w, grids, p, sim, adj, target, spacing, dims = T._setup()
    loss_c, _ = adj.forward(w.x, w.v, target)
    gx_c, gv_c, gnu_c = adj.backward()
    x0 = torch.tensor(w.x, device="cuda", requires_grad=True)
    v0 = torch.tensor(w.v, device="cuda", requires_grad=True)
    nu = torch.tensor(float(p.viscosity_strenght), dtype=torch.float64, requires_grad=True)
    t = torch.tensor(target, device="cuda")
    loss = mod.wcsph_density_loss(x0, v0, nu, t, sim.handle, adj.handle)
    loss.backward()
    torch.cuda.synchronize()
    gx_t, gv_t = x0.grad.cpu().numpy(), v0.grad.cpu().numpy()
    same = {
        "loss": float(loss.detach()) == loss_c,
        "grad_x0": np.array_equal(gx_t, gx_c),
        "grad_v0": np.array_equal(gv_t, gv_c),
        "grad_viscosity": float(nu.grad) == gnu_c,
    }
    print(f"torch autograd: loss {float(loss.detach())!r} (C ABI {loss_c!r}), d/d nu {float(nu.grad)!r} "
          f"(C ABI {gnu_c!r}), max |dx| {np.abs(gx_t - gx_c).max():.3e} of {np.abs(gx_c).max():.3e}, "
          f"max |dv| {np.abs(gv_t - gv_c).max():.3e} of {np.abs(gv_c).max():.3e}, identical: {same}")
    return 0 if all(same.values()) else 1


if __name__ == "__main__":
    sys.exit(main())
