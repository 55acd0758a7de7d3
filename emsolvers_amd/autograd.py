"""torch.autograd binding of the FDTD misfit gradient (C ABI ems_fdtd_gradient).

    J = fdtd_misfit(eps_r, sigma, sim, num_steps, d_obs)
    J.backward()          # eps_r.grad, sigma.grad = dJ/deps_r, dJ/dsigma

``sim`` is an ``emsolvers_amd.fdtd.FemSimulation`` with its sources, receivers,
CPML and wavelet set; eps_r and sigma are tensors on the (nx+1)(ny+1)(nz+1)
material grid (any device; they are copied to the host for the C ABI).  The
forward pass runs the discrete adjoint as well and keeps the gradient for the
backward pass.  torch is used only as the autograd front end: the
computation is the HIP library.
"""
import numpy as np
import torch


class FdtdMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, eps_r, sigma, sim, num_steps, d_obs):
        er = eps_r.detach().to("cpu", torch.float64).numpy().ravel()
        sg = sigma.detach().to("cpu", torch.float64).numpy().ravel()
        sim.set_material(er, sg)
        d = d_obs.detach().to("cpu", torch.float64).numpy() if torch.is_tensor(d_obs) else np.asarray(d_obs)
        J, ge, gs = sim.gradient(int(num_steps), d)
        ctx.save_for_backward(torch.from_numpy(ge).to(eps_r.device, eps_r.dtype).reshape(eps_r.shape),
                              torch.from_numpy(gs).to(sigma.device, sigma.dtype).reshape(sigma.shape))
        return torch.tensor(J, dtype=eps_r.dtype, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_out):
        ge, gs = ctx.saved_tensors
        return grad_out * ge, grad_out * gs, None, None, None


def fdtd_misfit(eps_r, sigma, sim, num_steps, d_obs):
    return FdtdMisfit.apply(eps_r, sigma, sim, num_steps, d_obs)


class TetMisfit(torch.autograd.Function):
    """J(eps_r) of a tetrahedral run (per-tetrahedron eps_r; lossless
    symplectic model), backward = the discrete adjoint of ems_tet_gradient."""

    @staticmethod
    def forward(ctx, eps_r, sim, num_steps, d_obs):
        er = eps_r.detach().to("cpu", torch.float64).numpy().ravel()
        sim.set_material(er, None)
        d = d_obs.detach().to("cpu", torch.float64).numpy() if torch.is_tensor(d_obs) else np.asarray(d_obs)
        J, ge, _, _ = sim.gradient(int(num_steps), d)
        ctx.save_for_backward(torch.from_numpy(ge).to(eps_r.device, eps_r.dtype).reshape(eps_r.shape))
        return torch.tensor(J, dtype=eps_r.dtype, device=eps_r.device)

    @staticmethod
    def backward(ctx, grad_out):
        (ge,) = ctx.saved_tensors
        return grad_out * ge, None, None, None


def tet_misfit(eps_r, sim, num_steps, d_obs):
    """``sim``: an ``emsolvers_amd.tet.TetSimulation`` that has not run yet
    (the material is set before each evaluation)."""
    return TetMisfit.apply(eps_r, sim, num_steps, d_obs)
