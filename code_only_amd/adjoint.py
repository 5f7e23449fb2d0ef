"""Differentiable periodic Poisson solve for torch: forward through the multigrid solver, backward
through its hand-written adjoint (sg_uniform_adjoint_dev), both behind libsg's C ABI.

    phi  = PoissonSolve.apply(rho, solver)          # phi = L^+ P (4 pi G rho), G = 1/(4 pi)
    gs   = SampledGradient.apply(phi, solver, cells) # (npts, 3) centred grad phi at cells
    loss = (gs * w).sum()
    loss.backward()                                  # rho.grad = P L^+ P (grad_h^T W)

The operator L (7-point, periodic) is symmetric, so the adjoint solve is the forward solver applied
to the adjoint source; P (mean removal, the ZERO_AVERAGE_DENSITY projection of
fmg_data_if.F90:53-78) is symmetric too.  SampledGradient's backward is the transpose kernel
sg_uniform_gradient_sample_adjoint_dev.  AmrPoissonSolve does the same for the sphere-boundary AMR
solve, whose composite operator is not symmetric: its backward solves the transposed problem
(sg_amr_adjoint_dev).
"""
import torch

from ._lib import SG_PREC_MIXED, SG_STOP_TRUE_RESIDUAL


class PoissonSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, solver, tol=1e-10, precision=SG_PREC_MIXED):
        src = (rho - rho.mean()).contiguous()
        phi = torch.zeros_like(src)
        solver.cycle(phi, src, tol=tol, maxit=60, stop_mode=SG_STOP_TRUE_RESIDUAL, precision=precision)
        ctx.solver, ctx.tol, ctx.precision = solver, tol, precision
        return phi

    @staticmethod
    def backward(ctx, gphi):
        grad, _ = ctx.solver.adjoint(gphi.contiguous(), tol=ctx.tol, maxit=60, precision=ctx.precision)
        return grad, None, None, None


class SampledGradient(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi, solver, cells):
        ctx.solver = solver
        ctx.save_for_backward(cells)
        return solver.gradient_sample(phi.contiguous(), cells)

    @staticmethod
    def backward(ctx, gout):
        (cells,) = ctx.saved_tensors
        return ctx.solver.gradient_sample_adjoint(cells, gout.contiguous()), None, None


def sampled_gradient_loss(rho, solver, cells, w, tol=1e-10, precision=SG_PREC_MIXED):
    """L = sum_p w_p . grad_h phi(c_p) with phi the periodic potential of rho."""
    phi = PoissonSolve.apply(rho, solver, tol, precision)
    return (SampledGradient.apply(phi, solver, cells) * w).sum()


class AmrPoissonSolve(torch.autograd.Function):
    """psi = AmrPoissonSolve.apply(src, solver): the sphere-boundary AMR solve (AmrSolver.solve from a zero
    guess, dense [nblk, 8, 8, 8] float64 GPU tensors); backward through sg_amr_adjoint_dev."""

    @staticmethod
    def forward(ctx, src, solver, tol=1e-10, precision=SG_PREC_MIXED):
        psi = torch.zeros_like(src)
        solver.solve(psi, src.contiguous(), tol=tol, maxit=60, stop_mode=SG_STOP_TRUE_RESIDUAL, precision=precision)
        ctx.solver, ctx.tol, ctx.precision = solver, tol, precision
        return psi

    @staticmethod
    def backward(ctx, gpsi):
        grad, _ = ctx.solver.adjoint(gpsi.contiguous(), tol=ctx.tol, precision=ctx.precision)
        return grad, None, None, None
