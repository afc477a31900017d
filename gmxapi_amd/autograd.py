"""Differentiable nonbonded forces for torch.

`NonbondedForces` is a torch.autograd.Function whose forward is the pair
kernel (nbx_compute, forces only) and whose backward is the hand-written
adjoint kernel (nbx_compute_adjoint): for a loss L(F) it returns dL/dx (the
Hessian-vector product with dL/dF), dL/dq per atom and dL/dC6, dL/dC12 for
the type-pair matrices.  The pair list is the context's current one (built by
put_on_grid + construct_pairlist); the backward recomputes every pair from it
and stores nothing per pair.  Per-type sigma/epsilon enter through
`lj_matrices`, nblib's geometric combination written in torch, so autograd
carries the matrix gradients back to them.

The reference's kernels are forward only; there is no counterpart to cite for
the backward.  The forward is GmxForceCalculator::compute
(api/nblib/gmxcalculator.cpp:70-83) with parameters set as
NbvSetupUtil::setNonBondedParameters (api/nblib/gmxsetup.cpp:143-166) and
nonbonded_verlet_t::setAtomProperties (src/gromacs/nbnxm/nbnxm.h:268-272) do.
"""
from __future__ import annotations

import torch

from . import nbx


def _sqrt_zero_grad_at_zero(c: torch.Tensor) -> torch.Tensor:
    """sqrt(c) for c >= 0 whose gradient at c = 0 is taken as 0 instead of
    inf (the sqrt is evaluated on 1 where c = 0 and the result masked)."""
    pos = c > 0
    return torch.sqrt(torch.where(pos, c, torch.ones_like(c))) * pos


def lj_matrices(sigma: torch.Tensor, epsilon: torch.Tensor):
    """Per-type sigma, epsilon -> (C6, C12) type-pair matrices with nblib's
    geometric rule on C6 and C12 (api/nblib/interactions.cpp:83-93):
    C_ij = sqrt(C_i C_j), C6_i = 4 eps_i sigma_i^6, C12_i = 4 eps_i sigma_i^12.

    Written as the outer product of sqrt(C_i), so a type with epsilon = 0
    (water hydrogens) gives zero C rows without a sqrt of 0 on the gradient
    path: its dL/dsigma is 0 (C does not depend on it) and its dL/depsilon,
    infinite at epsilon = 0 (d sqrt(eps)/d eps), is reported as 0."""
    s6 = _sqrt_zero_grad_at_zero(4.0 * epsilon * sigma ** 6)
    s12 = _sqrt_zero_grad_at_zero(4.0 * epsilon * sigma ** 12)
    return torch.outer(s6, s6), torch.outer(s12, s12)


def _apply_parameters(nb: nbx.Context, q: torch.Tensor, c6: torch.Tensor, c12: torch.Tensor):
    nb.set_stream(torch.cuda.current_stream(q.device).cuda_stream)
    nb.set_nonbonded_parameters(c6.detach().double().cpu().numpy(), c12.detach().double().cpu().numpy())
    q32 = q.detach().to(torch.float32).contiguous()
    nb.set_charges_device(q32.data_ptr())
    return q32


class NonbondedForces(torch.autograd.Function):
    """F = NonbondedForces.apply(x, q, c6, c12, context).

    x: (N, 3) positions on the context's device (used as float32); q: (N,)
    charges on that device; c6, c12: (T, T) symmetric type-pair matrices
    (any device, used as float64 -> float32 as the forward kernel does).
    Returns the (N, 3) float32 forces in atom order."""

    @staticmethod
    def forward(ctx, x, q, c6, c12, nb: nbx.Context):
        x32 = x.detach().to(torch.float32).contiguous()
        q32 = _apply_parameters(nb, q, c6, c12)
        f = torch.empty_like(x32)
        nb.compute_device(x32.data_ptr(), f.data_ptr(), nbx.COMPUTE_FORCES)
        ctx.save_for_backward(x32, q32, c6.detach(), c12.detach())
        ctx.nb = nb
        ctx.q_dtype = q.dtype
        ctx.x_dtype = x.dtype
        return f

    @staticmethod
    def backward(ctx, grad_f):
        x32, q32, c6, c12 = ctx.saved_tensors
        nb = ctx.nb
        # the context may have been used with other parameters since forward
        nb.set_stream(torch.cuda.current_stream(x32.device).cuda_stream)
        nb.set_nonbonded_parameters(c6.double().cpu().numpy(), c12.double().cpu().numpy())
        nb.set_charges_device(q32.data_ptr())
        g = grad_f.detach().to(torch.float32).contiguous()
        need_x, need_q, need_c6, need_c12 = ctx.needs_input_grad[:4]
        gx = torch.empty_like(x32) if need_x else None
        gq = torch.empty(x32.shape[0], dtype=torch.float64, device=x32.device) if need_q else None
        t = c6.shape[0]
        g6 = torch.empty((t, t), dtype=torch.float64, device=x32.device) if need_c6 else None
        g12 = torch.empty((t, t), dtype=torch.float64, device=x32.device) if need_c12 else None
        nb.compute_adjoint_device(x32.data_ptr(), g.data_ptr(), gx.data_ptr() if need_x else None,
                                  gq.data_ptr() if need_q else None, g6.data_ptr() if need_c6 else None,
                                  g12.data_ptr() if need_c12 else None)
        return (gx.to(ctx.x_dtype) if need_x else None, gq.to(ctx.q_dtype) if need_q else None,
                g6.to(device=c6.device, dtype=c6.dtype) if need_c6 else None,
                g12.to(device=c12.device, dtype=c12.dtype) if need_c12 else None, None)


def nonbonded_forces(x, q, c6, c12, nb: nbx.Context):
    """Functional form of NonbondedForces.apply."""
    return NonbondedForces.apply(x, q, c6, c12, nb)


def force_matching_loss(x, q, sigma, epsilon, f_target, nb: nbx.Context):
    """L = sum_i |F_i(sigma, epsilon, q, x) - F_target,i|^2 in float64."""
    c6, c12 = lj_matrices(sigma, epsilon)
    f = nonbonded_forces(x, q, c6, c12, nb)
    d = f.to(torch.float64) - f_target.to(torch.float64)
    return (d * d).sum()


def context_for_system(system, device: int = 0, **options) -> nbx.Context:
    """A context on `system` (gmxapi_amd.systems.System) with grid and pair
    list built, for use with NonbondedForces."""
    return nbx.context_for(system, device=device, **options)


__all__ = ["NonbondedForces", "nonbonded_forces", "force_matching_loss", "lj_matrices", "context_for_system"]
