"""torch.autograd bridge for the receiver misfit of an IBIM solve.

    J(level, k) = sum_r |u(x_r) - t_r|^2,   u = layer potential of the density
    that solves the traditional IBIM system built from `level` at wavenumber k.

Forward runs the whole device path (WENO geometry, tube, operator, BiCG,
evaluation); backward runs the hand-written adjoint in libibim.so (an A^H
BiCG solve, operator-entry derivatives recomputed tile by tile from the tube,
then back through the quadrature weights, projection, normals, curvature
regularisation, WENO gradient and limited second derivatives). The level set
may live on the CPU or on the GPU; gradients come back on its device.
"""
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

import ibim


@dataclass
class MisfitSpec:
    """What the misfit is computed on (everything but the level set and k)."""
    receivers: np.ndarray                 # (m, 3)
    target: np.ndarray                    # (m,) complex
    eps: float                            # tube half-width
    tau: float                            # near-field radius
    lo: float = -2.0
    interior: int = 0
    dirichlet: int = 1
    one_sided: int = 0
    weno_order: int = 5
    rhs_kind: int = ibim.RHS_POINT_SOURCE
    x0: Sequence[float] = (0.0, 0.0, 0.0)
    amp: float = 1.0
    direction: Sequence[float] = (0.0, 0.0, 1.0)
    mode: int = ibim.EVAL_DOUBLE
    assemble: bool = True                 # store the FP64 operator (else matrix-free)
    tol: float = 1e-12                    # forward and adjoint BiCG tolerance
    maxit: int = 2000
    last_info: Optional[object] = field(default=None, repr=False)


def _forward(level_np, k, spec):
    pb = ibim.Problem()
    pb.set_level(level_np, lo=spec.lo)
    pb.geometry(spec.weno_order)
    pb.tube(ibim.params(k, spec.eps, spec.tau, interior=spec.interior, dirichlet=spec.dirichlet,
                        one_sided=spec.one_sided))
    if spec.assemble:
        pb.assemble(ibim.C128)
    pb.rhs(spec.rhs_kind, x0=spec.x0, amp=spec.amp, direction=spec.direction)
    info, _ = pb.solve(tol=spec.tol, maxit=spec.maxit)
    if info.status not in (0, 4):
        raise ibim.IbimError(f"forward BiCG did not converge (status {info.status})")
    spec.last_info = info
    pb.set_receivers(spec.receivers, spec.target)
    re, im = pb.evaluate(spec.receivers, spec.mode)
    u = re + 1j * im
    return pb, float(np.sum(np.abs(u - np.asarray(spec.target)) ** 2))


class IbimMisfit(torch.autograd.Function):
    """J = IbimMisfit.apply(level, k, spec): level (n+1)^3 float64, k a 0-d float64 tensor."""

    @staticmethod
    def forward(ctx, level, k, spec):
        level_np = level.detach().to("cpu", torch.float64).contiguous().numpy()
        pb, J = _forward(level_np, float(k), spec)
        ctx.pb, ctx.spec = pb, spec
        ctx.level_device, ctx.k_device = level.device, k.device
        return torch.tensor(J, dtype=torch.float64, device=level.device)

    @staticmethod
    def backward(ctx, gJ):
        spec = ctx.spec
        _, dlevel, dk = ctx.pb.gradient(mode=spec.mode, tol=spec.tol, maxit=spec.maxit)
        ctx.pb.close()
        g = float(gJ)
        gl = torch.from_numpy(dlevel * g).to(ctx.level_device)
        gk = torch.tensor(dk * g, dtype=torch.float64, device=ctx.k_device)
        return gl, gk, None


def misfit(level, k, spec):
    """Differentiable receiver misfit (see IbimMisfit)."""
    if not torch.is_tensor(k):
        k = torch.tensor(float(k), dtype=torch.float64)
    return IbimMisfit.apply(level, k, spec)
