"""Differentiable projectors: torch.autograd.Function wrappers whose
backward is the hand-written matched projector (no autodiff through the
ray march).

The cone forward projector is y = D A' x with D = diag(d_conv)
(c2D.cpp:616-623) and the cone backprojector is A'^T D y (c2D.cpp:1105-1117),
so the pair is an exact transpose pair of the same discrete model and the
backward of either one is the other.  Both run on torch's current stream.
"""

from __future__ import annotations

import torch

from .projectors import Geometry, backward_project, forward_project


class ForwardProjection(torch.autograd.Function):
    """y = A x; dL/dx = A^T dL/dy."""

    @staticmethod
    def forward(ctx, vol: torch.Tensor, geom: Geometry) -> torch.Tensor:
        ctx.geom = geom
        return forward_project(geom, vol.detach().contiguous())

    @staticmethod
    def backward(ctx, grad_y: torch.Tensor):
        return backward_project(ctx.geom, grad_y.detach().contiguous()), None


class BackProjection(torch.autograd.Function):
    """x = A^T y; dL/dy = A dL/dx."""

    @staticmethod
    def forward(ctx, sino: torch.Tensor, geom: Geometry) -> torch.Tensor:
        ctx.geom = geom
        return backward_project(geom, sino.detach().contiguous())

    @staticmethod
    def backward(ctx, grad_x: torch.Tensor):
        return forward_project(ctx.geom, grad_x.detach().contiguous()), None


def project(vol: torch.Tensor, geom: Geometry) -> torch.Tensor:
    return ForwardProjection.apply(vol, geom)


def backproject(sino: torch.Tensor, geom: Geometry) -> torch.Tensor:
    return BackProjection.apply(sino, geom)


def data_fidelity(vol: torch.Tensor, geom: Geometry, b: torch.Tensor) -> torch.Tensor:
    """0.5 ||A x - b||^2 with the residual norm accumulated in float64."""
    r = project(vol, geom) - b
    return 0.5 * (r.double() ** 2).sum()


def tv_isotropic(vol: torch.Tensor, eps: float = 1e-8) -> torch.Tensor:
    """Isotropic total variation sum_i sqrt(|grad x|_i^2 + eps) with forward
    differences and replicated (zero-gradient) boundary."""
    dx = torch.diff(vol, dim=0, append=vol[-1:, :, :])
    dy = torch.diff(vol, dim=1, append=vol[:, -1:, :])
    dz = torch.diff(vol, dim=2, append=vol[:, :, -1:])
    return torch.sqrt(dx.double() ** 2 + dy.double() ** 2 + dz.double() ** 2 + eps).sum()


def tikhonov(vol: torch.Tensor) -> torch.Tensor:
    """0.5 ||L x||^2 with L the 6-neighbour 3-D Laplacian (zero boundary): a
    smoothness penalty for gradient-based reconstruction.  The reference's
    Tikhonov term (Core/src/tikhonov.cpp:4-32: the 4-neighbour xy stencil
    with mirrored faces, used by cgls_tikhonov) is projectors.regularize."""
    p = torch.nn.functional.pad(vol[None, None], (1, 1, 1, 1, 1, 1))[0, 0]
    lap = (6.0 * vol - p[:-2, 1:-1, 1:-1] - p[2:, 1:-1, 1:-1] - p[1:-1, :-2, 1:-1]
           - p[1:-1, 2:, 1:-1] - p[1:-1, 1:-1, :-2] - p[1:-1, 1:-1, 2:])
    return 0.5 * (lap.double() ** 2).sum()
