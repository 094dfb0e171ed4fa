"""PyTorch autograd wrapper: rho(P, R_A) on grid points, differentiable with
respect to the density matrix P and the atomic centres R_A.

PyTorch is only the plumbing here: tensors carry device memory and the
autograd graph; forward and backward are the library's HIP kernels called
through the C ABI (``ct_eval`` / ``ct_backward``) with the tensors' device
pointers.

    dL/dP_ij = sum_g (dL/drho_g) phi_i(g) phi_j(g)
    dL/dR_A  = -2 sum_{i on A} sum_g (dL/drho_g) grad phi_i(g) (P phi(g))_i

(The reference has no differentiation; these are new entry points.)

Losses on differences of densities, e.g. sum_g w_g (rho_g - rho_t,g)^2 with
rho_t from P_t, are best written on ``density(P - P_t, R, points)``: rho is
linear in P, so this is the same function, and the contraction of the
difference keeps the mixed path's FP32-level accuracy relative to the residual
itself rather than to rho.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi
from .basis import BasisSet


class DensityModel:
    """Owns one C-ABI handle for a basis set; call ``density(P, R, points)``."""

    def __init__(self, basis: BasisSet, charges=None, precision: str = "fp64"):
        if precision not in capi.PREC:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.basis = basis
        self.precision = precision
        M = basis.nbasis
        Z = np.zeros(basis.natom) if charges is None else np.asarray(charges, dtype=np.float64)
        self.handle = capi.Handle(basis.flat(), np.zeros((M, M)), Z)

    def _upload(self, P: torch.Tensor, R: torch.Tensor) -> None:
        """Makes the handle's P and centres those of this call (always: a tensor
        address can be reused by the caching allocator for different values)."""
        if P.shape != (self.basis.nbasis, self.basis.nbasis) or R.shape != (self.basis.natom, 3):
            raise ValueError("P must be (nbasis, nbasis) and R (natom, 3)")
        if P.is_cuda or R.is_cuda:
            torch.cuda.current_stream().synchronize()
        self.handle.set_density_matrix(P.data_ptr())
        self.handle.set_coordinates(R.data_ptr())

    def density(self, P: torch.Tensor, R: torch.Tensor, points: torch.Tensor) -> torch.Tensor:
        return _Density.apply(P, R, points, self)


def _as_f64(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P, R, points, model: DensityModel):
        P64, R64, pts = _as_f64(P), _as_f64(R), _as_f64(points)
        if pts.dim() != 2 or pts.shape[1] != 3:
            raise ValueError("points must have shape (N, 3)")
        model._upload(P64, R64)
        n = pts.shape[0]
        rho = torch.empty(n, dtype=torch.float64, device=pts.device)
        if pts.is_cuda:
            torch.cuda.current_stream().synchronize()
        model.handle.eval(pts.data_ptr(), n, {"rho": rho.data_ptr()}, model.precision)
        ctx.model = model
        ctx.save_for_backward(P64, R64, pts)
        return rho

    @staticmethod
    def backward(ctx, grad_rho):
        P64, R64, pts = ctx.saved_tensors
        model = ctx.model
        model._upload(P64, R64)
        g = _as_f64(grad_rho)
        dev = P64.device
        M, natom = P64.shape[0], R64.shape[0]
        dP = torch.empty((M, M), dtype=torch.float64, device=dev)
        dR = torch.empty((natom, 3), dtype=torch.float64, device=R64.device)
        if g.is_cuda or pts.is_cuda:
            torch.cuda.current_stream().synchronize()
        model.handle.backward(pts.data_ptr(), pts.shape[0], g.data_ptr(), dP.data_ptr(), dR.data_ptr(),
                              model.precision)
        return dP, dR, None, None


def density(model: DensityModel, P: torch.Tensor, R: torch.Tensor, points: torch.Tensor) -> torch.Tensor:
    """Functional form of ``DensityModel.density``."""
    return _Density.apply(P, R, points, model)
