"""PyTorch autograd wrapper: rho(P, R_A) on grid points, differentiable with
respect to the density matrix P and the atomic centres R_A.

PyTorch is only the plumbing here: tensors carry device memory and the
autograd graph; forward and backward are the library's HIP kernels called
through the C ABI (``ct_eval`` / ``ct_backward``) with the tensors' device
pointers.

    dL/dP_ij = sum_g (dL/drho_g) phi_i(g) phi_j(g)
    dL/dR_A  = -2 sum_{i on A} sum_g (dL/drho_g) grad phi_i(g) (P phi(g))_i

(The reference has no differentiation; these are new entry points.)

Losses on residuals, e.g. sum_g w_g (rho_g - rho_t,g)^2: their gradients
need rho - rho_t, not rho, to the target accuracy.  Pass the target as
``density(P, R, points, rho_target=rho_t)``: the mixed forward then holds its
error to 1e-6 (``DensityModel.tol``) of |rho - rho_t| per point, re-evaluating
in FP64 the points whose error bound misses that (``ct_eval_params.rho_target``).

The handle keeps P and the centres between calls: setting values equal to the
current ones (the backward after its forward, a loop that passes the same
tensors) costs a copy and a comparison, not the factorisation of P.
"""
from __future__ import annotations

import numpy as np
import torch

from . import capi
from .basis import BasisSet


class DensityModel:
    """Owns one C-ABI handle for a basis set; call ``density(P, R, points)``."""

    def __init__(self, basis: BasisSet, charges=None, precision: str = "fp64", tol: float = 1e-6):
        if precision not in capi.PREC:
            raise ValueError("precision must be 'fp64' or 'mixed'")
        self.basis = basis
        self.precision = precision
        self.tol = float(tol)
        self.last_report = None  # ct_eval_report of the last forward (points re-evaluated in FP64, ...)
        M = basis.nbasis
        Z = np.zeros(basis.natom) if charges is None else np.asarray(charges, dtype=np.float64)
        self.handle = capi.Handle(basis.flat(), np.zeros((M, M)), Z)

    def _upload(self, P: torch.Tensor, R: torch.Tensor) -> None:
        """Makes the handle's P and centres those of this call (the library
        compares them with what it holds and skips the work when they are
        equal; a tensor address alone proves nothing, the caching allocator
        reuses them)."""
        if P.shape != (self.basis.nbasis, self.basis.nbasis) or R.shape != (self.basis.natom, 3):
            raise ValueError("P must be (nbasis, nbasis) and R (natom, 3)")
        if P.is_cuda or R.is_cuda:
            torch.cuda.current_stream().synchronize()
        self.handle.set_density_matrix(P.data_ptr())
        self.handle.set_coordinates(R.data_ptr())

    def density(self, P: torch.Tensor, R: torch.Tensor, points: torch.Tensor,
                rho_target: torch.Tensor | None = None) -> torch.Tensor:
        """rho at the points, differentiable in P and R.  rho_target (mixed
        precision): the error is held to tol * |rho - rho_target| per point."""
        return _Density.apply(P, R, points, self, rho_target)


def _as_f64(t: torch.Tensor) -> torch.Tensor:
    return t.detach().to(torch.float64).contiguous()


class _Density(torch.autograd.Function):
    @staticmethod
    def forward(ctx, P, R, points, model: DensityModel, rho_target=None):
        P64, R64, pts = _as_f64(P), _as_f64(R), _as_f64(points)
        if pts.dim() != 2 or pts.shape[1] != 3:
            raise ValueError("points must have shape (N, 3)")
        model._upload(P64, R64)
        n = pts.shape[0]
        rho = torch.empty(n, dtype=torch.float64, device=pts.device)
        params = {"mixed_tol": model.tol}
        if rho_target is not None:
            tgt = _as_f64(rho_target).to(pts.device)
            if tgt.shape != (n,):
                raise ValueError("rho_target must have one entry per point")
            params["rho_target"] = tgt.data_ptr()
        if pts.is_cuda:
            torch.cuda.current_stream().synchronize()
        model.last_report = model.handle.eval(pts.data_ptr(), n, {"rho": rho.data_ptr()}, model.precision, params)
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
        return dP, dR, None, None, None


def density(model: DensityModel, P: torch.Tensor, R: torch.Tensor, points: torch.Tensor,
            rho_target: torch.Tensor | None = None) -> torch.Tensor:
    """Functional form of ``DensityModel.density``."""
    return _Density.apply(P, R, points, model, rho_target)
