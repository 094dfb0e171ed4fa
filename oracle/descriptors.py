"""Density-based descriptor epilogues restated in numpy (FP64).

TEST INFRASTRUCTURE ONLY (the checker for the product's fused epilogues).
Each function follows the reference arithmetic exactly, including its masking
quirks, with the reference file:line it restates.
"""
from __future__ import annotations

import numpy as np

MASK = 1e-20  # divide_inplace mask, reference src/cuda_utils.cu:71-84


def masked_divide(num, den, mask=MASK):
    """num/den where den > mask, else the mask value itself (src/cuda_utils.cu:71-84)."""
    num = np.asarray(num, dtype=np.float64)
    den = np.asarray(den, dtype=np.float64)
    out = np.full(np.broadcast(num, den).shape, mask)
    ok = den > mask
    out[ok] = num[ok] / den[ok]
    return out


def norm(v):
    """|v| per row (src/evaluate_densbased.cu:10-41)."""
    return np.sqrt(np.sum(np.asarray(v) ** 2, axis=1))


def reduced_density_gradient(rho, grad):
    """s = 1/2 (3 pi^2)^{-1/3} |grad rho| / rho^{4/3} (src/evaluate_densbased.cu:44-89)."""
    pref = 0.5 / (3.0 * np.pi ** 2) ** (1.0 / 3.0)
    return pref * masked_divide(norm(grad), np.power(rho, 4.0 / 3.0))


def shannon(rho):
    """rho ln rho, no guard (src/evaluate_densbased.cu:210-223)."""
    return rho * np.log(rho)


def weizsacker(rho, grad):
    """|grad rho|^2 / (8 rho) (src/evaluate_densbased.cu:91-140)."""
    return masked_divide(np.sum(grad ** 2, axis=1), 8.0 * rho)


def thomas_fermi(rho):
    """0.3 (3 pi^2)^{2/3} rho^{5/3} (src/evaluate_densbased.cu:142-173)."""
    return 0.3 * (3.0 * np.pi ** 2) ** (2.0 / 3.0) * np.power(rho, 5.0 / 3.0)


def gradient_expansion_general(rho, grad, lap, a, b):
    """tau_TF + a tau_W + b lap (src/evaluate_densbased.cu:175-193)."""
    return thomas_fermi(rho) + a * weizsacker(rho, grad) + b * lap


def general_ked(tau_pd, lap, a):
    """Python-bound general(ish) KED: tau_PD + (a-1)/4 lap (src/evaluate_densbased.cu:195-208)."""
    return tau_pd + (a - 1.0) / 4.0 * lap


def general_kinetic_energy_density(tau_pd, lap, alpha):
    """C++ t_alpha = tau_PD + alpha lap (src/evaluate_kinetic_dens.cu:195-214)."""
    return tau_pd + alpha * lap
