"""Coherent potential approximation for a space-filling lattice of cubes.

Drop-in for cubic_scattering/cpa_iteration.py: starting from the Voigt
average, the effective medium (lambda*, mu*_off, mu*_diag, rho*) is moved by
the phase-averaged effective contrasts of the Rayleigh cube T-matrix computed
in the current medium until <Drho*> = <Dlambda*> = <Dmu*_off> = 0
(cpa_iteration.py:165-318; the cubic anisotropy mu*_diag - mu*_off is carried
along as a derived quantity).  Each step is a handful of cube T-matrices
(closed forms with the native frequency constants of msc_cube_coefficients),
so the iteration runs on the host.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .effective_contrasts import (
    CubeTMatrixResult,
    MaterialContrast,
    ReferenceMedium,
    compute_cube_tmatrix,
)


@dataclass
class Phase:
    """One material phase (cpa_iteration.py:31-62)."""

    lam: float
    mu: float
    rho: float
    volume_fraction: float

    @property
    def alpha(self) -> float:
        return float(np.sqrt((self.lam + 2.0 * self.mu) / self.rho))

    @property
    def beta(self) -> float:
        return float(np.sqrt(self.mu / self.rho))

    @property
    def bulk_modulus(self) -> float:
        return self.lam + 2.0 * self.mu / 3.0


@dataclass
class CubicEffectiveMedium:
    """Cubic effective medium (cpa_iteration.py:65-113)."""

    lam: float
    mu_off: float
    mu_diag: float
    rho: float

    @property
    def mu_iso(self) -> float:
        return (self.mu_off + self.mu_diag) / 2.0

    @property
    def bulk_modulus(self) -> float:
        return self.lam + 2.0 * self.mu_diag / 3.0

    @property
    def cubic_anisotropy(self) -> float:
        return self.mu_diag - self.mu_off

    def as_reference_medium(self) -> ReferenceMedium:
        """Isotropic reference with mu = mu_off."""
        return ReferenceMedium(alpha=float(np.sqrt((self.lam + 2.0 * self.mu_off) / self.rho)),
                               beta=float(np.sqrt(self.mu_off / self.rho)), rho=self.rho)


@dataclass
class CPAResult:
    effective_medium: CubicEffectiveMedium
    n_iterations: int
    converged: bool
    residual_history: list[float]
    tmatrix_results: list[CubeTMatrixResult]
    phases: list[Phase]
    omega: float
    a: float


def voigt_average(phases: list[Phase]) -> CubicEffectiveMedium:
    lam = sum(p.lam * p.volume_fraction for p in phases)
    mu = sum(p.mu * p.volume_fraction for p in phases)
    rho = sum(p.rho * p.volume_fraction for p in phases)
    return CubicEffectiveMedium(lam=lam, mu_off=mu, mu_diag=mu, rho=rho)


def _step(phases, eff: CubicEffectiveMedium, omega: float, a: float):
    """One update: T of every phase in the current medium, then <T> = 0 on the
    isotropic channels (cpa_iteration.py:165-240)."""
    ref = eff.as_reference_medium()
    s_rho = s_lam = s_off = s_diag = 0.0
    res = []
    for ph in phases:
        t = compute_cube_tmatrix(omega, a, ref, MaterialContrast(
            Dlambda=ph.lam - eff.lam, Dmu=ph.mu - eff.mu_off, Drho=ph.rho - eff.rho))
        res.append(t)
        f = ph.volume_fraction
        s_rho += f * t.Drho_star.real
        s_lam += f * t.Dlambda_star.real
        s_off += f * t.Dmu_star_off.real
        s_diag += f * t.Dmu_star_diag.real
    mu_off = eff.mu_off + s_off
    new = CubicEffectiveMedium(lam=eff.lam + s_lam, mu_off=mu_off,
                               mu_diag=mu_off + (s_diag - s_off) + eff.cubic_anisotropy,
                               rho=eff.rho + s_rho)
    scale_m = max(abs(eff.lam) + 2.0 * abs(eff.mu_off), 1.0)
    scale_r = max(abs(eff.rho), 1.0)
    resid = max(abs(s_rho) / scale_r, abs(s_lam) / scale_m, abs(s_off) / scale_m)
    return new, resid, res


def compute_cpa(phases: list[Phase], omega: float, a: float, max_iter: int = 100,
                tol: float = 1e-10, damping: float = 1.0,
                initial_medium: CubicEffectiveMedium | None = None) -> CPAResult:
    """Self-consistent cubic effective medium (cpa_iteration.py:243-318)."""
    total = sum(p.volume_fraction for p in phases)
    if abs(total - 1.0) > 1e-10:
        raise ValueError(f"Volume fractions must sum to 1.0, got {total}")
    eff = initial_medium if initial_medium is not None else voigt_average(phases)
    hist: list[float] = []
    res: list[CubeTMatrixResult] = []
    for it in range(max_iter):
        new, r, res = _step(phases, eff, omega, a)
        hist.append(r)
        if r < tol:
            return CPAResult(new, it + 1, True, hist, res, phases, omega, a)
        if damping < 1.0:
            eff = CubicEffectiveMedium(
                lam=eff.lam + damping * (new.lam - eff.lam),
                mu_off=eff.mu_off + damping * (new.mu_off - eff.mu_off),
                mu_diag=eff.mu_diag + damping * (new.mu_diag - eff.mu_diag),
                rho=eff.rho + damping * (new.rho - eff.rho))
        else:
            eff = new
    return CPAResult(eff, max_iter, False, hist, res, phases, omega, a)


def phases_from_two_phase(ref: ReferenceMedium, contrast: MaterialContrast,
                          volume_fraction: float) -> list[Phase]:
    """Matrix (ref, 1 - phi) and inclusion (ref + contrast, phi) (cpa_iteration.py:321-352)."""
    return [Phase(ref.lam, ref.mu, ref.rho, 1.0 - volume_fraction),
            Phase(ref.lam + contrast.Dlambda, ref.mu + contrast.Dmu, ref.rho + contrast.Drho,
                  volume_fraction)]


def compute_cpa_two_phase(ref: ReferenceMedium, contrast: MaterialContrast,
                          volume_fraction: float, omega: float, a: float, **kwargs) -> CPAResult:
    return compute_cpa(phases_from_two_phase(ref, contrast, volume_fraction), omega, a, **kwargs)
