"""CPU oracle of the CPA iteration -- TEST INFRASTRUCTURE ONLY.

Restates cubic_scattering/cpa_iteration.py:137-318 (Voigt start, per-phase
Rayleigh cube T-matrix in the current medium with mu = mu_off, update of
lambda*, mu*_off, rho* by the phase-averaged real effective contrasts,
mu*_diag carried as mu*_off plus the accumulated anisotropy, residual on the
three isotropic channels, optional under-relaxation) on the oracle's cube
T-matrix.  Pinned by the CPA condition itself (the converged medium makes the
averaged effective contrasts vanish), the single-phase fixed point and the
weak-contrast Voigt limit (tests/test_cpa.py).
"""

from __future__ import annotations

import numpy as np

from . import tmatrix as tm


def medium(lam, mu, rho):
    return tm.Medium(float(np.sqrt((lam + 2 * mu) / rho)), float(np.sqrt(mu / rho)), rho)


def cpa(phases, omega, a, max_iter=100, tol=1e-10, damping=1.0):
    """phases: [(lam, mu, rho, phi)]; returns (lam, mu_off, mu_diag, rho, iters, converged)."""
    lam = sum(p[0] * p[3] for p in phases)
    mu = sum(p[1] * p[3] for p in phases)
    rho = sum(p[2] * p[3] for p in phases)
    eff = [lam, mu, mu, rho]
    for it in range(max_iter):
        ref = medium(eff[0], eff[1], eff[3])
        s = np.zeros(4)
        for la, m, r, f in phases:
            t = tm.cube_tmatrix(omega, a, ref, la - eff[0], m - eff[1], r - eff[3])
            s += f * np.array([t.Dlambda_star.real, t.Dmu_star_off.real, t.Dmu_star_diag.real,
                               t.Drho_star.real])
        mo = eff[1] + s[1]
        new = [eff[0] + s[0], mo, mo + (s[2] - s[1]) + (eff[2] - eff[1]), eff[3] + s[3]]
        sm = max(abs(eff[0]) + 2 * abs(eff[1]), 1.0)
        sr = max(abs(eff[3]), 1.0)
        res = max(abs(s[3]) / sr, abs(s[0]) / sm, abs(s[1]) / sm)
        if res < tol:
            return (*new, it + 1, True)
        eff = [e + damping * (n - e) for e, n in zip(eff, new)] if damping < 1.0 else new
    return (*eff, max_iter, False)
