"""Oracle: Rayleigh (closed-form) cube T-matrix, complex128.  TEST INFRASTRUCTURE.

Restates cubic_scattering/effective_contrasts.py:32-537 (Gamma0, A/B/C,
T1/T2/T3, amplification factors, effective contrasts),
cubic_scattering/voigt_tmatrix.py:131-169 (6x6 effective stiffness) and
cubic_scattering/resonance_tmatrix.py:461-487 (9x9 block-diagonal T0).
"""

from __future__ import annotations

from dataclasses import dataclass
from math import factorial

import numpy as np

# effective_contrasts.py:32 -- number of Taylor terms of Phi and Psi.
N_TAYLOR = 8

# effective_contrasts.py:109 -- g0 = int_{[-1,1]^3} 1/|x| d^3x, written with the
# Pell identity 70226^2 - 3*40545^2 = 1.
G0_CUBE = (4.0 / 3.0) * np.log(70226.0 + 40545.0 * np.sqrt(3.0)) - 2.0 * np.pi


@dataclass
class Medium:
    """Isotropic background (effective_contrasts.py:41-57)."""

    alpha: float
    beta: float
    rho: float

    @property
    def lam(self) -> float:
        return self.rho * (self.alpha**2 - 2.0 * self.beta**2)

    @property
    def mu(self) -> float:
        return self.rho * self.beta**2


@dataclass
class CubeT:
    """All intermediate scalars of the cube T-matrix (effective_contrasts.py:69-100)."""

    Gamma0: complex
    Ac: complex
    Bc: complex
    Cc: complex
    T1c: complex
    T2c: complex
    T3c: complex
    amp_u: complex
    amp_theta: complex
    amp_e_off: complex
    amp_e_diag: complex
    Drho_star: complex
    Dlambda_star: complex
    Dmu_star_off: complex
    Dmu_star_diag: complex


def kelvin_a0_b0(alpha: float, beta: float, rho: float) -> tuple[float, float]:
    """Static Kelvin coefficients (effective_contrasts.py:128-129, 260-261)."""
    den = 8.0 * np.pi * rho * alpha**2 * beta**2
    return (alpha**2 + beta**2) / den, (alpha**2 - beta**2) / den


def taylor_phi_psi(omega, alpha, beta, rho, n_taylor=N_TAYLOR):
    """phi_n, psi_n of the smooth Green's tensor (effective_contrasts.py:155-199).

    phi_n = (i w/beta)^(2n+1) / ((2n+1)! 4 pi rho beta^2)
    psi_n = [(i w/alpha)^(2n+3)/alpha^2 - (i w/beta)^(2n+3)/beta^2] / ((2n+3)! 4 pi rho)
    (the second assignment at lines 193-197 is the one that is returned).
    """
    n = np.arange(n_taylor)
    kb = 1j * omega / beta
    ka = 1j * omega / alpha
    fact1 = np.array([float(factorial(2 * k + 1)) for k in n])
    fact3 = np.array([float(factorial(2 * k + 3)) for k in n])
    phi = kb ** (2 * n + 1) / fact1 / (4.0 * np.pi * rho * beta**2)
    psi = (ka ** (2 * n + 3) / alpha**2 - kb ** (2 * n + 3) / beta**2) / fact3 / (4.0 * np.pi * rho)
    return phi.astype(complex), psi.astype(complex)


def cube_moments(a: float, n_max: int):
    """S0, S1, S2, S11 monomial moments of [-a,a]^3 (effective_contrasts.py:202-238)."""
    mu = np.array([2.0 * a ** (2 * k + 1) / (2 * k + 1) for k in range(n_max + 3)])
    S0 = np.zeros(n_max + 1)
    S1 = np.zeros(n_max + 1)
    S2 = np.zeros(n_max + 1)
    S11 = np.zeros(n_max + 1)
    for m in range(n_max + 1):
        for p in range(m + 1):
            for q in range(m - p + 1):
                r = m - p - q
                c = factorial(m) / (factorial(p) * factorial(q) * factorial(r))
                S0[m] += c * mu[p] * mu[q] * mu[r]
                S1[m] += c * mu[p + 1] * mu[q] * mu[r]
                S2[m] += c * mu[p + 2] * mu[q] * mu[r]
                S11[m] += c * mu[p + 1] * mu[q + 1] * mu[r]
    return S0, S1, S2, S11


def gamma0(omega, a, alpha, beta, rho, n_taylor=N_TAYLOR) -> complex:
    """Gamma0 = int_cube G_11 (effective_contrasts.py:112-152)."""
    a0, b0 = kelvin_a0_b0(alpha, beta, rho)
    stat = a**2 * (a0 + b0 / 3.0) * G0_CUBE
    phi, psi = taylor_phi_psi(omega, alpha, beta, rho, n_taylor)
    S0, S1, _, _ = cube_moments(a, n_taylor - 1)
    return complex(stat + np.sum(phi * S0[:n_taylor]) + np.sum(psi * S1[:n_taylor]))


def eshelby_static_geom():
    """Surface-integral constants j1, j2, k1 (effective_contrasts.py:264-269)."""
    s3 = np.sqrt(3.0)
    return 2.0 * np.pi / 3.0, -2.0 * (s3 - np.pi) / 9.0, 2.0 * (2.0 * s3 + np.pi) / 9.0


def abc_static(alpha, beta, rho):
    """Static Eshelby A, B, C (effective_contrasts.py:241-275)."""
    a0, b0 = kelvin_a0_b0(alpha, beta, rho)
    j1, j2, k1 = eshelby_static_geom()
    return (2.0 * (-a0 * j1 - 3.0 * b0 * j2), 2.0 * b0 * (j1 - 3.0 * j2), 6.0 * b0 * (3.0 * j2 - k1))


def abc(omega, a, alpha, beta, rho, n_taylor=N_TAYLOR):
    """A^c, B^c, C^c = static + smooth radiation part (effective_contrasts.py:278-357)."""
    As, Bs, Cs = abc_static(alpha, beta, rho)
    phi, psi = taylor_phi_psi(omega, alpha, beta, rho, n_taylor)
    S0, S1, S2, S11 = cube_moments(a, n_taylor - 1)
    N = n_taylor
    A = 0j
    B = complex(np.sum(psi * S0[:N]))
    C = 0j
    for n in range(N):
        if n >= 2:
            A += 4 * n * (n - 1) * (phi[n] * S1[n - 2] + psi[n] * S11[n - 2])
            B += 4 * n * (n - 1) * psi[n] * S11[n - 2]
            C += 4 * n * (n - 1) * psi[n] * (S2[n - 2] - 3.0 * S11[n - 2])
        if n >= 1:
            A += 2 * n * (phi[n] * S0[n - 1] + psi[n] * S1[n - 1])
            B += 4 * n * psi[n] * S1[n - 1]
    return As + A, Bs + B, Cs + C


def t123(Ac, Bc, Cc, Dlambda, Dmu):
    """T1, T2, T3 by explicit 4-index tensor contraction (effective_contrasts.py:365-402).

    T_mnlp = sum_jk S_mnjk Dc_jklp with S_mnjk = (I_mjkn + I_njkm)/2 and
    I_ijkl = A d_ij d_kl + B (d_ik d_jl + d_il d_jk) + C E_ijkl.
    """
    d = np.eye(3)
    E = np.zeros((3, 3, 3, 3))
    for i in range(3):
        E[i, i, i, i] = 1.0
    I = (
        Ac * np.einsum("ij,kl->ijkl", d, d)
        + Bc * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("il,jk->ijkl", d, d))
        + Cc * E
    )
    S = 0.5 * (np.einsum("mjkn->mnjk", I) + np.einsum("njkm->mnjk", I))
    Dc = Dlambda * np.einsum("jk,lp->jklp", d, d) + Dmu * (
        np.einsum("jl,kp->jklp", d, d) + np.einsum("jp,kl->jklp", d, d)
    )
    T = np.einsum("mnjk,jklp->mnlp", S, Dc)
    T1 = T[0, 0, 1, 1]
    T2 = T[0, 1, 0, 1]
    T3 = T[0, 0, 0, 0] - T1 - 2.0 * T2
    return complex(T1), complex(T2), complex(T3)


def cube_tmatrix(omega, a, ref: Medium, Dlambda, Dmu, Drho, n_taylor=N_TAYLOR) -> CubeT:
    """Full Rayleigh cube T-matrix (effective_contrasts.py:454-537, 405-451)."""
    G0 = gamma0(omega, a, ref.alpha, ref.beta, ref.rho, n_taylor)
    Ac, Bc, Cc = abc(omega, a, ref.alpha, ref.beta, ref.rho, n_taylor)
    T1, T2, T3 = t123(Ac, Bc, Cc, Dlambda, Dmu)
    amp_u = 1.0 / (1.0 - omega**2 * Drho * G0)
    amp_t = 1.0 / (1.0 - 3.0 * T1 - 2.0 * T2 - T3)
    amp_o = 1.0 / (1.0 - 2.0 * T2)
    amp_d = 1.0 / (1.0 - 2.0 * T2 - T3)
    return CubeT(
        Gamma0=G0, Ac=Ac, Bc=Bc, Cc=Cc, T1c=T1, T2c=T2, T3c=T3,
        amp_u=amp_u, amp_theta=amp_t, amp_e_off=amp_o, amp_e_diag=amp_d,
        Drho_star=Drho * amp_u,
        Dlambda_star=(Dlambda + 2.0 / 3.0 * Dmu) * amp_t - 2.0 / 3.0 * Dmu * amp_d,
        Dmu_star_off=Dmu * amp_o,
        Dmu_star_diag=Dmu * amp_d,
    )


def stiffness_voigt(Dlam_s, Dmu_diag, Dmu_off) -> np.ndarray:
    """6x6 cubic Voigt stiffness contrast (voigt_tmatrix.py:131-169)."""
    Dc = np.zeros((6, 6), dtype=complex)
    Dc[:3, :3] = Dlam_s
    for i in range(3):
        Dc[i, i] = Dlam_s + 2.0 * Dmu_diag
        Dc[3 + i, 3 + i] = 2.0 * Dmu_off
    return Dc


def tmatrix_9x9(res: CubeT, omega, a) -> np.ndarray:
    """Block-diagonal 9x9 T0 (resonance_tmatrix.py:461-487), V = (2a)^3."""
    V = (2.0 * a) ** 3
    T = np.zeros((9, 9), dtype=complex)
    T[:3, :3] = omega**2 * complex(res.Drho_star) * V * np.eye(3)
    T[3:, 3:] = V * stiffness_voigt(res.Dlambda_star, res.Dmu_star_diag, res.Dmu_star_off)
    return T
