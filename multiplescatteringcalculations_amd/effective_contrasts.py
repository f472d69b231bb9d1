"""Background medium, contrasts and the Rayleigh cube T-matrix (host API).

Mirrors the dataclasses and ``compute_cube_tmatrix`` of
cubic_scattering/effective_contrasts.py:41-100, 454-537.  The frequency
constants Gamma0, A^c, B^c, C^c come from the native library
(``msc_cube_coefficients``); the contrast-dependent closed forms below are the
same ones the device T0 kernel evaluates for every voxel.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import _lib


@dataclass
class ReferenceMedium:
    """Isotropic elastic background (effective_contrasts.py:41-57)."""

    alpha: float  # P-wave velocity (m/s)
    beta: float  # S-wave velocity (m/s)
    rho: float  # density (kg/m^3)

    @property
    def lam(self) -> float:
        """Lame parameter lambda = rho (alpha^2 - 2 beta^2)."""
        return self.rho * (self.alpha**2 - 2 * self.beta**2)

    @property
    def mu(self) -> float:
        """Shear modulus mu = rho beta^2."""
        return self.rho * self.beta**2


@dataclass
class MaterialContrast:
    """Isotropic contrast relative to the background (effective_contrasts.py:60-66)."""

    Dlambda: float  # (Pa)
    Dmu: float  # (Pa)
    Drho: float  # (kg/m^3)


@dataclass
class CubeTMatrixResult:
    """All scalars of the cube T-matrix (effective_contrasts.py:69-100)."""

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

    @property
    def cubic_anisotropy(self) -> complex:
        return self.Dmu_star_diag - self.Dmu_star_off


def cube_coefficients(omega: float, a: float, ref: ReferenceMedium) -> tuple[complex, ...]:
    """(Gamma0, A^c, B^c, C^c) for half-width ``a`` at ``omega`` (native, FP64)."""
    lib = _lib.load()
    out = (_lib.c_double * 8)()
    _lib.check(lib.msc_cube_coefficients(float(omega), float(a), float(ref.alpha),
                                         float(ref.beta), float(ref.rho), out))
    return tuple(complex(out[2 * i], out[2 * i + 1]) for i in range(4))


def compute_cube_tmatrix(omega: float, a: float, ref: ReferenceMedium,
                         contrast: MaterialContrast, n_gauss: int = 32,
                         n_taylor: int = 8) -> CubeTMatrixResult:
    """Self-consistent cubic T-matrix of one cube (effective_contrasts.py:454-537).

    ``n_gauss`` is accepted for signature compatibility and unused (the
    reference also integrates analytically); ``n_taylor`` must be 8.
    """
    if n_taylor != 8:
        raise ValueError("this build evaluates the Taylor series with n_taylor = 8")
    G0, A, B, C = cube_coefficients(omega, a, ref)
    dl, dm, dr = float(contrast.Dlambda), float(contrast.Dmu), float(contrast.Drho)
    # T_mnlp = sum_jk S_mnjk Dc_jklp contracted in closed form
    T1 = dl * (A + 4 * B + C) + 2 * B * dm
    T2 = (A + B) * dm
    T3 = 2 * C * dm
    amp_u = 1.0 / (1.0 - omega**2 * dr * G0)
    amp_t = 1.0 / (1.0 - 3.0 * T1 - 2.0 * T2 - T3)
    amp_o = 1.0 / (1.0 - 2.0 * T2)
    amp_d = 1.0 / (1.0 - 2.0 * T2 - T3)
    return CubeTMatrixResult(
        Gamma0=G0, Ac=A, Bc=B, Cc=C, T1c=T1, T2c=T2, T3c=T3,
        amp_u=amp_u, amp_theta=amp_t, amp_e_off=amp_o, amp_e_diag=amp_d,
        Drho_star=dr * amp_u,
        Dlambda_star=(dl + 2.0 / 3.0 * dm) * amp_t - 2.0 / 3.0 * dm * amp_d,
        Dmu_star_off=dm * amp_o,
        Dmu_star_diag=dm * amp_d,
    )


def effective_stiffness_voigt(Dlambda_star: complex, Dmu_star_diag: complex,
                              Dmu_star_off: complex) -> np.ndarray:
    """6x6 cubic Voigt stiffness contrast (voigt_tmatrix.py:131-169)."""
    Dc = np.zeros((6, 6), dtype=complex)
    Dc[:3, :3] = Dlambda_star
    for i in range(3):
        Dc[i, i] = Dlambda_star + 2.0 * Dmu_star_diag
        Dc[3 + i, 3 + i] = 2.0 * Dmu_star_off
    return Dc


def sub_cell_tmatrix_9x9(res: CubeTMatrixResult, omega: float, a_sub: float) -> np.ndarray:
    """Block-diagonal 9x9 T0 (resonance_tmatrix.py:461-487)."""
    V = (2.0 * a_sub) ** 3
    T = np.zeros((9, 9), dtype=complex)
    T[:3, :3] = omega**2 * complex(res.Drho_star) * V * np.eye(3)
    T[3:, 3:] = V * effective_stiffness_voigt(res.Dlambda_star, res.Dmu_star_diag,
                                              res.Dmu_star_off)
    return T
