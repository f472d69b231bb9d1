"""Kennett reflectivity of a stack of plane layers (host side).

Mirrors the layer API of cubic_scattering/kennett_layers.py (FluidLayer,
IsotropicLayer, LayerStack 30-180; KennettResult 209-250; kennett_layers
739-760) that the ocean-bottom caller (ocean_bottom.py:116-156, 245-300)
wraps around the slab solves.  This is per-frequency 2x2 algebra on a few
layers, so it runs on the host in float64/complex128.

Interface coefficients are computed here from the boundary conditions
themselves rather than from closed forms: in each medium the four plane-wave
modes (P and SV, down- and upgoing, time dependence e^{-i w t}, z downwards)
are written as displacement-traction vectors
    b = (u_x, u_z, sigma_xz / (i w), sigma_zz / (i w))
with eigenvectors normalised to unit vertical energy flux,
    P: u = (p, +-eta) / sqrt(rho eta),   SV: u = (+-nu, -p) / sqrt(rho nu) * i,
and continuity of b (solid-solid) or of u_z, sigma_zz with sigma_xz = 0
(fluid over solid) gives a small linear solve for the reflection and
transmission of every incident mode.  In this basis the coefficients are the
"modified" (flux-normalised) ones of kennett_layers.py:290-475: unitary for
real slownesses, Tu = Td^T, and PP equal to the displacement reflection
coefficient.  The recursion is Kennett's upward sweep
    R_D <- R_d + T_u M (I - R_u M)^{-1} T_d,   M = E R_D E
(kennett_layers.py:549-665), E the one-way phase of the layer below.
"""

from __future__ import annotations

from collections.abc import Sequence
from dataclasses import dataclass, field

import numpy as np


@dataclass
class FluidLayer:
    """Acoustic layer (kennett_layers.py:30-62): no shear wave."""

    alpha: float
    rho: float
    thickness: float
    Q_alpha: float = np.inf
    beta: float = 0.0
    Q_beta: float = np.inf

    def __post_init__(self) -> None:
        if self.alpha <= 0:
            raise ValueError(f"alpha must be positive, got {self.alpha}")
        if self.rho <= 0:
            raise ValueError(f"rho must be positive, got {self.rho}")
        if self.thickness <= 0:
            raise ValueError(f"thickness must be positive, got {self.thickness}")
        if self.beta != 0.0:
            raise ValueError(f"FluidLayer beta must be 0, got {self.beta}")


@dataclass
class IsotropicLayer:
    """Isotropic elastic layer (kennett_layers.py:65-103)."""

    alpha: float
    beta: float
    rho: float
    thickness: float
    Q_alpha: float = np.inf
    Q_beta: float = np.inf

    def __post_init__(self) -> None:
        if self.alpha <= 0:
            raise ValueError(f"alpha must be positive, got {self.alpha}")
        if self.beta <= 0:
            raise ValueError(f"beta must be positive, got {self.beta}")
        if self.rho <= 0:
            raise ValueError(f"rho must be positive, got {self.rho}")
        if self.thickness <= 0:
            raise ValueError(f"thickness must be positive, got {self.thickness}")
        if not np.isinf(self.Q_alpha) and self.Q_alpha <= 0:
            raise ValueError(f"Q_alpha must be positive or inf, got {self.Q_alpha}")
        if not np.isinf(self.Q_beta) and self.Q_beta <= 0:
            raise ValueError(f"Q_beta must be positive or inf, got {self.Q_beta}")


@dataclass
class LayerStack:
    """Layers top to bottom; the last one is the half-space (kennett_layers.py:106-179)."""

    layers: Sequence[FluidLayer | IsotropicLayer] = field(default_factory=list)

    def __post_init__(self) -> None:
        if len(self.layers) < 2:
            raise ValueError(f"Need >= 2 layers, got {len(self.layers)}")
        if not np.isinf(self.layers[-1].thickness):
            raise ValueError("Last layer must be half-space (thickness=inf), "
                             f"got {self.layers[-1].thickness}")

    @property
    def n_layers(self) -> int:
        return len(self.layers)

    @classmethod
    def from_arrays(cls, alpha, beta, rho, thickness, Q_alpha=None, Q_beta=None) -> "LayerStack":
        n = len(alpha)
        qa = np.full(n, np.inf) if Q_alpha is None else Q_alpha
        qb = np.full(n, np.inf) if Q_beta is None else Q_beta
        return cls(layers=[IsotropicLayer(float(alpha[i]), float(beta[i]), float(rho[i]),
                                          float(thickness[i]), float(qa[i]), float(qb[i]))
                           for i in range(n)])

    @classmethod
    def homogeneous(cls, alpha, beta, rho, n_layers=3, thickness=100.0) -> "LayerStack":
        lays = [IsotropicLayer(alpha, beta, rho, thickness) for _ in range(n_layers - 1)]
        lays.append(IsotropicLayer(alpha, beta, rho, np.inf))
        return cls(layers=lays)


@dataclass
class PSVCoefficients:
    """Flux-normalised P-SV interface coefficients, index 0 = P, 1 = SV."""

    Rd: np.ndarray
    Ru: np.ndarray
    Td: np.ndarray
    Tu: np.ndarray


@dataclass
class SHCoefficients:
    Rd: complex
    Ru: complex
    Td: complex
    Tu: complex


@dataclass
class KennettResult:
    """Cumulative reflection at the top interface (kennett_layers.py:209-250)."""

    RD_psv: np.ndarray  # (nfreq, 2, 2)
    RD_sh: np.ndarray  # (nfreq,)
    omega: np.ndarray
    p: float
    stack: LayerStack

    @property
    def RPP(self) -> np.ndarray:
        return self.RD_psv[:, 0, 0]

    @property
    def RSS(self) -> np.ndarray:
        return self.RD_psv[:, 1, 1]

    @property
    def RPS(self) -> np.ndarray:
        return self.RD_psv[:, 0, 1]

    @property
    def RSP(self) -> np.ndarray:
        return self.RD_psv[:, 1, 0]

    @property
    def RSH(self) -> np.ndarray:
        return self.RD_sh


def complex_slowness(velocity: float, Q: float) -> complex:
    """1/v, or for finite Q the constant-Q slowness (2Q)^2/((1+(2Q)^2) v) (1 + i/(2Q))
    (kennett_layers.py:255-270)."""
    if np.isinf(Q):
        return complex(1.0 / velocity)
    q2 = 2.0 * Q
    return complex(q2 * q2, q2) / ((1.0 + q2 * q2) * velocity)


def vertical_slowness(slowness: complex, p: float) -> complex:
    """sqrt(s^2 - p^2) on the branch Im > 0 (Re > 0 when real) (kennett_layers.py:273-287)."""
    eta = np.sqrt(complex((slowness + p) * (slowness - p)))
    if eta.imag < 0.0 or (eta.imag == 0.0 and eta.real < 0.0):
        eta = -eta
    return eta


# reference spellings
_complex_slowness = complex_slowness
_vertical_slowness = vertical_slowness


# ── modes and interface solves ──────────────────────────────────────────


def _psv_modes(p, eta, nu, rho, beta):
    """(down, up) 4x2 matrices of flux-normalised P, SV displacement-traction vectors.

    A fluid has nu = 0: its SV columns are zero and unused."""
    mu = complex(rho) * beta * beta
    down = np.zeros((4, 2), complex)
    up = np.zeros((4, 2), complex)
    cp = 1.0 / np.sqrt(complex(rho) * eta)
    for sgn, out in ((1.0, down), (-1.0, up)):
        q = sgn * eta
        ux, uz = p * cp, q * cp
        # rho (p u_x + q u_z) / s_p^2 = rho c_P: lambda div + 2 mu u_z,z, with
        # lambda = rho alpha^2 - 2 mu and alpha^2 (p^2 + eta^2) = 1
        out[:, 0] = (ux, uz, mu * (q * ux + p * uz), complex(rho) * cp - 2.0 * mu * p * ux)
    if nu != 0:
        cs = 1j / np.sqrt(complex(rho) * nu)
        for sgn, out in ((1.0, down), (-1.0, up)):
            q = sgn * nu
            ux, uz = sgn * nu * cs, -p * cs
            # div u = i w (p u_x + q u_z) = 0 for SV
            out[:, 1] = (ux, uz, mu * (q * ux + p * uz), 2.0 * mu * q * uz)
    return down, up


def _medium(layer, p):
    sp = complex_slowness(layer.alpha, layer.Q_alpha)
    eta = vertical_slowness(sp, p)
    if isinstance(layer, FluidLayer):
        return eta, 0j, 0j
    ss = complex_slowness(layer.beta, layer.Q_beta)
    return eta, vertical_slowness(ss, p), 1.0 / ss


def _interface(p, above, below) -> PSVCoefficients:
    fluid_above = isinstance(above, FluidLayer)
    fluid_below = isinstance(below, FluidLayer)
    if fluid_below:
        raise ValueError("fluid-fluid or solid-fluid not supported. "
                         f"Got {type(above).__name__} over {type(below).__name__}.")
    e1, n1, b1 = _medium(above, p)
    e2, n2, b2 = _medium(below, p)
    d1, u1 = _psv_modes(p, e1, n1, above.rho, b1)
    d2, u2 = _psv_modes(p, e2, n2, below.rho, b2)
    if fluid_above:
        # u_z and sigma_zz continuous, sigma_xz = 0 below; u_x may slip
        rows = [1, 2, 3]
        m1 = 1
    else:
        rows = [0, 1, 2, 3]
        m1 = 2
    A = np.concatenate([u1[rows, :m1], -d2[rows, :]], axis=1)
    Rd = np.zeros((2, 2), complex)
    Td = np.zeros((2, 2), complex)
    Ru = np.zeros((2, 2), complex)
    Tu = np.zeros((2, 2), complex)
    sol_d = np.linalg.solve(A, -d1[rows, :m1])  # incident from above, modes 0..m1-1
    Rd[:m1, :m1] = sol_d[:m1]
    Td[:, :m1] = sol_d[m1:]
    sol_u = np.linalg.solve(A, u2[rows, :])  # incident from below
    Tu[:m1, :] = sol_u[:m1]
    Ru[:, :] = sol_u[m1:]
    return PSVCoefficients(Rd=Rd, Ru=Ru, Td=Td, Tu=Tu)


def psv_solid_solid(p, eta1, neta1, rho1, beta1, eta2, neta2, rho2, beta2) -> PSVCoefficients:
    """Flux-normalised P-SV coefficients of a solid-solid interface
    (replaces kennett_layers.py:290-386), from the boundary conditions."""
    d1, u1 = _psv_modes(p, eta1, neta1, rho1, beta1)
    d2, u2 = _psv_modes(p, eta2, neta2, rho2, beta2)
    A = np.concatenate([u1, -d2], axis=1)
    sd = np.linalg.solve(A, -d1)
    su = np.linalg.solve(A, u2)
    return PSVCoefficients(Rd=sd[:2].copy(), Td=sd[2:].copy(), Tu=su[:2].copy(),
                           Ru=su[2:].copy())


def psv_fluid_solid(p, eta1, rho1, eta2, neta2, rho2, beta2) -> PSVCoefficients:
    """Fluid over solid (replaces kennett_layers.py:389-474): Rd only [0,0],
    Td column 1 and Tu row 1 zero."""
    d1, u1 = _psv_modes(p, eta1, 0j, rho1, 0j)
    d2, u2 = _psv_modes(p, eta2, neta2, rho2, beta2)
    rows = [1, 2, 3]
    A = np.concatenate([u1[rows, :1], -d2[rows, :]], axis=1)
    sd = np.linalg.solve(A, -d1[rows, :1])
    su = np.linalg.solve(A, u2[rows, :])
    z = np.zeros((2, 2), complex)
    Rd, Td, Ru, Tu = z.copy(), z.copy(), z.copy(), z.copy()
    Rd[0, 0] = sd[0, 0]
    Td[:, 0] = sd[1:, 0]
    Tu[0, :] = su[0]
    Ru[:, :] = su[1:]
    return PSVCoefficients(Rd=Rd, Ru=Ru, Td=Td, Tu=Tu)


def sh_solid_solid(neta1, rho1, beta1, neta2, rho2, beta2) -> SHCoefficients:
    """SH coefficients (replaces kennett_layers.py:477-506) from continuity of u_y
    and sigma_yz with modes u_y = 1/sqrt(Z), Z = mu nu."""
    z1 = complex(rho1) * beta1 * beta1 * neta1
    z2 = complex(rho2) * beta2 * beta2 * neta2
    # [u_y; s] columns: down (1, Z)/sqrt Z, up (1, -Z)/sqrt Z
    r1, r2 = np.sqrt(z1), np.sqrt(z2)
    A = np.array([[1.0 / r1, -1.0 / r2], [-z1 / r1, -z2 / r2]])
    rd, td = np.linalg.solve(A, -np.array([1.0 / r1, z1 / r1]))
    tu, ru = np.linalg.solve(A, np.array([1.0 / r2, -z2 / r2]))
    return SHCoefficients(Rd=complex(rd), Ru=complex(ru), Td=complex(td), Tu=complex(tu))


# ── recursion ───────────────────────────────────────────────────────────


def _phases(layer, eta, nu, omega):
    if np.isinf(layer.thickness):
        one = np.ones_like(omega, dtype=complex)
        return one, one
    ea = np.exp(1j * omega * eta * layer.thickness)
    eb = (np.exp(1j * omega * nu * layer.thickness) if not isinstance(layer, FluidLayer)
          else np.ones_like(ea))
    return ea, eb


def kennett_layers(stack: LayerStack, p: float, omega) -> KennettResult:
    """P-SV (2x2) and SH reflection of the stack seen from its top medium
    (kennett_layers.py:739-760); omega may be complex (damped), shape (nfreq,)."""
    omega = np.asarray(omega)
    if np.isrealobj(omega):
        omega = omega.astype(np.float64)
    lays = stack.layers
    nf = omega.shape[0]
    media = [_medium(l, p) for l in lays]
    coefs = [_interface(p, lays[i], lays[i + 1]) for i in range(len(lays) - 1)]
    R = np.zeros((nf, 2, 2), complex)
    Rsh = np.zeros(nf, complex)
    eye = np.eye(2)
    for i in range(len(lays) - 2, -1, -1):
        below = lays[i + 1]
        eta, nu, beta = media[i + 1]
        ea, eb = _phases(below, eta, nu, omega)
        E = np.zeros((nf, 2, 2), complex)
        E[:, 0, 0] = ea
        E[:, 1, 1] = eb
        M = E @ R @ E
        c = coefs[i]
        R = c.Rd + c.Tu @ M @ np.linalg.solve(eye - c.Ru @ M, np.broadcast_to(c.Td, M.shape))
        if not (isinstance(lays[i], FluidLayer) or isinstance(below, FluidLayer)):
            e_a, n_a, b_a = media[i]
            sh = sh_solid_solid(n_a, lays[i].rho, b_a, nu, below.rho, beta)
            Msh = eb * eb * Rsh
            Rsh = sh.Rd + sh.Tu * Msh * sh.Td / (1.0 - sh.Ru * Msh)
    return KennettResult(RD_psv=R, RD_sh=Rsh, omega=omega, p=p, stack=stack)


# ── CPA bridge (kennett_layers.py:916-1097) ─────────────────────────────


def cubic_to_isotropic_layer(eff, thickness: float, Q_alpha: float = np.inf,
                             Q_beta: float = np.inf) -> IsotropicLayer:
    """Isotropic layer from a CPA medium with mu = mu_off."""
    return IsotropicLayer(alpha=float(np.sqrt((eff.lam + 2.0 * eff.mu_off) / eff.rho)),
                          beta=float(np.sqrt(eff.mu_off / eff.rho)), rho=float(eff.rho),
                          thickness=thickness, Q_alpha=Q_alpha, Q_beta=Q_beta)


def cpa_stack_from_phases(layer_phases, omega: float, a: float, thickness: float | None = None,
                          Q_alpha: float = np.inf, Q_beta: float = np.inf,
                          half_space_index: int = -1, **cpa_kwargs) -> LayerStack:
    """One CPA-homogenised layer per phase list; thickness defaults to one cube (2a)."""
    from .cpa_iteration import compute_cpa

    h0 = 2.0 * a if thickness is None else thickness
    n = len(layer_phases)
    lays = []
    for i, ph in enumerate(layer_phases):
        eff = compute_cpa(ph, omega, a, **cpa_kwargs).effective_medium
        half = (i == n - 1) if half_space_index == -1 else (i == half_space_index)
        lays.append(cubic_to_isotropic_layer(eff, np.inf if half else h0, Q_alpha, Q_beta))
    return LayerStack(layers=lays)


def random_heterogeneous_stack(ref_alpha, ref_beta, ref_rho, n_layers, a, omega, Dlambda, Dmu,
                               Drho, phi_mean=0.5, phi_std=0.1, seed=None,
                               **cpa_kwargs) -> LayerStack:
    """Two-phase layers with phi ~ N(phi_mean, phi_std^2) clipped to [0.01, 0.99],
    CPA-homogenised (same draw order as the reference)."""
    from .cpa_iteration import Phase

    rng = np.random.default_rng(seed)
    lam = ref_rho * (ref_alpha ** 2 - 2 * ref_beta ** 2)
    mu = ref_rho * ref_beta ** 2
    phases = []
    for _ in range(n_layers):
        phi = float(np.clip(rng.normal(phi_mean, phi_std), 0.01, 0.99))
        phases.append([Phase(lam, mu, ref_rho, 1.0 - phi),
                       Phase(lam + Dlambda, mu + Dmu, ref_rho + Drho, phi)])
    return cpa_stack_from_phases(phases, omega, a, thickness=2.0 * a, **cpa_kwargs)


def random_velocity_stack(alpha_mean, beta_mean, rho_mean, n_layers, thickness, dalpha_std=0.0,
                          dbeta_std=0.0, drho_std=0.0, seed=None, Q_alpha=np.inf,
                          Q_beta=np.inf) -> LayerStack:
    """Layers with normally perturbed velocities and density (floored at half the
    mean), the last one a half-space (same draw order as the reference)."""
    rng = np.random.default_rng(seed)
    lays = []
    for i in range(n_layers):
        al = max(alpha_mean + rng.normal(0, dalpha_std), alpha_mean * 0.5)
        be = max(beta_mean + rng.normal(0, dbeta_std), beta_mean * 0.5)
        rh = max(rho_mean + rng.normal(0, drho_std), rho_mean * 0.5)
        lays.append(IsotropicLayer(float(al), float(be), float(rh),
                                   np.inf if i == n_layers - 1 else thickness, Q_alpha, Q_beta))
    return LayerStack(layers=lays)
