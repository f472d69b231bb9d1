"""CPU oracle of the resonance-regime cube T-matrix -- TEST INFRASTRUCTURE ONLY.

Restates compute_resonance_tmatrix / _solve_coupled
(cubic_scattering/resonance_tmatrix.py:650-855): sub-cell centres (436-455),
the sub-cell Rayleigh 9x9 T (461-487), the 9 incident patterns (526-593), the
dense (I - P (I x T_loc)) psi = psi_inc solve (np.linalg.solve) or its
truncated Neumann series, and T_comp = sum_m T_loc psi_m.

Pins (tests/test_oracle_resonance.py): the assembled matrix equals the slab
oracle's dense operator of a uniform n x n x n slab of half-width a/n (whose
entries are pinned by the brute-force slab matvec), n = 1 reduces to the
Rayleigh T-matrix, a density-only contrast gives an isotropic force monopole,
the Neumann series converges to the direct solve at weak contrast, and the
incident patterns have the closed-form strain.
"""

from __future__ import annotations

import numpy as np

from . import propagator as prop
from . import tmatrix as tm

VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def centres(a, n):
    s = a / n
    c = np.linspace(-a + s, a - s, n)
    g = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([x.ravel() for x in g], axis=1)


def t_loc(omega, a_sub, ref: tm.Medium, Dl, Dm, Dr):
    return tm.tmatrix_9x9(tm.cube_tmatrix(omega, a_sub, ref, Dl, Dm, Dr), omega, a_sub)


def incident(cx, omega, ref: tm.Medium, k_hat=None, wave_type="S"):
    N = len(cx)
    kh = np.array([0.0, 0.0, 1.0] if k_hat is None else k_hat, dtype=float)
    kh = kh / np.linalg.norm(kh)
    k = omega / (ref.beta if wave_type == "S" else ref.alpha)
    ph = np.exp(1j * k * (cx @ kh))
    x0 = cx.mean(axis=0)
    U = np.zeros((9 * N, 9), dtype=complex)
    for m in range(N):
        U[9 * m:9 * m + 3, :3] = np.eye(3) * ph[m]
        dx = cx[m] - x0
        for al, (p, q) in enumerate(VOIGT):
            E = np.zeros((3, 3))
            if p == q:
                E[p, p] = 1.0
            else:
                E[p, q] = E[q, p] = 0.5
            U[9 * m:9 * m + 3, 3 + al] = (E @ dx) * ph[m]
            U[9 * m + 3 + al, 3 + al] = ph[m]
    return U


def operator(omega, a, ref: tm.Medium, Dl, Dm, Dr, n):
    """(A = I - P (I x T_loc), T_loc, centres)."""
    cx = centres(a, n)
    N = len(cx)
    T = t_loc(omega, a / n, ref, Dl, Dm, Dr)
    A = np.eye(9 * N, dtype=complex)
    for m in range(N):
        for s in range(N):
            if m != s:
                P = prop.propagator_9x9(cx[m] - cx[s], omega, ref.alpha, ref.beta, ref.rho)
                A[9 * m:9 * m + 9, 9 * s:9 * s + 9] = -P @ T
    return A, T, cx


def resonance_tmatrix(omega, a, ref: tm.Medium, Dl, Dm, Dr, n, k_hat=None, wave_type="S",
                      neumann_order=0, neumann_tol=1e-8):
    """(T_comp 9x9, psi_exc (9N, 9), converged Neumann order or None)."""
    A, T, cx = operator(omega, a, ref, Dl, Dm, Dr, n)
    U = incident(cx, omega, ref, k_hat, wave_type)
    conv = None
    if neumann_order == 0:
        psi = np.linalg.solve(A, U)
    else:
        B = np.eye(A.shape[0]) - A
        psi = U.copy()
        term = U.copy()
        for k in range(1, neumann_order + 1):
            term = B @ term
            psi = psi + term
            if np.linalg.norm(term) / (np.linalg.norm(psi) + 1e-300) < neumann_tol:
                conv = k
                break
    Tc = np.zeros((9, 9), dtype=complex)
    for m in range(len(cx)):
        Tc += T @ psi[9 * m:9 * m + 9, :]
    return Tc, psi, conv
