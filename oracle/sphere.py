"""CPU oracle of the voxelised-sphere Foldy-Lax T-matrix -- TEST
INFRASTRUCTURE ONLY.

Restates compute_sphere_foldy_lax (cubic_scattering/sphere_scattering.py:
116-223: cells of the bounding-cube grid with |x| <= R, dense
(I - P (I x T_loc)) psi = psi_inc solve, T_comp = sum_m T_loc psi_m) and the
cell selection of the FFT form (sphere_scattering_fft.py:42-80, |x| < R).
Pinned (tests/test_oracle_sphere.py) by the resonance oracle when every cell
of the bounding cube is inside (n = 2) and by the reference's FFT form, whose
matvec (circular embedding on (2n-1)^3, sphere_scattering_fft.py:83-226) is
restated here and must equal the dense matrix.
"""

from __future__ import annotations

import numpy as np

from . import propagator as prop
from . import resonance as ores
from . import tmatrix as tm


def cells(R, n, inclusive=True):
    a_sub = R / n
    c = np.array([(-n / 2.0 + 0.5 + i) * 2.0 * a_sub for i in range(n)])
    idx, cen = [], []
    for i0 in range(n):
        for i1 in range(n):
            for i2 in range(n):
                p = np.array([c[i0], c[i1], c[i2]])
                r = np.linalg.norm(p)
                if (r <= R) if inclusive else (r < R):
                    idx.append([i0, i1, i2])
                    cen.append(p)
    return np.array(idx), np.array(cen), a_sub


def operator(omega, R, ref: tm.Medium, Dl, Dm, Dr, n, inclusive=True):
    idx, cen, a_sub = cells(R, n, inclusive)
    N = len(cen)
    T = ores.t_loc(omega, a_sub, ref, Dl, Dm, Dr)
    A = np.eye(9 * N, dtype=complex)
    for m in range(N):
        for s in range(N):
            if m != s:
                P = prop.propagator_9x9(cen[m] - cen[s], omega, ref.alpha, ref.beta, ref.rho)
                A[9 * m:9 * m + 9, 9 * s:9 * s + 9] = -P @ T
    return A, T, cen, idx


def sphere_tmatrix(omega, R, ref: tm.Medium, Dl, Dm, Dr, n, k_hat=None, wave_type="S",
                   inclusive=True):
    """(T_comp 9x9, psi_exc (9N, 9), centres)."""
    A, T, cen, _ = operator(omega, R, ref, Dl, Dm, Dr, n, inclusive)
    U = ores.incident(cen, omega, ref, k_hat, wave_type)
    psi = np.linalg.solve(A, U)
    Tc = np.zeros((9, 9), dtype=complex)
    for m in range(len(cen)):
        Tc += T @ psi[9 * m:9 * m + 9, :]
    return Tc, psi, cen


def fft_matvec(w, omega, R, ref: tm.Medium, Dl, Dm, Dr, n):
    """(I - P T) w by circular convolution on the (2n-1)^3 grid."""
    idx, cen, a_sub = cells(R, n, inclusive=False)
    T = ores.t_loc(omega, a_sub, ref, Dl, Dm, Dr)
    nP = 2 * n - 1
    K = np.zeros((9, 9, nP, nP, nP), dtype=complex)
    for d0 in range(-(n - 1), n):
        for d1 in range(-(n - 1), n):
            for d2 in range(-(n - 1), n):
                if d0 == d1 == d2 == 0:
                    continue
                P = prop.propagator_9x9(np.array([d0, d1, d2]) * 2.0 * a_sub, omega, ref.alpha,
                                        ref.beta, ref.rho)
                K[:, :, d0 % nP, d1 % nP, d2 % nP] = -(P @ T)
    Kh = np.fft.fftn(K, axes=(2, 3, 4))
    g = np.zeros((9, nP, nP, nP), dtype=complex)
    wb = w.reshape(-1, 9)
    for c in range(9):
        g[c, idx[:, 0], idx[:, 1], idx[:, 2]] = wb[:, c]
    gh = np.fft.fftn(g, axes=(1, 2, 3))
    yh = np.einsum("ijxyz,jxyz->ixyz", Kh, gh)
    y = np.fft.ifftn(yh, axes=(1, 2, 3))
    out = np.stack([y[c, idx[:, 0], idx[:, 1], idx[:, 2]] for c in range(9)], axis=1)
    return w + out.ravel()
