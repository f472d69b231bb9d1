"""Oracle: 9x9 elastodynamic point propagator and D4h transforms.  TEST INFRASTRUCTURE.

Restates cubic_scattering/resonance_tmatrix.py:162-428 (Kupradze G, its first
and second derivatives, the Voigt contraction to [[G, C], [H, S]]) and the
9x9 D4h matrices of cubic_scattering/lattice_greens.py:497-566.

Index conventions (slab_scattering.py:7-8): Cartesian order (z, x, y);
Voigt order (zz, xx, yy, xy, zy, zx).
"""

from __future__ import annotations

import numpy as np

# resonance_tmatrix.py:130-138 -- Voigt index -> Cartesian pair, (z,x,y) = (0,1,2).
VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def greens(r_vec, omega, alpha, beta, rho) -> np.ndarray:
    """Kupradze G_ij (resonance_tmatrix.py:162-211); zero at r = 0."""
    r_vec = np.asarray(r_vec, dtype=float)
    r = float(np.linalg.norm(r_vec))
    if r < 1e-14:
        return np.zeros((3, 3), dtype=complex)
    kP, kS = omega / alpha, omega / beta
    g = r_vec / r
    eP, eS = np.exp(1j * kP * r), np.exp(1j * kS * r)
    nP = (1.0 - 1j * kP * r) * eP / r**3
    nS = (1.0 - 1j * kS * r) * eS / r**3
    phi = kS**2 * eS / r - nS + nP
    psi = 3.0 * nS - 3.0 * nP + kP**2 * eP / r - kS**2 * eS / r
    return (phi * np.eye(3) + psi * np.outer(g, g)) / (4.0 * np.pi * rho * omega**2)


def radial(r, kP, kS):
    """phi, psi and their first two r-derivatives (resonance_tmatrix.py:214-271).

    phi = kS^2 eS/r + i kS eS/r^2 - eS/r^3 - i kP eP/r^2 + eP/r^3
    psi = -kS^2 eS/r - 3i kS eS/r^2 + 3 eS/r^3 + kP^2 eP/r + 3i kP eP/r^2 - 3 eP/r^3
    Each term c e^{ikr} r^p is differentiated with
    d/dr = e^{ikr} r^{p-1}(ikr + p), d2/dr2 = e^{ikr} r^{p-2}((ikr)^2 + 2p ikr + p(p-1)).
    """
    phi_terms = [(kS**2, kS, -1), (1j * kS, kS, -2), (-1.0, kS, -3),
                 (-1j * kP, kP, -2), (1.0, kP, -3)]
    psi_terms = [(-(kS**2), kS, -1), (-3j * kS, kS, -2), (3.0, kS, -3),
                 (kP**2, kP, -1), (3j * kP, kP, -2), (-3.0, kP, -3)]

    def acc(terms):
        f = f1 = f2 = 0j
        for c, k, p in terms:
            e = np.exp(1j * k * r)
            ikr = 1j * k * r
            f += c * e * r**p
            f1 += c * e * r ** (p - 1) * (ikr + p)
            f2 += c * e * r ** (p - 2) * (ikr**2 + 2 * p * ikr + p * (p - 1))
        return f, f1, f2

    phi, phi1, phi2 = acc(phi_terms)
    psi, psi1, psi2 = acc(psi_terms)
    return phi, psi, phi1, psi1, phi2, psi2


def greens_deriv(r_vec, omega, alpha, beta, rho):
    """G_ij, G_ij,k, G_ij,kl (resonance_tmatrix.py:274-350); zeros at r = 0."""
    r_vec = np.asarray(r_vec, dtype=float)
    r = float(np.linalg.norm(r_vec))
    if r < 1e-14:
        return (np.zeros((3, 3), complex), np.zeros((3, 3, 3), complex),
                np.zeros((3, 3, 3, 3), complex))
    kP, kS = omega / alpha, omega / beta
    g = r_vec / r
    d = np.eye(3)
    P = 1.0 / (4.0 * np.pi * rho * omega**2)
    phi, psi, phi1, psi1, phi2, psi2 = radial(r, kP, kS)
    G = P * (phi * d + psi * np.outer(g, g))
    c1, c2, c3 = phi1, psi1 - 2.0 * psi / r, psi / r
    Gd = P * (c1 * np.einsum("k,ij->ijk", g, d)
              + c2 * np.einsum("i,j,k->ijk", g, g, g)
              + c3 * (np.einsum("ik,j->ijk", d, g) + np.einsum("jk,i->ijk", d, g)))
    t1 = phi1 / r
    t2 = phi2 - phi1 / r
    t3 = psi1 / r - 2.0 * psi / r**2
    t4 = psi / r**2
    t7 = psi2 - 5.0 * psi1 / r + 8.0 * psi / r**2
    Gdd = P * (t1 * np.einsum("ij,kl->ijkl", d, d)
               + t2 * np.einsum("ij,k,l->ijkl", d, g, g)
               + t3 * np.einsum("i,j,kl->ijkl", g, g, d)
               + t4 * (np.einsum("ik,jl->ijkl", d, d) + np.einsum("jk,il->ijkl", d, d))
               + t3 * (np.einsum("il,j,k->ijkl", d, g, g) + np.einsum("jl,i,k->ijkl", d, g, g))
               + t3 * (np.einsum("ik,j,l->ijkl", d, g, g) + np.einsum("jk,i,l->ijkl", d, g, g))
               + t7 * np.einsum("i,j,k,l->ijkl", g, g, g, g))
    return G, Gd, Gdd


def voigt_contract(Gd, Gdd):
    """C (3x6), H (6x3), S (6x6) blocks (resonance_tmatrix.py:353-409).

    Off-diagonal Voigt pairs sum both index orders; the engineering stress
    input halves columns beta >= 3 of C and S (lines 405-407).
    """
    C = np.zeros((3, 6), dtype=complex)
    H = np.zeros((6, 3), dtype=complex)
    S = np.zeros((6, 6), dtype=complex)
    for al, (p, q) in enumerate(VOIGT_PAIRS):
        pairs_a = [(p, q)] if p == q else [(p, q), (q, p)]
        for i in range(3):
            C[i, al] = sum(Gd[i, u, v] for u, v in pairs_a)
            H[al, i] = sum(Gd[u, i, v] for u, v in pairs_a)
        for be, (m, n) in enumerate(VOIGT_PAIRS):
            pairs_b = [(m, n)] if m == n else [(m, n), (n, m)]
            # S[al,be] = sum over (p,q) orders and (m,n) orders of Gdd[p, m, n, q]
            S[al, be] = sum(Gdd[u, s, t, v] for u, v in pairs_a for s, t in pairs_b)
    C[:, 3:] *= 0.5
    S[:, 3:] *= 0.5
    return C, H, S


def propagator_9x9(r_vec, omega, alpha, beta, rho) -> np.ndarray:
    """[[G, C], [H, S]] at r = x_field - x_source (resonance_tmatrix.py:412-428)."""
    G, Gd, Gdd = greens_deriv(r_vec, omega, alpha, beta, rho)
    C, H, S = voigt_contract(Gd, Gdd)
    P = np.zeros((9, 9), dtype=complex)
    P[:3, :3] = G
    P[:3, 3:] = C
    P[3:, :3] = H
    P[3:, 3:] = S
    return P


# lattice_greens.py:522-545 -- 9x9 point operations on the propagator:
# P' = R P R^T with R = blockdiag(R3, V_R).
REFL_X = np.diag([1.0, -1.0, 1.0, 1.0, 1.0, 1.0, -1.0, 1.0, -1.0])
REFL_Y = np.diag([1.0, 1.0, -1.0, 1.0, 1.0, 1.0, -1.0, -1.0, 1.0])
SWAP_XY = np.zeros((9, 9))
for _dst, _src in [(0, 0), (1, 2), (2, 1), (3, 3), (4, 5), (5, 4), (6, 6), (7, 8), (8, 7)]:
    SWAP_XY[_dst, _src] = 1.0
del _dst, _src


def refl_x(G):
    return REFL_X @ G @ REFL_X.T


def refl_y(G):
    return REFL_Y @ G @ REFL_Y.T


def swap_xy(G):
    return SWAP_XY @ G @ SWAP_XY.T


def d4h_orbit(dx: int, dy: int):
    """Offsets related to (dx, dy) and their transforms (slab_scattering.py:233-259)."""
    ident = lambda G: G.copy()  # noqa: E731
    if dx == 0 and dy == 0:
        return [(0, 0, ident)]
    orbit = [
        (dx, dy, ident),
        (-dx, dy, refl_x),
        (dx, -dy, refl_y),
        (-dx, -dy, lambda G: refl_x(refl_y(G))),
    ]
    if dx == dy:
        return orbit
    return orbit + [
        (dy, dx, swap_xy),
        (-dy, dx, lambda G: refl_x(swap_xy(G))),
        (dy, -dx, lambda G: refl_y(swap_xy(G))),
        (-dy, -dx, lambda G: refl_x(refl_y(swap_xy(G)))),
    ]
