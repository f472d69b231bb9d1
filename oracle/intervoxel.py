"""CPU oracle of the volume-averaged nearest-neighbour 9x9 propagator --
TEST INFRASTRUCTURE ONLY.

Restates cubic_scattering/inter_voxel_propagator.py (static potentials
40-110, dynamic orders 112-312, assembly 314-830, 9x9 blocks 831-1355) as used
by _build_slab_kernels(volume_averaged=True) (slab_scattering.py:316-329):
for each of the 26 neighbours R of a cube the block [[G, C], [H, S]] is

  S = Voigt contraction of P_ijkl(w) = sum_n w^2n P^(n),
      P^(n) = -1/(2 mu_n) [d_ik A_jl + d_jk A_il - 2 eta_n B_ijkl]   (n = 0..3)
  G = 1/(4 pi mu) [d_ij Phi - eta_s d2Psi_ij] + w^2, w^4, w^6 corrections
  C_i,(mn) = sym. part of dG_imn = d<G_im>/dR_n,   H = C^T

with the potential derivatives A, B, d3W, dW tabulated for the canonical
face (1,0,0), edge (1,1,0) and corner (1,1,1) directions and mapped to R by
the signed permutation the reference's greedy search picks.  The values are
for the unit cube: the reference does not scale them with the cube size.

Pins: the reference's Mathematica values and Laplacian identities
(tests/test_oracle_intervoxel.py) and, for the w^2..w^6 series the reference
tests leave open, Gauss cubature of the cube-averaged Kupradze Taylor terms
(tests/test_oracle_intervoxel_cubature.py).

Two reference behaviours are kept on purpose (parity):
* the corner B tensors are filled from the component list
  {0012, 0102, 0122} (inter_voxel_propagator.py:431-435 and the dynamic
  corner builders), which leaves the {0112} class at zero;
* the O_h map is the first signed permutation found by
  _oh_permutation (inter_voxel_propagator.py:721-750), which matters once the
  corner tensors are not S3-symmetric.
"""

from __future__ import annotations

import itertools
import math

import numpy as np

# Raw Mathematica values (unit cube, mu = 1 stripped), as the reference stores them
# before its normalisation factors; keyed (neighbour type, order).
A_RAW = {
    ("face", 0): (-0.13501718054449527, +0.06750859027224763, 0.0),
    ("face", 1): (0.30434003593387251229, 0.82871516563404197555, 0.0),
    ("face", 2): (6.14687033491410334021, 3.93095155409960859065, 0.0),
    ("face", 3): (35.09972390156100239, 13.54175081054303704, 0.0),
    ("face", 4): (166.040235936049047, 47.1604212749853438, 0.0),
    ("edge", 0): (-0.01378576204834812, +0.02757152409669624, -0.04556482263891465),
    ("edge", 1): (0.38529898485200793724, 0.64639228402098784773, -0.26788042772784072223),
    ("edge", 2): (6.72804661239907577216, 4.92607224313657099253, 1.80873224417671787438),
    ("edge", 3): (51.67064499234048322, 25.50019876734907121, 26.14480437842729518),
    ("edge", 4): (336.612687195747214, 128.501096007485975, 208.609329750912862),
    ("corner", 0): (0.0, 0.0, -0.01606212781050823),
    ("corner", 1): (0.38586466785236013460, 0.38586466785236013460, -0.16121187591918290103),
    ("corner", 2): (7.31149202498253105099, 7.31149202498253105099, 1.55428124313434973873),
    ("corner", 3): (69.71171346052524425, 69.71171346052524425, 30.10535779745780000),
    ("corner", 4): (569.489574041776214, 569.489574041776214, 311.882584477024042),
}
B_RAW = {
    ("face", 0): (-0.10518162135737388, -0.01491777959356069, +0.06226739135121906, +0.02015897851458926),
    ("face", 1): (1.41327701454403944664, 1.11940170833121535043, 6.61390374095420502775, 2.21127653832308332846),
    ("face", 2): (121.10629615723429259, 31.64990694509440380732, 64.72963275084629337, 21.54900692704756054521),
    ("face", 3): (1401.02708938036651, 282.27872455352481, 356.74707004533833, 119.31225079154694),
    ("edge", 0): (-0.02188873710833393, +0.00965705243872159, -0.00155407737873578, +0.03067967885416780, -0.01652993415886442, -0.01250495432118580),
    ("edge", 1): (2.27247134973510232851, 1.16949686073991690189, 1.18161960774907601651, 5.39346819275370213980, -1.29181426541309712008, -0.63093660190789442655),
    ("edge", 2): (126.69450210248536477560, 40.52235834689523028490, 34.62453792259167810431, 78.53309144891377356740, 22.88723612738029927492, 8.48749507054093768171),
    ("edge", 3): (1840.74246384076366, 651.07099896048098, 401.74265676982242, 624.52581743190315, 635.51239625893890, 193.08425267405073),
    ("corner", 0): (-0.00625536598256419, +0.00312768299128210, -0.00926615309635844, +0.00247017838220864),
    ("corner", 1): (2.53304140303690244565, 1.04866730559570958475, -0.91956804085889468740, -0.09540642931240543758),
    ("corner", 2): (133.00383894991481736747, 43.17046089978055708110, 20.48871726047062015519, 5.65100277308925185145),
    ("corner", 3): (2302.44036080125068, 800.70779649408150, 708.32238983951127, 269.25525697861426),
}
D3_RAW = {
    ("face", 0): (-0.79543834473319394, -0.52826208819206832),
    ("face", 1): (5.59066550438255537, 2.22178066763092212),
    ("face", 2): (74.7028469944845582, 21.5319534595037667),
    ("face", 3): (520.366848816268754, 119.391877525300900),
    ("edge", 0): (-0.50186397481765351, +0.04882866652718464, -0.26277092923552829),
    ("edge", 1): (4.79793940592048508, 1.17783327456444731, 1.80511344407289295),
    ("edge", 2): (86.9352238951433639, 34.6277658601723239, 26.1528244054226945),
    ("edge", 3): (818.181285566568028, 401.751494351498964, 208.558979721460787),
    ("corner", 0): (-0.34647399020181943, -0.02187587552538497, +0.14258930467014646),
    ("corner", 1): (4.25436485917712435, 1.14631500401061554, -0.41087473026598235),
    ("corner", 2): (97.7452814064066722, 37.5180442143966006, 7.42106273126978393),
    ("corner", 3): (1154.46537928822656, 531.429829090769075, 219.478429989378284),
}
DW_RAW = {
    ("face", 0): -0.92598126055866529,
    ("face", 1): 0.83618556997036663,
    ("face", 2): 3.92555846378306972,
    ("face", 3): 13.5562607833369742,
    ("edge", 0): -0.35790311876299858,
    ("edge", 1): 0.64840717704648545,
    ("edge", 2): 4.92386047202461274,
    ("edge", 3): 25.5087814221344246,
    ("corner", 0): -0.19511287062629469,
    ("corner", 1): 0.54558290559986295,
    ("corner", 2): 5.75937899450666245,
    ("corner", 3): 39.5950899548172270,
}


NORM_A = {1: 1.0 / (8.0 * math.pi), 2: -1.0 / (96.0 * math.pi), 3: 1.0 / (2880.0 * math.pi),
          4: -1.0 / (161280.0 * math.pi)}
NORM_B = {1: 1.0 / (96.0 * math.pi), 2: -1.0 / (2880.0 * math.pi), 3: 1.0 / (161280.0 * math.pi)}

# Index lists that carry each tabulated B value (every permutation of each tuple
# receives the value), in the reference's order.
B_TUPLES = {
    "face": [[(0, 0, 0, 0)], [(0, 0, 1, 1), (0, 0, 2, 2)], [(1, 1, 1, 1), (2, 2, 2, 2)],
             [(1, 1, 2, 2)]],
    "edge": [[(0, 0, 0, 0), (1, 1, 1, 1)], [(0, 0, 1, 1)], [(0, 0, 2, 2), (1, 1, 2, 2)],
             [(2, 2, 2, 2)], [(0, 0, 0, 1), (0, 1, 1, 1)], [(0, 1, 2, 2)]],
    "corner": [[(0, 0, 0, 0), (1, 1, 1, 1), (2, 2, 2, 2)],
               [(0, 0, 1, 1), (0, 0, 2, 2), (1, 1, 2, 2)],
               [(0, 0, 0, 1), (0, 0, 0, 2), (0, 1, 1, 1), (1, 1, 1, 2), (0, 2, 2, 2), (1, 2, 2, 2)],
               [(0, 0, 1, 2), (0, 1, 0, 2), (0, 1, 2, 2)]],
}
D3_TUPLES = {
    "face": [[(0, 0, 0)], [(0, 1, 1), (0, 2, 2)]],
    "edge": [[(0, 0, 0), (1, 1, 1)], [(0, 0, 1), (0, 1, 1)], [(0, 2, 2), (1, 2, 2)]],
    "corner": [[(0, 0, 0), (1, 1, 1), (2, 2, 2)],
               [(0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 2, 2), (1, 1, 2), (1, 2, 2)], [(0, 1, 2)]],
}
CANON = {"face": (1.0, 0.0, 0.0), "edge": (1.0, 1.0, 0.0), "corner": (1.0, 1.0, 1.0)}
VOIGT = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def neighbour_type(R) -> str:
    s = sorted((abs(int(v)) for v in R), reverse=True)
    if s == [1, 0, 0]:
        return "face"
    if s == [1, 1, 0]:
        return "edge"
    if s == [1, 1, 1]:
        return "corner"
    raise ValueError(f"R_lattice={tuple(R)} is not a nearest neighbour")


def a_matrix(t: str, v) -> np.ndarray:
    """Symmetric 3x3 from the (first diagonal, second diagonal, off-diagonal) triple."""
    d1, d2, off = v
    if t == "face":
        return np.diag([d1, d2, d2]).astype(float)
    if t == "edge":
        A = np.diag([d1, d1, d2]).astype(float)
        A[0, 1] = A[1, 0] = off
        return A
    A = np.full((3, 3), float(off))
    np.fill_diagonal(A, d1)
    return A


def fill_symmetric(shape_rank: int, tuples, values) -> np.ndarray:
    T = np.zeros((3,) * shape_rank)
    for group, val in zip(tuples, values):
        for tup in group:
            for p in set(itertools.permutations(tup)):
                T[p] = val
    return T


def b_tensor(t: str, order: int) -> np.ndarray:
    vals = B_RAW[(t, order)]
    if order > 0:
        vals = [x * NORM_B[order] for x in vals]
    return fill_symmetric(4, B_TUPLES[t], vals)


def a_order(t: str, order: int) -> np.ndarray:
    A = a_matrix(t, A_RAW[(t, order)])
    return A if order == 0 else A * NORM_A[order]


def assemble_p(A, B, mu, eta) -> np.ndarray:
    d = np.eye(3)
    return -1.0 / (2.0 * mu) * (np.einsum("ik,jl->ijkl", d, A) + np.einsum("jk,il->ijkl", d, A)
                                - 2.0 * eta * B)


def p_canonical(t: str, order: int, alpha, beta, rho) -> np.ndarray:
    """Static (order 0) or the order-n dynamic strain propagator, canonical direction."""
    if order == 0:
        mu = rho * beta ** 2
        nu = (alpha ** 2 - 2 * beta ** 2) / (2 * (alpha ** 2 - beta ** 2))
        return assemble_p(a_order(t, 0), b_tensor(t, 0), mu, 1.0 / (2.0 * (1.0 - nu)))
    mu_n = rho * beta ** (2 * order + 2)
    eta_n = 1.0 - (beta / alpha) ** (2 * order + 2)
    return assemble_p(a_order(t, order), b_tensor(t, order), mu_n, eta_n)


def p_canonical_static_unit(t: str, mu: float, nu: float) -> np.ndarray:
    return assemble_p(a_order(t, 0), b_tensor(t, 0), mu, 1.0 / (2.0 * (1.0 - nu)))


def oh_map(R) -> np.ndarray:
    """The signed permutation the reference's greedy search maps the canonical
    direction onto R with (first match per row, zero entries matched to zeros)."""
    t = np.asarray(R, dtype=float)
    s = np.asarray(CANON[neighbour_type(R)])
    Pm = np.zeros((3, 3))
    used = [False] * 3
    for i in range(3):
        for j in range(3):
            if not used[j] and abs(abs(t[i]) - abs(s[j])) < 1e-10 and abs(s[j]) > 1e-10:
                Pm[i, j] = np.sign(t[i]) * np.sign(s[j])
                used[j] = True
                break
            if not used[j] and abs(s[j]) < 1e-10 and abs(t[i]) < 1e-10:
                Pm[i, j] = 1.0
                used[j] = True
                break
    assert np.allclose(Pm @ s, t) and abs(np.linalg.det(Pm)) > 0.5
    return Pm


def rot(T: np.ndarray, Q: np.ndarray) -> np.ndarray:
    for ax in range(T.ndim):
        T = np.moveaxis(np.tensordot(Q, T, axes=([1], [ax])), 0, ax)
    return T


def strain_propagator(R, alpha, beta, rho, omega, n_orders) -> np.ndarray:
    """P_ijkl(R, w) = sum_{n <= n_orders} w^2n P^(n), rotated to R."""
    t = neighbour_type(R)
    P = p_canonical(t, 0, alpha, beta, rho)
    for n in range(1, min(n_orders, 3) + 1):
        P = P + omega ** (2 * n) * p_canonical(t, n, alpha, beta, rho)
    return rot(P, oh_map(R))


def voigt_s(P: np.ndarray) -> np.ndarray:
    S = np.zeros((6, 6), dtype=P.dtype)
    for a, (p, q) in enumerate(VOIGT):
        ma = 2.0 if p != q else 1.0
        for b, (m, n) in enumerate(VOIGT):
            v = ma * P[p, q, m, n]
            if m != n:
                v = v + ma * P[p, q, n, m]
            S[a, b] = v * (0.5 if b >= 3 else 1.0)
    return S


def _series(alpha, beta, rho, omega, n_orders):
    """(coefficient, eta/denominator, order) of the w^2n terms of G and dG."""
    out = []
    cs, cp = beta, alpha
    if n_orders >= 1:
        out.append((-(omega ** 2) / (8.0 * math.pi * rho * cs ** 4), (1 - (cs / cp) ** 4) / 12.0, 1, 12.0))
    if n_orders >= 2:
        out.append((omega ** 4 / (96.0 * math.pi * rho * cs ** 6), (1 - (cs / cp) ** 6) / 30.0, 2, 30.0))
    if n_orders >= 3:
        out.append((-(omega ** 6) / (2880.0 * math.pi * rho * cs ** 8), (1 - (cs / cp) ** 8) / 56.0, 3, 56.0))
    return out


def g_canonical(t, alpha, beta, rho, omega, n_orders) -> np.ndarray:
    mu = rho * beta ** 2
    nu = (alpha ** 2 - 2 * beta ** 2) / (2 * (alpha ** 2 - beta ** 2))
    d2 = a_matrix(t, A_RAW[(t, 1)])  # raw d2Psi
    G = (np.eye(3) * 0.5 * np.trace(d2) - d2 / (4.0 * (1.0 - nu))) / (4.0 * math.pi * mu)
    for coeff, dev, n, lap in _series(alpha, beta, rho, omega, min(n_orders, 3)):
        d2 = a_matrix(t, A_RAW[(t, n + 1)])  # raw d2X, d2Omega, d2H
        G = G + coeff * (np.eye(3) * np.trace(d2) / lap - dev * d2)
    return G


def d3_tensor(t, order):
    return fill_symmetric(3, D3_TUPLES[t], D3_RAW[(t, order)])


def dw_vector(t, order):
    v = DW_RAW[(t, order)]
    return np.array({"face": [v, 0, 0], "edge": [v, v, 0], "corner": [v, v, v]}[t], float)


def dg_canonical(t, alpha, beta, rho, omega, n_orders) -> np.ndarray:
    mu = rho * beta ** 2
    nu = (alpha ** 2 - 2 * beta ** 2) / (2 * (alpha ** 2 - beta ** 2))
    d = np.eye(3)
    dG = (np.einsum("ij,k->ijk", d, dw_vector(t, 0))
          - d3_tensor(t, 0) / (4.0 * (1.0 - nu))) / (4.0 * math.pi * mu)
    if omega != 0.0:
        for coeff, dev, n, _ in _series(alpha, beta, rho, omega, min(n_orders, 3)):
            dG = dG + coeff * (np.einsum("ij,k->ijk", d, dw_vector(t, n)) - dev * d3_tensor(t, n))
    return dG


def c_block(dG: np.ndarray) -> np.ndarray:
    C = np.zeros((3, 6), dtype=dG.dtype)
    for a, (m, n) in enumerate(VOIGT):
        C[:, a] = dG[:, m, m] if m == n else 0.5 * (dG[:, m, n] + dG[:, n, m])
    return C


def propagator_9x9(R, alpha, beta, rho, omega, n_orders=2) -> np.ndarray:
    """The reference's inter_voxel_propagator_9x9(R, alpha, beta, rho, omega, n_orders)."""
    t = neighbour_type(R)
    Q = oh_map(R)
    S = voigt_s(strain_propagator(R, alpha, beta, rho, omega, n_orders))
    G = Q @ g_canonical(t, alpha, beta, rho, omega, n_orders) @ Q.T
    C = c_block(rot(dg_canonical(t, alpha, beta, rho, omega, n_orders), Q))
    P9 = np.zeros((9, 9), dtype=complex)
    P9[:3, :3] = G
    P9[:3, 3:] = C
    P9[3:, :3] = C.T
    P9[3:, 3:] = S
    return P9
