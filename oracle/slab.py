"""Oracle: slab geometry, T0 per voxel, G0 spectral table, matvec, solve, R_PP.

TEST INFRASTRUCTURE.  Restates cubic_scattering/slab_scattering.py:40-852 and
the 3-D-FFT formula of cubic_scattering/slab_scattering_gpu.py:34-155 in
complex128 NumPy/SciPy.  Index order is (z, x, y, component), C-contiguous.
"""

from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import LinearOperator, gmres

from . import propagator as prop
from . import tmatrix as tm


# ── geometry (slab_scattering.py:40-99) ─────────────────────────────────


def centres(M: int, N_z: int, a: float) -> np.ndarray:
    """Cube centres (N_z, M, M, 3): z = (l+1/2)d, x,y = (i-(M-1)/2)d, d = 2a."""
    d = 2.0 * a
    z = (np.arange(N_z) + 0.5) * d
    x = (np.arange(M) - (M - 1) / 2.0) * d
    zz, xx, yy = np.meshgrid(z, x, x, indexing="ij")
    return np.stack([zz, xx, yy], axis=-1)


# ── T0 per voxel (slab_scattering.py:169-205) ───────────────────────────


def slab_tmatrices(M, N_z, a, Dlambda, Dmu, Drho, omega, ref: tm.Medium) -> np.ndarray:
    """(N_z, M, M, 9, 9) T0, one closed-form evaluation per unique contrast triple."""
    T = np.zeros((N_z, M, M, 9, 9), dtype=complex)
    cache: dict = {}
    for idx in np.ndindex(N_z, M, M):
        key = (float(Dlambda[idx]), float(Dmu[idx]), float(Drho[idx]))
        if key not in cache:
            cache[key] = tm.tmatrix_9x9(tm.cube_tmatrix(omega, a, ref, *key), omega, a)
        T[idx] = cache[key]
    return T


def tparams_from_T(T: np.ndarray) -> np.ndarray:
    """Compress block-diagonal T0 to (t_u, t_d, t_o, t_s) per voxel.

    t_u = w^2 V Drho*, t_d = V(Dlam* + 2 Dmu*_diag), t_o = V Dlam*,
    t_s = 2 V Dmu*_off  (resonance_tmatrix.py:482-486, voigt_tmatrix.py:151-167).
    """
    return np.stack([T[..., 0, 0], T[..., 3, 3], T[..., 3, 4], T[..., 6, 6]], axis=-1)


# ── G0 spectral table (slab_scattering.py:267-351) ──────────────────────


def slab_kernels(M, N_z, a, omega, ref: tm.Medium, *, periodic=False, volume_averaged=False,
                 n_orders=2) -> np.ndarray:
    """xy-FFT'd kernels (2N_z-1, H, H, 9, 9), H = M if periodic else 2M-1.

    For each dz, the propagator is evaluated on the fundamental domain
    0 <= dy <= dx < M, mapped to its D4h images, the self term is left zero;
    periodic mode folds the (2M-1)^2 spatial kernel mod M without a Bloch phase.
    ``volume_averaged`` replaces the 26 nearest-neighbour offsets by the
    volume-averaged block of each signed offset (slab_scattering.py:316-329).
    """
    from . import intervoxel

    d = 2.0 * a
    S = 2 * M - 1
    H = M if periodic else S
    out = np.zeros((2 * N_z - 1, H, H, 9, 9), dtype=complex)
    for k in range(2 * N_z - 1):
        dzv = k - (N_z - 1)
        sp = np.zeros((S, S, 9, 9), dtype=complex)
        for dx in range(M):
            for dy in range(dx + 1):
                if dx == 0 and dy == 0 and dzv == 0:
                    continue
                if volume_averaged and max(abs(dzv), dx, dy) <= 1:
                    for sdx, sdy, _ in prop.d4h_orbit(dx, dy):
                        sp[sdx + M - 1, sdy + M - 1] = intervoxel.propagator_9x9(
                            (dzv, sdx, sdy), ref.alpha, ref.beta, ref.rho, omega, n_orders)
                    continue
                G0 = prop.propagator_9x9([dzv * d, dx * d, dy * d], omega, ref.alpha, ref.beta, ref.rho)
                for sdx, sdy, f in prop.d4h_orbit(dx, dy):
                    sp[sdx + M - 1, sdy + M - 1] = f(G0)
        if periodic:
            circ = np.zeros((M, M, 9, 9), dtype=complex)
            for ix in range(S):
                for iy in range(S):
                    circ[(ix - (M - 1)) % M, (iy - (M - 1)) % M] += sp[ix, iy]
            out[k] = np.fft.fft2(circ, axes=(0, 1))
        else:
            out[k] = np.fft.fft2(sp, axes=(0, 1))
    return out


def spatial_kernel(M, N_z, a, omega, ref: tm.Medium) -> np.ndarray:
    """Dense spatial kernel K[dz+N_z-1, dx+M-1, dy+M-1] (9x9), direct evaluation."""
    d = 2.0 * a
    K = np.zeros((2 * N_z - 1, 2 * M - 1, 2 * M - 1, 9, 9), dtype=complex)
    for iz in range(2 * N_z - 1):
        for ix in range(2 * M - 1):
            for iy in range(2 * M - 1):
                r = [(iz - N_z + 1) * d, (ix - M + 1) * d, (iy - M + 1) * d]
                K[iz, ix, iy] = prop.propagator_9x9(r, omega, ref.alpha, ref.beta, ref.rho)
    return K


# ── matvec (slab_scattering.py:359-418; slab_scattering_gpu.py:95-155) ──


def slab_matvec(psi_flat, T, kernel_hat, M, N_z, *, periodic=False) -> np.ndarray:
    """(I - G T) psi with 2-D FFTs per layer and the N_z^2 layer-pair loop."""
    psi = psi_flat.reshape(N_z, M, M, 9)
    tau = np.einsum("lmnab,lmnb->lmna", T, psi)
    if periodic:
        H = M
        tau_hat = np.fft.fft2(tau, axes=(1, 2))
    else:
        H = 2 * M - 1
        pad = np.zeros((N_z, H, H, 9), dtype=complex)
        pad[:, :M, :M] = tau
        tau_hat = np.fft.fft2(pad, axes=(1, 2))
    acc = np.zeros((N_z, H, H, 9), dtype=complex)
    for m in range(N_z):
        for n in range(N_z):
            acc[m] += np.einsum("xyij,xyj->xyi", kernel_hat[m - n + N_z - 1], tau_hat[n])
    acc = np.fft.ifft2(acc, axes=(1, 2))
    if not periodic:
        acc = acc[:, M - 1:H, M - 1:H]
    return (psi - acc).ravel()


def kernel_hat_3d(kernel_hat_xy, M, N_z, pad=None, *, periodic=False) -> np.ndarray:
    """3-D FFT table (9, 9, Pz, Px, Py) (slab_scattering_gpu.py:34-92).

    The reference stores offset d at index d + N - 1 of a (2N-1)-point grid and
    extracts the window [N-1, 2N-1) (slab_scattering_gpu.py:143-150).  Here the
    offset d is stored circularly at d mod P in a grid P >= 2N-1, so the valid
    window is [0, N); both are the same linear convolution.  Periodic mode keeps
    the M x M circular xy grid and pads z only.
    """
    sp = np.fft.ifft2(kernel_hat_xy, axes=(1, 2))  # (2N_z-1, H, H, 9, 9)
    nz, H = sp.shape[0], sp.shape[1]
    if pad is None:
        pad = (nz, H, H)
    Pz, Px, Py = pad
    if periodic and (Px != M or Py != M):
        raise ValueError("periodic mode keeps the M x M circular xy grid")
    grid = np.zeros((Pz, Px, Py, 9, 9), dtype=complex)
    for iz in range(nz):
        for ix in range(H):
            for iy in range(H):
                if periodic:
                    jx, jy = ix, iy
                else:
                    jx, jy = (ix - (M - 1)) % Px, (iy - (M - 1)) % Py
                grid[(iz - (N_z - 1)) % Pz, jx, jy] = sp[iz, ix, iy]
    return np.fft.fftn(grid.transpose(3, 4, 0, 1, 2), axes=(2, 3, 4))


def slab_matvec_3d(psi_flat, T, khat3, M, N_z) -> np.ndarray:
    """(I - G T) psi by one 3-D FFT convolution (slab_scattering_gpu.py:95-155),
    circular kernel storage: the valid window is [0, N_z) x [0, M) x [0, M)."""
    Pz, Px, Py = khat3.shape[2:]
    psi = psi_flat.reshape(N_z, M, M, 9)
    tau = np.einsum("lmnab,lmnb->lmna", T, psi).transpose(3, 0, 1, 2)
    pad = np.zeros((9, Pz, Px, Py), dtype=complex)
    pad[:, :N_z, :M, :M] = tau
    acc = np.fft.ifftn(np.einsum("ijzxy,jzxy->izxy", khat3, np.fft.fftn(pad, axes=(1, 2, 3))),
                       axes=(1, 2, 3))
    acc = acc[:, :N_z, :M, :M]
    return (psi - acc.transpose(1, 2, 3, 0)).ravel()


def direct_matvec(psi_flat, T, M, N_z, a, omega, ref: tm.Medium) -> np.ndarray:
    """Brute-force O(N^2) dense sum over all voxel pairs (test_slab_scattering.py:225-244)."""
    psi = psi_flat.reshape(N_z, M, M, 9)
    tau = np.einsum("lmnab,lmnb->lmna", T, psi)
    c = centres(M, N_z, a)
    acc = np.zeros_like(psi)
    idx = list(np.ndindex(N_z, M, M))
    for i1 in idx:
        for i2 in idx:
            if i1 == i2:
                continue
            P = prop.propagator_9x9(c[i1] - c[i2], omega, ref.alpha, ref.beta, ref.rho)
            acc[i1] += P @ tau[i2]
    return (psi - acc).ravel()


def dense_operator(T, M, N_z, a, omega, ref: tm.Medium) -> np.ndarray:
    """Dense (9N x 9N) matrix of I - G T, for direct solves on tiny slabs."""
    n = N_z * M * M * 9
    A = np.zeros((n, n), dtype=complex)
    c = centres(M, N_z, a)
    idx = list(np.ndindex(N_z, M, M))
    Tf = T.reshape(-1, 9, 9)
    for p1, i1 in enumerate(idx):
        for p2, i2 in enumerate(idx):
            if p1 == p2:
                continue
            P = prop.propagator_9x9(c[i1] - c[i2], omega, ref.alpha, ref.beta, ref.rho)
            A[9 * p1:9 * p1 + 9, 9 * p2:9 * p2 + 9] = -P @ Tf[p2]
    return A + np.eye(n)


# ── incident field (slab_scattering.py:426-478; sphere_scattering.py:1256-1287) ──


def plane_wave_strain_voigt(k_hat, pol, k_mag) -> np.ndarray:
    """eps_lm = (ik/2)(k_l pol_m + k_m pol_l), engineering (doubled) shear."""
    eps = 0.5j * k_mag * (np.outer(k_hat, pol) + np.outer(pol, k_hat))
    out = np.zeros(6, dtype=complex)
    for i, (p, q) in enumerate(prop.VOIGT_PAIRS):
        out[i] = eps[p, q] if p == q else 2.0 * eps[p, q]
    return out


def incident_field(M, N_z, a, omega, ref: tm.Medium, k_hat, wave_type) -> np.ndarray:
    k_hat = np.asarray(k_hat, dtype=float)
    k_hat = k_hat / np.linalg.norm(k_hat)
    if wave_type == "P":
        k = omega / ref.alpha
        pol = k_hat.copy()
    elif wave_type == "S":
        k = omega / ref.beta
        cr = np.cross(k_hat, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(cr) < 1e-10:
            pol = np.array([0.0, 1.0, 0.0])
        else:
            pol = np.cross(cr, k_hat)
            pol = pol / np.linalg.norm(pol)
    else:
        raise ValueError(f"wave_type must be 'P' or 'S', got '{wave_type}'")
    v = np.zeros(9, dtype=complex)
    v[:3] = pol
    v[3:] = plane_wave_strain_voigt(k_hat, pol, k)
    phase = np.exp(1j * k * np.einsum("k,lmnk->lmn", k_hat, centres(M, N_z, a)))
    return v[None, None, None, :] * phase[..., None]


# ── solve (slab_scattering.py:486-565) ──────────────────────────────────


def solve(M, N_z, a, Dlambda, Dmu, Drho, omega, ref, k_hat, wave_type="P",
          tol=1e-6, max_iter=500, *, periodic=False, restart=20, volume_averaged=False,
          n_orders=2):
    """SciPy GMRES(restart), x0 = b, followed by the true residual.

    Returns (psi (N_z,M,M,9), psi0, T, n_matvec, rel_residual).
    """
    T = slab_tmatrices(M, N_z, a, Dlambda, Dmu, Drho, omega, ref)
    kh = slab_kernels(M, N_z, a, omega, ref, periodic=periodic, volume_averaged=volume_averaged,
                      n_orders=n_orders)
    psi0 = incident_field(M, N_z, a, omega, ref, k_hat, wave_type)
    n = N_z * M * M * 9
    count = [0]

    def mv(x):
        count[0] += 1
        return slab_matvec(x, T, kh, M, N_z, periodic=periodic)

    b = psi0.ravel()
    x, _info = gmres(LinearOperator((n, n), matvec=mv, dtype=complex), b, x0=b.copy(),
                     rtol=tol, atol=0.0, maxiter=max_iter, restart=restart)
    res = np.linalg.norm(mv(x) - b) / np.linalg.norm(b)
    return x.reshape(N_z, M, M, 9), psi0, T, count[0], float(res)


# ── outputs (slab_scattering.py:573-735) ────────────────────────────────


def voigt_to_tensor(v6) -> np.ndarray:
    """Voigt stress (doubled shear) -> symmetric 3x3 (slab_scattering.py:573-586)."""
    S = np.zeros((3, 3), dtype=complex)
    S[0, 0], S[1, 1], S[2, 2] = v6[0], v6[1], v6[2]
    S[1, 2] = S[2, 1] = v6[3] / 2.0
    S[0, 2] = S[2, 0] = v6[4] / 2.0
    S[0, 1] = S[1, 0] = v6[5] / 2.0
    return S


def vertical_slowness(s, p) -> complex:
    """eta = sqrt(s^2 - p^2), Im >= 0 branch (kennett_layers.py:273-287)."""
    eta = np.sqrt(complex((s + p) * (s - p)))
    if eta.imag < 0.0 or (eta.imag == 0.0 and eta.real < 0.0):
        eta = -eta
    return eta


def rpp_periodic(psi, T, M, N_z, a, omega, ref: tm.Medium, p=0.0) -> complex:
    """Weyl specular R_PP (slab_scattering.py:672-735)."""
    d = 2.0 * a
    eta = vertical_slowness(complex(1.0 / ref.alpha), p)
    kz = omega * eta
    r_hat = np.array([-eta * ref.alpha, p * ref.alpha, 0.0])
    src = np.einsum("lmnab,lmnb->lmna", T, psi)
    c = centres(M, N_z, a)
    total = 0j
    for l in range(N_z):
        f = src[l, :, :, :3].mean(axis=(0, 1))
        sig = voigt_to_tensor(src[l, :, :, 3:].mean(axis=(0, 1)))
        Q = -np.dot(r_hat, f) - 1j * kz * np.dot(r_hat, sig @ r_hat)
        total += Q * np.exp(1j * kz * c[l, 0, 0, 0])
    return complex(-1j / (2.0 * kz * d**2 * ref.rho * ref.alpha**2) * total)


def reflected_field(psi, T, M, N_z, a, omega, ref: tm.Medium, k_hat):
    """Far-field (R_PP, R_PS, R_SP) per unit area (slab_scattering.py:589-669)."""
    k_hat = np.asarray(k_hat, float) / np.linalg.norm(k_hat)
    kP, kS = omega / ref.alpha, omega / ref.beta
    r_hat = -k_hat
    gP = 1.0 / (4.0 * np.pi * ref.rho * ref.alpha**2)
    gS = 1.0 / (4.0 * np.pi * ref.rho * ref.beta**2)
    src = np.einsum("lmnab,lmnb->lmna", T, psi)
    c = centres(M, N_z, a)
    uP = np.zeros(3, dtype=complex)
    uS = np.zeros(3, dtype=complex)
    for idx in np.ndindex(N_z, M, M):
        f = src[idx][:3]
        sig = voigt_to_tensor(src[idx][3:])
        rc = c[idx]
        sr = sig @ r_hat
        QP = np.dot(r_hat, f) - 1j * kP * np.dot(r_hat, sr)
        uP -= gP * np.exp(-1j * kP * np.dot(r_hat, rc)) * QP * r_hat
        QS = f - 1j * kS * sr
        uS -= gS * np.exp(-1j * kS * np.dot(r_hat, rc)) * (QS - np.dot(r_hat, QS) * r_hat)
    area = (M * 2.0 * a) ** 2
    uP /= area
    uS /= area
    RPP = complex(np.dot(r_hat, uP))
    cr = np.cross(r_hat, np.array([1.0, 0.0, 0.0]))
    if np.linalg.norm(cr) < 1e-10:
        sv = np.array([0.0, 1.0, 0.0])
    else:
        sv = np.cross(cr, r_hat)
        sv = sv / np.linalg.norm(sv)
    return RPP, complex(np.dot(sv, uS)), RPP


def kennett_layer_rpp_normal(ref: tm.Medium, Dlambda, Dmu, Drho, H, omega) -> complex:
    """Normal-incidence R_PP of one uniform layer between background half-spaces.

    Closed form of the Kennett recursion (kennett_layers.py:549-640) for the
    3-layer stack of slab_scattering.py:738-783 with interface coefficient
    r = (Z_layer - Z_bg)/(Z_layer + Z_bg) (test_kennett_layers.py:443-461),
    Ru = -Rd and Tu Td = 1 - r^2:  R = r (1 - E) / (1 - r^2 E), E = exp(2 i w H / alpha_layer).
    """
    lam = ref.lam + Dlambda
    mu = ref.mu + Dmu
    rho = ref.rho + Drho
    al = np.sqrt((lam + 2.0 * mu) / rho)
    r = (rho * al - ref.rho * ref.alpha) / (rho * al + ref.rho * ref.alpha)
    E = np.exp(2j * omega * H / al)
    return complex(r * (1.0 - E) / (1.0 - r * r * E))
