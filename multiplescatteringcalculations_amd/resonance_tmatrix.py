"""Resonance-regime cube T-matrix: internal Foldy-Lax over n^3 sub-cells.

Drop-in for cubic_scattering/resonance_tmatrix.py's public path
(sub_cell_centres 436-455, suggest_n_subcells 495-523, ResonanceTmatrixResult
596-647, compute_resonance_tmatrix 650-749 with _solve_coupled 752-855,
voigt_tmatrix_from_resonance_result 906-965, validate_rayleigh_limit 968-1011,
scattering_order_decomposition 1014-1097).  Every solve runs in the HIP
library (``msc_resonance_tmatrices``): the sub-cell propagator table, the
9N x 9N systems, their elimination with partial pivoting and the composite
T-matrix are built on the device, for one cube or for a batch of contrasts
(``compute_resonance_tmatrices``) in one call.

``condition_number`` of ``compute_resonance_tmatrix`` is the reference's:
the 2-norm condition number of the device-assembled 9N x 9N matrix
(``msc_resonance_system``), from its singular values on the host
(np.linalg.cond, as resonance_tmatrix.py:788 does), with the same 1e10
warning.  The batch ``compute_resonance_tmatrices`` returns the pivot ratios
max|u_ii| / min|u_ii| of the device eliminations instead (an SVD per cube
would cost more than the batch); a pivot ratio bounds the condition number
from below.
"""

from __future__ import annotations

import ctypes
import warnings
from dataclasses import dataclass

import numpy as np
import torch
from numpy.typing import NDArray

from . import _lib
from .effective_contrasts import (
    CubeTMatrixResult,
    MaterialContrast,
    ReferenceMedium,
    compute_cube_tmatrix,
    sub_cell_tmatrix_9x9,
)

_KA_SUB_WARN = 0.50
_KA_SUB_MAX = 0.30
VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


def sub_cell_centres(a: float, n: int) -> NDArray[np.floating]:
    """Centres of the n^3 sub-cells of [-a, a]^3, index (i0 n + i1) n + i2, shape (n^3, 3)."""
    s = a / n
    c = np.linspace(-a + s, a - s, n)
    g = np.meshgrid(c, c, c, indexing="ij")
    return np.stack([x.ravel() for x in g], axis=1)


def suggest_n_subcells(omega: float, a: float, ref: ReferenceMedium,
                       ka_threshold: float = _KA_SUB_MAX) -> int:
    """Smallest n with k_S a / n <= ka_threshold (at least 1)."""
    return max(1, int(np.ceil((omega / ref.beta) * a / ka_threshold)))


@dataclass
class ResonanceTmatrixResult:
    """Composite T-matrix of one cube (resonance_tmatrix.py:596-647)."""

    T3x3: NDArray
    ka_cube: float
    ka_sub: float
    n_sub: int
    omega: float
    a: float
    ref: ReferenceMedium
    contrast: MaterialContrast
    condition_number: float
    rayleigh_result: CubeTMatrixResult
    T_comp_9x9: NDArray
    force_monopole: NDArray
    stress_dipole_voigt: NDArray
    psi_exc: NDArray
    centres: NDArray
    T_loc_9x9: NDArray
    n_iterations_converged: int | None = None


def _run(omega, a, ref, Dl, Dm, Dr, n_sub, k_hat, wave_type, neumann_order, neumann_tol,
         want_psi, device):
    if wave_type not in ("P", "S"):
        raise ValueError(f"wave_type must be 'P' or 'S', got '{wave_type}'")
    if device is None:
        device = _lib.require_device()
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError("the resonance solver runs on a HIP device only")
    lib = _lib.load()
    kh = np.array([0.0, 0.0, 1.0] if k_hat is None else k_hat, dtype=np.float64).reshape(3)
    count = len(Dl)
    N = n_sub ** 3
    dl = torch.as_tensor(np.ascontiguousarray(Dl, dtype=np.float64), device=device)
    dm = torch.as_tensor(np.ascontiguousarray(Dm, dtype=np.float64), device=device)
    dr = torch.as_tensor(np.ascontiguousarray(Dr, dtype=np.float64), device=device)
    tc = torch.empty((count, 9, 9), dtype=torch.complex128, device=device)
    px = (torch.empty((count, 9 * N, 9), dtype=torch.complex128, device=device)
          if want_psi else None)
    ratio = np.empty(count, dtype=np.float64)
    iters = np.empty(count, dtype=np.int32)
    khc = (ctypes.c_double * 3)(*kh)
    with torch.cuda.device(device):
        torch.cuda.current_stream(device).synchronize()
        rc = lib.msc_resonance_tmatrices(
            float(omega), float(a), float(ref.alpha), float(ref.beta), float(ref.rho), int(n_sub),
            ctypes.cast(khc, ctypes.c_void_p), 1 if wave_type == "S" else 0, int(count),
            dl.data_ptr(), dm.data_ptr(), dr.data_ptr(), int(neumann_order), float(neumann_tol),
            tc.data_ptr(), px.data_ptr() if px is not None else None,
            ratio.ctypes.data, iters.ctypes.data, None)
    _lib.check(rc)
    return tc, px, ratio, iters


def compute_resonance_tmatrices(omega: float, a: float, ref: ReferenceMedium, Dlambda, Dmu, Drho,
                                n_sub: int | None = None, k_hat: NDArray | None = None,
                                wave_type: str = "S", device=None,
                                keep_on_device: bool = False):
    """Composite 9x9 T-matrices of a batch of cubes with contrasts (Dlambda[i],
    Dmu[i], Drho[i]), all solved in one device call.  Returns (count, 9, 9)
    complex128 (a torch tensor on the device with ``keep_on_device``) and the
    pivot ratios."""
    if n_sub is None:
        n_sub = suggest_n_subcells(omega, a, ref)
    Dl, Dm, Dr = (np.atleast_1d(np.asarray(x, dtype=np.float64)) for x in (Dlambda, Dmu, Drho))
    if not (Dl.shape == Dm.shape == Dr.shape) or Dl.ndim != 1:
        raise ValueError("Dlambda, Dmu, Drho must be 1-D arrays of equal length")
    tc, _, ratio, _ = _run(omega, a, ref, Dl, Dm, Dr, n_sub, k_hat, wave_type, 0, 0.0, False,
                           device)
    return (tc if keep_on_device else _lib.to_host(tc)), ratio


def _condition_number(omega: float, a: float, ref: ReferenceMedium, contrast: MaterialContrast,
                      n_sub: int) -> float:
    """2-norm condition number of the cube's internal Foldy-Lax matrix
    I - P (I x T_loc) (resonance_tmatrix.py:783-788): the matrix is assembled
    on the device and its singular values are taken on the host."""
    device = torch.device(_lib.require_device())
    lib = _lib.load()
    n9 = 9 * n_sub ** 3
    A = torch.empty((n9, n9), dtype=torch.complex128, device=device)
    with torch.cuda.device(device):
        torch.cuda.current_stream(device).synchronize()
        rc = lib.msc_resonance_system(float(omega), float(a), float(ref.alpha), float(ref.beta),
                                      float(ref.rho), int(n_sub), float(contrast.Dlambda),
                                      float(contrast.Dmu), float(contrast.Drho), A.data_ptr(),
                                      None)
    _lib.check(rc)
    # column-major on the device; the condition number of A^T is that of A
    return float(np.linalg.cond(_lib.to_host(A)))


def compute_resonance_tmatrix(omega: float, a: float, ref: ReferenceMedium,
                              contrast: MaterialContrast, n_sub: int | None = None,
                              k_hat: NDArray | None = None, wave_type: str = "S",
                              neumann_order: int = 0,
                              neumann_tol: float = 1.0e-8) -> ResonanceTmatrixResult:
    """Full-wave T-matrix of one cube (resonance_tmatrix.py:650-855)."""
    if n_sub is None:
        n_sub = suggest_n_subcells(omega, a, ref)
    a_sub = a / n_sub
    kS = omega / ref.beta
    ka_cube, ka_sub = kS * a, kS * a_sub
    if ka_sub > _KA_SUB_WARN:
        warnings.warn(f"ka_sub = {ka_sub:.3f} > {_KA_SUB_WARN}. Sub-cells are outside the "
                      f"Rayleigh regime. Use n_sub >= {suggest_n_subcells(omega, a, ref)} "
                      f"(currently n_sub = {n_sub}).", UserWarning, stacklevel=2)
    ray_sub = compute_cube_tmatrix(omega, a_sub, ref, contrast)
    ray_full = compute_cube_tmatrix(omega, a, ref, contrast)
    tc, px, ratio, iters = _run(omega, a, ref, [contrast.Dlambda], [contrast.Dmu],
                                [contrast.Drho], n_sub, k_hat, wave_type, neumann_order,
                                neumann_tol, True, None)
    T = tc[0].cpu().numpy()
    cond = _condition_number(omega, a, ref, contrast, n_sub)
    if cond > 1.0e10:
        warnings.warn(f"Coupled Foldy-Lax condition number = {cond:.2e}. Near-resonance or "
                      "strong-coupling instability.", UserWarning, stacklevel=2)
    n_conv = None
    if neumann_order > 0:
        n_conv = int(iters[0]) if iters[0] >= 0 else None
        if n_conv is None:
            warnings.warn(f"Coupled Neumann series did not converge in {neumann_order} orders "
                          f"(ka_cube = {ka_cube:.3f}).", UserWarning, stacklevel=2)
    return ResonanceTmatrixResult(
        T3x3=T[:3, :3].copy(), ka_cube=ka_cube, ka_sub=ka_sub, n_sub=n_sub, omega=omega, a=a,
        ref=ref, contrast=contrast, condition_number=cond, rayleigh_result=ray_full,
        T_comp_9x9=T, force_monopole=T[:3, :3].copy(), stress_dipole_voigt=T[3:, 3:].copy(),
        psi_exc=_lib.to_host(px[0]), centres=sub_cell_centres(a, n_sub),
        T_loc_9x9=sub_cell_tmatrix_9x9(ray_sub, omega, a_sub), n_iterations_converged=n_conv)


def strain_from_displacement_traction(kx: float, ky: float, ref: ReferenceMedium) -> NDArray:
    """6x6 map (u_z, u_x, u_y, t_zz, t_xz, t_yz) -> Voigt strain of a plane wave with
    horizontal wavenumbers (kx, ky) (voigt_tmatrix.py:177-240)."""
    lam, mu = ref.lam, ref.mu
    l2m = lam + 2.0 * mu
    S = np.zeros((6, 6), dtype=complex)
    S[0, 1], S[0, 2], S[0, 3] = -lam * 1j * kx / l2m, -lam * 1j * ky / l2m, 1.0 / l2m
    S[1, 1] = 1j * kx
    S[2, 2] = 1j * ky
    S[3, 1], S[3, 2] = 1j * ky, 1j * kx
    S[4, 5] = 1.0 / mu
    S[5, 4] = 1.0 / mu
    return S


def _volume_integral_voigt(Ac: complex, Bc: complex, Cc: complex) -> NDArray:
    """Voigt form of A d_ij d_kl + B (d_ik d_jl + d_il d_jk) + C E_ijkl (resonance_tmatrix.py:858-882)."""
    S = np.zeros((6, 6), dtype=complex)
    S[:3, :3] = Bc
    for i in range(3):
        S[i, i] = Ac + 2.0 * Bc + Cc
        S[3 + i, 3 + i] = 2.0 * (Ac + Bc)
    return S


def _traction_from_voigt_strain(ref: ReferenceMedium) -> NDArray:
    """(t_zz, t_xz, t_yz) from Voigt strain by the background stiffness (885-903)."""
    C = np.zeros((3, 6), dtype=complex)
    C[0, 0], C[0, 1], C[0, 2] = ref.lam + 2.0 * ref.mu, ref.lam, ref.lam
    C[1, 5] = ref.mu
    C[2, 4] = ref.mu
    return C


def voigt_tmatrix_from_resonance_result(result: ResonanceTmatrixResult, kx: float = 0.0,
                                        ky: float = 0.0) -> NDArray:
    """6x6 T-matrix in the (u_z, u_x, u_y, t_zz, t_xz, t_yz) basis: G_self T_comp input_conv
    (resonance_tmatrix.py:906-965)."""
    ref = result.ref
    conv = np.zeros((9, 6), dtype=complex)
    conv[:3, :3] = np.eye(3)
    conv[3:, :] = strain_from_displacement_traction(kx, ky, ref)
    ray = result.rayleigh_result
    G = np.zeros((6, 9), dtype=complex)
    G[:3, :3] = ray.Gamma0 * np.eye(3)
    G[3:, 3:] = _traction_from_voigt_strain(ref) @ _volume_integral_voigt(ray.Ac, ray.Bc, ray.Cc)
    return G @ result.T_comp_9x9 @ conv


def validate_rayleigh_limit(omega: float, a: float, ref: ReferenceMedium,
                            contrast: MaterialContrast, rtol: float = 0.05) -> dict:
    """n_sub = 1 against the density-only Rayleigh T-matrix w^2 Drho* V I3 (968-1011)."""
    res = compute_resonance_tmatrix(omega, a, ref, contrast, n_sub=1)
    ray = compute_cube_tmatrix(omega, a, ref, contrast)
    T_ray = (2.0 * a) ** 3 * omega ** 2 * complex(ray.Drho_star) * np.eye(3, dtype=complex)
    err = float(np.linalg.norm(res.T3x3 - T_ray)) / max(float(np.linalg.norm(T_ray)), 1.0e-300)
    return {"ka_cube": res.ka_cube, "T3x3_resonance": res.T3x3, "T3x3_rayleigh": T_ray,
            "relative_error": err, "passed": err < rtol}


def scattering_order_decomposition(omega: float, a: float, ref: ReferenceMedium,
                                   contrast: MaterialContrast, n_sub: int, max_order: int = 10,
                                   tol: float = 1.0e-8) -> dict:
    """T3x3 by Neumann order next to the direct solve (resonance_tmatrix.py:1014-1097).

    As in the reference, order k is compute_resonance_tmatrix(neumann_order=k):
    k = 0 selects the direct solve, so entry 0 is the full T-matrix and entry
    1 is the first-order partial sum minus it."""
    full = compute_resonance_tmatrix(omega, a, ref, contrast, n_sub=n_sub)
    nf = float(np.linalg.norm(full.T3x3))
    orders, by, part, norms = [], [], [], []
    conv = None
    prev = None
    with warnings.catch_warnings():
        warnings.simplefilter("ignore", UserWarning)
        for k in range(max_order + 1):
            Tk = compute_resonance_tmatrix(omega, a, ref, contrast, n_sub=n_sub, neumann_order=k,
                                           neumann_tol=0.0).T3x3
            contrib = Tk.copy() if prev is None else Tk - prev
            orders.append(k)
            by.append(contrib)
            part.append(Tk.copy())
            nk = float(np.linalg.norm(contrib))
            norms.append(nk)
            if conv is None and nk < tol * max(nf, 1.0e-300):
                conv = k
            prev = Tk.copy()
    return {"orders": orders, "T3x3_by_order": by, "T3x3_partial": part, "T3x3_full": full.T3x3,
            "norm_by_order": norms, "converged_order": conv, "ka_cube": full.ka_cube}
