"""Foldy-Lax T-matrix of a voxelised sphere (FFT-accelerated, on the device).

Drop-in for cubic_scattering/sphere_scattering_fft.py:228-326 and
sphere_scattering_fft_gpu.py:132-244 (compute_sphere_foldy_lax_fft[_gpu]) and
for the dense compute_sphere_foldy_lax (sphere_scattering.py:140-223).

A sphere voxelised on an n x n x n grid of sub-cubes is the uniform
n x n x n slab whose contrast is zero outside the sphere: outside cells carry
no source (T0 = 0), so the equations of the inside cells are exactly the
sphere's 9N x 9N Foldy-Lax system.  The solve therefore runs on the slab
plan of this package (spectral G0 table of the sub-cube lattice, five-pass
FFT matvec, device GMRES), once per incident pattern (9 right-hand sides,
resonance_tmatrix.py:526-593, built at the inside centres), and the composite
T-matrix column T_loc sum_m psi_m is the device sum of tau = T0 psi over the
slab.  The 9 solves are independent and share the operator: by default they
run as one batch (msc_gmres_batched on nine plans of one grid, the first
building G0 and the others copying it), so every operator pass and Arnoldi
kernel is one launch for all nine right-hand sides; with ``workers`` > 0 host
threads instead each run a plan on a stream of their own.
"""

from __future__ import annotations

import os
import threading
import warnings
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np
import torch
from numpy.typing import NDArray

from . import _lib
from .effective_contrasts import (
    MaterialContrast,
    ReferenceMedium,
    compute_cube_tmatrix,
    sub_cell_tmatrix_9x9,
)
from .slab_scattering import SlabGeometry, SlabMaterial, SlabPlan

VOIGT_PAIRS = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]


@dataclass
class SphereDecompositionResult:
    """Composite T-matrix of a voxelised sphere (sphere_scattering.py:46-76)."""

    T3x3: NDArray
    T_comp_9x9: NDArray
    centres: NDArray
    n_sub: int
    n_cells: int
    a_sub: float
    condition_number: float
    psi_exc: NDArray
    omega: float
    radius: float
    ref: ReferenceMedium
    contrast: MaterialContrast


def _grid(radius: float, n_sub: int, inclusive: bool):
    a_sub = radius / n_sub
    c = (np.arange(n_sub) - n_sub / 2.0 + 0.5) * (2.0 * a_sub)
    g = np.stack(np.meshgrid(c, c, c, indexing="ij"), axis=-1).reshape(-1, 3)
    r = np.linalg.norm(g, axis=1)
    keep = r <= radius if inclusive else r < radius
    idx = np.stack(np.meshgrid(*(np.arange(n_sub),) * 3, indexing="ij"), axis=-1).reshape(-1, 3)
    return idx[keep], g[keep], a_sub, keep


def _build_grid_index_map(radius: float, n_sub: int):
    """(grid indices (N, 3), centres (N, 3), a_sub) of the cells with |x| < radius
    (sphere_scattering_fft.py:42-80)."""
    idx, cen, a_sub, _ = _grid(radius, n_sub, inclusive=False)
    return idx, cen, a_sub


def sphere_sub_cell_centres(radius: float, n_sub: int):
    """(centres (N, 3), a_sub) of the cells with |x| <= radius (sphere_scattering.py:116-137)."""
    _, cen, a_sub, _ = _grid(radius, n_sub, inclusive=True)
    return cen, a_sub


def _incident_patterns(centres, omega, ref, k_hat, wave_type):
    """(N, 9 components, 9 patterns) incident field of resonance_tmatrix.py:526-593."""
    kh = np.array([0.0, 0.0, 1.0] if k_hat is None else k_hat, dtype=float)
    kh = kh / np.linalg.norm(kh)
    k = omega / (ref.beta if wave_type == "S" else ref.alpha)
    ph = np.exp(1j * k * (centres @ kh))
    dx = centres - centres.mean(axis=0)
    N = len(centres)
    U = np.zeros((N, 9, 9), dtype=complex)
    for p in range(3):
        U[:, p, p] = ph
    for al, (p, q) in enumerate(VOIGT_PAIRS):
        if p == q:
            U[:, p, 3 + al] = dx[:, p] * ph
        else:
            U[:, p, 3 + al] = 0.5 * dx[:, q] * ph
            U[:, q, 3 + al] = 0.5 * dx[:, p] * ph
        U[:, 3 + al, 3 + al] = ph
    return U


def _solve(omega, radius, ref, contrast, n_sub, k_hat, wave_type, inclusive, tol, max_iter,
           restart, precision, x0, device, workers=0):
    if wave_type not in ("P", "S"):
        raise ValueError(f"wave_type must be 'P' or 'S', got '{wave_type}'")
    if n_sub < 1:
        raise ValueError("n_sub must be >= 1")
    if int(workers) < 0:
        raise ValueError("workers must be >= 0 (0: one batch of the 9 right-hand sides)")
    if device is None:
        device = _lib.require_device()
    _, cen, a_sub, keep = _grid(radius, n_sub, inclusive)
    mask = keep.reshape(n_sub, n_sub, n_sub)
    N = len(cen)
    g = SlabGeometry(M=n_sub, N_z=n_sub, a=a_sub)
    mat = SlabMaterial(Dlambda=np.where(mask, float(contrast.Dlambda), 0.0),
                       Dmu=np.where(mask, float(contrast.Dmu), 0.0),
                       Drho=np.where(mask, float(contrast.Drho), 0.0), ref=ref)
    U = _incident_patterns(cen, omega, ref, k_hat, wave_type)
    flat = np.flatnonzero(keep)
    psi = np.zeros((N, 9, 9), dtype=complex)
    T = np.zeros((9, 9), dtype=complex)
    n_bad = []
    if int(workers) == 0:
        return _solve_batched(g, mat, omega, ref, U, flat, N, n_sub, tol, max_iter, restart,
                              precision, x0, device, cen, a_sub)
    nw = min(int(workers), 9)

    def make_plan(src):
        plan = SlabPlan(g, device=device)
        if nw > 1:  # no graph capture beside the other workers' allocations
            _lib.check(plan.lib.msc_plan_use_graphs(plan._h, 0))
        if src is None:
            plan.build_g0(omega, ref)
        else:
            plan.copy_g0_from(src)
        plan.set_material(mat, omega)
        return plan

    # The first plan builds G0 (and, with several workers, tunes its tiles
    # once); the others copy both.
    with torch.cuda.device(device):
        plan0 = make_plan(None)
        if nw > 1 and os.environ.get("MSC_AUTOTUNE", "1") != "0":
            plan0.autotune("c128")
            if precision == "mixed":
                plan0.autotune("c64")
        torch.cuda.current_stream(device).synchronize()

    def run(cols, stream, plan):
        # One plan per worker on its own stream; the 9 right-hand sides are
        # independent solves, so the workers' launches overlap on the device.
        with torch.cuda.device(device), torch.cuda.stream(stream):
            if plan is None:
                plan = make_plan(plan0)
            idx = torch.as_tensor(flat, device=plan.device)
            for col in cols:
                b = np.zeros((n_sub ** 3, 9), dtype=complex)
                b[flat] = U[:, :, col]
                bt = torch.as_tensor(b.reshape(-1)).to(plan.device)
                x, st = plan.gmres(bt, x0=x0, tol=tol, max_iter=max_iter, restart=restart,
                                   precision=precision)
                if st["rel_residual"] > tol:
                    n_bad.append((col, st["rel_residual"], st["iterations"]))
                ls, _ = plan.source_sums(x, 0.0, 0.0, np.zeros(3))
                T[:, col] = ls.sum(axis=0)
                psi[:, :, col] = x.reshape(-1, 9)[idx].cpu().numpy()
            stream.synchronize()

    streams = [torch.cuda.current_stream(device)] if nw == 1 else \
        [torch.cuda.Stream(device) for _ in range(nw)]
    groups = [list(range(w, 9, nw)) for w in range(nw)]
    if nw == 1:
        run(groups[0], streams[0], plan0)
    else:
        with ThreadPoolExecutor(max_workers=nw) as ex:
            futs = [ex.submit(run, grp, st, plan0 if w == 0 else None)
                    for w, (grp, st) in enumerate(zip(groups, streams))]
            for f in futs:
                f.result()
    n_bad.sort()
    return cen, a_sub, N, psi.reshape(9 * N, 9), T, n_bad


# The nine plans of the last batched sphere solve, kept for the next solve on
# the same sub-cube grid (their buffers and GMRES workspaces are reused);
# slab_scattering.release_plans() drops them.
_BATCH_PLANS: dict = {}
_BATCH_LOCK = threading.Lock()


def _solve_batched(g, mat, omega, ref, U, flat, N, n_sub, tol, max_iter, restart, precision, x0,
                   device, cen, a_sub):
    """The 9 right-hand sides as one batch: nine plans of the sub-cube grid
    (G0 built once and copied), one msc_gmres_batched call."""
    from .slab_scattering import gmres_batched

    key = (str(torch.device(device)), g.M, g.N_z, float(g.a))
    with _BATCH_LOCK:
        plans = _BATCH_PLANS.pop(key, None)
    with torch.cuda.device(device):
        if plans is None:
            plans = [SlabPlan(g, device=device) for _ in range(9)]
        plans[0].build_g0(omega, ref)
        for plan in plans[1:]:
            plan.copy_g0_from(plans[0])
        for plan in plans:
            plan.set_material(mat, omega)
        # the nine incident patterns in one upload, scattered onto the grid
        # on the device; the nine solutions gathered back in one download
        dev = plans[0].device
        idx = torch.as_tensor(flat, device=dev)
        U_d = _lib.to_device(U, dev)
        bs = []
        for col in range(9):
            b = torch.zeros((n_sub ** 3, 9), dtype=torch.complex128, device=dev)
            b[idx] = U_d[:, :, col]
            bs.append(b.reshape(-1))
        m = int(max_iter) if restart is None else int(restart)
        xs, sts = gmres_batched(plans, bs, x0, tol=tol, max_iter=max_iter, restart=m,
                                precision=precision)
        T = np.zeros((9, 9), dtype=complex)
        n_bad = []
        for col, (plan, x, st) in enumerate(zip(plans, xs, sts)):
            if st["rel_residual"] > tol:
                n_bad.append((col, st["rel_residual"], st["iterations"]))
            ls, _ = plan.source_sums(x, 0.0, 0.0, np.zeros(3))
            T[:, col] = ls.sum(axis=0)
        psi = _lib.to_host(torch.stack([x.reshape(-1, 9)[idx] for x in xs], dim=-1))
    with _BATCH_LOCK:
        _BATCH_PLANS.clear()  # nine plans hold the grid's workspaces: keep one set
        _BATCH_PLANS[key] = plans
    return cen, a_sub, N, psi.reshape(9 * N, 9), T, n_bad


def release_batch_plans() -> None:
    """Drop the nine plans kept by the batched solve."""
    with _BATCH_LOCK:
        _BATCH_PLANS.clear()


def _result(omega, radius, ref, contrast, n_sub, cen, a_sub, N, psi, T):
    return SphereDecompositionResult(T3x3=T[:3, :3].copy(), T_comp_9x9=T, centres=cen,
                                     n_sub=n_sub, n_cells=N, a_sub=a_sub,
                                     condition_number=float("nan"), psi_exc=psi, omega=omega,
                                     radius=radius, ref=ref, contrast=contrast)


def compute_sphere_foldy_lax_fft(omega: float, radius: float, ref: ReferenceMedium,
                                 contrast: MaterialContrast, n_sub: int,
                                 k_hat: NDArray | None = None, wave_type: str = "S",
                                 gmres_tol: float = 1e-8,
                                 gmres_maxiter: int = 200) -> SphereDecompositionResult:
    """Sphere T-matrix by FFT Foldy-Lax: GMRES(20) from the incident field per
    pattern to ``gmres_tol`` (sphere_scattering_fft.py:228-326)."""
    cen, a_sub, N, psi, T, bad = _solve(omega, radius, ref, contrast, n_sub, k_hat, wave_type,
                                        False, gmres_tol, gmres_maxiter * 20, 20, "fp64", "born",
                                        None, 0)
    for col, rr, it in bad:
        warnings.warn(f"GMRES did not converge for column {col} (rel_res={rr:.2e} after {it} "
                      "iterations)", UserWarning, stacklevel=2)
    return _result(omega, radius, ref, contrast, n_sub, cen, a_sub, N, psi, T)


def compute_sphere_foldy_lax_fft_gpu(omega: float, radius: float, ref: ReferenceMedium,
                                     contrast: MaterialContrast, n_sub: int,
                                     k_hat: NDArray | None = None, wave_type: str = "S",
                                     gmres_tol: float = 1e-8, gmres_maxiter: int = 200,
                                     initial_guess: str = "born", device=None,
                                     dtype: torch.dtype | None = None,
                                     restart: int | None = None,
                                     workers: int = 0) -> SphereDecompositionResult:
    """Sphere T-matrix on the device (sphere_scattering_fft_gpu.py:132-244).

    ``dtype`` complex128 (default) solves in FP64; complex64 selects the mixed
    form (complex64 operator inside FP64 refinement, same FP64 residual
    target).  ``restart=None`` keeps the reference's unrestarted GMRES.
    ``workers=0`` (default) solves the 9 right-hand sides as one batch (one
    launch per operator pass and Arnoldi kernel for all nine); ``workers`` > 0
    runs that many concurrent solves on plans of their own streams instead."""
    if initial_guess not in ("born", "zero"):
        raise ValueError(f"Invalid initial_guess '{initial_guess}'. Must be 'born' or 'zero'.")
    precision = "mixed" if dtype == torch.complex64 else "fp64"
    cen, a_sub, N, psi, T, bad = _solve(omega, radius, ref, contrast, n_sub, k_hat, wave_type,
                                        False, gmres_tol, gmres_maxiter, restart, precision,
                                        initial_guess, device, workers)
    for col, rr, it in bad:
        warnings.warn(f"GPU GMRES did not converge for column {col}: rel_res={rr:.2e} after "
                      f"{it} iterations", UserWarning, stacklevel=2)
    return _result(omega, radius, ref, contrast, n_sub, cen, a_sub, N, psi, T)


def compute_sphere_foldy_lax(omega: float, radius: float, ref: ReferenceMedium,
                             contrast: MaterialContrast, n_sub: int,
                             k_hat: NDArray | None = None,
                             wave_type: str = "S") -> SphereDecompositionResult:
    """The dense solve of sphere_scattering.py:140-223 (cells with |x| <= radius),
    here by the same device GMRES driven to 1e-13 FP64 relative residual;
    ``condition_number`` is not computed (NaN)."""
    cen, a_sub, N, psi, T, bad = _solve(omega, radius, ref, contrast, n_sub, k_hat, wave_type,
                                        True, 1e-13, 2000, 60, "fp64", "born", None, 3)
    for col, rr, it in bad:
        warnings.warn(f"solve did not reach 1e-13 for column {col} (rel_res={rr:.2e})",
                      UserWarning, stacklevel=2)
    return _result(omega, radius, ref, contrast, n_sub, cen, a_sub, N, psi, T)


def sphere_tloc(omega: float, radius: float, ref: ReferenceMedium, contrast: MaterialContrast,
                n_sub: int) -> NDArray:
    """The sub-cell 9x9 T-matrix the sphere solve uses (resonance_tmatrix.py:461-487)."""
    a_sub = radius / n_sub
    return sub_cell_tmatrix_9x9(compute_cube_tmatrix(omega, a_sub, ref, contrast), omega, a_sub)
