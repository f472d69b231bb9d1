"""Resonance-regime slab: every cube carries its full-wave composite T-matrix.

The reference solves the slab with the Rayleigh (quasi-static) 9x9 T0 of each
cube (cubic_scattering/slab_scattering.py:169-205) and computes, separately,
the full-wave T-matrix of one cube split into n^3 sub-cells
(compute_resonance_tmatrix, cubic_scattering/resonance_tmatrix.py:650-850).
Here the two meet: the composite T_comp of every distinct inclusion comes from
one batched device solve (msc_resonance_tmatrices), the slab operator applies
it per voxel through an index table (msc_set_tgeneral), and the slab's
specular R_PP is embedded between a water and an elastic half-space exactly as
the ocean-bottom caller embeds it (ocean_bottom.py:116-156, 234-323).  The
reference value is the Kennett reflectivity of the CPA effective medium of the
same contrast distribution (cpa_iteration.py:243-318 -> kennett_layers.py:
916-993, 549-760).

The same medium resolved into its n^3 sub-cells per cube (an ordinary slab of
half-width a/n with the Rayleigh T0 of each sub-cell, k_S a/n in the Rayleigh
range) is the cross-check of the composite solve: the two differ only in how
neighbouring cubes couple (centre-to-centre propagators between composite
T-matrices, or sub-cell to sub-cell).  It also converges where the
point-coupled composite system of a thick slab does not (the ordinary
Rayleigh slab of the same thickness stalls there as well).
"""

from __future__ import annotations

import time
import warnings
from dataclasses import dataclass, field

import numpy as np
import torch
from numpy.typing import NDArray

from . import _lib
from .cpa_iteration import Phase, compute_cpa
from .effective_contrasts import ReferenceMedium
from .kennett_layers import (
    FluidLayer,
    IsotropicLayer,
    LayerStack,
    complex_slowness,
    cubic_to_isotropic_layer,
    kennett_layers,
    psv_fluid_solid,
    vertical_slowness,
)
from .resonance_tmatrix import compute_resonance_tmatrices
from .slab_scattering import (
    SlabGeometry,
    SlabMaterial,
    SlabPlan,
    SlabResult,
    _incident_vector,
    compute_slab_scattering,
    slab_rpp_periodic,
)


def resonance_slab_tmatrices(geometry: SlabGeometry, material: SlabMaterial, omega: float,
                             k_hat, wave_type: str = "P", n_sub: int = 4, device=None):
    """Index table of the slab's composite T-matrices: (tidx (N_z, M, M) int32,
    table (K, 9, 9) complex128 on the device).  Entry 0 is the zero matrix of
    the background cubes; every distinct contrast triple of the inclusions
    gets one full-wave T_comp (n_sub^3 sub-cells, incident patterns along
    ``k_hat``), all solved in one device call."""
    if device is None:
        device = _lib.require_device()
    k_hat = np.asarray(k_hat, dtype=float)
    trip = np.stack([material.Dlambda.ravel(), material.Dmu.ravel(), material.Drho.ravel()], axis=1)
    incl = np.any(trip != 0.0, axis=1)
    tidx = np.zeros(trip.shape[0], dtype=np.int32)
    if incl.any():
        uniq, inv = np.unique(trip[incl], axis=0, return_inverse=True)
        tc, _ = compute_resonance_tmatrices(omega, geometry.a, material.ref, uniq[:, 0], uniq[:, 1],
                                            uniq[:, 2], n_sub=n_sub, k_hat=k_hat,
                                            wave_type=wave_type, device=device,
                                            keep_on_device=True)
        table = torch.cat([torch.zeros((1, 9, 9), dtype=torch.complex128, device=device), tc])
        tidx[incl] = 1 + inv.reshape(-1).astype(np.int32)
    else:
        table = torch.zeros((1, 9, 9), dtype=torch.complex128, device=device)
    g = geometry
    return tidx.reshape(g.N_z, g.M, g.M), table


def compute_resonance_slab_scattering(geometry: SlabGeometry, material: SlabMaterial,
                                      omega: float, k_hat, wave_type: str = "P", *,
                                      n_sub: int = 4, gmres_tol: float = 1e-8,
                                      max_iter: int = 3000, restart: int = 30,
                                      precision: str = "mixed", periodic: bool = True,
                                      preconditioner=None, plan: SlabPlan | None = None,
                                      device=None, warn: bool = True) -> SlabResult:
    """Solve (I - G0 T) psi = psi0 with the full-wave composite T of every cube.

    Returns a ``SlabResult`` whose ``slab_rpp_periodic`` and
    ``slab_reflected_field`` use the composite T (their source sums are
    reduced on the device with the solution); ``stats`` adds
    ``tmatrix_seconds`` and ``distinct_tmatrices``.  ``plan`` (a SlabPlan of
    the same geometry) is reused when given; its G0 is rebuilt when ``omega``
    or the medium changes."""
    k_hat_u, _, _ = _incident_vector(omega, material.ref, k_hat, wave_type)
    if plan is None:
        plan = SlabPlan(geometry, periodic=periodic, device=device)
    elif plan.geometry != geometry or plan.periodic != bool(periodic):
        raise ValueError("plan was made for another slab geometry or boundary")
    ref = material.ref
    if plan.omega != float(omega) or plan.ref is None or (plan.ref.alpha, plan.ref.beta,
                                                          plan.ref.rho) != (ref.alpha, ref.beta,
                                                                            ref.rho):
        plan.build_g0(omega, ref)
    t0 = time.perf_counter()
    tidx, table = resonance_slab_tmatrices(geometry, material, omega, k_hat_u, wave_type, n_sub,
                                           plan.device)
    plan.set_tgeneral(tidx, table)
    torch.cuda.synchronize(plan.device)
    t_tm = time.perf_counter() - t0
    psi0 = plan.incident_field(omega, ref, k_hat_u, wave_type)
    x, st = plan.gmres(psi0, "born", tol=gmres_tol, max_iter=max_iter, restart=restart,
                       precision=precision, precond=preconditioner)
    if warn and not st["converged"]:
        warnings.warn(f"GMRES did not converge: rel_res={st['rel_residual']:.2e} after "
                      f"{st['iterations']} iterations", stacklevel=2)
    st["tmatrix_seconds"] = t_tm
    st["distinct_tmatrices"] = int(table.shape[0]) - 1
    g = geometry
    kP, kS = omega / ref.alpha, omega / ref.beta
    ls, ph = plan.source_sums(x, kP, kS, -k_hat_u)
    res = SlabResult(psi=_lib.to_host(x.reshape(g.N_z, g.M, g.M, 9)), psi0=_lib.to_host(psi0),
                     geometry=geometry, material=material, omega=float(omega),
                     k_hat=np.asarray(k_hat, dtype=float), wave_type=wave_type,
                     n_gmres_iter=int(st["iterations"]), gmres_residual=float(st["rel_residual"]),
                     periodic=periodic, stats=st)
    res._sums = (ls, ph, (kP, kS), -k_hat_u)
    res._plan = plan
    res._plan_token = None
    return res


def subcell_slab(geometry: SlabGeometry, material: SlabMaterial,
                 n_sub: int) -> tuple[SlabGeometry, SlabMaterial]:
    """The slab resolved into n_sub^3 sub-cells per cube: an
    (n_sub M) x (n_sub M) x (n_sub N_z) slab of half-width a / n_sub whose
    sub-cells carry their cube's contrasts (sub-cell (i0 n + j0, ...) of cube
    (i0, ...), the sub-cell ordering of the composite T)."""
    n = int(n_sub)
    if n < 1:
        raise ValueError(f"n_sub must be >= 1, got {n_sub}")
    g = SlabGeometry(M=geometry.M * n, N_z=geometry.N_z * n, a=geometry.a / n)

    def rep(x):
        return np.ascontiguousarray(np.asarray(x, dtype=np.float64).repeat(n, 0).repeat(n, 1)
                                    .repeat(n, 2))

    return g, SlabMaterial(Dlambda=rep(material.Dlambda), Dmu=rep(material.Dmu),
                           Drho=rep(material.Drho), ref=material.ref)


def compute_subcell_slab_scattering(geometry: SlabGeometry, material: SlabMaterial,
                                    omega: float, k_hat, wave_type: str = "P", *,
                                    n_sub: int = 4, gmres_tol: float = 1e-8,
                                    max_iter: int = 3000, restart: int = 30,
                                    precision: str = "mixed", periodic: bool = True,
                                    preconditioner=None, device=None) -> SlabResult:
    """The periodic slab solve of ``subcell_slab(geometry, material, n_sub)``
    (Rayleigh T0 per sub-cell); its R_PP per unit area is directly comparable
    with the composite solve's."""
    gf, mf = subcell_slab(geometry, material, n_sub)
    return compute_slab_scattering(gf, mf, omega, k_hat, wave_type, gmres_tol, max_iter,
                                   periodic=periodic, restart=restart, precision=precision,
                                   device=device, preconditioner=preconditioner)


def uniform_contrast_phases(ref: ReferenceMedium, phi: float, rel_contrast: float,
                            order: int = 3) -> list[Phase]:
    """CPA phases of the i.i.d. contrast distribution of random_iid_slab_material:
    the background with fraction 1 - phi, and the inclusions' relative contrasts
    U[-c, c]^3 in (lambda, mu, rho) as a tensor Gauss-Legendre rule of ``order``
    points per axis (order^3 phases, fraction phi times the rule's weight)."""
    xg, wg = np.polynomial.legendre.leggauss(order)
    wg = wg / 2.0
    phases = [Phase(ref.lam, ref.mu, ref.rho, 1.0 - phi)]
    for i, a in enumerate(xg):
        for j, b in enumerate(xg):
            for k, c in enumerate(xg):
                phases.append(Phase(ref.lam * (1 + rel_contrast * a), ref.mu * (1 + rel_contrast * b),
                                    ref.rho * (1 + rel_contrast * c), phi * wg[i] * wg[j] * wg[k]))
    return phases


@dataclass
class ResonanceOceanBottomResult:
    """Per frequency: the slab's R_PP, the embedded R_PP, the homogeneous
    background's and the CPA medium's Kennett R_PP, solver statistics."""

    omega: NDArray
    ka: NDArray
    R_slab: NDArray
    R_total: NDArray
    R_bg: NDArray
    R_cpa: NDArray
    iterations: list[int] = field(default_factory=list)
    converged: list[bool] = field(default_factory=list)
    tmatrix_seconds: list[float] = field(default_factory=list)
    solve_seconds: list[float] = field(default_factory=list)
    distinct_tmatrices: list[int] = field(default_factory=list)
    cpa_media: list = field(default_factory=list)
    R_slab_subcell: NDArray | None = None
    R_total_subcell: NDArray | None = None
    subcell_iterations: list[int] = field(default_factory=list)
    subcell_converged: list[bool] = field(default_factory=list)
    subcell_seconds: list[float] = field(default_factory=list)


def resonance_ocean_bottom(geometry: SlabGeometry, material: SlabMaterial, omegas, *,
                           water_alpha: float = 1500.0, water_rho: float = 1000.0,
                           hs_alpha: float = 4000.0, hs_beta: float = 2300.0,
                           hs_rho: float = 2600.0, p: float = 0.0, n_sub: int = 4,
                           cpa_phases: list[Phase] | None = None, gmres_tol: float = 1e-8,
                           precision: str = "mixed", restart: int = 30, max_iter: int = 3000,
                           subcell: bool = False, preconditioners: dict | None = None,
                           subcell_preconditioner=None,
                           device=None) -> ResonanceOceanBottomResult:
    """A resonance-regime slab of the background medium ``material.ref``
    embedded between a water half-space (above) and an elastic half-space
    (below), over the frequencies ``omegas``.

    For each frequency: the composite T of every cube (n_sub^3 sub-cells), the
    periodic slab solve and its specular R_PP; the sub-water reflectivity
    MT = e^{2 i w eta H} R[background | half-space] + R_slab dressed by the
    water-background interface, Rd + Tu MT Td / (1 - Ru MT) (the composition
    of ocean_bottom.py:116-156, 290-323); and the reference: the Kennett R_PP
    of [water | CPA medium of ``cpa_phases`` (thickness H) | half-space] and of
    the homogeneous background.  With ``subcell`` the sub-cell resolved slab
    (``subcell_slab``) is solved too and embedded the same way.
    ``preconditioners`` maps a frequency index to a LearnedPreconditioner for
    that composite solve; ``subcell_preconditioner`` serves every sub-cell
    solve.  A solve that stops at ``max_iter`` is reported through
    ``converged`` / ``subcell_converged``, without a warning."""
    ref = material.ref
    g = geometry
    H = g.N_z * g.d
    omegas = np.asarray(omegas, dtype=float)
    if p >= 1.0 / water_alpha:
        raise ValueError(f"Horizontal slowness p={p:.6e} s/m exceeds critical angle for water "
                         f"(1/alpha_water = {1.0 / water_alpha:.6e} s/m).")
    eta_w = vertical_slowness(complex_slowness(water_alpha, np.inf), p)
    eta_s = vertical_slowness(complex_slowness(ref.alpha, np.inf), p)
    nu_s = vertical_slowness(complex_slowness(ref.beta, np.inf), p)
    c = psv_fluid_solid(p, eta_w, water_rho, eta_s, nu_s, ref.rho, ref.beta)
    k_hat = np.array([float(np.real(eta_s * ref.alpha)), float(p * ref.alpha), 0.0])
    water = FluidLayer(water_alpha, water_rho, 1.0)
    hs = IsotropicLayer(hs_alpha, hs_beta, hs_rho, np.inf)
    sed = IsotropicLayer(ref.alpha, ref.beta, ref.rho, H)
    R_sed_hs = kennett_layers(LayerStack([sed, hs]), p, omegas[:1]).RPP[0]
    R_bg = kennett_layers(LayerStack([water, sed, hs]), p, omegas).RPP
    nf = len(omegas)
    out = ResonanceOceanBottomResult(omega=omegas, ka=omegas * g.a / ref.beta,
                                     R_slab=np.zeros(nf, complex), R_total=np.zeros(nf, complex),
                                     R_bg=R_bg, R_cpa=np.zeros(nf, complex))

    def embed(R_slab, w):
        MT = np.exp(2j * w * eta_s * H) * R_sed_hs + R_slab
        return c.Rd[0, 0] + c.Tu[0, 0] * MT * c.Td[0, 0] / (1.0 - c.Ru[0, 0] * MT)

    if subcell:
        out.R_slab_subcell = np.zeros(nf, complex)
        out.R_total_subcell = np.zeros(nf, complex)
    plan = SlabPlan(g, periodic=True, device=device)
    for i, w in enumerate(omegas):
        pc = (preconditioners or {}).get(i)
        t0 = time.perf_counter()
        res = compute_resonance_slab_scattering(g, material, w, k_hat, "P", n_sub=n_sub,
                                                gmres_tol=gmres_tol, max_iter=max_iter,
                                                restart=restart, precision=precision,
                                                preconditioner=pc, plan=plan, warn=False)
        torch.cuda.synchronize(plan.device)
        out.solve_seconds.append(time.perf_counter() - t0)
        out.tmatrix_seconds.append(res.stats["tmatrix_seconds"])
        out.distinct_tmatrices.append(res.stats["distinct_tmatrices"])
        out.iterations.append(res.n_gmres_iter)
        out.converged.append(bool(res.stats["converged"]))
        out.R_slab[i] = slab_rpp_periodic(res, p=p)
        out.R_total[i] = embed(out.R_slab[i], w)
        if subcell:
            t0 = time.perf_counter()
            with warnings.catch_warnings():
                warnings.simplefilter("ignore")
                rf = compute_subcell_slab_scattering(g, material, w, k_hat, "P", n_sub=n_sub,
                                                     gmres_tol=gmres_tol, max_iter=max_iter,
                                                     restart=restart, precision=precision,
                                                     preconditioner=subcell_preconditioner,
                                                     device=plan.device)
            torch.cuda.synchronize(plan.device)
            out.subcell_seconds.append(time.perf_counter() - t0)
            out.subcell_iterations.append(rf.n_gmres_iter)
            out.subcell_converged.append(bool(rf.stats["converged"]))
            out.R_slab_subcell[i] = slab_rpp_periodic(rf, p=p)
            out.R_total_subcell[i] = embed(out.R_slab_subcell[i], w)
        if cpa_phases is not None:
            eff = compute_cpa(cpa_phases, w, g.a).effective_medium
            out.cpa_media.append(eff)
            lay = cubic_to_isotropic_layer(eff, H)
            out.R_cpa[i] = kennett_layers(LayerStack([water, lay, hs]), p, np.array([w])).RPP[0]
    return out


__all__ = ["resonance_slab_tmatrices", "compute_resonance_slab_scattering", "subcell_slab",
           "compute_subcell_slab_scattering", "uniform_contrast_phases", "resonance_ocean_bottom",
           "ResonanceOceanBottomResult"]
