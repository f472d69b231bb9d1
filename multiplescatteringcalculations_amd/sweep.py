"""Batched sweep of independent slab solves (frequencies x incidences x slabs).

The reference solves one slab per call and loops over frequencies on the host
(compute_ocean_bottom_reflection, cubic_scattering/ocean_bottom.py:290-307:
compute_slab_scattering + compute_slab_tmatrices + slab_rpp_periodic per
frequency).  ``compute_slab_sweep`` hands a whole list of cases to the
library's native sweep driver (``msc_sweep``): worker threads with their own
HIP streams and one re-shapable plan per FFT grid run the solves concurrently
on the device, G0 tables are rebuilt only when a case needs a new one, and
the layer / far-field source sums come back with each solution, so
``slab_rpp_periodic`` and ``slab_reflected_field`` of the results need no
further device work.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch
from numpy.typing import NDArray

from . import _lib
from .slab_scattering import SlabGeometry, SlabMaterial, SlabResult, _incident_vector


@dataclass
class SweepCase:
    """One solve of a sweep: slab, material, frequency and incidence."""

    geometry: SlabGeometry
    material: SlabMaterial
    omega: float
    k_hat: NDArray
    wave_type: str = "P"
    periodic: bool = False
    volume_averaged: bool = False
    n_orders: int = 2


def compute_slab_sweep(cases: list[SweepCase], *, gmres_tol: float = 1e-8, max_iter: int = 500,
                       restart: int = 30, precision: str = "mixed", inner_tol: float = 1e-4,
                       orth: str = "adaptive", workers: int = 4,
                       device: torch.device | None = None,
                       keep_on_device: bool = False) -> list[SlabResult]:
    """Solve every case (GMRES(restart) from the Born guess to ``gmres_tol`` FP64
    relative residual) and return one ``SlabResult`` per case, in order.

    ``workers`` host threads run cases concurrently on the device.  With
    ``keep_on_device`` the fields stay torch tensors on the device.
    Non-convergence warns per case (as the reference's solver does); invalid
    cases raise ``ValueError`` before anything runs.
    """
    import warnings

    if device is None:
        device = _lib.require_device()
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError("the sweep runs on a HIP device only")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got '{precision}'")
    orth_code = {"adaptive": 0, "cgs": 1, "cgs2": 2}.get(orth)
    if orth_code is None:
        raise ValueError(f"orth must be 'adaptive', 'cgs' or 'cgs2', got '{orth}'")
    if workers < 1:
        raise ValueError("workers must be >= 1")
    lib = _lib.load()
    n = len(cases)
    inst = (_lib.Instance * max(n, 1))()
    keep = []  # device tensors referenced by the instances
    outs = []
    for i, c in enumerate(cases):
        g, mat = c.geometry, c.material
        k_hat, k_mag, inc = _incident_vector(c.omega, mat.ref, c.k_hat, c.wave_type)
        shape = (g.N_z, g.M, g.M)
        arrs = [torch.tensor(np.broadcast_to(np.asarray(a, dtype=np.float64), shape),
                             device=device) for a in (mat.Dlambda, mat.Dmu, mat.Drho)]
        psi = torch.empty((*shape, 9), dtype=torch.complex128, device=device)
        psi0 = torch.empty_like(psi)
        ls = torch.empty((g.N_z, 9), dtype=torch.complex128, device=device)
        ph = torch.empty((g.N_z, 2, 9), dtype=torch.complex128, device=device)
        keep.extend(arrs)
        outs.append((psi, psi0, ls, ph, k_hat))
        I = inst[i]
        I.M, I.N_z, I.periodic, I.a = g.M, g.N_z, int(bool(c.periodic)), float(g.a)
        I.volume_averaged, I.n_orders = int(bool(c.volume_averaged)), int(c.n_orders)
        I.omega = float(c.omega)
        I.alpha, I.beta, I.rho = mat.ref.alpha, mat.ref.beta, mat.ref.rho
        I.k_mag = float(k_mag)
        for k in range(3):
            I.khat[k] = float(k_hat[k])
        for k in range(9):
            I.inc18[2 * k] = float(inc[k].real)
            I.inc18[2 * k + 1] = float(inc[k].imag)
        I.dlambda, I.dmu, I.drho = (a.data_ptr() for a in arrs)
        I.psi, I.psi0 = psi.data_ptr(), psi0.data_ptr()
        I.layer_sum, I.phased = ls.data_ptr(), ph.data_ptr()
    m = max(1, int(restart))
    opts = _lib.GmresOpts(m, int(max_iter), float(gmres_tol), 1,
                          1 if precision == "mixed" else 0, float(inner_tol), orth_code, None)
    torch.cuda.synchronize(device)  # the uploads are on torch's stream; the workers use their own
    with torch.cuda.device(device):
        rc = lib.msc_sweep(ctypes.cast(inst, ctypes.c_void_p), n, ctypes.byref(opts), int(workers))
    if rc in (_lib.MSC_ERR_INVALID, _lib.MSC_ERR_UNSUPPORTED):
        bad = [i for i in range(n) if inst[i].status == rc]
        msg = inst[bad[0]].error.decode(errors="replace") if bad else "invalid sweep arguments"
        raise ValueError(f"sweep: {msg}")
    results = []
    for i, c in enumerate(cases):
        I = inst[i]
        if I.status != _lib.MSC_OK:
            msg = I.error.decode(errors="replace")
            if I.status in (_lib.MSC_ERR_INVALID, _lib.MSC_ERR_UNSUPPORTED):
                raise ValueError(f"sweep case {i}: {msg}")
            raise RuntimeError(f"sweep case {i}: libmsc_slab error {I.status}: {msg}")
        st = {k: getattr(I.stats, k) for k, _ in _lib.GmresStats._fields_}
        psi, psi0, ls, ph, k_hat = outs[i]
        if not st["converged"]:
            warnings.warn(f"sweep case {i} did not converge: rel_res={st['rel_residual']:.2e} "
                          f"after {st['iterations']} iterations", stacklevel=2)
        res = SlabResult(psi=psi if keep_on_device else _lib.to_host(psi),
                         psi0=psi0 if keep_on_device else _lib.to_host(psi0),
                         geometry=c.geometry, material=c.material, omega=float(c.omega),
                         k_hat=np.asarray(c.k_hat, dtype=float), wave_type=c.wave_type,
                         n_gmres_iter=int(st["iterations"]),
                         gmres_residual=float(st["rel_residual"]), periodic=bool(c.periodic),
                         stats=st)
        ref = c.material.ref
        res._sums = (ls.cpu().numpy(), ph.cpu().numpy(),
                     (c.omega / ref.alpha, c.omega / ref.beta), -k_hat)
        results.append(res)
    return results
