"""Batched sweep of independent slab solves (frequencies x incidences x slabs).

The reference solves one slab per call and loops over frequencies on the host
(compute_ocean_bottom_reflection, cubic_scattering/ocean_bottom.py:290-307:
compute_slab_scattering + compute_slab_tmatrices + slab_rpp_periodic per
frequency).  ``compute_slab_sweep`` hands a whole list of cases to the
library's native sweep driver.  The default (``mode="batched"``,
``msc_sweep_batched``) groups the cases by padded FFT grid and solves each
group as one batch: every operator pass and Arnoldi kernel of a GMRES step is
a single launch for all slabs of the group, whatever their sizes.  The
alternative (``mode="workers"``, ``msc_sweep``) runs the cases one by one on
worker threads with their own HIP streams, concurrently on the device.  Either
way G0 tables are built on the device per case, plans are re-shaped rather
than re-created, and the layer / far-field source sums come back with each
solution, so ``slab_rpp_periodic`` and ``slab_reflected_field`` of the results
need no further device work.
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
                       keep_on_device: bool = False, mode: str = "batched",
                       max_batch: int = 64) -> list[SlabResult]:
    """Solve every case (GMRES(restart) from the Born guess to ``gmres_tol`` FP64
    relative residual) and return one ``SlabResult`` per case, in order.

    ``mode="batched"``: the cases of one padded FFT grid are solved as batches
    of at most ``max_batch`` (one launch per operator pass and Arnoldi kernel
    for the whole batch).  ``mode="workers"``: ``workers`` host threads run
    cases one at a time, concurrently on the device.  With ``keep_on_device``
    the fields stay torch tensors on the device.
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
    if mode not in ("batched", "workers"):
        raise ValueError(f"mode must be 'batched' or 'workers', got '{mode}'")
    if workers < 1:
        raise ValueError("workers must be >= 1")
    if max_batch < 1:
        raise ValueError("max_batch must be >= 1")
    lib = _lib.load()
    n = len(cases)
    inst = (_lib.Instance * max(n, 1))()
    # every case's contrasts in one pinned host buffer and one upload; every
    # case's layer / far-field sums in one device buffer and one download
    nv = [c.geometry.N_z * c.geometry.M * c.geometry.M for c in cases]
    voff = np.concatenate([[0], np.cumsum(nv)]).astype(np.int64)
    nzs = [c.geometry.N_z for c in cases]
    soff = np.concatenate([[0], np.cumsum(nzs)]).astype(np.int64)
    host = torch.empty(3 * int(voff[-1]), dtype=torch.float64, pin_memory=True)
    hv = host.numpy()
    for i, c in enumerate(cases):
        g, mat = c.geometry, c.material
        for k, a in enumerate((mat.Dlambda, mat.Dmu, mat.Drho)):
            seg = hv[3 * voff[i] + k * nv[i]:3 * voff[i] + (k + 1) * nv[i]]
            seg[:] = np.broadcast_to(np.asarray(a, dtype=np.float64),
                                     (g.N_z, g.M, g.M)).reshape(-1)
    contrasts = host.to(device, non_blocking=True)
    sums = torch.empty(27 * int(soff[-1]), dtype=torch.complex128, device=device)
    keep = [host, contrasts]  # buffers referenced by the instances
    outs = []
    for i, c in enumerate(cases):
        g, mat = c.geometry, c.material
        k_hat, k_mag, inc = _incident_vector(c.omega, mat.ref, c.k_hat, c.wave_type)
        shape = (g.N_z, g.M, g.M)
        arrs = [contrasts[3 * voff[i] + k * nv[i]:3 * voff[i] + (k + 1) * nv[i]]
                for k in range(3)]
        psi = torch.empty((*shape, 9), dtype=torch.complex128, device=device)
        psi0 = torch.empty_like(psi)
        # contiguous (N_z, 9) layer sums, then (N_z, 2, 9) phased sums
        o = 27 * int(soff[i])
        ls = sums[o:o + 9 * g.N_z].view(g.N_z, 9)
        ph = sums[o + 9 * g.N_z:o + 27 * g.N_z].view(g.N_z, 2, 9)
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
    torch.cuda.synchronize(device)  # the uploads are on torch's stream; the sweep uses its own
    with torch.cuda.device(device):
        if mode == "batched":
            rc = lib.msc_sweep_batched(ctypes.cast(inst, ctypes.c_void_p), n, ctypes.byref(opts),
                                       int(max_batch),
                                       ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream))
        else:
            rc = lib.msc_sweep(ctypes.cast(inst, ctypes.c_void_p), n, ctypes.byref(opts),
                               int(workers))
    if rc in (_lib.MSC_ERR_INVALID, _lib.MSC_ERR_UNSUPPORTED):
        bad = [i for i in range(n) if inst[i].status == rc]
        msg = inst[bad[0]].error.decode(errors="replace") if bad else "invalid sweep arguments"
        raise ValueError(f"sweep: {msg}")
    sums_h = _lib.to_host(sums)
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
        o, nz = 27 * int(soff[i]), c.geometry.N_z
        res._sums = (sums_h[o:o + 9 * nz].reshape(nz, 9).copy(),
                     sums_h[o + 9 * nz:o + 27 * nz].reshape(nz, 2, 9).copy(),
                     (c.omega / ref.alpha, c.omega / ref.beta), -k_hat)
        results.append(res)
    return results
