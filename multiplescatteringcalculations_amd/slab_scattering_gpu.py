"""GPU slab Foldy-Lax solve on MI355X (host mirror of
cubic_scattering/slab_scattering_gpu.py:158-266).

The reference builds T0, G0 and psi0 on the host and runs a torch-op 3-D FFT
matvec inside an unrestarted torch GMRES.  Here every step runs in
libmsc_slab.so: the G0 spectral table and T0 are built on the device, the
operator is the five-pass HIP FFT pipeline and GMRES is device-resident.
"""

from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .slab_scattering import SlabGeometry, SlabMaterial, SlabPlan, SlabResult, _solve

__all__ = ["compute_slab_scattering_gpu", "SlabOperator"]


def compute_slab_scattering_gpu(geometry: SlabGeometry, material: SlabMaterial, omega: float,
                                k_hat: np.ndarray, wave_type: str = "P", gmres_tol: float = 1e-6,
                                max_iter: int = 500, initial_guess: str = "born",
                                device: torch.device | None = None,
                                dtype: torch.dtype | None = None, *,
                                volume_averaged: bool = False, n_orders: int = 2,
                                periodic: bool = False, restart: int | None = None,
                                precision: str | None = None, inner_tol: float = 1e-4,
                                orth: str = "adaptive",
                                pad: tuple[int, int] | None = None,
                                preconditioner=None) -> SlabResult:
    """Solve (I - G0 T0) psi = psi0 with device GMRES (slab_scattering_gpu.py:158-266).

    Reference-compatible arguments: ``device`` must be a HIP device (``cuda``
    on ROCm PyTorch; a CPU device raises, there is no CPU path); ``dtype``
    complex128 (default) selects ``precision="fp64"``, complex64 selects the
    mixed form (complex64 operator and basis inside FP64 refinement, so the
    returned residual is still an FP64 one).  ``restart=None`` keeps the
    reference's unrestarted GMRES.  ``preconditioner`` (a LearnedPreconditioner)
    selects flexible GMRES.  Non-convergence warns, as the reference does.
    """
    if device is None:
        device = _lib.require_device()
    device = torch.device(device)
    if device.type != "cuda":
        raise ValueError(f"device must be a HIP device ('cuda' on ROCm), got '{device}'")
    if device.index is None:
        device = torch.device("cuda", torch.cuda.current_device())
    if precision is None:
        if dtype is None or dtype == torch.complex128:
            precision = "fp64"
        elif dtype == torch.complex64:
            precision = "mixed"
        else:
            raise ValueError(f"dtype must be complex64 or complex128, got {dtype}")
    return _solve(geometry, material, omega, k_hat, wave_type, gmres_tol, max_iter, initial_guess,
                  device, volume_averaged=volume_averaged, n_orders=n_orders, periodic=periodic,
                  restart=restart, precision=precision, inner_tol=inner_tol, orth=orth, pad=pad,
                  warn_label="GPU GMRES", precond=preconditioner)


class SlabOperator:
    """The operator x -> (I - G0 T0) x of one slab, for use with torch_gmres
    or direct calls (replaces _build_slab_kernels_gpu + _slab_matvec_gpu,
    slab_scattering_gpu.py:34-155)."""

    def __init__(self, geometry: SlabGeometry, material: SlabMaterial, omega: float, *,
                 periodic: bool = False, pad: tuple[int, int] | None = None,
                 device: torch.device | None = None):
        self.plan = SlabPlan(geometry, periodic=periodic, pad=pad, device=device)
        self.plan.build_g0(omega, material.ref)
        self.plan.set_material(material, omega)

    def __call__(self, x: torch.Tensor) -> torch.Tensor:
        return self.plan.matvec(x)
