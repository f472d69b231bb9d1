"""Device helpers and a generic device-resident GMRES (host mirror of
cubic_scattering/torch_gmres.py:27-205).

``torch_gmres(matvec, b, x0, tol, maxiter)`` keeps the reference contract for
any Python operator: the Arnoldi orthogonalisation (block CGS with FP64
reductions), the Givens QR, back-substitution and the solution update run in
libmsc_slab.so (``msc_arnoldi_*``); only ``matvec`` is the caller's.  Like the
reference it is unrestarted and synchronises once per step to test
convergence.
"""

from __future__ import annotations

import ctypes
from collections.abc import Callable

import numpy as np
import torch
from numpy.typing import NDArray

from . import _lib

__all__ = ["get_device", "select_dtype", "to_numpy", "to_torch", "torch_gmres"]


def get_device() -> torch.device:
    """The HIP device (``cuda`` on ROCm PyTorch); CPU only when no GPU exists."""
    if torch.cuda.is_available():
        return torch.device("cuda")
    return torch.device("cpu")


def select_dtype(device: torch.device, prefer_double: bool = False) -> torch.dtype:
    """complex128 (the reference default on CUDA/CPU, torch_gmres.py:39-56)."""
    return torch.complex128


def to_torch(arr: NDArray[np.complexfloating], device: torch.device, dtype: torch.dtype) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device=device, dtype=dtype)


def to_numpy(tensor: torch.Tensor) -> NDArray[np.complexfloating]:
    return _lib.to_host(tensor.detach().to(torch.complex128))


def _complex_norm(x: torch.Tensor) -> torch.Tensor:
    return torch.sqrt(torch.sum(x.real**2 + x.imag**2))


class _Arnoldi:
    def __init__(self, n: int, m: int, dtype: torch.dtype):
        self.lib = _lib.load()
        self.prec = _lib.MSC_C128 if dtype == torch.complex128 else _lib.MSC_C64
        h = ctypes.c_void_p()
        _lib.check(self.lib.msc_arnoldi_create(ctypes.byref(h), int(n), int(m), self.prec))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and self.h.value:
            try:
                self.lib.msc_arnoldi_destroy(self.h)
            except Exception:
                pass
            self.h = None


def torch_gmres(matvec: Callable[[torch.Tensor], torch.Tensor], b: torch.Tensor,
                x0: torch.Tensor | None = None, tol: float = 1e-6,
                maxiter: int = 500) -> tuple[torch.Tensor, int, float]:
    """Solve A x = b (A given by ``matvec``) with unrestarted GMRES on the device.

    Returns (x, n_iter, rel_residual) with the true residual ||b - A x||/||b||
    recomputed at the end (torch_gmres.py:83-205).  ``b`` must be a complex64 or
    complex128 tensor on a HIP device.
    """
    if b.device.type != "cuda":
        raise ValueError("torch_gmres runs on a HIP device; move b to 'cuda'")
    if b.dtype not in (torch.complex64, torch.complex128):
        raise ValueError(f"b must be complex64 or complex128, got {b.dtype}")
    n = b.shape[0]
    b = b.contiguous()
    x0 = torch.zeros_like(b) if x0 is None else x0.to(b.dtype).contiguous()
    stream = _lib.stream_ptr(b.device)
    r = (b - matvec(x0)).contiguous()
    b_norm = float(_complex_norm(b))
    if b_norm < 1e-30:
        return x0.clone(), 0, 0.0
    r_norm = float(_complex_norm(r))
    if r_norm / b_norm < tol:
        return x0.clone(), 0, r_norm / b_norm
    m = max(1, int(maxiter))
    ark = _Arnoldi(n, m, b.dtype)
    ldv = (n + 63) // 64 * 64
    V = torch.zeros((m + 1, ldv), dtype=b.dtype, device=b.device)
    lib = ark.lib
    with torch.cuda.device(b.device):
        _lib.check(lib.msc_arnoldi_start(ark.h, _lib.ptr(V), ldv, _lib.ptr(r), tol * b_norm, stream))
        st = (ctypes.c_double * 4)()
        n_iter = 0
        for j in range(m):
            V[j + 1, :n] = matvec(V[j, :n])
            _lib.check(lib.msc_arnoldi_step(ark.h, _lib.ptr(V), ldv, j, 0, stream))
            _lib.check(lib.msc_arnoldi_state(ark.h, st, stream))
            n_iter = int(st[0])
            if st[1]:
                break
        x = x0.clone()
        xs = torch.zeros((ldv,), dtype=b.dtype, device=b.device)
        _lib.check(lib.msc_arnoldi_solution(ark.h, _lib.ptr(V), ldv, _lib.ptr(xs), stream))
        x = x + xs[:n]
    r_final = b - matvec(x)
    return x, n_iter, float(_complex_norm(r_final) / b_norm)
