"""ctypes binding of libmsc_slab.so (C ABI declared in include/msc_slab.h).

The library is loaded after ``import torch`` so that it binds to the HIP runtime
PyTorch already loaded.  There is no CPU fallback: if the library is missing or
no HIP device is present, every solver entry point raises.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_double, c_int, c_longlong, c_void_p
from pathlib import Path

import numpy as np
import torch  # noqa: F401  (loads the HIP runtime first)

LIB_PATH = Path(__file__).resolve().parent / "libmsc_slab.so"

MSC_OK = 0
MSC_ERR_INVALID = 1
MSC_ERR_UNSUPPORTED = 2
MSC_ERR_HIP = 3
MSC_ERR_NOT_READY = 4

MSC_C128 = 0
MSC_C64 = 1

# Every symbol the header declares, with (restype, argtypes).
SIGNATURES = {
    "msc_last_error": (ctypes.c_char_p, []),
    "msc_version": (ctypes.c_char_p, []),
    "msc_plan_create": (c_int, [POINTER(c_void_p), c_int, c_int, c_double, c_int, c_int, c_int]),
    "msc_plan_destroy": (c_int, [c_void_p]),
    "msc_plan_reshape": (c_int, [c_void_p, c_int, c_int, c_double]),
    "msc_plan_use_graphs": (c_int, [c_void_p, c_int]),
    "msc_plan_copy_g0": (c_int, [c_void_p, c_void_p, c_void_p]),
    "msc_sweep": (c_int, [c_void_p, c_int, c_void_p, c_int]),
    "msc_sweep_batched": (c_int, [c_void_p, c_int, c_void_p, c_int, c_void_p]),
    "msc_sweep_release": (c_int, []),
    "msc_resonance_system": (c_int, [c_double, c_double, c_double, c_double, c_double, c_int,
                                     c_double, c_double, c_double, c_void_p, c_void_p]),
    "msc_resonance_tmatrices": (c_int, [c_double, c_double, c_double, c_double, c_double, c_int,
                                        c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                        c_int, c_double, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p]),
    "msc_resonance_tmatrices_noalloc": (c_int, [c_double, c_double, c_double, c_double, c_double,
                                                c_int, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                                c_void_p, c_int, c_double, c_void_p, c_void_p,
                                                c_void_p, c_void_p, c_void_p]),
    "msc_plan_info": (c_int, [c_void_p, POINTER(c_longlong)]),
    "msc_plan_autotune": (c_int, [c_void_p, c_int, c_void_p]),
    "msc_plan_tiles": (c_int, [c_void_p, POINTER(c_int)]),
    "msc_plan_set_tiles": (c_int, [c_void_p, POINTER(c_int)]),
    "msc_cube_coefficients": (c_int, [c_double, c_double, c_double, c_double, c_double,
                                      POINTER(c_double)]),
    "msc_build_g0": (c_int, [c_void_p, c_double, c_double, c_double, c_double, c_int, c_int,
                             c_void_p]),
    "msc_set_tmatrix": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_double,
                                POINTER(c_double), c_void_p]),
    "msc_set_tparams": (c_int, [c_void_p, c_void_p, c_void_p]),
    "msc_get_tparams": (c_int, [c_void_p, c_void_p, c_void_p]),
    "msc_set_tgeneral": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "msc_incident_field": (c_int, [c_void_p, c_double, POINTER(c_double), POINTER(c_double),
                                   c_void_p, c_void_p]),
    "msc_matvec": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "msc_residual": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
    "msc_gmres": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    "msc_gmres_batched": (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                  c_void_p]),
    "msc_source_sums": (c_int, [c_void_p, c_void_p, c_double, c_double, POINTER(c_double),
                                c_void_p, c_void_p, c_void_p]),
    "msc_fft_axis": (c_int, [c_void_p, c_void_p, c_int, c_int, c_longlong, c_int, c_int,
                             c_void_p]),
    "msc_arnoldi_create": (c_int, [POINTER(c_void_p), c_longlong, c_int, c_int]),
    "msc_arnoldi_destroy": (c_int, [c_void_p]),
    "msc_arnoldi_start": (c_int, [c_void_p, c_void_p, c_longlong, c_void_p, c_double, c_void_p]),
    "msc_arnoldi_step": (c_int, [c_void_p, c_void_p, c_longlong, c_int, c_int, c_void_p]),
    "msc_arnoldi_state": (c_int, [c_void_p, POINTER(c_double), c_void_p]),
    "msc_arnoldi_solution": (c_int, [c_void_p, c_void_p, c_longlong, c_void_p, c_void_p]),
    "msc_timing_enable": (c_int, [c_void_p, c_int]),
    "msc_timing_read": (c_int, [c_void_p, POINTER(c_double)]),
    "msc_timing_durations": (c_int, [c_void_p, c_int, c_void_p, c_int, POINTER(c_int)]),
    # learned preconditioner (include/msc_precond.h)
    "msc_net_create": (c_int, [POINTER(c_void_p), c_int]),
    "msc_net_create_unet": (c_int, [POINTER(c_void_p), POINTER(c_int)]),
    "msc_net_destroy": (c_int, [c_void_p]),
    "msc_net_num_params": (c_int, [c_void_p, POINTER(c_longlong)]),
    "msc_net_set_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "msc_net_get_params": (c_int, [c_void_p, c_void_p, c_void_p]),
    "msc_net_input": (c_int, [c_void_p, c_void_p, c_int, c_double, c_void_p, c_void_p]),
    "msc_net_forward": (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p]),
    "msc_net_gradients": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                  c_void_p, POINTER(c_double), c_void_p]),
    "msc_net_train_step": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                   c_double, POINTER(c_double), c_void_p]),
    "msc_gmres_basis": (c_int, [c_void_p, c_int, c_int, c_void_p, c_void_p]),
    "msc_net_apply": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p]),
}


class GmresOpts(ctypes.Structure):
    _fields_ = [("restart", c_int), ("max_iter", c_int), ("tol", c_double), ("x0", c_int),
                ("mode", c_int), ("inner_tol", c_double), ("orth", c_int),
                ("precond", c_void_p)]


class GmresStats(ctypes.Structure):
    _fields_ = [("iterations", c_int), ("cycles", c_int), ("matvecs_c128", c_int),
                ("matvecs_c64", c_int), ("converged", c_int), ("fell_back", c_int),
                ("precond_applies", c_int), ("rel_residual", c_double), ("bnorm", c_double)]


class Instance(ctypes.Structure):
    """msc_instance: one solve of msc_sweep."""
    _fields_ = [("M", c_int), ("N_z", c_int), ("periodic", c_int), ("volume_averaged", c_int),
                ("n_orders", c_int), ("a", c_double),
                ("omega", c_double), ("alpha", c_double), ("beta", c_double), ("rho", c_double),
                ("k_mag", c_double), ("khat", c_double * 3), ("inc18", c_double * 18),
                ("dlambda", c_void_p), ("dmu", c_void_p), ("drho", c_void_p), ("psi", c_void_p),
                ("psi0", c_void_p), ("layer_sum", c_void_p), ("phased", c_void_p), ("stats", GmresStats),
                ("status", c_int), ("error", ctypes.c_char * 160)]


_lib = None


def load() -> ctypes.CDLL:
    """Load (once) and return the library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH.name} is not built; run `python -m multiplescatteringcalculations_amd.build` "
            "(this package has no CPU fallback)")
    lib = ctypes.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    """Map a status code to the reference's exception types."""
    if rc == MSC_OK:
        return
    msg = load().msc_last_error().decode(errors="replace")
    if rc in (MSC_ERR_INVALID, MSC_ERR_UNSUPPORTED):
        raise ValueError(msg)
    raise RuntimeError(f"libmsc_slab error {rc}: {msg}")


def to_host(t: torch.Tensor) -> np.ndarray:
    """A device tensor as a host NumPy array, copied through pinned memory on
    the current stream: a copy into pageable memory runs at a fraction of the
    link's rate (75.5 MB: about 20 ms against 1.4 ms pinned on MI355X)."""
    t = t.detach()
    if t.device.type != "cuda" or t.numel() * t.element_size() < (1 << 20):
        return t.cpu().numpy()
    out = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    with torch.cuda.device(t.device):
        out.copy_(t, non_blocking=True)
        torch.cuda.current_stream(t.device).synchronize()
    return out.numpy()


def to_device(a, device: torch.device) -> torch.Tensor:
    """A host array on the device, copied through pinned memory on the current
    stream (the counterpart of to_host; pageable copies of a few MB cost
    milliseconds each)."""
    a = np.ascontiguousarray(a)
    t = torch.from_numpy(a)
    if t.numel() * t.element_size() < (1 << 20):
        return t.to(device)
    pinned = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    pinned.copy_(t)
    with torch.cuda.device(device):
        out = pinned.to(device, non_blocking=True)
        torch.cuda.current_stream(device).synchronize()
    return out


def require_device() -> torch.device:
    """The HIP device every solve runs on; raises when there is none."""
    if not torch.cuda.is_available():
        raise RuntimeError(
            "no HIP device available: the slab solver runs on AMD GPUs only (no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr(device: torch.device | None = None) -> c_void_p:
    return c_void_p(torch.cuda.current_stream(device).cuda_stream)


def ptr(t: torch.Tensor) -> c_void_p:
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")
    return c_void_p(t.data_ptr())


def dbl_array(vals) -> ctypes.Array:
    arr = (c_double * len(vals))()
    for i, v in enumerate(vals):
        arr[i] = float(v)
    return arr
