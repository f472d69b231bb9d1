"""3-D slab Foldy-Lax multiple scattering on MI355X: (I - G0 T0) psi = psi0.

Host mirror of cubic_scattering/slab_scattering.py (data types 40-161, T0
40-205, incident field 426-478, solver 486-565, outputs 573-735, material
generators 791-852) and of slab_scattering_gpu.py.  Every numerical step runs
in libmsc_slab.so on the HIP device (plan, G0 table, T0, incident field,
FFT matvec, GMRES, source sums); this module only moves arrays and keeps the
reference's signatures, errors and warnings.

Coordinate system: z = 0 (down), x = 1, y = 2.  Voigt order
(zz, xx, yy, xy, zy, zx) with engineering (doubled) shear strain.
"""

from __future__ import annotations

import atexit
import ctypes
import threading
import warnings
from dataclasses import dataclass, field
from typing import Callable

import numpy as np
import torch
from numpy.typing import NDArray

from . import _lib
from .effective_contrasts import (
    MaterialContrast,
    ReferenceMedium,
    compute_cube_tmatrix,
    cube_coefficients,
    sub_cell_tmatrix_9x9,
)

# ═══════════════════════════════════════════════════════════════
#  Data structures (slab_scattering.py:40-161)
# ═══════════════════════════════════════════════════════════════


@dataclass
class SlabGeometry:
    """Geometry of the M x M x N_z slab lattice; ``a`` is the cube half-width."""

    M: int
    N_z: int
    a: float

    def __post_init__(self) -> None:
        if self.M < 1:
            raise ValueError(f"M must be >= 1, got {self.M}")
        if self.N_z < 1:
            raise ValueError(f"N_z must be >= 1, got {self.N_z}")
        if self.a <= 0:
            raise ValueError(f"a must be > 0, got {self.a}")

    @property
    def d(self) -> float:
        """Cube side length (= 2a)."""
        return 2.0 * self.a

    @property
    def n_cubes(self) -> int:
        return self.M * self.M * self.N_z

    def cube_centre(self, lz: int, i: int, j: int) -> NDArray[np.floating]:
        """Centre [z, x, y] of cube (lz, i, j)."""
        d = self.d
        return np.array([(lz + 0.5) * d, (i - (self.M - 1) / 2.0) * d,
                         (j - (self.M - 1) / 2.0) * d])

    def all_centres(self) -> NDArray[np.floating]:
        """Centres of all cubes, shape (N_z, M, M, 3)."""
        d = self.d
        z = (np.arange(self.N_z) + 0.5) * d
        x = (np.arange(self.M) - (self.M - 1) / 2.0) * d
        zz, xx, yy = np.meshgrid(z, x, x, indexing="ij")
        return np.stack([zz, xx, yy], axis=-1)


@dataclass
class SlabMaterial:
    """Per-cube contrasts Dlambda, Dmu (Pa), Drho (kg/m^3), shape (N_z, M, M)."""

    Dlambda: NDArray
    Dmu: NDArray
    Drho: NDArray
    ref: ReferenceMedium

    def __post_init__(self) -> None:
        if self.Dlambda.shape != self.Dmu.shape or self.Dlambda.shape != self.Drho.shape:
            raise ValueError(
                f"Contrast array shapes must match: Dlambda={self.Dlambda.shape}, "
                f"Dmu={self.Dmu.shape}, Drho={self.Drho.shape}")
        if self.Dlambda.ndim != 3:
            raise ValueError(f"Contrast arrays must be 3D (N_z, M, M), got ndim={self.Dlambda.ndim}")


@dataclass
class SlabResult:
    """Solution of the slab problem (slab_scattering.py:136-161).

    ``n_gmres_iter`` counts Arnoldi steps (as the reference GPU path does);
    ``gmres_residual`` is the true FP64 ||b - A psi|| / ||b||.  ``stats`` holds
    the solver's counters (cycles, matvecs per precision, fallback flag).
    """

    psi: NDArray
    psi0: NDArray
    geometry: SlabGeometry
    material: SlabMaterial
    omega: float
    k_hat: NDArray
    wave_type: str
    n_gmres_iter: int
    gmres_residual: float
    periodic: bool = False
    stats: dict = field(default_factory=dict)


# ═══════════════════════════════════════════════════════════════
#  Device plan
# ═══════════════════════════════════════════════════════════════


def _vec_c128(x, n: int, device) -> torch.Tensor:
    t = torch.as_tensor(np.ascontiguousarray(x).reshape(-1) if isinstance(x, np.ndarray) else x)
    t = t.to(device=device, dtype=torch.complex128).reshape(-1).contiguous()
    if t.numel() != n:
        raise ValueError(f"vector has {t.numel()} entries, expected {n}")
    return t


class SlabPlan:
    """One slab geometry on one HIP device: padded FFT grid, spectral G0 table,
    per-voxel T0 parameters and solver workspaces (``msc_plan``).

    ``pad`` = (Pz, Pxy) padded FFT lengths (powers of two, default the
    smallest >= 2N-1); ``periodic`` selects the circular lateral convolution of
    the reference's ``periodic=True``: on the M-point lateral grid when M is a
    power of two, otherwise on the zero-padded grid with the periodically
    folded kernel (the same operator for any M).
    """

    def __init__(self, geometry: SlabGeometry, *, periodic: bool = False,
                 pad: tuple[int, int] | None = None, device: torch.device | None = None):
        self.lib = _lib.load()
        self.device = device if device is not None else _lib.require_device()
        self.geometry = geometry
        self.periodic = bool(periodic)
        pz, pxy = pad if pad is not None else (0, 0)
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_plan_create(ctypes.byref(h), geometry.M, geometry.N_z,
                                                float(geometry.a), int(self.periodic), int(pz),
                                                int(pxy)))
        self._h = h
        info = (ctypes.c_longlong * 6)()
        _lib.check(self.lib.msc_plan_info(self._h, info))
        self.Pz, self.Pxy, self.NUV, self.NC = int(info[2]), int(info[3]), int(info[4]), int(info[5])
        self.n = geometry.n_cubes * 9
        self.omega = None
        self.ref = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                self.lib.msc_plan_destroy(h)
            except Exception:  # interpreter shutdown
                pass
            self._h = None

    @property
    def stream(self) -> ctypes.c_void_p:
        return _lib.stream_ptr(self.device)

    def reshape(self, geometry: SlabGeometry) -> None:
        """Re-target the plan to a slab no larger than the one it was created
        for (same FFT grid) without reallocating; G0 and T0 must be set again."""
        torch.cuda.current_stream(self.device).synchronize()
        _lib.check(self.lib.msc_plan_reshape(self._h, geometry.M, geometry.N_z,
                                             float(geometry.a)))
        self.geometry = geometry
        self.n = geometry.n_cubes * 9
        self.omega = None
        self.ref = None

    def build_g0(self, omega: float, ref: ReferenceMedium, *, volume_averaged: bool = False,
                 n_orders: int = 2) -> None:
        """Spectral G0 table on the device (replaces _build_slab_kernels[_gpu])."""
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_build_g0(self._h, float(omega), float(ref.alpha),
                                             float(ref.beta), float(ref.rho),
                                             int(bool(volume_averaged)), int(n_orders),
                                             self.stream))
        self.omega = float(omega)
        self.ref = ref

    def set_material(self, material: SlabMaterial, omega: float) -> None:
        """Per-voxel T0 on the device from the contrast arrays."""
        g = self.geometry
        shape = (g.N_z, g.M, g.M)
        if material.Dlambda.shape != shape:
            raise ValueError(f"material shape {material.Dlambda.shape} != geometry {shape}")
        coef = cube_coefficients(omega, g.a, material.ref)
        c8 = _lib.dbl_array([v for c in coef for v in (c.real, c.imag)])
        arrs = [torch.as_tensor(np.ascontiguousarray(x, dtype=np.float64)).to(self.device)
                for x in (material.Dlambda, material.Dmu, material.Drho)]
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_set_tmatrix(self._h, _lib.ptr(arrs[0]), _lib.ptr(arrs[1]),
                                                _lib.ptr(arrs[2]), float(omega), c8, self.stream))
            torch.cuda.current_stream(self.device).synchronize()

    def set_tparams(self, tparams) -> None:
        """T0 as (t_u, t_d, t_o, t_s) per voxel, shape (N_z, M, M, 4)."""
        t = _vec_c128(tparams, self.geometry.n_cubes * 4, self.device)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_set_tparams(self._h, _lib.ptr(t), self.stream))
            torch.cuda.current_stream(self.device).synchronize()

    def set_tgeneral(self, tidx, table) -> None:
        """A general 9x9 T-matrix per voxel: voxel v applies table[tidx[v]]
        (tidx: (N_z, M, M) integers, table: (K, 9, 9) complex, tau = T psi)."""
        g = self.geometry
        idx = torch.as_tensor(np.asarray(tidx) if not isinstance(tidx, torch.Tensor) else tidx)
        idx = idx.to(device=self.device, dtype=torch.int32).reshape(-1).contiguous()
        if idx.numel() != g.n_cubes:
            raise ValueError(f"tidx must have {g.n_cubes} entries, got {idx.numel()}")
        tab = torch.as_tensor(table).to(device=self.device, dtype=torch.complex128).contiguous()
        if tab.dim() != 3 or tab.shape[1:] != (9, 9):
            raise ValueError(f"table must be (K, 9, 9), got {tuple(tab.shape)}")
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_set_tgeneral(self._h, _lib.ptr(idx), _lib.ptr(tab),
                                                 int(tab.shape[0]), self.stream))
        self._material_token = None

    def tparams(self) -> torch.Tensor:
        g = self.geometry
        out = torch.empty((g.N_z, g.M, g.M, 4), dtype=torch.complex128, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_get_tparams(self._h, _lib.ptr(out), self.stream))
        return out

    def incident_field(self, omega: float, ref: ReferenceMedium, k_hat, wave_type: str) -> torch.Tensor:
        """psi0 on the device, (N_z, M, M, 9) complex128."""
        k_hat, k_mag, inc = _incident_vector(omega, ref, k_hat, wave_type)
        g = self.geometry
        out = torch.empty((g.N_z, g.M, g.M, 9), dtype=torch.complex128, device=self.device)
        inc18 = _lib.dbl_array([v for c in inc for v in (c.real, c.imag)])
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_incident_field(self._h, float(k_mag), _lib.dbl_array(k_hat),
                                                   inc18, _lib.ptr(out), self.stream))
        return out

    def _prec(self, t: torch.Tensor) -> int:
        if t.dtype == torch.complex128:
            return _lib.MSC_C128
        if t.dtype == torch.complex64:
            return _lib.MSC_C64
        raise ValueError(f"dtype must be complex128 or complex64, got {t.dtype}")

    def matvec(self, x: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """(I - G0 T0) x for a flat complex128/complex64 device vector."""
        x = x.reshape(-1)
        if x.numel() != self.n or x.device != self.device:
            raise ValueError(f"x must hold {self.n} entries on {self.device}")
        x = x.contiguous()
        y = torch.empty_like(x) if out is None else out
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_matvec(self._h, _lib.ptr(x), _lib.ptr(y), self._prec(x),
                                           self.stream))
        return y

    def residual(self, b: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        """b - (I - G0 T0) x."""
        b = b.reshape(-1).contiguous()
        x = x.reshape(-1).contiguous()
        if b.dtype != x.dtype:
            raise ValueError("b and x must share a dtype")
        r = torch.empty_like(x)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_residual(self._h, _lib.ptr(b), _lib.ptr(x), _lib.ptr(r),
                                             self._prec(x), self.stream))
        return r

    def gmres(self, b: torch.Tensor, x0: str | torch.Tensor = "born", *, tol: float = 1e-6,
              max_iter: int = 500, restart: int | None = None, precision: str = "fp64",
              inner_tol: float = 1e-4, orth: str = "adaptive",
              precond=None) -> tuple[torch.Tensor, dict]:
        """GMRES(restart) on the device; returns (x, stats).

        ``precision`` "fp64" (complex128 throughout) or "mixed" (complex64
        cycles inside FP64 iterative refinement; the FP64 true residual judges
        every cycle and a stalled complex64 cycle falls back to FP64).
        ``restart=None`` keeps the reference's unrestarted GMRES (restart =
        max_iter).  ``precond`` (a LearnedPreconditioner) switches to flexible
        GMRES with that right preconditioner; a stalled flexible cycle hands
        over to plain GMRES cycles.
        """
        b = _vec_c128(b, self.n, self.device)
        if isinstance(x0, torch.Tensor):
            x = _vec_c128(x0, self.n, self.device).clone()
            mode_x0 = 2
        elif x0 == "born":
            x = torch.empty_like(b)
            mode_x0 = 1
        elif x0 == "zero":
            x = torch.empty_like(b)
            mode_x0 = 0
        else:
            raise ValueError(f"Invalid initial_guess '{x0}'. Must be 'born' or 'zero'.")
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got '{precision}'")
        orth_code = {"adaptive": 0, "cgs": 1, "cgs2": 2}.get(orth)
        if orth_code is None:
            raise ValueError(f"orth must be 'adaptive', 'cgs' or 'cgs2', got '{orth}'")
        m = int(max_iter) if restart is None else int(restart)
        if precond is not None and torch.device(precond.device) != torch.device(self.device):
            raise ValueError("preconditioner and plan are on different devices")
        opts = _lib.GmresOpts(max(1, m), int(max_iter), float(tol), mode_x0,
                              1 if precision == "mixed" else 0, float(inner_tol), orth_code,
                              precond.handle if precond is not None else None)
        st = _lib.GmresStats()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_gmres(self._h, _lib.ptr(b), _lib.ptr(x), ctypes.byref(opts),
                                          ctypes.byref(st), self.stream))
        stats = {k: getattr(st, k) for k, _ in _lib.GmresStats._fields_}
        return x, stats

    def source_sums(self, psi: torch.Tensor, kP: float, kS: float, r_hat) -> tuple[np.ndarray, np.ndarray]:
        """Per-layer sums of tau = T0 psi (N_z, 9) and phased sums (N_z, 2, 9)."""
        psi = _vec_c128(psi, self.n, self.device)
        g = self.geometry
        ls = torch.empty((g.N_z, 9), dtype=torch.complex128, device=self.device)
        ph = torch.empty((g.N_z, 2, 9), dtype=torch.complex128, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_source_sums(self._h, _lib.ptr(psi), float(kP), float(kS),
                                                _lib.dbl_array(np.asarray(r_hat, float)),
                                                _lib.ptr(ls), _lib.ptr(ph), self.stream))
        return ls.cpu().numpy(), ph.cpu().numpy()

    def copy_g0_from(self, other: "SlabPlan") -> None:
        """Take G0 (and its near-neighbour correction) and the tuned tiles from
        ``other``, a plan of the same geometry, grid and device, by a device
        copy instead of a table build."""
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_plan_copy_g0(self._h, other._h, self.stream))
        self.omega = other.omega
        self.ref = other.ref

    def autotune(self, precision: str = "c64") -> None:
        """Time every tile candidate of the FFT passes in whole operator
        applications and keep the fastest."""
        prec = {"c128": _lib.MSC_C128, "c64": _lib.MSC_C64}[precision]
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_plan_autotune(self._h, prec, self.stream))

    def tiles(self) -> dict:
        """Tiles per precision: (y-fwd rows, x-fwd cols, z cols, x-inv cols, y-inv rows)."""
        out = (ctypes.c_int * 10)()
        _lib.check(self.lib.msc_plan_tiles(self._h, out))
        return {"c128": tuple(out[0:5]), "c64": tuple(out[5:10])}

    def set_tiles(self, c128: tuple, c64: tuple) -> None:
        """Fix the tiles (5 per precision, order as in ``tiles()``)."""
        arr = (ctypes.c_int * 10)(*[int(v) for v in tuple(c128) + tuple(c64)])
        _lib.check(self.lib.msc_plan_set_tiles(self._h, arr))

    def timing_durations(self, precision: str = "c64") -> list[float]:
        """Every recorded pass-3 launch duration (ms) of one precision, in
        launch order (no-op launches of a finished cycle included)."""
        eb = {"c64": 8, "c128": 16}[precision]
        n = ctypes.c_int(0)
        _lib.check(self.lib.msc_timing_durations(self._h, eb, None, 0, ctypes.byref(n)))
        out = (ctypes.c_double * max(n.value, 1))()
        _lib.check(self.lib.msc_timing_durations(self._h, eb, ctypes.cast(out, ctypes.c_void_p),
                                                 n.value, ctypes.byref(n)))
        return [float(out[i]) for i in range(n.value)]

    def timing(self, enable: bool | None = None) -> dict:
        """Enable/disable (and reset) the HIP-event timing of the block-multiply
        pass, or read it: {"c64": (launches, ms), "c128": (launches, ms)}."""
        if enable is not None:
            _lib.check(self.lib.msc_timing_enable(self._h, int(bool(enable))))
            return {}
        out = (ctypes.c_double * 4)()
        _lib.check(self.lib.msc_timing_read(self._h, out))
        return {"c64": (int(out[0]), float(out[1])), "c128": (int(out[2]), float(out[3]))}


# ═══════════════════════════════════════════════════════════════
#  T0 and incident field (slab_scattering.py:169-205, 426-478)
# ═══════════════════════════════════════════════════════════════


def gmres_batched(plans: list["SlabPlan"], bs, x0: str = "born", *, tol: float = 1e-6,
                  max_iter: int = 500, restart: int = 30, precision: str = "fp64",
                  inner_tol: float = 1e-4, orth: str = "adaptive"):
    """``SlabPlan.gmres`` on several plans of one padded grid as one batch
    (msc_gmres_batched): each operator pass and Arnoldi kernel is a single
    launch for all of them.  ``bs`` holds one right-hand side per plan;
    returns ([x_i], [stats_i])."""
    if len(plans) != len(bs) or not plans:
        raise ValueError("one right-hand side per plan, at least one plan")
    dev = plans[0].device
    if any(torch.device(p.device) != torch.device(dev) for p in plans):
        raise ValueError("the plans of a batch must be on one device")
    if precision not in ("fp64", "mixed"):
        raise ValueError(f"precision must be 'fp64' or 'mixed', got '{precision}'")
    orth_code = {"adaptive": 0, "cgs": 1, "cgs2": 2}.get(orth)
    if orth_code is None:
        raise ValueError(f"orth must be 'adaptive', 'cgs' or 'cgs2', got '{orth}'")
    b_t = [_vec_c128(b, p.n, dev) for p, b in zip(plans, bs)]
    if isinstance(x0, str):
        mode_x0 = {"zero": 0, "born": 1}.get(x0)
        if mode_x0 is None:
            raise ValueError(f"Invalid initial_guess '{x0}'. Must be 'born' or 'zero'.")
        xs = [torch.empty_like(b) for b in b_t]
    else:
        mode_x0 = 2
        xs = [_vec_c128(v, p.n, dev).clone() for p, v in zip(plans, x0)]
    k = len(plans)
    opts = _lib.GmresOpts(max(1, int(restart)), int(max_iter), float(tol), mode_x0,
                          1 if precision == "mixed" else 0, float(inner_tol), orth_code, None)
    sts = (_lib.GmresStats * k)()
    hp = (ctypes.c_void_p * k)(*[p._h.value for p in plans])
    hb = (ctypes.c_void_p * k)(*[b.data_ptr() for b in b_t])
    hx = (ctypes.c_void_p * k)(*[x.data_ptr() for x in xs])
    lib = plans[0].lib
    with torch.cuda.device(dev):
        _lib.check(lib.msc_gmres_batched(hp, hb, hx, k, ctypes.byref(opts), sts,
                                         plans[0].stream))
    stats = [{f: getattr(s, f) for f, _ in _lib.GmresStats._fields_} for s in sts]
    return xs, stats


def tparams_to_tmatrices(tp: NDArray) -> NDArray:
    """Expand (..., 4) T0 parameters to (..., 9, 9) block-diagonal matrices."""
    tp = np.asarray(tp)
    T = np.zeros(tp.shape[:-1] + (9, 9), dtype=complex)
    for c in range(3):
        T[..., c, c] = tp[..., 0]
        T[..., 6 + c, 6 + c] = tp[..., 3]
        for c2 in range(3):
            T[..., 3 + c, 3 + c2] = tp[..., 1] if c == c2 else tp[..., 2]
    return T


def tmatrices_to_tparams(T: NDArray) -> NDArray:
    """Compress block-diagonal cubic T0 (..., 9, 9) to (..., 4); raises if T0
    does not have the structure of compute_slab_tmatrices."""
    T = np.asarray(T)
    tp = np.stack([T[..., 0, 0], T[..., 3, 3], T[..., 3, 4], T[..., 6, 6]], axis=-1)
    if not np.allclose(tparams_to_tmatrices(tp), T, rtol=1e-13, atol=1e-300):
        raise ValueError("T_local must be the block-diagonal cubic T0 of compute_slab_tmatrices")
    return tp


def compute_slab_tmatrices(geometry: SlabGeometry, material: SlabMaterial, omega: float,
                           *, device: torch.device | None = None) -> NDArray:
    """9x9 T0 per cube, (N_z, M, M, 9, 9) complex (slab_scattering.py:169-205).

    Evaluated on the device (4 complex numbers per voxel) and expanded on the
    host for the reference's return type.
    """
    plan = SlabPlan(geometry, device=device)
    plan.set_material(material, omega)
    return tparams_to_tmatrices(_lib.to_host(plan.tparams()))


def _plane_wave_strain_voigt(k_hat, pol, k_mag) -> NDArray:
    """Voigt strain of a plane wave, engineering shear (sphere_scattering.py:1256-1287)."""
    eps = 0.5j * k_mag * (np.outer(k_hat, pol) + np.outer(pol, k_hat))
    pairs = [(0, 0), (1, 1), (2, 2), (1, 2), (0, 2), (0, 1)]
    return np.array([eps[p, q] if p == q else 2.0 * eps[p, q] for p, q in pairs], dtype=complex)


def _incident_vector(omega, ref: ReferenceMedium, k_hat, wave_type):
    k_hat = np.asarray(k_hat, dtype=float)
    k_hat = k_hat / np.linalg.norm(k_hat)
    if wave_type == "P":
        k_mag = omega / ref.alpha
        pol = k_hat.copy()
    elif wave_type == "S":
        k_mag = omega / ref.beta
        cross = np.cross(k_hat, np.array([1.0, 0.0, 0.0]))
        if np.linalg.norm(cross) < 1e-10:
            pol = np.array([0.0, 1.0, 0.0])
        else:
            pol = np.cross(cross, k_hat)
            pol = pol / np.linalg.norm(pol)
    else:
        raise ValueError(f"wave_type must be 'P' or 'S', got '{wave_type}'")
    inc = np.zeros(9, dtype=complex)
    inc[:3] = pol
    inc[3:] = _plane_wave_strain_voigt(k_hat, pol, k_mag)
    return k_hat, k_mag, inc


def _build_slab_incident_field(geometry: SlabGeometry, omega: float, ref: ReferenceMedium,
                               k_hat, wave_type: str, *, device=None) -> NDArray:
    """Incident field (N_z, M, M, 9) (slab_scattering.py:426-478), built on the device."""
    _incident_vector(omega, ref, k_hat, wave_type)  # validates wave_type first
    plan = SlabPlan(geometry, device=device)
    return _lib.to_host(plan.incident_field(omega, ref, k_hat, wave_type))


# ═══════════════════════════════════════════════════════════════
#  Solver (slab_scattering.py:486-565)
# ═══════════════════════════════════════════════════════════════


# One plan is kept between solves of the same slab shape: a new plan pays its
# buffer allocations and the tile autotuning again.  A solve takes the plan out of the cache while it runs.
_PLAN_CACHE: dict = {}
_PLAN_LOCK = threading.Lock()


def _take_plan(geometry, periodic, pad, device):
    key = (str(device), geometry.M, geometry.N_z, float(geometry.a), bool(periodic),
           None if pad is None else tuple(int(v) for v in pad))
    with _PLAN_LOCK:
        plan = _PLAN_CACHE.pop(key, None)
    if plan is None:
        plan = SlabPlan(geometry, periodic=periodic, pad=pad, device=device)
    else:
        plan._material_token = None  # its T0 is about to change
    return plan, key


def _give_plan(key, plan) -> None:
    with _PLAN_LOCK:
        _PLAN_CACHE.clear()  # a plan holds the slab's workspaces: keep one
        _PLAN_CACHE[key] = plan


def release_plans() -> None:
    """Drop the plan kept between solves (its device memory is freed once no
    result refers to it), the nine plans of the batched sphere solve and the
    plan pools the batched sweep keeps."""
    with _PLAN_LOCK:
        _PLAN_CACHE.clear()
    from .sphere_scattering_fft import release_batch_plans

    release_batch_plans()
    if _lib._lib is not None:
        _lib._lib.msc_sweep_release()


# free the cached plan while the HIP runtime is still up
atexit.register(release_plans)


def _solve(geometry, material, omega, k_hat, wave_type, gmres_tol, max_iter, initial_guess,
           device, *, volume_averaged, n_orders, periodic, restart, precision, inner_tol,
           orth, pad, warn_label, precond=None) -> SlabResult:
    if initial_guess not in ("born", "zero"):
        raise ValueError(f"Invalid initial_guess '{initial_guess}'. Must be 'born' or 'zero'.")
    _incident_vector(omega, material.ref, k_hat, wave_type)
    plan, key = _take_plan(geometry, periodic, pad, device)
    r = material.ref
    g0_key = (float(omega), float(r.alpha), float(r.beta), float(r.rho), bool(volume_averaged),
              int(n_orders))
    if getattr(plan, "_g0_key", None) != g0_key:
        plan._g0_key = None
        plan.build_g0(omega, r, volume_averaged=volume_averaged, n_orders=n_orders)
        plan._g0_key = g0_key
    plan.set_material(material, omega)
    psi0 = plan.incident_field(omega, material.ref, k_hat, wave_type)
    x, st = plan.gmres(psi0, initial_guess, tol=gmres_tol, max_iter=max_iter, restart=restart,
                       precision=precision, inner_tol=inner_tol, orth=orth, precond=precond)
    if not st["converged"]:
        warnings.warn(f"{warn_label} did not converge: rel_res={st['rel_residual']:.2e} "
                      f"after {st['iterations']} iterations", stacklevel=3)
    g = geometry
    res = SlabResult(
        psi=_lib.to_host(x.reshape(g.N_z, g.M, g.M, 9)),
        psi0=_lib.to_host(psi0),
        geometry=geometry,
        material=material,
        omega=omega,
        k_hat=np.asarray(k_hat, dtype=float),
        wave_type=wave_type,
        n_gmres_iter=int(st["iterations"]),
        gmres_residual=float(st["rel_residual"]),
        periodic=periodic,
        stats=st,
    )
    # the output functions use the plan's T0 while no later solve has replaced it
    token = object()
    plan._material_token = token
    res._plan, res._plan_token = plan, token
    _give_plan(key, plan)
    return res


def compute_slab_scattering(geometry: SlabGeometry, material: SlabMaterial, omega: float,
                            k_hat: NDArray[np.floating], wave_type: str = "P",
                            gmres_tol: float = 1e-6, max_iter: int = 500, *,
                            volume_averaged: bool = False, n_orders: int = 2,
                            periodic: bool = False, restart: int | None = 20,
                            precision: str = "fp64", inner_tol: float = 1e-4,
                            orth: str = "adaptive", pad: tuple[int, int] | None = None,
                            device: torch.device | None = None,
                            preconditioner=None) -> SlabResult:
    """Solve (I - G0 T0) psi = psi0 (slab_scattering.py:486-565) on the HIP device.

    Same signature and result as the reference CPU solver: GMRES(20) (SciPy's
    default restart) from the Born guess x0 = psi0, with the true residual
    reported.  ``max_iter`` caps Arnoldi steps.  ``preconditioner`` (a
    LearnedPreconditioner) selects flexible GMRES.
    """
    return _solve(geometry, material, omega, k_hat, wave_type, gmres_tol, max_iter, "born",
                  device, volume_averaged=volume_averaged, n_orders=n_orders, periodic=periodic,
                  restart=restart, precision=precision, inner_tol=inner_tol, orth=orth, pad=pad,
                  warn_label="GMRES", precond=preconditioner)


# ═══════════════════════════════════════════════════════════════
#  Outputs (slab_scattering.py:573-735)
# ═══════════════════════════════════════════════════════════════


def _voigt_to_tensor(voigt_6: NDArray) -> NDArray:
    """Voigt stress [zz, xx, yy, 2xy, 2zy, 2zx] -> symmetric 3x3 (slab_scattering.py:573-586)."""
    T = np.zeros((3, 3), dtype=complex)
    T[0, 0], T[1, 1], T[2, 2] = voigt_6[0], voigt_6[1], voigt_6[2]
    T[1, 2] = T[2, 1] = voigt_6[3] / 2.0
    T[0, 2] = T[2, 0] = voigt_6[4] / 2.0
    T[0, 1] = T[1, 0] = voigt_6[5] / 2.0
    return T


def _source_sums(result: SlabResult, T_local, kP: float, kS: float, r_hat):
    sums = getattr(result, "_sums", None)
    if sums is not None and T_local is None:
        # reduced on the device by the sweep: layer sums always apply, the
        # phased sums only for the far-field direction and wavenumbers they used
        layer_sum, phased, k2, rh = sums
        if kP == 0.0 and kS == 0.0:
            return layer_sum, None
        if np.allclose([kP, kS], k2, rtol=1e-14, atol=0.0) and np.allclose(r_hat, rh, rtol=0, atol=1e-14):
            return layer_sum, phased
    plan = getattr(result, "_plan", None)
    if plan is not None and getattr(plan, "_material_token", None) is not getattr(
            result, "_plan_token", None):
        plan = None  # a later solve has reused the plan
    g = result.geometry
    if plan is None or T_local is not None:
        plan = SlabPlan(g)
        if T_local is None:
            plan.set_material(result.material, result.omega)
        else:
            plan.set_tparams(tmatrices_to_tparams(np.asarray(T_local)))
    return plan.source_sums(torch.as_tensor(np.ascontiguousarray(result.psi)), kP, kS, r_hat)


def _vertical_slowness(s: complex, p: float) -> complex:
    """eta = sqrt(s^2 - p^2) with Im(eta) > 0 (kennett_layers.py:273-287)."""
    eta = np.sqrt(complex((s + p) * (s - p)))
    if eta.imag < 0.0 or (eta.imag == 0.0 and eta.real < 0.0):
        eta = -eta
    return eta


def slab_rpp_periodic(result: SlabResult, T_local: NDArray | None = None, *, p: float = 0.0) -> complex:
    """Specular P->P reflection of a periodic slab (slab_scattering.py:672-735).

    R_PP = -i/(2 k_z d^2 rho alpha^2) sum_l Q_P,l exp(i k_z z_l) with
    Q_P,l = -r.f_l - i k_z r.sigma_l.r from the layer means of tau = T0 psi
    (reduced on the device).
    """
    g = result.geometry
    ref = result.material.ref
    d = g.d
    eta = _vertical_slowness(complex(1.0 / ref.alpha), p)
    kz = result.omega * eta
    r_hat = np.array([-eta * ref.alpha, p * ref.alpha, 0.0])
    layer_sum, _ = _source_sums(result, T_local, 0.0, 0.0, np.real(r_hat))
    mean = layer_sum / (g.M * g.M)
    total = 0j
    for lz in range(g.N_z):
        sig = _voigt_to_tensor(mean[lz, 3:])
        Q = -np.dot(r_hat, mean[lz, :3]) - 1j * kz * np.dot(r_hat, sig @ r_hat)
        total += Q * np.exp(1j * kz * (lz + 0.5) * d)
    return complex(-1j / (2.0 * kz * d**2 * ref.rho * ref.alpha**2) * total)


def slab_reflected_field(result: SlabResult, T_local: NDArray | None = None) -> tuple[complex, complex, complex]:
    """Far-field (R_PP, R_PS, R_SP) per unit area (slab_scattering.py:589-669).

    The voxel sums exp(-i k r.x_v) tau_v for k = kP, kS run on the device; the
    per-source scalar projections are linear in tau and applied on the host.
    """
    g = result.geometry
    ref = result.material.ref
    k_hat = result.k_hat / np.linalg.norm(result.k_hat)
    kP, kS = result.omega / ref.alpha, result.omega / ref.beta
    r_hat = -k_hat
    _, phased = _source_sums(result, T_local, kP, kS, r_hat)
    SP = phased[:, 0, :].sum(axis=0)
    SS = phased[:, 1, :].sum(axis=0)
    GP = 1.0 / (4.0 * np.pi * ref.rho * ref.alpha**2)
    GS = 1.0 / (4.0 * np.pi * ref.rho * ref.beta**2)
    sigP = _voigt_to_tensor(SP[3:])
    QP = np.dot(r_hat, SP[:3]) - 1j * kP * np.dot(r_hat, sigP @ r_hat)
    u_P = -GP * QP * r_hat
    sigS = _voigt_to_tensor(SS[3:])
    QS = SS[:3] - 1j * kS * (sigS @ r_hat)
    u_S = -GS * (QS - np.dot(r_hat, QS) * r_hat)
    area = (g.M * g.d) ** 2
    u_P = u_P / area
    u_S = u_S / area
    R_PP = complex(np.dot(r_hat, u_P))
    cross = np.cross(r_hat, np.array([1.0, 0.0, 0.0]))
    if np.linalg.norm(cross) < 1e-10:
        sv = np.array([0.0, 1.0, 0.0])
    else:
        sv = np.cross(cross, r_hat)
        sv = sv / np.linalg.norm(sv)
    return R_PP, complex(np.dot(sv, u_S)), R_PP


def kennett_reference_rpp(ref: ReferenceMedium, contrast: MaterialContrast, H: float,
                          omega: float) -> complex:
    """Normal-incidence R_PP of a uniform layer of thickness H (slab_scattering.py:738-783).

    The layer (background moduli plus ``contrast``) sits between two half-spaces
    of the background medium; the reflection is the Kennett recursion at p = 0
    over that three-layer stack (the top layer's thickness does not enter R_PP
    at its lower interface), evaluated by ``kennett_layers``.
    """
    from .kennett_layers import IsotropicLayer, LayerStack, kennett_layers

    lam = ref.rho * (ref.alpha**2 - 2.0 * ref.beta**2) + contrast.Dlambda
    mu = ref.rho * ref.beta**2 + contrast.Dmu
    rho = ref.rho + contrast.Drho
    layer = IsotropicLayer(alpha=float(np.sqrt((lam + 2.0 * mu) / rho)),
                           beta=float(np.sqrt(mu / rho)), rho=rho, thickness=H)
    stack = LayerStack(layers=[
        IsotropicLayer(alpha=ref.alpha, beta=ref.beta, rho=ref.rho, thickness=100.0),
        layer,
        IsotropicLayer(alpha=ref.alpha, beta=ref.beta, rho=ref.rho, thickness=np.inf),
    ])
    return complex(kennett_layers(stack, p=0.0, omega=np.array([omega])).RPP[0])


# ═══════════════════════════════════════════════════════════════
#  Material generators (slab_scattering.py:791-852)
# ═══════════════════════════════════════════════════════════════


def random_slab_material(geometry: SlabGeometry, ref: ReferenceMedium, contrast: MaterialContrast,
                         phi: float | Callable[[int], float], seed: int | None = None) -> SlabMaterial:
    """Binary Bernoulli(phi) inclusions with one contrast (slab_scattering.py:791-828).

    Same draw order as the reference: one ``rng.random((M, M))`` per layer.
    """
    rng = np.random.default_rng(seed)
    M, N_z = geometry.M, geometry.N_z
    shape = (N_z, M, M)
    Dl, Dm, Dr = np.zeros(shape), np.zeros(shape), np.zeros(shape)
    for lz in range(N_z):
        p = phi(lz) if callable(phi) else phi
        mask = rng.random((M, M)) < p
        Dl[lz, mask] = contrast.Dlambda
        Dm[lz, mask] = contrast.Dmu
        Dr[lz, mask] = contrast.Drho
    return SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=ref)


def uniform_slab_material(geometry: SlabGeometry, ref: ReferenceMedium,
                          contrast: MaterialContrast) -> SlabMaterial:
    """All cubes share one contrast (slab_scattering.py:831-852)."""
    shape = (geometry.N_z, geometry.M, geometry.M)
    return SlabMaterial(Dlambda=np.full(shape, contrast.Dlambda), Dmu=np.full(shape, contrast.Dmu),
                        Drho=np.full(shape, contrast.Drho), ref=ref)


def random_iid_slab_material(geometry: SlabGeometry, ref: ReferenceMedium, phi: float,
                             rel_contrast: float, seed: int | None = None) -> SlabMaterial:
    """Bernoulli(phi) inclusions with i.i.d. relative contrasts.

    Each inclusion draws Dlambda/lambda, Dmu/mu, Drho/rho independently from
    U[-rel_contrast, rel_contrast].  Draw order with ``default_rng(seed)``:
    mask = rng.random(shape) < phi, then three rng.uniform(shape) arrays
    (lambda, mu, rho), each over the whole (N_z, M, M) grid.
    """
    rng = np.random.default_rng(seed)
    shape = (geometry.N_z, geometry.M, geometry.M)
    mask = rng.random(shape) < phi
    rl = rng.uniform(-rel_contrast, rel_contrast, shape)
    rm = rng.uniform(-rel_contrast, rel_contrast, shape)
    rr = rng.uniform(-rel_contrast, rel_contrast, shape)
    return SlabMaterial(Dlambda=np.where(mask, rl * ref.lam, 0.0), Dmu=np.where(mask, rm * ref.mu, 0.0),
                        Drho=np.where(mask, rr * ref.rho, 0.0), ref=ref)


__all__ = [
    "SlabGeometry", "SlabMaterial", "SlabResult", "SlabPlan", "compute_slab_tmatrices",
    "compute_slab_scattering", "slab_rpp_periodic", "slab_reflected_field", "kennett_reference_rpp",
    "random_slab_material", "uniform_slab_material", "random_iid_slab_material",
    "tparams_to_tmatrices", "tmatrices_to_tparams", "compute_cube_tmatrix",
    "sub_cell_tmatrix_9x9", "MaterialContrast", "ReferenceMedium",
]
