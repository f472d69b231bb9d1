"""YAML configuration of the slab solver (host mirror of
cubic_scattering/solver_config.py:35-475).

Accepts the reference's ``device / solver / problem`` schema
(configs/example_slab.yml) unchanged and adds, as a strict superset:

* ``device.backend: hip`` (same as ``cuda`` on ROCm PyTorch);
* ``solver.restart`` (GMRES restart length, default: unrestarted as in the
  reference GPU path), ``solver.precision: fp64 | mixed``,
  ``solver.inner_tol`` (mixed: reduction per complex64 cycle),
  ``solver.orth: adaptive | cgs | cgs2``;
* ``problem.slab.material_type: random_iid`` with ``phi``, ``rel_contrast``,
  ``seed`` (Bernoulli inclusions with i.i.d. uniform relative contrasts);
* ``problem.slab.periodic`` and ``problem.slab.pad: [Pz, Pxy]``.

Problems are solved on the HIP device only; ``backend: cpu`` is accepted by
the validator (the reference schema) but ``run_from_config`` raises for it.
``type: sphere`` runs the FFT Foldy-Lax sphere solve
(``compute_sphere_foldy_lax_fft_gpu``) as the reference's GPU branch does.
"""

from __future__ import annotations

from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
import torch
import yaml

from .effective_contrasts import MaterialContrast, ReferenceMedium
from .slab_scattering import (
    SlabGeometry,
    SlabResult,
    random_iid_slab_material,
    random_slab_material,
    uniform_slab_material,
)
from .torch_gmres import get_device, select_dtype

__all__ = ["DeviceConfig", "ProblemConfig", "ScatteringConfig", "SolverConfig", "load_config",
           "run_from_config", "validate_config"]


@dataclass
class DeviceConfig:
    """Device and dtype (solver_config.py:35-99); backend adds ``hip``."""

    backend: str = "auto"
    dtype: str = "auto"
    memory_limit_gb: float = 8.0

    def resolve_device(self) -> torch.device:
        if self.backend == "auto":
            return get_device()
        if self.backend == "mps":
            if not torch.backends.mps.is_available():
                raise ValueError("MPS backend requested but not available. "
                                 "Requires Apple Silicon with macOS 12.3+.")
            return torch.device("mps")
        if self.backend in ("cuda", "hip"):
            if not torch.cuda.is_available():
                raise ValueError(f"{self.backend.upper()} backend requested but no GPU found.")
            return torch.device("cuda")
        if self.backend == "cpu":
            return torch.device("cpu")
        raise ValueError(f"Invalid backend '{self.backend}'. Valid options: auto, mps, cuda, hip, cpu.")

    def resolve_dtype(self, device: torch.device) -> torch.dtype:
        if self.dtype == "auto":
            return select_dtype(device)
        if self.dtype == "complex64":
            return torch.complex64
        if self.dtype == "complex128":
            return torch.complex128
        raise ValueError(f"Invalid dtype '{self.dtype}'. Valid options: auto, complex64, complex128.")


@dataclass
class SolverConfig:
    """GMRES parameters (solver_config.py:102-114) plus the build's extensions."""

    gmres_tol: float = 1e-6
    max_iter: int = 500
    initial_guess: str = "born"
    restart: int | None = None
    precision: str | None = None
    inner_tol: float = 1e-4
    orth: str = "adaptive"


@dataclass
class ProblemConfig:
    """Problem definition (solver_config.py:117-163)."""

    type: str
    geometry: dict = field(default_factory=dict)
    reference: dict = field(default_factory=dict)
    contrast: dict = field(default_factory=dict)
    frequency: dict = field(default_factory=dict)
    incident: dict = field(default_factory=dict)
    sphere: dict | None = None
    slab: dict | None = None

    def to_reference_medium(self) -> ReferenceMedium:
        return ReferenceMedium(alpha=float(self.reference["alpha"]), beta=float(self.reference["beta"]),
                               rho=float(self.reference["rho"]))

    def to_material_contrast(self) -> MaterialContrast:
        return MaterialContrast(Dlambda=float(self.contrast["Dlambda"]), Dmu=float(self.contrast["Dmu"]),
                                Drho=float(self.contrast["Drho"]))

    def to_slab_geometry(self) -> SlabGeometry:
        return SlabGeometry(M=int(self.geometry["M"]), N_z=int(self.geometry["N_z"]),
                            a=float(self.geometry["a"]))


@dataclass
class ScatteringConfig:
    device: DeviceConfig
    solver: SolverConfig
    problem: ProblemConfig


def load_config(path: Path) -> ScatteringConfig:
    """Load and validate a YAML config (solver_config.py:181-253)."""
    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(
            f"Config file not found: {path}\n"
            f"Expected a YAML file with device, solver, and problem sections.\n"
            f"See configs/example_slab.yml for a template.")
    with path.open() as f:
        raw = yaml.safe_load(f)
    if not isinstance(raw, dict):
        raise ValueError(f"Config file must contain a YAML mapping, got {type(raw).__name__}")
    dev = raw.get("device", {}) or {}
    device_config = DeviceConfig(backend=dev.get("backend", "auto"), dtype=dev.get("dtype", "auto"),
                                 memory_limit_gb=float(dev.get("memory_limit_gb", 8.0)))
    sol = raw.get("solver", {}) or {}
    restart = sol.get("restart")
    solver_config = SolverConfig(
        gmres_tol=float(sol.get("gmres_tol", 1e-6)), max_iter=int(sol.get("max_iter", 500)),
        initial_guess=sol.get("initial_guess", "born"),
        restart=None if restart is None else int(restart), precision=sol.get("precision"),
        inner_tol=float(sol.get("inner_tol", 1e-4)), orth=sol.get("orth", "adaptive"))
    prob = raw.get("problem")
    if prob is None:
        raise ValueError("Config missing required 'problem' section.\n"
                         "Must contain: type, geometry, reference, contrast, frequency, incident.")
    problem_config = ProblemConfig(
        type=prob.get("type", "slab"), geometry=prob.get("geometry", {}),
        reference=prob.get("reference", {}), contrast=prob.get("contrast", {}),
        frequency=prob.get("frequency", {}), incident=prob.get("incident", {}),
        sphere=prob.get("sphere"), slab=prob.get("slab"))
    config = ScatteringConfig(device=device_config, solver=solver_config, problem=problem_config)
    validate_config(config)
    return config


def validate_config(config: ScatteringConfig) -> None:
    """Actionable ValueErrors, same messages as solver_config.py:256-372."""
    if config.problem.type not in ("slab", "sphere"):
        raise ValueError(f"Invalid problem type '{config.problem.type}'. Must be 'slab' or 'sphere'.")
    ref = config.problem.reference
    for key in ("alpha", "beta", "rho"):
        if key not in ref:
            raise ValueError(f"Missing 'problem.reference.{key}'. Required fields: alpha (P-velocity m/s), "
                             f"beta (S-velocity m/s), rho (density kg/m³).")
        if float(ref[key]) <= 0:
            raise ValueError(f"problem.reference.{key} must be positive, got {ref[key]}")
    slab = config.problem.slab or {}
    needs_contrast = not (config.problem.type == "slab" and slab.get("material_type") == "random_iid")
    if needs_contrast:
        for key in ("Dlambda", "Dmu", "Drho"):
            if key not in config.problem.contrast:
                raise ValueError(f"Missing 'problem.contrast.{key}'. "
                                 f"Required fields: Dlambda (Pa), Dmu (Pa), Drho (kg/m³).")
    if "omega" not in config.problem.frequency:
        raise ValueError("Missing 'problem.frequency.omega'. Must be a positive float (rad/s).")
    if float(config.problem.frequency["omega"]) <= 0:
        raise ValueError(f"problem.frequency.omega must be positive, got {config.problem.frequency['omega']}")
    geom = config.problem.geometry
    for key in ("M", "N_z", "a"):
        if key not in geom:
            raise ValueError(f"Missing 'problem.geometry.{key}'. Required: M (int, horizontal grid), "
                             f"N_z (int, vertical layers), a (float, half-width m).")
    inc = config.problem.incident
    if "wave_type" not in inc:
        raise ValueError("Missing 'problem.incident.wave_type'. Must be 'P' or 'S'.")
    if inc["wave_type"] not in ("P", "S"):
        raise ValueError(f"Invalid wave_type '{inc['wave_type']}'. Must be 'P' or 'S'.")
    if "k_hat" not in inc:
        raise ValueError("Missing 'problem.incident.k_hat'. Must be a 3-element list [z, x, y] direction.")
    if config.solver.initial_guess not in ("born", "zero"):
        raise ValueError(f"Invalid initial_guess '{config.solver.initial_guess}'. Must be 'born' or 'zero'.")
    if config.solver.precision not in (None, "fp64", "mixed"):
        raise ValueError(f"Invalid precision '{config.solver.precision}'. Must be 'fp64' or 'mixed'.")
    if config.solver.restart is not None and config.solver.restart < 1:
        raise ValueError(f"solver.restart must be >= 1, got {config.solver.restart}")
    if config.solver.orth not in ("adaptive", "cgs", "cgs2"):
        raise ValueError(f"Invalid orth '{config.solver.orth}'. Must be 'adaptive', 'cgs' or 'cgs2'.")
    valid_backends = ("auto", "mps", "cuda", "hip", "cpu")
    if config.device.backend not in valid_backends:
        raise ValueError(f"Invalid backend '{config.device.backend}'. Valid options: {', '.join(valid_backends)}.")
    valid_dtypes = ("auto", "complex64", "complex128")
    if config.device.dtype not in valid_dtypes:
        raise ValueError(f"Invalid dtype '{config.device.dtype}'. Valid options: {', '.join(valid_dtypes)}.")
    if config.problem.type == "slab" and slab.get("material_type", "uniform") not in (
            "uniform", "random", "random_iid"):
        raise ValueError(f"Invalid slab.material_type '{slab.get('material_type')}'. "
                         f"Must be 'uniform', 'random' or 'random_iid'.")
    if config.problem.type == "sphere":
        sp = config.problem.sphere
        if sp is None:
            raise ValueError("Problem type 'sphere' requires a 'sphere' section with 'radius' and 'n_sub'.")
        if "radius" not in sp or "n_sub" not in sp:
            raise ValueError(f"Sphere config requires 'radius' (m) and 'n_sub' (int). Got keys: {list(sp.keys())}")


def build_material(config: ScatteringConfig):
    """The SlabMaterial the config describes (seeded, reproducible)."""
    prob = config.problem
    geometry = prob.to_slab_geometry()
    ref = prob.to_reference_medium()
    slab_cfg = prob.slab or {}
    kind = slab_cfg.get("material_type", "uniform")
    seed = slab_cfg.get("seed")
    seed = None if seed is None else int(seed)
    if kind == "random":
        return geometry, random_slab_material(geometry, ref, prob.to_material_contrast(),
                                              float(slab_cfg.get("phi", 0.1)), seed)
    if kind == "random_iid":
        return geometry, random_iid_slab_material(geometry, ref, float(slab_cfg.get("phi", 0.3)),
                                                  float(slab_cfg.get("rel_contrast", 0.1)), seed)
    return geometry, uniform_slab_material(geometry, ref, prob.to_material_contrast())


def run_from_config(config: ScatteringConfig):
    """Dispatch a config to the HIP solver (solver_config.py:375-475): a
    ``SlabResult`` for ``type: slab``, a ``SphereDecompositionResult`` for
    ``type: sphere``."""
    device = config.device.resolve_device()
    if device.type != "cuda":
        raise ValueError(f"backend '{config.device.backend}' resolved to '{device}': this build runs "
                         f"the solve on HIP devices only (no CPU solver)")
    dtype = config.device.resolve_dtype(device)
    prob = config.problem
    if prob.type == "sphere":
        from .sphere_scattering_fft import compute_sphere_foldy_lax_fft_gpu

        sp = prob.sphere or {}
        return compute_sphere_foldy_lax_fft_gpu(
            omega=float(prob.frequency["omega"]), radius=float(sp["radius"]),
            ref=prob.to_reference_medium(), contrast=prob.to_material_contrast(),
            n_sub=int(sp["n_sub"]), k_hat=np.array(prob.incident["k_hat"], dtype=float),
            wave_type=prob.incident["wave_type"], gmres_tol=config.solver.gmres_tol,
            gmres_maxiter=config.solver.max_iter, initial_guess=config.solver.initial_guess,
            device=device, dtype=dtype, restart=config.solver.restart)
    geometry, material = build_material(config)
    slab_cfg = prob.slab or {}
    pad = slab_cfg.get("pad")
    from .slab_scattering_gpu import compute_slab_scattering_gpu

    return compute_slab_scattering_gpu(
        geometry=geometry, material=material, omega=float(prob.frequency["omega"]),
        k_hat=np.array(prob.incident["k_hat"], dtype=float), wave_type=prob.incident["wave_type"],
        gmres_tol=config.solver.gmres_tol, max_iter=config.solver.max_iter,
        initial_guess=config.solver.initial_guess, device=device, dtype=dtype,
        periodic=bool(slab_cfg.get("periodic", False)), restart=config.solver.restart,
        precision=config.solver.precision, inner_tol=config.solver.inner_tol,
        orth=config.solver.orth, pad=None if pad is None else tuple(int(v) for v in pad))
