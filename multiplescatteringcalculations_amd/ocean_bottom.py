"""Ocean-bottom reflection through a heterogeneous sediment slab.

Drop-in for cubic_scattering/ocean_bottom.py: water (acoustic) over an
M x M x N_z sediment slab over an elastic half-space.  The homogeneous
background comes from the Kennett recursion of [water | sediment |
half-space]; the slab's specular R_PP (periodic slab, one solve per active
frequency) is added to the sub-ocean reflectivity and dressed by one Kennett
step at the water-sediment interface (ocean_bottom.py:116-156, 234-323),
then the water column (and optionally its free-surface multiples) and a
Ricker source give time traces (324-368).

Where the reference loops over frequencies on the host, solving one slab
after another (ocean_bottom.py:290-307), every active frequency here is one
instance of a single batched sweep (``compute_slab_sweep``: concurrent solves
on the device, one G0 table per frequency, R_PP's layer sums returned with
each solution).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np
from numpy.typing import NDArray

from .effective_contrasts import MaterialContrast, ReferenceMedium
from .kennett_layers import (
    FluidLayer,
    IsotropicLayer,
    LayerStack,
    complex_slowness,
    kennett_layers,
    psv_fluid_solid,
    vertical_slowness,
)
from .slab_scattering import (
    SlabGeometry,
    SlabMaterial,
    random_slab_material,
    slab_rpp_periodic,
    uniform_slab_material,
)


@dataclass
class OceanBottomConfig:
    """Water / slab / half-space model and recording (ocean_bottom.py:42-80), SI units."""

    water_alpha: float
    water_rho: float
    water_depth: float
    sed_ref: ReferenceMedium
    hs_alpha: float
    hs_beta: float
    hs_rho: float
    geometry: SlabGeometry
    material: SlabMaterial
    f_peak: float
    T: float
    nw: int
    gamma: float | None = None
    f_min: float = 1.0
    f_max: float = 200.0
    free_surface: bool = False
    p: float = 0.0


@dataclass
class OceanBottomResult:
    """Traces and reflectivities (ocean_bottom.py:83-113).

    ``freq_elapsed`` holds, for each active frequency, the sweep's wall time
    divided evenly over its instances (the solves run concurrently)."""

    time: NDArray[np.floating]
    trace_total: NDArray[np.floating]
    trace_homogeneous: NDArray[np.floating]
    R_bg: NDArray[np.complexfloating]
    R_slab: NDArray[np.complexfloating]
    R_total: NDArray[np.complexfloating]
    omega_real: NDArray[np.floating]
    config: OceanBottomConfig
    n_gmres_iters: list[int] = field(default_factory=list)
    freq_elapsed: list[float] = field(default_factory=list)
    elapsed_seconds: float = 0.0


def ricker_source_spectrum(omega: NDArray, f_peak: float) -> NDArray:
    """S(w) = (w/w_p)^2 exp(-(w/w_p)^2), w_p = 2 pi f_peak (seismic_survey.py:217-234)."""
    x = (np.asarray(omega, dtype=float) / (2.0 * np.pi * f_peak)) ** 2
    return x * np.exp(-x)


def _slownesses(cfg: OceanBottomConfig):
    p = cfg.p
    eta_w = vertical_slowness(complex_slowness(cfg.water_alpha, np.inf), p)
    s_sp = complex_slowness(cfg.sed_ref.alpha, np.inf)
    s_ss = complex_slowness(cfg.sed_ref.beta, np.inf)
    return eta_w, vertical_slowness(s_sp, p), vertical_slowness(s_ss, p), 1.0 / s_ss


def _kennett_water_step(MT_pp: NDArray, cfg: OceanBottomConfig) -> NDArray:
    """PP reflection at the water-sediment interface of a sub-ocean PP
    reflectivity MT: Rd + Tu MT Td / (1 - Ru MT) (ocean_bottom.py:116-156)."""
    eta_w, eta_s, nu_s, beta_s = _slownesses(cfg)
    c = psv_fluid_solid(cfg.p, eta_w, cfg.water_rho, eta_s, nu_s, cfg.sed_ref.rho, beta_s)
    MT_pp = np.asarray(MT_pp)
    return c.Rd[0, 0] + c.Tu[0, 0] * MT_pp * c.Td[0, 0] / (1.0 - c.Ru[0, 0] * MT_pp)


def compute_ocean_bottom_reflection(config: OceanBottomConfig, *, volume_averaged: bool = False,
                                    n_orders: int = 2, gmres_tol: float = 1e-6,
                                    progress: bool = True, precision: str = "fp64",
                                    restart: int = 30, workers: int = 4) -> OceanBottomResult:
    """Ocean-bottom reflection (ocean_bottom.py:159-368).

    Same arguments and result as the reference; ``progress`` is accepted and
    has nothing to report (the frequencies run as one sweep).  ``precision``
    ("fp64" or "mixed"), ``restart`` and ``workers`` tune the sweep; every
    slab solve reaches ``gmres_tol`` FP64 relative residual either way.
    """
    del progress
    from .sweep import SweepCase, compute_slab_sweep

    t0 = time.perf_counter()
    cfg = config
    p = cfg.p
    if p >= 1.0 / cfg.water_alpha:
        raise ValueError(
            f"Horizontal slowness p={p:.6e} s/m exceeds critical angle for water "
            f"(1/α_water = {1.0 / cfg.water_alpha:.6e} s/m). "
            "P-wave becomes evanescent in the water column.")
    gamma = cfg.gamma if cfg.gamma is not None else np.pi / cfg.T
    nwm = cfg.nw - 1
    omega_real = np.arange(1, nwm + 1, dtype=np.float64) * (2.0 * np.pi / cfg.T)
    omega_damped = omega_real + 1j * gamma
    nt = 2 * cfg.nw
    time_axis = np.arange(nt, dtype=np.float64) * (cfg.T / float(nt))
    f_hz = omega_real / (2.0 * np.pi)
    active = np.nonzero((f_hz >= cfg.f_min) & (f_hz <= cfg.f_max))[0]

    eta_w, eta_s, _, _ = _slownesses(cfg)
    H = cfg.geometry.N_z * cfg.geometry.d
    sed = IsotropicLayer(cfg.sed_ref.alpha, cfg.sed_ref.beta, cfg.sed_ref.rho, H)
    hs = IsotropicLayer(cfg.hs_alpha, cfg.hs_beta, cfg.hs_rho, np.inf)
    w_act = omega_damped[active]
    R_bg = np.zeros(nwm, dtype=complex)
    R_total = np.zeros(nwm, dtype=complex)
    R_slab = np.zeros(nwm, dtype=complex)
    n_iters: list[int] = []
    freq_elapsed: list[float] = []
    if active.size:
        water = FluidLayer(cfg.water_alpha, cfg.water_rho, cfg.water_depth)
        R_bg[active] = kennett_layers(LayerStack([water, sed, hs]), p, w_act).RPP
        # the top medium's thickness never enters: a constant interface coefficient
        R_sed_hs = kennett_layers(LayerStack([sed, hs]), p, w_act[:1]).RPP[0]
        k_hat = np.array([float(np.real(eta_s * cfg.sed_ref.alpha)),
                          float(p * cfg.sed_ref.alpha), 0.0])
        cases = [SweepCase(cfg.geometry, cfg.material, float(omega_real[i]), k_hat, "P",
                           periodic=True, volume_averaged=volume_averaged, n_orders=n_orders)
                 for i in active]
        ts = time.perf_counter()
        results = compute_slab_sweep(cases, gmres_tol=gmres_tol, restart=restart,
                                     max_iter=500, precision=precision, workers=workers)
        per = (time.perf_counter() - ts) / len(cases)
        for i, res in zip(active, results):
            R_slab[i] = slab_rpp_periodic(res, p=p)
            n_iters.append(int(res.n_gmres_iter))
            freq_elapsed.append(per)
        MT = np.exp(2j * omega_damped * eta_s * H) * R_sed_hs + R_slab
        R_total[active] = _kennett_water_step(MT[active], cfg)

    water_phase = np.exp(2j * omega_damped * eta_w * cfg.water_depth)

    def observed(R):
        x = water_phase * R
        return x / (1.0 + x) if cfg.free_surface else x

    S = ricker_source_spectrum(omega_real, cfg.f_peak)
    decay = np.exp(-gamma * time_axis)

    def trace(R):
        spec = np.zeros(nt, dtype=complex)
        U = S * R
        spec[1:cfg.nw] = U
        spec[cfg.nw + 1:] = np.conj(U[::-1])
        return np.real(np.fft.fft(spec)) * decay

    return OceanBottomResult(
        time=time_axis, trace_total=trace(observed(R_total)),
        trace_homogeneous=trace(observed(R_bg)), R_bg=R_bg, R_slab=R_slab, R_total=R_total,
        omega_real=omega_real, config=config, n_gmres_iters=n_iters, freq_elapsed=freq_elapsed,
        elapsed_seconds=time.perf_counter() - t0)


def load_ocean_bottom_config(path: str | Path) -> OceanBottomConfig:
    """YAML in seismic units (km/s, g/cm^3, GPa, km, s/km) to an SI config
    (ocean_bottom.py:484-616): sections ocean, sediment, halfspace, slab,
    recording, optional slowness; the slab gives exactly one of ``contrast``
    (inclusion values) or ``mean`` (ensemble mean, inclusion = mean / phi for
    a random material)."""
    import yaml

    path = Path(path)
    if not path.exists():
        raise FileNotFoundError(f"Config file not found: {path}")
    with path.open() as f:
        cfg = yaml.safe_load(f)
    required = ["ocean", "sediment", "halfspace", "slab", "recording"]
    missing = [s for s in required if s not in cfg]
    if missing:
        raise ValueError(f"Missing required config sections: {missing}. "
                         f"Expected sections: {required}. "
                         "See configs/example_ocean_bottom.yml for an example.")
    oc, sd, hs, sl, rec = (cfg[k] for k in required)
    slow = cfg.get("slowness", {})
    ref = ReferenceMedium(alpha=sd["alpha"] * 1e3, beta=sd["beta"] * 1e3, rho=sd["rho"] * 1e3)
    geom = SlabGeometry(M=sl["M"], N_z=sl.get("N_z", 1), a=sl["a"] * 1e3)
    if ("contrast" in sl) == ("mean" in sl):
        raise ValueError("slab section must contain exactly one of 'contrast' or 'mean'. "
                         "Use 'contrast' for raw inclusion values, 'mean' for ensemble-mean "
                         "parameterization (inclusion contrast = mean / phi).")
    kind = sl.get("material", "uniform")
    phi = sl.get("phi", 1.0 if kind == "uniform" else 0.3)
    if "mean" in sl:
        m = sl["mean"]
        scale = 1.0 if kind == "uniform" else 1.0 / phi
        contrast = MaterialContrast(Dlambda=m["Dlambda"] * 1e9 * scale,
                                    Dmu=m["Dmu"] * 1e9 * scale, Drho=m["Drho"] * 1e3 * scale)
    else:
        c = sl["contrast"]
        contrast = MaterialContrast(Dlambda=c["Dlambda"] * 1e9, Dmu=c["Dmu"] * 1e9,
                                    Drho=c["Drho"] * 1e3)
    if kind == "random":
        material = random_slab_material(geom, ref, contrast, phi=phi, seed=sl.get("seed", 42))
    else:
        material = uniform_slab_material(geom, ref, contrast)
    return OceanBottomConfig(
        water_alpha=oc["water_alpha"] * 1e3, water_rho=oc["water_rho"] * 1e3,
        water_depth=oc["water_depth"] * 1e3, sed_ref=ref, hs_alpha=hs["alpha"] * 1e3,
        hs_beta=hs["beta"] * 1e3, hs_rho=hs["rho"] * 1e3, geometry=geom, material=material,
        f_peak=rec["f_peak"], T=rec["T"], nw=rec["nw"], f_min=rec.get("f_min", 5.0),
        f_max=rec.get("f_max", 100.0), free_surface=oc.get("free_surface", False),
        p=slow.get("p", 0.0) * 1e-3)


def write_log(result: OceanBottomResult, path: str | Path) -> None:
    """Plain-text run log: configuration, material statistics, per-frequency
    |R_slab|, |R_total|, GMRES iterations and time, summary (ocean_bottom.py:371-481)."""
    from datetime import datetime, timezone

    cfg = result.config
    g = cfg.geometry
    H = g.N_z * g.d
    Dl, Dm, Dr = cfg.material.Dlambda, cfg.material.Dmu, cfg.material.Drho
    n = Dl.size
    n_inc = int(np.sum(Dl > 0))
    f_hz = result.omega_real / (2.0 * np.pi)
    act = np.nonzero((f_hz >= cfg.f_min) & (f_hz <= cfg.f_max))[0]
    out = ["# Ocean-Bottom Reflection Log",
           f"# Generated: {datetime.now(timezone.utc).isoformat()}", "", "## Configuration", "",
           f"Water:        α = {cfg.water_alpha / 1e3:.3f} km/s, ρ = {cfg.water_rho / 1e3:.3f} g/cm³, "
           f"depth = {cfg.water_depth / 1e3:.4f} km",
           f"Sediment:     α = {cfg.sed_ref.alpha / 1e3:.3f} km/s, β = {cfg.sed_ref.beta / 1e3:.3f} km/s, "
           f"ρ = {cfg.sed_ref.rho / 1e3:.3f} g/cm³",
           f"Halfspace:    α = {cfg.hs_alpha / 1e3:.3f} km/s, β = {cfg.hs_beta / 1e3:.3f} km/s, "
           f"ρ = {cfg.hs_rho / 1e3:.3f} g/cm³",
           f"Slab:         {g.M}×{g.M}×{g.N_z}, a = {g.a:.3f} m, d = {g.d:.3f} m, H = {H:.3f} m",
           f"Slowness:     p = {cfg.p * 1e3:.6f} s/km"]
    if cfg.p > 0:
        out.append(f"              θ_water = {np.degrees(np.arcsin(cfg.p * cfg.water_alpha)):.2f}°")
    out += [f"Free surface: {cfg.free_surface}", f"Wavelet:      Ricker f_peak = {cfg.f_peak} Hz",
            f"Recording:    T = {cfg.T} s, nw = {cfg.nw}, f = [{cfg.f_min}–{cfg.f_max}] Hz", "",
            "## Material statistics", "", f"Cubes:        {n} ({g.M}×{g.M}×{g.N_z})",
            f"Inclusions:   {n_inc} (φ_realised = {n_inc / n if n else 0.0:.3f})"]
    if 0 < n_inc < n:
        out += [f"Inclusion Δλ: {Dl[Dl > 0][0] / 1e9:.4f} GPa",
                f"Inclusion Δμ: {Dm[Dm > 0][0] / 1e9:.4f} GPa",
                f"Inclusion Δρ: {Dr[Dr > 0][0] / 1e3:.4f} g/cm³"]
    out += [f"Mean Δλ:      {np.mean(Dl) / 1e9:.6f} GPa", f"Mean Δμ:      {np.mean(Dm) / 1e9:.6f} GPa",
            f"Mean Δρ:      {np.mean(Dr) / 1e3:.6f} g/cm³", f"Var  Δλ:      {np.var(Dl) / 1e18:.6f} GPa²",
            f"Var  Δμ:      {np.var(Dm) / 1e18:.6f} GPa²", f"Var  Δρ:      {np.var(Dr) / 1e6:.6f} (g/cm³)²"]
    if np.mean(Dl) != 0:
        out.append(f"σ/μ (Δλ):     {np.std(Dl) / np.mean(Dl):.3f}")
    out += ["", "## Per-frequency diagnostics", "",
            f"{'freq_Hz':>10s}  {'|R_slab|':>10s}  {'|R_total|':>10s}  {'GMRES_iter':>10s}  {'elapsed_ms':>10s}",
            "-" * 58]
    for k, i in enumerate(act):
        it = result.n_gmres_iters[k] if k < len(result.n_gmres_iters) else -1
        ms = result.freq_elapsed[k] * 1e3 if k < len(result.freq_elapsed) else -1
        out.append(f"{f_hz[i]:10.2f}  {abs(result.R_slab[i]):10.2e}  {abs(result.R_total[i]):10.2e}  "
                   f"{it:10d}  {ms:10.2f}")
    out += ["", "## Summary", "", f"Active frequencies: {len(act)}"]
    if result.n_gmres_iters:
        it = result.n_gmres_iters
        out.append(f"GMRES iters:  mean = {np.mean(it):.1f}, min = {min(it)}, max = {max(it)}")
    if result.freq_elapsed:
        ft = result.freq_elapsed
        out += [f"Per-freq (ms): mean = {np.mean(ft) * 1e3:.1f}, min = {min(ft) * 1e3:.1f}, "
                f"max = {max(ft) * 1e3:.1f}", f"Slab total:   {sum(ft):.2f} s"]
    out += [f"Total elapsed: {result.elapsed_seconds:.2f} s",
            f"Peak |R_bg|:   {np.max(np.abs(result.R_bg)):.6f}",
            f"Peak |R_slab|: {np.max(np.abs(result.R_slab)):.6f}",
            f"Peak |R_total|: {np.max(np.abs(result.R_total)):.6f}"]
    Path(path).write_text("\n".join(out) + "\n")
