"""CPU tests of the ocean-bottom caller's host parts: the YAML loader in
seismic units (cubic_scattering/tests/test_ocean_bottom.py:394-530, same
dictionary), the critical-angle check, the Ricker spectrum, the water step's
decomposition identity and the run log."""

import numpy as np
import pytest
import yaml

import multiplescatteringcalculations_amd as msc
import multiplescatteringcalculations_amd.kennett_layers  # noqa: F401  (the package root exports the function of that name)
import sys

kl = sys.modules["multiplescatteringcalculations_amd.kennett_layers"]
from multiplescatteringcalculations_amd import ocean_bottom as ob


def _dict():
    return {
        "ocean": {"water_alpha": 1.5, "water_rho": 1.025, "water_depth": 0.07,
                  "free_surface": False},
        "sediment": {"alpha": 2.0, "beta": 0.8, "rho": 1.8},
        "halfspace": {"alpha": 3.0, "beta": 1.7, "rho": 2.2},
        "slab": {"M": 2, "N_z": 1, "a": 0.001, "material": "uniform",
                 "mean": {"Dlambda": 1.3244, "Dmu": 0.3116, "Drho": 0.09}},
        "recording": {"f_peak": 30.0, "T": 0.5, "nw": 64, "f_min": 10.0, "f_max": 60.0},
        "slowness": {"p": 0.0},
    }


def _write(tmp_path, d):
    p = tmp_path / "config.yml"
    p.write_text(yaml.dump(d))
    return p


def test_yaml_load_converts_seismic_units(tmp_path):
    cfg = msc.load_ocean_bottom_config(_write(tmp_path, _dict()))
    assert (cfg.water_alpha, cfg.water_rho, cfg.water_depth) == (1500.0, 1025.0, 70.0)
    assert (cfg.sed_ref.alpha, cfg.sed_ref.beta, cfg.sed_ref.rho) == (2000.0, 800.0, 1800.0)
    assert (cfg.hs_alpha, cfg.hs_beta, cfg.hs_rho) == (3000.0, 1700.0, 2200.0)
    assert (cfg.geometry.M, cfg.geometry.N_z, cfg.geometry.a) == (2, 1, 1.0)
    assert cfg.f_peak == 30.0 and cfg.p == 0.0 and cfg.free_surface is False
    np.testing.assert_allclose(cfg.material.Dlambda, 1.3244e9)
    d = _dict()
    d["slowness"]["p"] = 0.2
    assert msc.load_ocean_bottom_config(_write(tmp_path, d)).p == 0.0002


def test_yaml_mean_random_and_contrast(tmp_path):
    d = _dict()
    d["slab"].update(material="random", phi=0.4, seed=7,
                     mean={"Dlambda": 0.5, "Dmu": 0.1, "Drho": 0.02})
    cfg = msc.load_ocean_bottom_config(_write(tmp_path, d))
    inc = cfg.material.Dlambda[cfg.material.Dlambda > 0]
    assert inc.size and inc[0] == pytest.approx(0.5e9 / 0.4, rel=1e-12)
    d = _dict()
    del d["slab"]["mean"]
    d["slab"]["contrast"] = {"Dlambda": 1.0, "Dmu": 0.5, "Drho": 0.05}
    cfg = msc.load_ocean_bottom_config(_write(tmp_path, d))
    assert cfg.material.Dlambda[0, 0, 0] == 1.0e9 and cfg.material.Dmu[0, 0, 0] == 0.5e9


def test_yaml_errors(tmp_path):
    d = _dict()
    del d["sediment"]
    with pytest.raises(ValueError, match="Missing required config sections"):
        msc.load_ocean_bottom_config(_write(tmp_path, d))
    with pytest.raises(FileNotFoundError):
        msc.load_ocean_bottom_config("/nonexistent/path.yml")
    d = _dict()
    d["slab"]["contrast"] = {"Dlambda": 1.0, "Dmu": 0.5, "Drho": 0.05}
    with pytest.raises(ValueError, match="exactly one of"):
        msc.load_ocean_bottom_config(_write(tmp_path, d))
    d = _dict()
    del d["slab"]["mean"]
    with pytest.raises(ValueError, match="exactly one of"):
        msc.load_ocean_bottom_config(_write(tmp_path, d))


def test_critical_angle_is_refused_before_any_solve(tmp_path):
    cfg = msc.load_ocean_bottom_config(_write(tmp_path, _dict()))
    cfg.p = 1.0 / cfg.water_alpha
    with pytest.raises(ValueError, match="exceeds critical angle"):
        msc.compute_ocean_bottom_reflection(cfg, progress=False)


def test_ricker_spectrum():
    w = 2 * np.pi * np.array([0.0, 15.0, 30.0, 60.0])
    S = msc.ricker_source_spectrum(w, 30.0)
    np.testing.assert_allclose(S, [0.0, 0.25 * np.exp(-0.25), np.exp(-1.0), 4 * np.exp(-4.0)],
                               rtol=1e-15)


@pytest.mark.parametrize("p", [0.0, 2e-4, 4e-4])
def test_water_step_decomposition_equals_full_recursion(tmp_path, p):
    """At normal incidence [water | sed(H) | hs] = water step of
    (E_sed^2 R[sed | hs]), the identity the caller's slab injection rests on
    (ocean_bottom.py:270-323).  At p > 0 the reference's scalar PP step drops
    the P-S conversions, so the two differ (kept for parity)."""
    cfg = msc.load_ocean_bottom_config(_write(tmp_path, _dict()))
    cfg.p = p
    H = 7.0
    w = 2 * np.pi * np.linspace(5, 90, 12) + 1j * np.pi / 0.5
    sed = kl.IsotropicLayer(2000.0, 800.0, 1800.0, H)
    hs = kl.IsotropicLayer(3000.0, 1700.0, 2200.0, np.inf)
    water = kl.FluidLayer(1500.0, 1025.0, 70.0)
    full = kl.kennett_layers(kl.LayerStack([water, sed, hs]), p, w).RPP
    sub = kl.kennett_layers(kl.LayerStack([sed, hs]), p, w).RPP
    eta = kl.vertical_slowness(kl.complex_slowness(2000.0, np.inf), p)
    got = ob._kennett_water_step(np.exp(2j * w * eta * H) * sub, cfg)
    if p == 0.0:
        np.testing.assert_allclose(got, full, rtol=1e-12)
    else:
        assert np.abs(got - full).max() > 1e-3 * np.abs(full).max()


def test_write_log(tmp_path):
    cfg = msc.load_ocean_bottom_config(_write(tmp_path, _dict()))
    nwm = cfg.nw - 1
    om = np.arange(1, nwm + 1) * 2 * np.pi / cfg.T
    f = om / (2 * np.pi)
    n_act = int(np.sum((f >= cfg.f_min) & (f <= cfg.f_max)))
    res = msc.OceanBottomResult(time=np.zeros(2 * cfg.nw), trace_total=np.zeros(2 * cfg.nw),
                                trace_homogeneous=np.zeros(2 * cfg.nw),
                                R_bg=np.full(nwm, 0.3 + 0j), R_slab=np.full(nwm, 1e-3 + 0j),
                                R_total=np.full(nwm, 0.31 + 0j), omega_real=om, config=cfg,
                                n_gmres_iters=[3] * n_act, freq_elapsed=[0.01] * n_act,
                                elapsed_seconds=1.5)
    path = tmp_path / "run.log"
    msc.write_log(res, path)
    text = path.read_text()
    assert "## Per-frequency diagnostics" in text and f"Active frequencies: {n_act}" in text
    assert "GMRES iters:  mean = 3.0, min = 3, max = 3" in text
    assert text.count("\n") > n_act + 20
