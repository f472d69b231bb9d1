"""GPU tests of compute_ocean_bottom_reflection (ocean_bottom.py:159-368):
the reference's own checks (cubic_scattering/tests/test_ocean_bottom.py:
74-430, same parameters) and parity with the oracle composition, whose
R_slab comes from the oracle's periodic slab solve and R_PP, and whose
Kennett and water steps restate the reference's closed forms.
Tolerances: R_bg, R_total, traces vs oracle 1e-9 relative (slab solves at
tol 1e-12); zero-contrast decomposition 1e-10 (reference's bound)."""

import numpy as np
import pytest

from oracle import kennett as ok
from oracle import slab as osl
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

SED = (2000.0, 800.0, 1800.0)
HS = (3000.0, 1700.0, 2200.0)
WATER = (1500.0, 1025.0, 70.0)
CONTRAST = (1.3244e9, 0.3116e9, 90.0)
WEAK = (1.3244e6, 0.3116e6, 0.09)


def _cfg(M=2, Nz=1, contrast=CONTRAST, nw=64, T=0.5, f_min=10.0, f_max=60.0, f_peak=30.0,
         free_surface=False, p=0.0, a=1.0):
    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(*SED)
    g = msc.SlabGeometry(M=M, N_z=Nz, a=a)
    mat = msc.uniform_slab_material(g, ref, msc.MaterialContrast(*contrast))
    return msc.OceanBottomConfig(water_alpha=WATER[0], water_rho=WATER[1], water_depth=WATER[2],
                                 sed_ref=ref, hs_alpha=HS[0], hs_beta=HS[1], hs_rho=HS[2],
                                 geometry=g, material=mat, f_peak=f_peak, T=T, nw=nw,
                                 f_min=f_min, f_max=f_max, free_surface=free_surface, p=p)


def _run(cfg, **kw):
    import multiplescatteringcalculations_amd as msc

    return msc.compute_ocean_bottom_reflection(cfg, progress=False, **kw)


def test_zero_contrast_traces_and_decomposition(hip_device):
    r = _run(_cfg(contrast=(0.0, 0.0, 0.0)))
    np.testing.assert_allclose(r.trace_total, r.trace_homogeneous, atol=1e-15)
    np.testing.assert_allclose(r.R_slab, 0.0, atol=1e-15)
    f = r.omega_real / (2 * np.pi)
    act = (f >= 10.0) & (f <= 60.0)
    np.testing.assert_allclose(r.R_total[act], r.R_bg[act], rtol=1e-10)


@pytest.mark.parametrize("p,free_surface,M,Nz", [(0.0, False, 2, 1), (2e-4, True, 4, 2),
                                                   (4e-4, False, 2, 2), (2e-4, False, 3, 2),
                                                   (1e-4, True, 6, 1)])
def test_matches_oracle_composition(hip_device, p, free_surface, M, Nz):
    cfg = _cfg(M=M, Nz=Nz, p=p, free_surface=free_surface)
    r = _run(cfg, gmres_tol=1e-12)
    ref = otm.Medium(*SED)
    mat = cfg.material
    eta = ok.vslow(ok.slowness(SED[0]), p)
    kh = [float(np.real(eta * SED[0])), p * SED[0], 0.0]
    gamma = np.pi / cfg.T
    f = r.omega_real / (2 * np.pi)
    act = np.nonzero((f >= cfg.f_min) & (f <= cfg.f_max))[0]
    R_slab = np.zeros(r.omega_real.shape, complex)
    for i in act:
        w = float(r.omega_real[i])
        x, _, T, _, res = osl.solve(M, Nz, 1.0, mat.Dlambda, mat.Dmu, mat.Drho, w, ref, kh, "P",
                                    tol=1e-13, max_iter=200, periodic=True)
        R_slab[i] = osl.rpp_periodic(x, T, M, Nz, 1.0, w, ref, p=p)
    H = Nz * 2.0
    mask = np.zeros(r.omega_real.shape, bool)
    mask[act] = True
    R_bg, R_tot, tr_t, tr_h, t = ok.ocean_bottom(WATER, SED, HS, H, p, r.omega_real, gamma, mask,
                                                 R_slab, free_surface, cfg.nw, cfg.T, cfg.f_peak)

    def rel(a, b):
        return np.abs(a - b).max() / np.abs(b).max()

    assert rel(r.R_slab, R_slab) < 1e-9
    assert rel(r.R_bg, R_bg) < 1e-12
    assert rel(r.R_total, R_tot) < 1e-9
    assert rel(r.trace_total, tr_t) < 1e-9 and rel(r.trace_homogeneous, tr_h) < 1e-12
    np.testing.assert_allclose(r.time, t, rtol=0, atol=0)
    assert len(r.n_gmres_iters) == len(act) == len(r.freq_elapsed)


def test_born_scaling_and_weak_perturbation(hip_device):
    r1 = _run(_cfg(contrast=WEAK))
    r2 = _run(_cfg(contrast=tuple(2 * c for c in WEAK)))
    act = np.abs(r1.R_slab) > 1e-20
    assert act.any()
    ratio = np.mean(np.abs(r2.R_slab[act]) / np.abs(r1.R_slab[act]))
    assert 1.5 < ratio < 2.5
    diff = np.abs(r1.trace_total - r1.trace_homogeneous).max()
    assert diff / np.abs(r1.trace_homogeneous).max() < 0.01


def test_energy_causality_coupling_and_shapes(hip_device):
    cfg = _cfg()
    r = _run(cfg)
    f = r.omega_real / (2 * np.pi)
    act = (f >= cfg.f_min) & (f <= cfg.f_max)
    assert np.abs(r.R_total[act]).max() < 1.0
    slab = np.abs(r.R_slab) > 1e-15
    assert np.mean(np.abs(r.R_total[slab] - r.R_bg[slab]) / np.abs(r.R_slab[slab])) < 1.0
    nt, nwm = 2 * cfg.nw, cfg.nw - 1
    assert r.time.shape == r.trace_total.shape == r.trace_homogeneous.shape == (nt,)
    assert r.R_bg.shape == r.R_slab.shape == r.R_total.shape == r.omega_real.shape == (nwm,)
    assert np.isrealobj(r.trace_total) and np.iscomplexobj(r.R_total)
    assert r.elapsed_seconds > 0 and r.config is cfg
    rc = _run(_cfg(nw=128, T=1.0, f_min=5.0, f_max=80.0))
    t_w = 2.0 * WATER[2] / WATER[0]
    early = rc.time < 0.5 * t_w
    assert np.abs(rc.trace_total[early]).max() / np.abs(rc.trace_total).max() < 0.15


def test_free_surface_multiples_and_mixed_precision(hip_device):
    """Free surface adds late water-column multiples; the mixed sweep gives the
    same R_slab as FP64 to its tolerance."""
    off = _run(_cfg(nw=128, T=1.0, f_min=5.0, f_max=80.0))
    on = _run(_cfg(nw=128, T=1.0, f_min=5.0, f_max=80.0, free_surface=True))
    t_mult = 3.0 * 2.0 * WATER[2] / WATER[0]
    late = (on.time > 0.7 * t_mult) & (on.time < 1.3 * t_mult)
    assert np.sum(on.trace_total[late] ** 2) > 2.0 * np.sum(off.trace_total[late] ** 2)
    quiet = (off.time > 0.8 * t_mult) & (off.time < 1.2 * t_mult)
    assert np.abs(off.trace_total[quiet]).max() < 0.15 * np.abs(off.trace_total).max()
    cfg = _cfg(p=2e-4)
    a = _run(cfg, gmres_tol=1e-10)
    b = _run(cfg, gmres_tol=1e-10, precision="mixed")
    assert np.abs(a.R_slab - b.R_slab).max() <= 1e-7 * np.abs(a.R_slab).max()
