"""GPU parity of the voxelised-sphere Foldy-Lax solve (slab plan with the
contrast masked to the sphere) against the oracle's dense complex128 solve:
T_comp and psi_exc <= 1e-9 relative at GMRES tol 1e-12; complex64 (mixed)
<= 1e-6; the two device paths (sphere and resonance cube) agree at n = 2."""

import numpy as np
import pytest
import torch

from oracle import sphere as osph
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

REF = otm.Medium(5000.0, 3000.0, 2500.0)


def _mods():
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import sphere_scattering_fft as sph

    return msc, sph


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


@pytest.mark.parametrize("n,R,om,contrast,wave,kh", [
    (3, 12.0, 2 * np.pi * 40.0, (2e9, 1e9, 100.0), "S", None),
    (4, 12.0, 2 * np.pi * 50.0, (-3e9, 2e9, -200.0), "P", [0.3, 0.0, 1.0]),
    (5, 20.0, 2 * np.pi * 30.0, (4e9, 1.5e9, 300.0), "S", [1.0, 1.0, 1.0]),
])
def test_fft_sphere_matches_dense_oracle(hip_device, n, R, om, contrast, wave, kh):
    msc, sph = _mods()
    r = msc.ReferenceMedium(REF.alpha, REF.beta, REF.rho)
    c = msc.MaterialContrast(*contrast)
    res = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, k_hat=kh, wave_type=wave,
                                               gmres_tol=1e-12, gmres_maxiter=400)
    Tc, psi, cen = osph.sphere_tmatrix(om, R, REF, *contrast, n, k_hat=kh, wave_type=wave,
                                       inclusive=False)
    assert res.n_cells == len(cen) and res.psi_exc.shape == (9 * len(cen), 9)
    np.testing.assert_allclose(res.centres, cen, rtol=0, atol=1e-12 * R)
    assert _rel(res.T_comp_9x9, Tc) < 1e-9
    assert _rel(res.psi_exc, psi) < 1e-9
    np.testing.assert_array_equal(res.T3x3, res.T_comp_9x9[:3, :3])
    cpu = sph.compute_sphere_foldy_lax_fft(om, R, r, c, n, k_hat=kh, wave_type=wave,
                                           gmres_tol=1e-12)
    assert _rel(cpu.T_comp_9x9, Tc) < 1e-9


def test_dense_api_and_mixed_precision(hip_device):
    msc, sph = _mods()
    r = msc.ReferenceMedium(REF.alpha, REF.beta, REF.rho)
    c = msc.MaterialContrast(2e9, 1e9, 100.0)
    om, R, n = 2 * np.pi * 40.0, 12.0, 4
    d = sph.compute_sphere_foldy_lax(om, R, r, c, n, wave_type="P")
    Tc, _, _ = osph.sphere_tmatrix(om, R, REF, 2e9, 1e9, 100.0, n, wave_type="P")
    assert _rel(d.T_comp_9x9, Tc) < 1e-10 and np.isnan(d.condition_number)
    m = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P", gmres_tol=1e-9,
                                             dtype=torch.complex64)
    f = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P", gmres_tol=1e-9,
                                             initial_guess="zero")
    assert _rel(m.T_comp_9x9, Tc) < 1e-6 and _rel(f.T_comp_9x9, Tc) < 1e-7
    with pytest.raises(ValueError, match="initial_guess"):
        sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, initial_guess="x")
    with pytest.raises(ValueError, match="wave_type"):
        sph.compute_sphere_foldy_lax_fft(om, R, r, c, n, wave_type="Q")


@pytest.mark.parametrize("dtype", [None, torch.complex64])
def test_workers_give_the_serial_result(hip_device, dtype):
    """Concurrent solves (plans on their own streams, G0 copied from the first
    plan) return exactly the one-worker result."""
    msc, sph = _mods()
    r = msc.ReferenceMedium(REF.alpha, REF.beta, REF.rho)
    c = msc.MaterialContrast(-3e9, 2e9, -200.0)
    om, R, n = 2 * np.pi * 50.0, 12.0, 6
    one = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P", gmres_tol=1e-10,
                                               restart=20, dtype=dtype, workers=1)
    for w in (2, 4, 9):
        res = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P",
                                                   gmres_tol=1e-10, restart=20, dtype=dtype,
                                                   workers=w)
        np.testing.assert_array_equal(res.T_comp_9x9, one.T_comp_9x9)
        np.testing.assert_array_equal(res.psi_exc, one.psi_exc)
    with pytest.raises(ValueError, match="workers"):
        sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, workers=-1)


@pytest.mark.parametrize("dtype", [None, torch.complex64])
def test_batched_right_hand_sides_equal_the_serial_result(hip_device, dtype):
    """The default solve (the 9 right-hand sides as one msc_gmres_batched
    batch) agrees with one worker solving them one after another to the
    solver tolerance (the batched cycles normalise the right-hand side, so
    not bit for bit)."""
    msc, sph = _mods()
    r = msc.ReferenceMedium(REF.alpha, REF.beta, REF.rho)
    c = msc.MaterialContrast(-3e9, 2e9, -200.0)
    om, R, n = 2 * np.pi * 50.0, 12.0, 6
    one = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P", gmres_tol=1e-11,
                                               restart=20, dtype=dtype, workers=1)
    bat = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P", gmres_tol=1e-11,
                                               restart=20, dtype=dtype)
    assert _rel(bat.T_comp_9x9, one.T_comp_9x9) < 1e-9
    assert _rel(bat.psi_exc, one.psi_exc) < 1e-9
    # unrestarted (the reference's GPU default) through the batch as well
    unr = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, n, wave_type="P", gmres_tol=1e-11,
                                               dtype=dtype)
    assert _rel(unr.T_comp_9x9, one.T_comp_9x9) < 1e-9


def test_two_cells_per_edge_equals_the_resonance_solver(hip_device):
    msc, sph = _mods()
    from multiplescatteringcalculations_amd import resonance_tmatrix as rt

    r = msc.ReferenceMedium(REF.alpha, REF.beta, REF.rho)
    c = msc.MaterialContrast(2e9, 1e9, 100.0)
    om, R = 2 * np.pi * 40.0, 10.0
    s = sph.compute_sphere_foldy_lax_fft_gpu(om, R, r, c, 2, wave_type="P", gmres_tol=1e-13)
    q = rt.compute_resonance_tmatrix(om, R, r, c, n_sub=2, wave_type="P")
    assert _rel(s.T_comp_9x9, q.T_comp_9x9) < 1e-10


SPHERE_CFG = """\
device:
  backend: hip
  dtype: {dtype}
solver:
  gmres_tol: 1.0e-11
  max_iter: 300
  initial_guess: born
problem:
  type: sphere
  geometry: {{M: 4, N_z: 4, a: 1.0}}
  reference: {{alpha: 5000.0, beta: 3000.0, rho: 2500.0}}
  contrast: {{Dlambda: 2.0e9, Dmu: 1.0e9, Drho: 100.0}}
  frequency: {{omega: 251.32741228718345}}
  incident:
    k_hat: [0.0, 0.0, 1.0]
    wave_type: S
  sphere: {{radius: 12.0, n_sub: 4}}
"""


@pytest.mark.parametrize("dtype,tol", [("complex128", 1e-9), ("complex64", 1e-6)])
def test_run_from_config_sphere(hip_device, tmp_path, dtype, tol):
    """type: sphere in a YAML config runs the FFT sphere solve
    (solver_config.py:440-475) and matches the oracle's dense solve.  The
    reference's validator requires problem.geometry for spheres too
    (solver_config.py:307-317), so the config carries one."""
    from multiplescatteringcalculations_amd.solver_config import load_config, run_from_config

    p = tmp_path / "sphere.yml"
    p.write_text(SPHERE_CFG.format(dtype=dtype))
    res = run_from_config(load_config(p))
    Tc, _, cen = osph.sphere_tmatrix(251.32741228718345, 12.0, REF, 2e9, 1e9, 100.0, 4,
                                     k_hat=[0.0, 0.0, 1.0], wave_type="S", inclusive=False)
    assert res.n_cells == len(cen) and res.n_sub == 4
    assert _rel(res.T_comp_9x9, Tc) < tol
