"""GPU parity of the resonance T-matrix solver (msc_resonance_tmatrices)
against the oracle's dense complex128 solve: composite 9x9 T-matrix and
exciting field <= 1e-10 relative (direct solves), Neumann partial sums
<= 1e-11; batched contrasts each equal their own oracle solve; and the
reference's own checks (cubic_scattering/tests/test_cubic_tmatrix.py:591-719)."""

import numpy as np
import pytest

from oracle import resonance as ores
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

REF_T = otm.Medium(5000.0, 3000.0, 2500.0)
OM_T = 2.0 * np.pi * 10.0


def _msc():
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import resonance_tmatrix as rt

    return msc, rt


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(b).max())


@pytest.mark.parametrize("n,a,om,contrast,wave,kh", [
    (1, 1.0, OM_T, (2e9, 1e9, 100.0), "S", None),
    (2, 1.0, OM_T, (2e9, 1e9, 100.0), "S", None),
    (2, 30.0, 2 * np.pi * 40.0, (-4e9, 3e9, -300.0), "P", [0.4, -0.2, 1.0]),
    (3, 20.0, 2 * np.pi * 45.0, (6e9, -2e9, 400.0), "S", [1.0, 1.0, 0.5]),
    (4, 25.0, 2 * np.pi * 50.0, (3e9, 2e9, 250.0), "P", None),
    # 9N = 1125: the 8-column panels of the unblocked elimination
    (5, 15.0, 2 * np.pi * 30.0, (2e9, -1e9, 150.0), "S", [0.3, 0.5, 1.0]),
])
def test_single_cube_matches_oracle(hip_device, n, a, om, contrast, wave, kh):
    msc, rt = _msc()
    r = msc.ReferenceMedium(REF_T.alpha, REF_T.beta, REF_T.rho)
    c = msc.MaterialContrast(*contrast)
    with np.errstate(all="ignore"):
        res = rt.compute_resonance_tmatrix(om, a, r, c, n_sub=n, k_hat=kh, wave_type=wave)
    Tc, psi, _ = ores.resonance_tmatrix(om, a, REF_T, *contrast, n, k_hat=kh, wave_type=wave)
    assert _rel(res.T_comp_9x9, Tc) < 1e-10
    assert _rel(res.psi_exc, psi) < 1e-10
    np.testing.assert_array_equal(res.T3x3, res.T_comp_9x9[:3, :3])
    assert res.n_sub == n and res.centres.shape == (n ** 3, 3)
    np.testing.assert_allclose(res.centres, ores.centres(a, n), rtol=0, atol=1e-12 * a)
    # the reference's condition number: 2-norm of the assembled 9N x 9N matrix
    A_o, _, _ = ores.operator(om, a, REF_T, *contrast, n)
    assert abs(res.condition_number - np.linalg.cond(A_o)) <= 1e-8 * np.linalg.cond(A_o)
    T_loc = ores.t_loc(om, a / n, REF_T, *contrast)
    assert _rel(res.T_loc_9x9, T_loc) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3])
def test_assembled_system_matches_oracle(hip_device, n):
    """msc_resonance_system writes the oracle's A = I - P (I x T_loc)
    (column-major); a wrong sign, transposed block or dropped T_loc term would
    fail here before it reached the condition number."""
    import ctypes

    import torch

    from multiplescatteringcalculations_amd import _lib

    a, om, contrast = 20.0, 2 * np.pi * 45.0, (6e9, -2e9, 400.0)
    A_o, _, _ = ores.operator(om, a, REF_T, *contrast, n)
    n9 = 9 * n ** 3
    A = torch.empty((n9, n9), dtype=torch.complex128, device=hip_device)
    lib = _lib.load()
    _lib.check(lib.msc_resonance_system(om, a, REF_T.alpha, REF_T.beta, REF_T.rho, n,
                                        *map(float, contrast), ctypes.c_void_p(A.data_ptr()),
                                        None))
    A_h = A.cpu().numpy().T  # column-major on the device
    assert _rel(A_h, A_o) < 1e-12
    with pytest.raises(ValueError, match="n_sub"):
        _lib.check(lib.msc_resonance_system(om, a, REF_T.alpha, REF_T.beta, REF_T.rho, 9,
                                            *map(float, contrast), ctypes.c_void_p(A.data_ptr()),
                                            None))


@pytest.mark.parametrize("n,count", [(2, 64), (3, 12)])
def test_batch_matches_oracle_per_contrast(hip_device, n, count):
    msc, rt = _msc()
    rng = np.random.default_rng(n)
    lam = REF_T.rho * (REF_T.alpha ** 2 - 2 * REF_T.beta ** 2)
    mu = REF_T.rho * REF_T.beta ** 2
    Dl = rng.uniform(-0.3, 0.3, count) * lam
    Dm = rng.uniform(-0.3, 0.3, count) * mu
    Dr = rng.uniform(-0.3, 0.3, count) * REF_T.rho
    om, a = 2 * np.pi * 60.0, 12.0
    r = msc.ReferenceMedium(REF_T.alpha, REF_T.beta, REF_T.rho)
    T, ratio = rt.compute_resonance_tmatrices(om, a, r, Dl, Dm, Dr, n_sub=n, wave_type="P",
                                              k_hat=[0.0, 0.6, 0.8])
    assert T.shape == (count, 9, 9) and ratio.shape == (count,)
    check = range(count) if n == 2 else range(0, count, 3)
    for i in check:
        Tc, _, _ = ores.resonance_tmatrix(om, a, REF_T, Dl[i], Dm[i], Dr[i], n, wave_type="P",
                                          k_hat=[0.0, 0.6, 0.8])
        assert _rel(T[i], Tc) < 1e-10, i


def test_neumann_orders_match_oracle(hip_device):
    msc, rt = _msc()
    om, a = 2 * np.pi * 200.0, 2.0
    r = msc.ReferenceMedium(REF_T.alpha, REF_T.beta, REF_T.rho)
    c = msc.MaterialContrast(2e8, 1e8, 10.0)
    for k in (1, 3, 7):
        with pytest.warns(UserWarning, match="did not converge"):
            res = rt.compute_resonance_tmatrix(om, a, r, c, n_sub=2, wave_type="P",
                                               neumann_order=k, neumann_tol=0.0)
        Tn, psi, _ = ores.resonance_tmatrix(om, a, REF_T, 2e8, 1e8, 10.0, 2, wave_type="P",
                                            neumann_order=k, neumann_tol=0.0)
        assert _rel(res.T_comp_9x9, Tn) < 1e-11 and _rel(res.psi_exc, psi) < 1e-11
        assert res.n_iterations_converged is None
    res = rt.compute_resonance_tmatrix(om, a, r, c, n_sub=2, wave_type="P", neumann_order=80,
                                       neumann_tol=1e-13)
    _, _, kconv = ores.resonance_tmatrix(om, a, REF_T, 2e8, 1e8, 10.0, 2, wave_type="P",
                                         neumann_order=80, neumann_tol=1e-13)
    assert res.n_iterations_converged == kconv
    direct = rt.compute_resonance_tmatrix(om, a, r, c, n_sub=2, wave_type="P")
    assert _rel(res.T_comp_9x9, direct.T_comp_9x9) < 1e-11
    dec = rt.scattering_order_decomposition(om, a, r, c, n_sub=2, max_order=3)
    assert dec["orders"] == [0, 1, 2, 3]
    np.testing.assert_allclose(dec["T3x3_partial"][0], dec["T3x3_full"], rtol=1e-13)


def test_reference_checks(hip_device):
    """Rayleigh limit at n = 1 (1e-10), isotropy and O_h symmetry of the force
    monopole, stress dipole, near-field 6x6 (test_cubic_tmatrix.py:591-719)."""
    msc, rt = _msc()
    r = msc.ReferenceMedium(REF_T.alpha, REF_T.beta, REF_T.rho)
    c = msc.MaterialContrast(2.0e9, 1.0e9, 100.0)
    v = rt.validate_rayleigh_limit(OM_T, 1.0, r, c)
    assert v["relative_error"] < 1e-10 and v["passed"]
    res = rt.compute_resonance_tmatrix(OM_T, 1.0, r, msc.MaterialContrast(0.0, 0.0, 200.0),
                                       n_sub=2)
    d = np.diag(res.T3x3)
    assert np.abs(res.T3x3 - np.diag(d)).max() < 1e-10 * np.abs(d).max()
    np.testing.assert_allclose(d, d[0], rtol=1e-10)
    assert np.linalg.norm(res.stress_dipole_voigt) < 1e-10 * np.linalg.norm(res.T3x3)
    res = rt.compute_resonance_tmatrix(OM_T, 1.0, r, c, n_sub=2)
    d = np.diag(res.force_monopole)
    assert np.abs(res.force_monopole - np.diag(d)).max() < 1e-10 * np.abs(d).max()
    np.testing.assert_allclose(d, d[0], rtol=1e-10)
    assert np.linalg.norm(rt.compute_resonance_tmatrix(OM_T, 1.0, r, c,
                                                       n_sub=1).stress_dipole_voigt) > 0
    # near-field 6x6 at n = 1
    om, a, kx = 2 * np.pi * 1.0, 0.5, 0.001
    ray = msc.compute_cube_tmatrix(om, a, r, c)
    V = (2 * a) ** 3
    T6 = rt.voigt_tmatrix_from_resonance_result(
        rt.compute_resonance_tmatrix(om, a, r, c, n_sub=1), kx=kx, ky=0.0)
    want_u = complex(ray.Gamma0 * om ** 2 * ray.Drho_star * V)
    for i in range(3):
        assert abs(T6[i, i] - want_u) <= 1e-10 * abs(want_u)
    np.testing.assert_allclose(T6[4, 4], T6[5, 5], rtol=1e-10)
    Dc = msc.effective_stiffness_voigt(ray.Dlambda_star, ray.Dmu_star_diag, ray.Dmu_star_off)
    S = rt.strain_from_displacement_traction(kx, 0.0, r)
    want_t = V * rt._traction_from_voigt_strain(r) @ rt._volume_integral_voigt(
        ray.Ac, ray.Bc, ray.Cc) @ Dc @ S
    assert np.linalg.norm(T6[3:, :] - want_t) <= 1e-10 * np.linalg.norm(want_t)


def test_argument_errors(hip_device):
    msc, rt = _msc()
    r = msc.ReferenceMedium(REF_T.alpha, REF_T.beta, REF_T.rho)
    c = msc.MaterialContrast(2.0e9, 1.0e9, 100.0)
    with pytest.raises(ValueError, match="wave_type"):
        rt.compute_resonance_tmatrix(OM_T, 1.0, r, c, n_sub=1, wave_type="X")
    with pytest.raises(ValueError, match="n_sub too large"):
        rt.compute_resonance_tmatrices(OM_T, 1.0, r, [1e9], [1e9], [1.0], n_sub=9)
    with pytest.warns(UserWarning, match="ka_sub"):
        rt.compute_resonance_tmatrix(2 * np.pi * 400.0, 1.0, r, c, n_sub=1)
