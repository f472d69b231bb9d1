"""CPU pins of the resonance T-matrix oracle (oracle/resonance.py) against the
slab oracle, the Rayleigh limit and the symmetry checks of
cubic_scattering/tests/test_cubic_tmatrix.py:591-755."""

import numpy as np
import pytest

from oracle import resonance as ores
from oracle import slab as osl
from oracle import tmatrix as otm

REF = otm.Medium(5000.0, 3000.0, 2500.0)
OM = 2.0 * np.pi * 10.0


@pytest.mark.parametrize("n,a", [(2, 1.0), (3, 40.0)])
def test_internal_operator_is_the_uniform_slab_operator(n, a):
    """A = I - P (I x T_loc) equals the slab oracle's dense operator of an
    n x n x n uniform slab of half-width a/n (same centre ordering and frame:
    axis 0 = slab z); the slab's own FFT matvec agrees too."""
    Dl, Dm, Dr = 2e9, 1e9, 100.0
    A, T, cx = ores.operator(OM, a, REF, Dl, Dm, Dr, n)
    Ts = osl.slab_tmatrices(n, n, a / n, np.full((n, n, n), Dl), np.full((n, n, n), Dm),
                            np.full((n, n, n), Dr), OM, REF)
    D = osl.dense_operator(Ts, n, n, a / n, OM, REF)
    np.testing.assert_allclose(A, D, rtol=0, atol=1e-12 * np.abs(D).max())
    x = np.random.default_rng(0).standard_normal(A.shape[0]) * (1 + 1j)
    kh = osl.slab_kernels(n, n, a / n, OM, REF)
    y = osl.slab_matvec(x, Ts, kh, n, n)
    np.testing.assert_allclose(A @ x, y, rtol=0, atol=1e-10 * np.abs(y).max())


def test_n1_is_the_rayleigh_tmatrix():
    Tc, psi, _ = ores.resonance_tmatrix(OM, 1.0, REF, 2e9, 1e9, 100.0, 1)
    ray = otm.cube_tmatrix(OM, 1.0, REF, 2e9, 1e9, 100.0)
    T_ray = 8.0 * OM ** 2 * ray.Drho_star * np.eye(3)
    assert np.linalg.norm(Tc[:3, :3] - T_ray) / np.linalg.norm(T_ray) < 1e-12
    np.testing.assert_allclose(Tc, otm.tmatrix_9x9(ray, OM, 1.0) @ ores.incident(
        ores.centres(1.0, 1), OM, REF)[:9], rtol=1e-13)


def test_density_only_is_isotropic_and_stiffness_gives_a_stress_dipole():
    Tc, _, _ = ores.resonance_tmatrix(OM, 1.0, REF, 0.0, 0.0, 200.0, 2)
    d = np.diag(Tc[:3, :3])
    assert np.abs(Tc[:3, :3] - np.diag(d)).max() < 1e-10 * np.abs(d).max()
    np.testing.assert_allclose(d, d[0], rtol=1e-10)
    assert np.linalg.norm(Tc[3:, 3:]) < 1e-10 * np.linalg.norm(Tc[:3, :3])
    Tc, _, _ = ores.resonance_tmatrix(OM, 1.0, REF, 2e9, 1e9, 100.0, 2)
    d = np.diag(Tc[:3, :3])
    assert np.abs(Tc[:3, :3] - np.diag(d)).max() < 1e-10 * np.abs(d).max()
    np.testing.assert_allclose(d, d[0], rtol=1e-10)
    assert np.linalg.norm(Tc[3:, 3:]) > 0


def test_neumann_series_converges_to_the_direct_solve():
    om = 2 * np.pi * 200.0
    Td, _, _ = ores.resonance_tmatrix(om, 2.0, REF, 2e8, 1e8, 10.0, 2, wave_type="P",
                                      k_hat=[0.3, 0.2, 1.0])
    Tn, _, k = ores.resonance_tmatrix(om, 2.0, REF, 2e8, 1e8, 10.0, 2, wave_type="P",
                                      k_hat=[0.3, 0.2, 1.0], neumann_order=60, neumann_tol=1e-14)
    assert k is not None
    np.testing.assert_allclose(Tn, Td, rtol=0, atol=1e-12 * np.abs(Td).max())


def test_incident_patterns_closed_form():
    cx = ores.centres(1.0, 2)
    U = ores.incident(cx, OM, REF)
    m = len(cx) // 2
    assert abs(U[9 * m, 0]) > 0 and np.allclose(U[9 * m + 3:9 * m + 9, 0], 0.0)
    ph = U[9 * m + 3, 3]
    want = np.zeros(6, complex)
    want[0] = ph
    np.testing.assert_allclose(U[9 * m + 3:9 * m + 9, 3], want, atol=1e-14)
    k = OM / REF.beta
    np.testing.assert_allclose(U[9 * m, 0], np.exp(1j * k * cx[m, 2]), rtol=1e-14)
