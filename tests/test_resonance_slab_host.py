"""Host checks of the resonance-regime slab pipeline (resonance_slab.py): the
sub-cell resolution of a slab, the CPA phases of the uniform contrast
distribution and argument validation (no device work)."""

import numpy as np
import pytest

import multiplescatteringcalculations_amd as msc
from multiplescatteringcalculations_amd import resonance_slab as rs

REF = msc.ReferenceMedium(3000.0, 1732.0, 2500.0)


@pytest.mark.parametrize("n", [1, 2, 4])
def test_subcell_slab_keeps_the_medium(n):
    g = msc.SlabGeometry(M=5, N_z=3, a=1.5)
    mat = msc.random_iid_slab_material(g, REF, 0.4, 0.3, seed=7)
    gf, mf = rs.subcell_slab(g, mat, n)
    assert (gf.M, gf.N_z) == (5 * n, 3 * n)
    assert gf.a == pytest.approx(1.5 / n)
    # same thickness and lateral period
    assert gf.N_z * gf.d == pytest.approx(g.N_z * g.d)
    assert gf.M * gf.d == pytest.approx(g.M * g.d)
    # sub-cell (z, x, y) of cube (z // n, x // n, y // n) carries its contrasts
    rng = np.random.default_rng(0)
    for _ in range(50):
        z, x, y = rng.integers(0, 3 * n), rng.integers(0, 5 * n), rng.integers(0, 5 * n)
        for a, b in ((mf.Dlambda, mat.Dlambda), (mf.Dmu, mat.Dmu), (mf.Drho, mat.Drho)):
            assert a[z, x, y] == b[z // n, x // n, y // n]
    # every sub-cell centre lies inside its cube
    cf = gf.all_centres().reshape(3 * n, 5 * n, 5 * n, 3)
    cc = g.all_centres().reshape(3, 5, 5, 3)
    for z, x, y in [(0, 0, 0), (3 * n - 1, 5 * n - 1, 2), (n, 2 * n + 1 if n > 1 else 2, 0)]:
        assert np.all(np.abs(cf[z, x, y] - cc[z // n, x // n, y // n]) < g.a)
    with pytest.raises(ValueError):
        rs.subcell_slab(g, mat, 0)


@pytest.mark.parametrize("order", [2, 3, 4])
def test_uniform_contrast_phases_moments(order):
    """Fractions sum to one; the contrasts' mean is zero and their second
    moments equal those of U[-c, c] (c^2 / 3 per axis, no cross terms), which a
    Gauss-Legendre rule of two or more points integrates exactly."""
    phi, c = 0.3, 0.3
    ph = rs.uniform_contrast_phases(REF, phi, c, order=order)
    assert len(ph) == 1 + order ** 3
    f = np.array([q.volume_fraction for q in ph])
    assert f.sum() == pytest.approx(1.0, abs=1e-14)
    assert f[0] == pytest.approx(1 - phi)
    x = np.array([[q.lam / REF.lam - 1, q.mu / REF.mu - 1, q.rho / REF.rho - 1] for q in ph[1:]])
    w = f[1:] / phi
    assert np.abs(w @ x).max() < 1e-14
    m2 = np.einsum("k,ki,kj->ij", w, x, x)
    assert np.allclose(m2, np.eye(3) * c * c / 3.0, atol=1e-14)


def test_resonance_ocean_bottom_rejects_evanescent_water():
    g = msc.SlabGeometry(M=2, N_z=1, a=1.0)
    mat = msc.uniform_slab_material(g, REF, msc.MaterialContrast(0.0, 0.0, 0.0))
    with pytest.raises(ValueError, match="critical angle"):
        rs.resonance_ocean_bottom(g, mat, [1000.0], p=1.0 / 1500.0)
