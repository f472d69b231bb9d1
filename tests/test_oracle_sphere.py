"""CPU pins of the voxelised-sphere oracle (oracle/sphere.py)."""

import numpy as np

from oracle import resonance as ores
from oracle import sphere as osph
from oracle import tmatrix as otm

REF = otm.Medium(5000.0, 3000.0, 2500.0)
OM = 2 * np.pi * 40.0


def test_two_cells_per_edge_is_the_resonance_cube():
    """n = 2: all 8 cells are inside the sphere; the system is the resonance
    cube of half-width R split into 2^3 sub-cells."""
    Tc, psi, cen = osph.sphere_tmatrix(OM, 10.0, REF, 2e9, 1e9, 100.0, 2, wave_type="P",
                                       k_hat=[0.2, 0.1, 1.0])
    Tr, psir, _ = ores.resonance_tmatrix(OM, 10.0, REF, 2e9, 1e9, 100.0, 2, wave_type="P",
                                         k_hat=[0.2, 0.1, 1.0])
    assert len(cen) == 8
    np.testing.assert_allclose(Tc, Tr, rtol=0, atol=1e-12 * np.abs(Tr).max())
    np.testing.assert_allclose(psi, psir, rtol=0, atol=1e-12 * np.abs(psir).max())


def test_fft_matvec_equals_dense_and_cell_counts():
    n, R = 4, 12.0
    A, _, cen, idx = osph.operator(OM, R, REF, 2e9, 1e9, 100.0, n, inclusive=False)
    w = np.random.default_rng(1).standard_normal(A.shape[0]) * (1 - 0.5j)
    y = osph.fft_matvec(w, OM, R, REF, 2e9, 1e9, 100.0, n)
    np.testing.assert_allclose(y, A @ w, rtol=0, atol=1e-11 * np.abs(A @ w).max())
    assert len(osph.cells(R, 3)[1]) == 19 and len(osph.cells(R, 4)[1]) == 32
    assert len(osph.cells(R, 3, inclusive=False)[1]) == 19
