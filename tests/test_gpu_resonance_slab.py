"""GPU parity of the resonance-regime slab (resonance_slab.py): the full-wave
composite T of every cube (n_sub^3 sub-cells) in the periodic slab solve, the
slab's specular R_PP, its embedding between a water and an elastic half-space
and the Kennett reflectivity of the CPA medium, against the oracle: the
oracle's composite T per inclusion (oracle/resonance.py), the periodic slab
operator with that dense T per voxel and its direct solve and R_PP
(oracle/slab.py), the fluid-solid step and Kennett recursion
(oracle/kennett.py) and the CPA iteration (oracle/cpa.py).

Tolerances: slab R_PP 1e-8 relative (GMRES to 1e-12 against a direct solve),
the embedded R_PP 1e-8, the Kennett values 1e-10; the device field 1e-9.
"""

import numpy as np
import pytest

from oracle import cpa as ocpa
from oracle import kennett as ok
from oracle import resonance as ores
from oracle import slab as osl
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

VP, VS, RHO = 3000.0, 1732.0, 2500.0
WATER = (1500.0, 1000.0)
HS = (4000.0, 2300.0, 2600.0)


def _mods():
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import resonance_slab as rs

    return msc, rs


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _oracle_slab(mat, M, Nz, om, n_sub, p):
    """Dense-T periodic slab solve of the oracle: (psi, T, k_hat)."""
    oref = otm.Medium(VP, VS, RHO)
    eta = ok.vslow(ok.slowness(VP), p)
    k_hat = np.array([float(np.real(eta * VP)), p * VP, 0.0])
    T = np.zeros((Nz, M, M, 9, 9), complex)
    for iz, ix, iy in np.argwhere((mat.Dlambda != 0) | (mat.Dmu != 0) | (mat.Drho != 0)):
        T[iz, ix, iy] = ores.resonance_tmatrix(om, 1.0, oref, mat.Dlambda[iz, ix, iy],
                                               mat.Dmu[iz, ix, iy], mat.Drho[iz, ix, iy], n_sub,
                                               k_hat=k_hat, wave_type="P")[0]
    kh = osl.slab_kernels(M, Nz, 1.0, om, oref, periodic=True)
    n = Nz * M * M * 9
    A = np.stack([osl.slab_matvec(e, T, kh, M, Nz, periodic=True) for e in np.eye(n, dtype=complex)],
                 axis=1)
    psi0 = osl.incident_field(M, Nz, 1.0, om, oref, k_hat, "P").ravel()
    psi = np.linalg.solve(A, psi0).reshape(Nz, M, M, 9)
    return psi, T, k_hat


@pytest.mark.parametrize("M,Nz,n_sub,ka,p", [(4, 2, 4, 1.5, 0.0), (3, 2, 2, 2.0, 1.0e-4)])
def test_resonance_slab_matches_oracle(hip_device, M, Nz, n_sub, ka, p):
    msc, rs = _mods()
    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    mat = msc.random_iid_slab_material(g, ref, 0.3, 0.3, seed=11)
    om = ka * VS
    res = rs.compute_resonance_slab_scattering(g, mat, om, [1.0, 0.0, 0.0] if p == 0 else
                                               [float(np.sqrt(1 - (p * VP) ** 2)), p * VP, 0.0],
                                               "P", n_sub=n_sub, gmres_tol=1e-12,
                                               precision="fp64")
    psi, T, k_hat = _oracle_slab(mat, M, Nz, om, n_sub, p)
    assert _rel(res.psi, psi) < 1e-9
    R_or = osl.rpp_periodic(psi, T, M, Nz, 1.0, om, otm.Medium(VP, VS, RHO), p)
    assert abs(msc.slab_rpp_periodic(res, p=p) - R_or) < 1e-8 * abs(R_or)
    # the whole embedding over two frequencies, with the CPA reference
    phases = rs.uniform_contrast_phases(ref, 0.3, 0.3, order=2)
    oms = np.array([om, 1.1 * om])
    out = rs.resonance_ocean_bottom(g, mat, oms, p=p, n_sub=n_sub, cpa_phases=phases,
                                    gmres_tol=1e-12, precision="fp64")
    H = Nz * 2.0
    R_slab = np.array([R_or, osl.rpp_periodic(*_oracle_slab(mat, M, Nz, oms[1], n_sub, p)[:2], M, Nz,
                                              1.0, oms[1], otm.Medium(VP, VS, RHO), p)])
    assert _rel(out.R_slab, R_slab) < 1e-8
    # undamped frequencies (gamma 0); nw = 3 makes the oracle's trace spectrum hold two bins
    R_bg, R_tot, *_ = ok.ocean_bottom((WATER[0], WATER[1], 1.0), (VP, VS, RHO), HS, H, p, oms, 0.0,
                                      np.ones(2, bool), R_slab, False, 3, 1.0, 10.0)
    assert _rel(out.R_bg, R_bg) < 1e-10
    assert _rel(out.R_total, R_tot) < 1e-8
    for i, w in enumerate(oms):
        lam, mo, _, rho, _, conv = ocpa.cpa([(q.lam, q.mu, q.rho, q.volume_fraction) for q in phases],
                                            w, 1.0)
        assert conv
        layers = [(WATER[0], 0.0, WATER[1], 1.0),
                  (float(np.sqrt((lam + 2 * mo) / rho)), float(np.sqrt(mo / rho)), rho, H),
                  (*HS, np.inf)]
        R_cpa = ok.reflectivity(layers, p, np.array([w]))[0][0, 0, 0]
        assert abs(out.R_cpa[i] - R_cpa) < 1e-10 * abs(R_cpa)
    assert all(it > 0 for it in out.iterations)


def test_resonance_slab_n1_is_rayleigh_slab(hip_device):
    """One sub-cell per cube: the composite T is the cube's Rayleigh T0, so the
    resonance slab solve equals the ordinary periodic slab solve (the mixed
    path, to the FP64 tolerance of both)."""
    msc, rs = _mods()
    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=8, N_z=4, a=1.0)
    mat = msc.random_iid_slab_material(g, ref, 0.3, 0.3, seed=4)
    om = 0.8 * VS
    kh = [float(np.sqrt(1 - 0.45 ** 2)), 0.45, 0.0]
    a = rs.compute_resonance_slab_scattering(g, mat, om, kh, "P", n_sub=1, gmres_tol=1e-11,
                                             precision="mixed")
    b = msc.compute_slab_scattering(g, mat, om, kh, "P", gmres_tol=1e-11, periodic=True,
                                    restart=30)
    assert _rel(a.psi, b.psi) < 1e-9
    ra, rb = msc.slab_rpp_periodic(a), msc.slab_rpp_periodic(b)
    assert abs(ra - rb) < 1e-9 * abs(rb)
    fa, fb = msc.slab_reflected_field(a), msc.slab_reflected_field(b)
    assert np.allclose(fa, fb, rtol=1e-8, atol=1e-9 * max(abs(v) for v in fb))


def test_resonance_slab_background_only(hip_device):
    """No inclusions: the T table holds only the zero matrix, the solve returns
    the incident field and R_slab = 0, so the embedded R_PP is the homogeneous
    background's Kennett value."""
    msc, rs = _mods()
    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=4, N_z=2, a=1.0)
    z = np.zeros((2, 4, 4))
    mat = msc.SlabMaterial(Dlambda=z, Dmu=z.copy(), Drho=z.copy(), ref=ref)
    out = rs.resonance_ocean_bottom(g, mat, [1.2 * VS], n_sub=2, gmres_tol=1e-10)
    assert out.R_slab[0] == 0
    assert abs(out.R_total[0] - out.R_bg[0]) < 1e-12 * abs(out.R_bg[0])


def test_subcell_embedding_matches_oracle(hip_device):
    """The sub-cell resolved slab (``subcell=True``): its R_PP equals the
    oracle's ordinary periodic slab of the expanded medium (Rayleigh T0 per
    sub-cell, direct solve), embedded between the half-spaces the same way."""
    msc, rs = _mods()
    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=2, N_z=1, a=1.0)
    mat = msc.random_iid_slab_material(g, ref, 0.7, 0.3, seed=3)
    om, p = 1.4 * VS, 1.0e-4
    out = rs.resonance_ocean_bottom(g, mat, [om], p=p, n_sub=2, subcell=True, gmres_tol=1e-12,
                                    precision="fp64")
    assert out.subcell_converged == [True] and out.subcell_iterations[0] > 0
    gf, mf = rs.subcell_slab(g, mat, 2)
    oref = otm.Medium(VP, VS, RHO)
    eta = ok.vslow(ok.slowness(VP), p)
    k_hat = np.array([float(np.real(eta * VP)), p * VP, 0.0])
    T = osl.slab_tmatrices(4, 2, 0.5, mf.Dlambda, mf.Dmu, mf.Drho, om, oref)
    kh = osl.slab_kernels(4, 2, 0.5, om, oref, periodic=True)
    A = np.stack([osl.slab_matvec(e, T, kh, 4, 2, periodic=True)
                  for e in np.eye(4 * 4 * 2 * 9, dtype=complex)], axis=1)
    psi0 = osl.incident_field(4, 2, 0.5, om, oref, k_hat, "P").ravel()
    psi = np.linalg.solve(A, psi0).reshape(2, 4, 4, 9)
    R_or = osl.rpp_periodic(psi, T, 4, 2, 0.5, om, oref, p)
    assert abs(out.R_slab_subcell[0] - R_or) < 1e-8 * abs(R_or)
    _, R_tot, *_ = ok.ocean_bottom((WATER[0], WATER[1], 1.0), (VP, VS, RHO), HS, 2.0, p,
                                   np.array([om]), 0.0, np.ones(1, bool), np.array([R_or]), False,
                                   2, 1.0, 10.0)
    assert abs(out.R_total_subcell[0] - R_tot[0]) < 1e-8 * abs(R_tot[0])
