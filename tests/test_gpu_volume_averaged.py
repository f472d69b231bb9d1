"""GPU parity of the volume-averaged nearest-neighbour propagator.

``volume_averaged=True`` replaces the point propagator at the 26 nearest
neighbours by the inter-voxel (volume-averaged) block (slab_scattering.py:
316-329).  The HIP path keeps the spectral table of the point propagator and
adds the 26-block difference as a stencil after the FFT passes; these tests
compare the whole operator and the solve with the oracle, whose kernels put
the volume-averaged blocks into the spatial kernel before the FFT, as the
reference does.  Tolerances: complex128 operator 1e-11 relative on the
scattered part, complex64 operator 2e-5, FP64 solve 1e-10.
"""

import numpy as np
import pytest
import torch

from oracle import slab as osl
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

REF_C = otm.Medium(3000.0, 1732.0, 2500.0)


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _iid(M, Nz, ref, phi, c, seed):
    rng = np.random.default_rng(seed)
    shape = (Nz, M, M)
    mask = rng.random(shape) < phi
    lam = ref.rho * (ref.alpha**2 - 2 * ref.beta**2)
    mu = ref.rho * ref.beta**2
    Dl = np.where(mask, rng.uniform(-c, c, shape) * lam, 0.0)
    Dm = np.where(mask, rng.uniform(-c, c, shape) * mu, 0.0)
    Dr = np.where(mask, rng.uniform(-c, c, shape) * ref.rho, 0.0)
    return Dl, Dm, Dr


def _objects(M, Nz, a, Dl, Dm, Dr):
    import multiplescatteringcalculations_amd as msc

    g = msc.SlabGeometry(M=M, N_z=Nz, a=a)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = msc.SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=r)
    return msc, g, r, mat


@pytest.mark.parametrize("M,Nz,periodic,n_orders,a", [
    (1, 3, False, 2, 1.0),
    (3, 2, False, 2, 1.0),
    (5, 3, False, 1, 1.0),
    (4, 3, False, 3, 0.5),
    (2, 2, True, 2, 1.0),
    (4, 2, True, 2, 1.0),
    (8, 3, True, 0, 1.0),
])
def test_va_matvec_matches_oracle(hip_device, M, Nz, periodic, n_orders, a):
    from multiplescatteringcalculations_amd import slab_scattering as ss

    om = 0.3 * REF_C.beta / a
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.6, 0.3, seed=100 * M + Nz)
    msc, g, r, mat = _objects(M, Nz, a, Dl, Dm, Dr)
    plan = ss.SlabPlan(g, periodic=periodic)
    plan.build_g0(om, r, volume_averaged=True, n_orders=n_orders)
    plan.set_material(mat, om)
    T = osl.slab_tmatrices(M, Nz, a, Dl, Dm, Dr, om, REF_C)
    kh = osl.slab_kernels(M, Nz, a, om, REF_C, periodic=periodic, volume_averaged=True,
                          n_orders=n_orders)
    kp = osl.slab_kernels(M, Nz, a, om, REF_C, periodic=periodic)
    rng = np.random.default_rng(17)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz, periodic=periodic)
    y_pt = osl.slab_matvec(x, T, kp, M, Nz, periodic=periodic)
    # the correction is a visible change of the operator, not round-off
    assert _rel(x - y_pt, x - y_ref) > 1e-6
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y, x - y_ref) < 1e-11
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    assert _rel(x - y64, x - y_ref) < 2e-5
    # the fused residual b - A x carries the same correction
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r_dev = plan.residual(torch.as_tensor(b).to(hip_device),
                          torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(r_dev, b - y_ref) < 1e-11
    # the point table comes back when the same plan is rebuilt without averaging
    plan.build_g0(om, r)
    y2 = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y2, x - y_pt) < 1e-11


@pytest.mark.parametrize("periodic", [False, True])
def test_va_solve_matches_oracle(hip_device, periodic):
    M, Nz, a = 8, 3, 1.0
    om = 0.4 * REF_C.beta / a
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=11)
    msc, g, r, mat = _objects(M, Nz, a, Dl, Dm, Dr)
    kh = [0.893, 0.45, 0.0]
    kh = np.asarray(kh) / np.linalg.norm(kh)
    x_o, psi0_o, T_o, _, res_o = osl.solve(M, Nz, a, Dl, Dm, Dr, om, REF_C, kh, "P", tol=1e-13,
                                           max_iter=400, periodic=periodic, volume_averaged=True)
    assert res_o < 1e-12
    res = msc.compute_slab_scattering_gpu(g, mat, om, kh, "P", gmres_tol=1e-13, max_iter=400,
                                          restart=30, periodic=periodic, volume_averaged=True)
    assert res.gmres_residual < 1e-13
    assert _rel(res.psi, x_o) < 1e-10
    mixed = msc.compute_slab_scattering(g, mat, om, kh, "P", gmres_tol=1e-10, periodic=periodic,
                                        volume_averaged=True, precision="mixed")
    assert mixed.gmres_residual < 1e-10
    assert _rel(mixed.psi, x_o) < 1e-8
