"""The volume-averaged slab kernels of the oracle (slab.slab_kernels with
volume_averaged=True, following slab_scattering.py:316-329) equal the point
kernels plus a 26-offset stencil of P_VA(R) - P_point(R), aperiodic and
periodic (lateral images |d| <= M - 1).  The HIP path is built on this
decomposition; here it is checked on the oracle by a direct voxel loop."""

import numpy as np
import pytest

from oracle import intervoxel as iv
from oracle import propagator as prop
from oracle import slab as osl
from oracle import tmatrix as otm

REF = otm.Medium(3000.0, 1732.0, 2500.0)


@pytest.mark.parametrize("M,Nz,periodic,n_orders", [
    (1, 3, False, 2), (3, 2, False, 2), (4, 3, False, 3), (2, 2, True, 2), (4, 2, True, 1)])
def test_va_kernels_are_point_kernels_plus_stencil(M, Nz, periodic, n_orders):
    a = 1.0
    om = 0.3 * REF.beta / a
    rng = np.random.default_rng(M * 7 + Nz)
    sh = (Nz, M, M)
    Dl = rng.uniform(-0.3, 0.3, sh) * REF.rho * (REF.alpha ** 2 - 2 * REF.beta ** 2)
    Dm = rng.uniform(-0.3, 0.3, sh) * REF.rho * REF.beta ** 2
    Dr = rng.uniform(-0.3, 0.3, sh) * REF.rho
    T = osl.slab_tmatrices(M, Nz, a, Dl, Dm, Dr, om, REF)
    kv = osl.slab_kernels(M, Nz, a, om, REF, periodic=periodic, volume_averaged=True,
                          n_orders=n_orders)
    kp = osl.slab_kernels(M, Nz, a, om, REF, periodic=periodic)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    yv = osl.slab_matvec(x, T, kv, M, Nz, periodic=periodic)
    yp = osl.slab_matvec(x, T, kp, M, Nz, periodic=periodic)
    tau = np.einsum("lmnab,lmnb->lmna", T, x.reshape(Nz, M, M, 9))
    acc = np.zeros_like(tau)
    d = 2 * a
    for dz in (-1, 0, 1):
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                if dz == dx == dy == 0:
                    continue
                D = (iv.propagator_9x9((dz, dx, dy), REF.alpha, REF.beta, REF.rho, om, n_orders)
                     - prop.propagator_9x9([dz * d, dx * d, dy * d], om, REF.alpha, REF.beta,
                                           REF.rho))
                for z in range(Nz):
                    for X in range(M):
                        for Y in range(M):
                            uz, ux, uy = z - dz, X - dx, Y - dy
                            if not 0 <= uz < Nz:
                                continue
                            if periodic:
                                if M == 1 and (dx or dy):
                                    continue
                                ux, uy = ux % M, uy % M
                            elif not (0 <= ux < M and 0 <= uy < M):
                                continue
                            acc[z, X, Y] += D @ tau[uz, ux, uy]
    scale = np.abs(x - yv).max()
    assert np.abs(yp.reshape(acc.shape) - acc - yv.reshape(acc.shape)).max() < 1e-13 * scale
    assert np.abs(yp - yv).max() > 1e-3 * scale  # the correction is not negligible
