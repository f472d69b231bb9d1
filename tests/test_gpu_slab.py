"""GPU parity tests: the HIP path (through the C ABI) against the complex128 oracle.

Tolerances: complex128 operator and solve 1e-10 relative (written per test);
complex64 operator 1e-5; mixed-precision solve 1e-6 relative to the FP64 one,
with the same FP64 residual target.
"""

import numpy as np
import pytest
import torch

from oracle import slab as osl
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

REF_T = otm.Medium(5000.0, 3000.0, 2500.0)
OMEGA_T = 150.0  # ka = 0.05 with a = 1 (reference test constants)
REF_C = otm.Medium(3000.0, 1732.0, 2500.0)


def _mods():
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import slab_scattering as ss

    return msc, ss


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


def _plan(M, Nz, a, ref, omega, Dl, Dm, Dr, periodic=False, pad=None):
    msc, ss = _mods()
    g = msc.SlabGeometry(M=M, N_z=Nz, a=a)
    r = msc.ReferenceMedium(ref.alpha, ref.beta, ref.rho)
    mat = msc.SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=r)
    plan = ss.SlabPlan(g, periodic=periodic, pad=pad)
    plan.build_g0(omega, r)
    plan.set_material(mat, omega)
    return plan, g, r, mat


@pytest.mark.parametrize("prec", ["c128", "c64"])
def test_fft_axis_matches_numpy(hip_device, prec):
    from multiplescatteringcalculations_amd import _lib

    lib = _lib.load()
    dt = torch.complex128 if prec == "c128" else torch.complex64
    tol = 1e-13 if prec == "c128" else 3e-6
    rng = np.random.default_rng(1)
    for n in [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024]:
        for outer, inner in [(3, 1), (2, 9), (1, 45)]:
            x = rng.standard_normal((outer, n, inner)) + 1j * rng.standard_normal((outer, n, inner))
            t = torch.as_tensor(x).to(hip_device, dt)
            out = torch.empty_like(t)
            for inv in (0, 1):
                _lib.check(lib.msc_fft_axis(_lib.ptr(t), _lib.ptr(out), outer, n, inner, inv,
                                            0 if prec == "c128" else 1, _lib.stream_ptr()))
                ref = np.fft.fft(x, axis=1) if not inv else np.fft.ifft(x, axis=1) * n
                assert _rel(out.cpu().numpy(), ref) < tol, (n, outer, inner, inv)


def test_tparams_match_oracle(hip_device):
    M, Nz = 4, 3
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.6, 0.4, seed=3)
    plan, g, r, mat = _plan(M, Nz, 1.0, REF_C, 346.4, Dl, Dm, Dr)
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, 346.4, REF_C)
    tp = plan.tparams().cpu().numpy()
    assert _rel(tp, osl.tparams_from_T(T)) < 1e-12


def test_cube_tmatrix_host_matches_oracle():
    msc, _ = _mods()
    r = msc.ReferenceMedium(5.0, 3.0, 2.5)
    res = msc.compute_cube_tmatrix(2 * np.pi * 10, 0.01, r, msc.MaterialContrast(1.25, 0.75, 0.3))
    o = otm.cube_tmatrix(2 * np.pi * 10, 0.01, otm.Medium(5.0, 3.0, 2.5), 1.25, 0.75, 0.3)
    for k in ("Gamma0", "Ac", "Bc", "Cc", "T1c", "T2c", "T3c", "Drho_star", "Dlambda_star",
              "Dmu_star_off", "Dmu_star_diag"):
        assert abs(getattr(res, k) - getattr(o, k)) <= 1e-12 * abs(getattr(o, k)), k


@pytest.mark.parametrize("wave,khat", [("P", [1.0, 0.0, 0.0]), ("P", [0.893, 0.45, 0.0]),
                                       ("S", [0.8, 0.36, 0.48]), ("S", [1.0, 0.0, 0.0])])
def test_incident_matches_oracle(hip_device, wave, khat):
    _, ss = _mods()
    M, Nz = 5, 3
    g = ss.SlabGeometry(M=M, N_z=Nz, a=1.0)
    plan = ss.SlabPlan(g)
    r = ss.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    psi0 = plan.incident_field(692.8, r, khat, wave).cpu().numpy()
    ref = osl.incident_field(M, Nz, 1.0, 692.8, REF_C, khat, wave)
    assert _rel(psi0, ref) < 1e-13


@pytest.mark.parametrize("M,Nz,contrast,pad", [
    (1, 1, "iid", None),  # FFT grid 2 x 2 x 2 (table build's degenerate octant)
    (1, 3, "iid", None),
    (4, 1, "iid", None),
    (2, 2, "uniform", None),
    (3, 2, "iid", None),
    (3, 2, "iid", (8, 16)),
    (5, 3, "iid", None),
    (13, 5, "iid", None),
])
def test_matvec_matches_oracle(hip_device, M, Nz, contrast, pad):
    if contrast == "uniform":
        Dl = np.full((Nz, M, M), 2e9)
        Dm = np.full((Nz, M, M), 1e9)
        Dr = np.full((Nz, M, M), 100.0)
        ref, om = REF_T, OMEGA_T
    else:
        ref, om = REF_C, 692.8
        Dl, Dm, Dr = _iid(M, Nz, ref, 0.5, 0.4, seed=M * 10 + Nz)
    plan, *_ = _plan(M, Nz, 1.0, ref, om, Dl, Dm, Dr, pad=pad)
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, ref)
    kh = osl.slab_kernels(M, Nz, 1.0, om, ref)
    rng = np.random.default_rng(123)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    # compare the scattered part (y - x) too, which carries all of G0 T0
    assert _rel(y, y_ref) < 1e-12
    assert _rel(x - y, x - y_ref) < 1e-11
    if M * M * Nz <= 18:
        y_dir = osl.direct_matvec(x, T, M, Nz, 1.0, om, ref)
        assert _rel(x - y, x - y_dir) < 1e-11
    # complex64 operator
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    # relative to the scattered part, plus the input's complex64 rounding, which
    # enters x - y64 directly (and is all of it where a single voxel has no
    # coupling)
    tol = 2e-5 * np.abs(x - y_ref).max() + 1e-7 * np.abs(x).max()
    assert np.abs((x - y64) - (x - y_ref)).max() < tol


@pytest.mark.parametrize("M,Nz,pad,lengths", [
    (260, 1, None, (2, 1024)),  # 2M - 1 = 519: the 1024-point lateral transforms
    (2, 130, None, (512, 4)),  # 2N_z - 1 = 259: the 512-point z transforms
    (3, 2, (512, 8), (512, 8)),
    (3, 2, (4, 1024), (4, 1024)),
])
def test_matvec_matches_oracle_at_the_longest_transforms(hip_device, M, Nz, pad, lengths):
    """The largest plans: lateral FFT length 1024 (M <= 512) and vertical 512
    (N_z <= 256), on natural and on explicitly padded grids."""
    ref, om = REF_C, 692.8
    Dl, Dm, Dr = _iid(M, Nz, ref, 0.5, 0.4, seed=M + Nz)
    plan, *_ = _plan(M, Nz, 1.0, ref, om, Dl, Dm, Dr, pad=pad)
    assert (plan.Pz, plan.Pxy) == lengths
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, ref)
    kh = osl.slab_kernels(M, Nz, 1.0, om, ref)
    rng = np.random.default_rng(11)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y, x - y_ref) < 1e-11
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    tol = 2e-5 * np.abs(x - y_ref).max() + 1e-7 * np.abs(x).max()
    assert np.abs((x - y64) - (x - y_ref)).max() < tol


@pytest.mark.parametrize("M,Nz", [(2, 1), (2, 2), (4, 2), (8, 3), (3, 2), (5, 1), (6, 3),
                                  (10, 2), (12, 3)])
def test_matvec_periodic_matches_oracle(hip_device, M, Nz):
    """Circular lateral convolution: on the M-point grid for M a power of two,
    on the zero-padded grid with the folded kernel otherwise (the reference's
    fold accepts any M, slab_scattering.py:337-346)."""
    ref, om = REF_C, 692.8
    Dl, Dm, Dr = _iid(M, Nz, ref, 0.5, 0.3, seed=7)
    plan, *_ = _plan(M, Nz, 1.0, ref, om, Dl, Dm, Dr, periodic=True)
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, ref)
    kh = osl.slab_kernels(M, Nz, 1.0, om, ref, periodic=True)
    rng = np.random.default_rng(5)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz, periodic=True)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert plan.Pxy == (M if M & (M - 1) == 0 else 1 << (2 * M - 2).bit_length())
    assert _rel(x - y, x - y_ref) < 1e-11
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    tol = 2e-5 * np.abs(x - y_ref).max() + 1e-7 * np.abs(x).max()
    assert np.abs((x - y64) - (x - y_ref)).max() < tol


def test_matvec_identity_when_T_zero(hip_device):
    M, Nz = 6, 3
    z = np.zeros((Nz, M, M))
    plan, *_ = _plan(M, Nz, 1.0, REF_T, OMEGA_T, z, z, z)
    x = torch.randn(Nz * M * M * 9, dtype=torch.complex128, device=hip_device)
    assert torch.equal(plan.matvec(x), x)


def test_residual_is_fused_b_minus_Ax(hip_device):
    M, Nz = 7, 4
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=11)
    plan, *_ = _plan(M, Nz, 1.0, REF_C, 692.8, Dl, Dm, Dr)
    n = Nz * M * M * 9
    x = torch.randn(n, dtype=torch.complex128, device=hip_device)
    b = torch.randn(n, dtype=torch.complex128, device=hip_device)
    r = plan.residual(b, x)
    assert torch.allclose(r, b - plan.matvec(x), rtol=0, atol=1e-13)


def test_gmres_matches_dense_solve(hip_device):
    M, Nz = 3, 2
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.7, 0.4, seed=21)
    plan, g, r, mat = _plan(M, Nz, 1.0, REF_C, 692.8, Dl, Dm, Dr)
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, 692.8, REF_C)
    A = osl.dense_operator(T, M, Nz, 1.0, 692.8, REF_C)
    psi0 = osl.incident_field(M, Nz, 1.0, 692.8, REF_C, [0.893, 0.45, 0.0], "P").ravel()
    x_ref = np.linalg.solve(A, psi0)
    for restart in (None, 5):
        x, st = plan.gmres(torch.as_tensor(psi0), "born", tol=1e-13, max_iter=200, restart=restart)
        assert st["converged"] and st["rel_residual"] < 1e-13
        assert _rel(x.cpu().numpy(), x_ref) < 1e-10


def test_solve_matches_oracle_and_outputs(hip_device):
    """End-to-end: psi, R_PP (periodic formula) and far-field R vs the oracle."""
    msc, ss = _mods()
    M, Nz = 8, 4
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=0)
    om = 0.4 * REF_C.beta / 1.0
    p = 1.5e-4
    kh = [np.sqrt(1 - (p * REF_C.alpha) ** 2), p * REF_C.alpha, 0.0]
    x_o, psi0_o, T_o, _, res_o = osl.solve(M, Nz, 1.0, Dl, Dm, Dr, om, REF_C, kh, "P", tol=1e-13,
                                           max_iter=400)
    assert res_o < 1e-12
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = msc.SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=r)
    res = msc.compute_slab_scattering_gpu(g, mat, om, np.array(kh), "P", gmres_tol=1e-13,
                                          max_iter=400, restart=30)
    assert res.gmres_residual < 1e-13
    assert _rel(res.psi0, psi0_o) < 1e-13
    assert _rel(res.psi, x_o) < 1e-10
    rpp = msc.slab_rpp_periodic(res, p=p)
    rpp_o = osl.rpp_periodic(x_o, T_o, M, Nz, 1.0, om, REF_C, p=p)
    assert abs(rpp - rpp_o) <= 1e-10 * abs(rpp_o)
    ff = msc.slab_reflected_field(res)
    ff_o = osl.reflected_field(x_o, T_o, M, Nz, 1.0, om, REF_C, kh)
    for a, b in zip(ff, ff_o):
        assert abs(a - b) <= 1e-10 * max(abs(b), 1e-300)
    # explicit T_local path of the outputs (reference signature)
    T_loc = msc.compute_slab_tmatrices(g, mat, om)
    assert _rel(T_loc, T_o) < 1e-12
    assert abs(msc.slab_rpp_periodic(res, T_loc, p=p) - rpp_o) <= 1e-10 * abs(rpp_o)


def test_mixed_precision_solve(hip_device):
    msc, ss = _mods()
    M, Nz = 16, 8
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=4)
    om = 0.4 * REF_C.beta
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = msc.SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=r)
    kh = np.array([0.893, 0.45, 0.0])
    r64 = msc.compute_slab_scattering_gpu(g, mat, om, kh, gmres_tol=1e-12, restart=30)
    rmx = msc.compute_slab_scattering_gpu(g, mat, om, kh, gmres_tol=1e-8, restart=30,
                                          precision="mixed")
    assert rmx.gmres_residual <= 1e-8 and r64.gmres_residual <= 1e-12
    assert rmx.stats["matvecs_c64"] > 0
    assert _rel(rmx.psi, r64.psi) < 1e-6
    assert abs(msc.slab_rpp_periodic(rmx) - msc.slab_rpp_periodic(r64)) <= 1e-6 * abs(
        msc.slab_rpp_periodic(r64))


def test_zero_contrast_gives_incident_field(hip_device):
    msc, _ = _mods()
    g = msc.SlabGeometry(M=4, N_z=2, a=1.0)
    r = msc.ReferenceMedium(5000.0, 3000.0, 2500.0)
    mat = msc.uniform_slab_material(g, r, msc.MaterialContrast(0.0, 0.0, 0.0))
    res = msc.compute_slab_scattering(g, mat, OMEGA_T, np.array([1.0, 0.0, 0.0]))
    assert np.allclose(res.psi, res.psi0, atol=1e-12)
    assert abs(msc.slab_rpp_periodic(res)) < 1e-20
    R_PP, R_PS, _ = msc.slab_reflected_field(res)
    assert abs(R_PP) < 1e-20 and abs(R_PS) < 1e-20


def test_nonconvergence_warns(hip_device):
    msc, _ = _mods()
    g = msc.SlabGeometry(M=4, N_z=2, a=1.0)
    r = msc.ReferenceMedium(5000.0, 3000.0, 2500.0)
    mat = msc.uniform_slab_material(g, r, msc.MaterialContrast(2e9, 1e9, 100.0))
    with pytest.warns(UserWarning, match="did not converge"):
        res = msc.compute_slab_scattering_gpu(g, mat, OMEGA_T, np.array([1.0, 0.0, 0.0]),
                                              gmres_tol=1e-15, max_iter=1)
    assert res.n_gmres_iter == 1 and res.gmres_residual > 1e-15


def test_torch_gmres_dense(hip_device):
    msc, _ = _mods()
    rng = np.random.default_rng(42)
    n = 20
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    A = A @ A.conj().T + 5 * np.eye(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    At = torch.as_tensor(A).to(hip_device)
    bt = torch.as_tensor(b).to(hip_device)
    x, it, rr = msc.torch_gmres(lambda v: At @ v, bt, tol=1e-12, maxiter=50)
    assert rr < 1e-12
    assert _rel(x.cpu().numpy(), np.linalg.solve(A, b)) < 1e-9
    d = torch.tensor([1.0, 2.0, 3.0], dtype=torch.complex128, device=hip_device)
    bb = torch.tensor([1.0, 4.0, 9.0], dtype=torch.complex128, device=hip_device)
    _, it0, rr0 = msc.torch_gmres(lambda v: d * v, bb, x0=bb / d, tol=1e-10)
    assert it0 == 0 and rr0 < 1e-10


def test_full_size_operator_properties(hip_device):
    """128 x 128 x 32 slab (Pxy 256, Pz 64): linearity, c64 vs c128, T=0 identity."""
    _, ss = _mods()
    M, Nz = 128, 32
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=0)
    plan, *_ = _plan(M, Nz, 1.0, REF_C, 0.4 * REF_C.beta, Dl, Dm, Dr)
    assert (plan.Pz, plan.Pxy) == (64, 256)
    n = plan.n
    gen = torch.Generator(device=hip_device).manual_seed(0)
    x = torch.randn(n, dtype=torch.complex128, device=hip_device, generator=gen)
    y = torch.randn(n, dtype=torch.complex128, device=hip_device, generator=gen)
    a, b = 0.7 - 0.2j, -1.3 + 0.5j
    lhs = plan.matvec(a * x + b * y)
    rhs = a * plan.matvec(x) + b * plan.matvec(y)
    assert float((lhs - rhs).abs().max() / rhs.abs().max()) < 1e-13
    y128 = plan.matvec(x)
    y64 = plan.matvec(x.to(torch.complex64)).to(torch.complex128)
    assert float(((x - y64) - (x - y128)).norm() / (x - y128).norm()) < 1e-5


@pytest.mark.parametrize("M,Nz,pad", [(3, 2, (1024, 8)), (3, 2, (4, 2048)), (513, 1, None),
                                      (2, 257, None)])
def test_transform_lengths_past_the_limits_are_refused(hip_device, M, Nz, pad):
    """Vertical lengths above 512 and lateral above 1024 (N_z > 256, M > 512)
    raise ValueError before any allocation."""
    msc, ss = _mods()
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    with pytest.raises(ValueError, match="power of two"):
        ss.SlabPlan(g, pad=pad)


def test_largest_lateral_slab_solves(hip_device):
    """M = 512 (lateral FFT length 1024), N_z = 2, through the public solve:
    it converges, and the FP64 residual of its solution recomputed by a fresh
    plan's operator (pinned to the oracle at this length above) meets the
    tolerance."""
    msc, ss = _mods()
    M, Nz = 512, 2
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=4)
    om = 0.4 * REF_C.beta
    kh = [1.0, 0.0, 0.0]
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = msc.SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=r)
    res = msc.compute_slab_scattering_gpu(g, mat, om, np.array(kh), "P", gmres_tol=1e-8,
                                          max_iter=300, restart=30)
    assert res.gmres_residual <= 1e-8
    ss.release_plans()
    plan, *_ = _plan(M, Nz, 1.0, REF_C, om, Dl, Dm, Dr)
    assert (plan.Pz, plan.Pxy) == (4, 1024)
    b = plan.incident_field(om, r, kh, "P").reshape(-1)
    x = torch.as_tensor(np.ascontiguousarray(res.psi).reshape(-1)).to(hip_device)
    rr = plan.residual(b, x)
    assert float(torch.linalg.vector_norm(rr) / torch.linalg.vector_norm(b)) <= 1.01e-8


ZDMA = 256  # z tile code flag of the persistent direct-to-LDS z pass
ZPRUNE = 512  # z tile code flag of the pruned, HBM-direct z pass


@pytest.mark.parametrize("form", [ZDMA, ZPRUNE])
@pytest.mark.parametrize("q", [2, 4, 8])
@pytest.mark.parametrize("M,Nz,periodic", [(9, 4, False), (7, 3, False), (8, 5, True),
                                           (16, 1, False), (5, 2, False), (5, 17, False),
                                           (4, 32, True), (3, 20, False), (2, 130, False)])
def test_persistent_z_pass_matches_oracle(hip_device, form, q, M, Nz, periodic):
    """The persistent z pass (tile code Q | 256: the next tile's rows prefetched
    into LDS by direct-to-LDS loads) and the pruned one (Q | 512: the first
    forward pass reads the non-zero rows from HBM, the last inverse pass writes
    the kept rows back), both with z FFTs pruned to the Nz input and output
    rows, give the oracle's operator, with Nz at and below half the padded z
    length, Pz from 2 to 512 (radix 2, 4 and 8 first passes), periodic and
    aperiodic."""
    ref, om = REF_C, 692.8
    Dl, Dm, Dr = _iid(M, Nz, ref, 0.5, 0.3, seed=M + 31 * Nz)
    plan, *_ = _plan(M, Nz, 1.0, ref, om, Dl, Dm, Dr, periodic=periodic)
    tiles = (1, 2, q | form, 2, 1)
    plan.set_tiles(tiles, tiles)
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, ref)
    kh = osl.slab_kernels(M, Nz, 1.0, om, ref, periodic=periodic)
    rng = np.random.default_rng(q + M)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz, periodic=periodic)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y, x - y_ref) < 1e-11
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    tol = 2e-5 * np.abs(x - y_ref).max() + 1e-7 * np.abs(x).max()
    assert np.abs((x - y64) - (x - y_ref)).max() < tol


def test_persistent_z_pass_full_size(hip_device):
    """At 128 x 128 x 32 (16384 tiles of Q = 4, several per workgroup) the
    persistent z pass equals the one-tile-per-workgroup pass, in complex128 to
    rounding and in complex64 to its precision, and a mixed solve with it
    reaches the same FP64 solution."""
    M, Nz = 128, 32
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=0)
    plan, g, r, mat = _plan(M, Nz, 1.0, REF_C, 0.4 * REF_C.beta, Dl, Dm, Dr)
    gen = torch.Generator(device=hip_device).manual_seed(3)
    x = torch.randn(plan.n, dtype=torch.complex128, device=hip_device, generator=gen)
    out = {}
    codes = (4 | ZDMA, 2 | ZDMA, 8 | ZDMA, 2 | ZPRUNE, 4 | ZPRUNE, 8 | ZPRUNE)
    for code in (4,) + codes:
        plan.set_tiles((1, 1, code, 1, 1), (1, 2, code, 2, 1))
        out[code] = (plan.matvec(x), plan.matvec(x.to(torch.complex64)).to(torch.complex128))
    base = x - out[4][0]
    for code in codes:
        assert float(((x - out[code][0]) - base).abs().max() / base.abs().max()) < 1e-13
        assert float(((x - out[code][1]) - base).norm() / base.norm()) < 1e-5
    b = plan.incident_field(0.4 * REF_C.beta, r, [0.893, 0.45, 0.0], "P").reshape(-1)
    plan.set_tiles((1, 1, 4, 1, 1), (1, 2, 4, 2, 1))
    x0, st0 = plan.gmres(b, "born", tol=1e-10, max_iter=200, restart=30, precision="mixed")
    plan.set_tiles((1, 1, 4 | ZDMA, 1, 1), (1, 2, 4 | ZDMA, 2, 1))
    x1, st1 = plan.gmres(b, "born", tol=1e-10, max_iter=200, restart=30, precision="mixed")
    assert st0["converged"] and st1["converged"]
    assert float((x1 - x0).abs().max() / x0.abs().max()) < 1e-8


@pytest.mark.parametrize("tiles", [(1, 1, 1, 1, 1), (2, 2, 2, 2, 2), (1, 4, 4, 4, 1), (2, 1, 8, 2, 1),
                                   (1, 2, 16, 2, 1), (1, 8, 4, 16, 1), (2, 32, 8, 8, 2),
                                   (1, 16, 4 | 512, 32, 1)])
def test_every_tile_variant_matches_oracle(hip_device, tiles):
    """Each tile instantiation of the five passes gives the same operator."""
    M, Nz = 9, 4
    ref, om = REF_C, 692.8
    Dl, Dm, Dr = _iid(M, Nz, ref, 0.5, 0.3, seed=17)
    plan, *_ = _plan(M, Nz, 1.0, ref, om, Dl, Dm, Dr)
    plan.set_tiles(tiles, tiles)
    assert plan.tiles() == {"c128": tiles, "c64": tiles}
    T = osl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, ref)
    kh = osl.slab_kernels(M, Nz, 1.0, om, ref)
    rng = np.random.default_rng(8)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y, x - y_ref) < 1e-11
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    # relative to the scattered part, plus the input's complex64 rounding, which
    # enters x - y64 directly (and is all of it where a single voxel has no
    # coupling)
    tol = 2e-5 * np.abs(x - y_ref).max() + 1e-7 * np.abs(x).max()
    assert np.abs((x - y64) - (x - y_ref)).max() < tol


def test_x_pass_line_tiles_match_column_tiles_on_the_256_point_grid(hip_device):
    """At the 256-point lateral grid of the 128 x 128 slab, x passes over
    whole-cache-line tiles (8, 16, 32 complex of a row) give the operator of
    the ky-column tiles to rounding in both precisions (each instantiation is
    compiled on its own, so the floating-point contraction may differ)."""
    M, Nz = 128, 3
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=23)
    plan, *_ = _plan(M, Nz, 1.0, REF_C, 0.4 * REF_C.beta, Dl, Dm, Dr)
    assert plan.Pxy == 256
    rng = np.random.default_rng(4)
    n = Nz * M * M * 9
    x = torch.as_tensor(rng.standard_normal(n) + 1j * rng.standard_normal(n)).to(hip_device)
    for dt, tol in ((torch.complex128, 1e-14), (torch.complex64, 2e-6)):
        outs = []
        for q in (1, 2, 4, 8, 16, 32):
            t = (1, q, 8 | 512, q, 1)
            plan.set_tiles(t, t)
            outs.append(plan.matvec(x.to(dt)).cpu())
        # the scattered part x - y carries the passes' rounding
        ref = (x.to(dt).cpu() - outs[0])
        for q, o in zip((2, 4, 8, 16, 32), outs[1:]):
            rel = float((o - outs[0]).abs().max() / ref.abs().max())
            print(f"{dt} x tile {q}: max diff vs tile 1 = {rel:.2e} of the scattered part")
            assert rel <= tol


def test_autotune_and_graph_solves_are_reproducible(hip_device):
    """Autotuned tiles stay valid, and repeated solves give bit-identical
    results with direct launches and with graph-replayed steps (deterministic
    reductions, replay-safe graphs)."""
    msc, ss = _mods()
    M, Nz = 16, 6
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.3, 0.2, seed=5)
    plan, g, r, mat = _plan(M, Nz, 1.0, REF_C, 0.4 * REF_C.beta, Dl, Dm, Dr)
    plan.autotune("c128")
    plan.autotune("c64")
    t = plan.tiles()
    for tiles in (t["c128"], t["c64"]):
        assert tiles[0] in (1, 2) and tiles[4] in (1, 2)
        assert tiles[1] in (1, 2, 4, 8, 16, 32) and tiles[3] in (1, 2, 4, 8, 16, 32)
        assert tiles[2] in (1, 2, 4, 8, 16, 2 | ZDMA, 4 | ZDMA, 8 | ZDMA, 2 | ZPRUNE, 4 | ZPRUNE,
                            8 | ZPRUNE)
    psi0 = plan.incident_field(0.4 * REF_C.beta, r, [0.893, 0.45, 0.0], "P")
    outs = []
    for prec in ("fp64", "mixed", "fp64", "mixed"):
        x, st = plan.gmres(psi0, "born", tol=1e-10, max_iter=200, restart=30, precision=prec)
        assert st["converged"], st
        outs.append(x.cpu().numpy())
    assert np.array_equal(outs[0], outs[2]) and np.array_equal(outs[1], outs[3])
    assert _rel(outs[1], outs[0]) < 1e-6
    # the same steps as instantiated HIP graphs (captured on the first solve,
    # replayed on the second) give bit-identical solves
    from multiplescatteringcalculations_amd import _lib

    _lib.check(plan.lib.msc_plan_use_graphs(plan._h, 1))
    for prec, ref_x in (("fp64", outs[0]), ("mixed", outs[1]), ("fp64", outs[0]), ("mixed", outs[1])):
        x, st = plan.gmres(psi0, "born", tol=1e-10, max_iter=200, restart=30, precision=prec)
        assert np.array_equal(x.cpu().numpy(), ref_x)
    _lib.check(plan.lib.msc_plan_use_graphs(plan._h, 0))


@pytest.mark.parametrize("va", [False, True])
def test_copied_g0_gives_the_same_operator(hip_device, va):
    """msc_plan_copy_g0: a plan whose G0 (and near-neighbour correction and
    tiles) is copied from another plan applies exactly the same operator."""
    msc, ss = _mods()
    M, Nz = 9, 4
    Dl, Dm, Dr = _iid(M, Nz, REF_C, 0.4, 0.3, seed=11)
    om = 692.8
    g = ss.SlabGeometry(M=M, N_z=Nz, a=1.0)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = ss.SlabMaterial(Dlambda=Dl, Dmu=Dm, Drho=Dr, ref=r)
    src = ss.SlabPlan(g, device=hip_device)
    src.build_g0(om, r, volume_averaged=va, n_orders=2)
    src.set_material(mat, om)
    src.set_tiles((1, 2, 4, 2, 1), (2, 1, 2, 1, 2))
    dst = ss.SlabPlan(g, device=hip_device)
    dst.set_material(mat, om)
    dst.copy_g0_from(src)
    assert dst.tiles() == src.tiles()
    rng = np.random.default_rng(7)
    x = torch.as_tensor(rng.standard_normal(dst.n) + 1j * rng.standard_normal(dst.n)).to(hip_device)
    assert torch.equal(dst.matvec(x), src.matvec(x))
    assert torch.equal(dst.matvec(x.to(torch.complex64)), src.matvec(x.to(torch.complex64)))
    other = ss.SlabPlan(ss.SlabGeometry(M=M, N_z=Nz + 1, a=1.0), device=hip_device)
    with pytest.raises(ValueError, match="differ"):
        other.copy_g0_from(src)
    with pytest.raises(RuntimeError, match="no G0"):
        other.copy_g0_from(ss.SlabPlan(ss.SlabGeometry(M=M, N_z=Nz + 1, a=1.0), device=hip_device))


def test_reused_plan_solves_and_outputs(hip_device):
    """Repeated solves of one slab shape reuse a cached plan: each result equals
    the same solve on a fresh plan, and an earlier result's outputs stay those
    of its own material after a later solve has reused the plan."""
    msc, ss = _mods()
    M, Nz = 8, 3
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    om, kh = 692.8, np.array([0.893, 0.45, 0.0])
    mats = [msc.SlabMaterial(*_iid(M, Nz, REF_C, 0.5, 0.3, seed=s), ref=r) for s in (1, 2)]
    ss._PLAN_CACHE.clear()
    fresh = []
    for m in mats:
        ss._PLAN_CACHE.clear()
        res = msc.compute_slab_scattering_gpu(g, m, om, kh, "P", gmres_tol=1e-12, restart=20,
                                              periodic=True)
        fresh.append((res.psi.copy(), msc.slab_rpp_periodic(res, p=1e-4)))
    ss._PLAN_CACHE.clear()
    a = msc.compute_slab_scattering_gpu(g, mats[0], om, kh, "P", gmres_tol=1e-12, restart=20,
                                        periodic=True)
    b = msc.compute_slab_scattering_gpu(g, mats[1], om, kh, "P", gmres_tol=1e-12, restart=20,
                                        periodic=True)
    assert b._plan is a._plan  # the plan was reused
    # (a fresh plan tunes its own tiles: agreement to rounding, not bitwise)
    assert _rel(a.psi, fresh[0][0]) < 1e-12 and _rel(b.psi, fresh[1][0]) < 1e-12
    assert abs(msc.slab_rpp_periodic(b, p=1e-4) - fresh[1][1]) <= 1e-12 * abs(fresh[1][1])
    # a's plan now holds b's T0: a's outputs come from its own material
    assert abs(msc.slab_rpp_periodic(a, p=1e-4) - fresh[0][1]) <= 1e-12 * abs(fresh[0][1])
    # a different frequency on the cached plan rebuilds G0
    c = msc.compute_slab_scattering_gpu(g, mats[0], 1.3 * om, kh, "P", gmres_tol=1e-12,
                                        restart=20, periodic=True)
    ss._PLAN_CACHE.clear()
    d = msc.compute_slab_scattering_gpu(g, mats[0], 1.3 * om, kh, "P", gmres_tol=1e-12,
                                        restart=20, periodic=True)
    assert _rel(c.psi, d.psi) < 1e-12
    ss.release_plans()
    assert not ss._PLAN_CACHE


def test_volume_averaged_toggle_on_one_plan(hip_device):
    """Switching the near-neighbour correction on or off on a plan that has
    already solved gives the solve of a fresh plan (the captured Arnoldi steps
    are dropped when the correction switches)."""
    msc, ss = _mods()
    M, Nz = 8, 3
    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    r = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = msc.SlabMaterial(*_iid(M, Nz, REF_C, 0.5, 0.3, seed=4), ref=r)
    om = 692.8

    def solve(plan, va):
        plan.build_g0(om, r, volume_averaged=va, n_orders=2)
        plan.set_material(mat, om)
        psi0 = plan.incident_field(om, r, [0.893, 0.45, 0.0], "P")
        x, st = plan.gmres(psi0, "born", tol=1e-12, max_iter=200, restart=20)
        assert st["converged"]
        return x.cpu().numpy()

    shared = ss.SlabPlan(g, device=hip_device)
    seq = [solve(shared, va) for va in (False, True, False, True)]
    for va, x in zip((False, True, False, True), seq):
        assert _rel(x, solve(ss.SlabPlan(g, device=hip_device), va)) < 1e-10
    assert _rel(seq[1], seq[0]) > 1e-6  # the correction changes the field


def test_general_tmatrix_operator_matches_oracle(hip_device):
    """A general 9x9 T per voxel (msc_set_tgeneral): the operator, the fused
    residual, the complex64 operator and the layer / far-field source sums
    against the oracle's dense-T formulas, with random full 9x9 matrices
    shared through an index table (some voxels zero)."""
    M, Nz = 7, 3
    om = 692.8
    plan, *_ = _plan(M, Nz, 1.0, REF_C, om, *_iid(M, Nz, REF_C, 0.5, 0.3, seed=2))
    rng = np.random.default_rng(31)
    K = 5
    scale = np.abs(osl.tparams_from_T(osl.slab_tmatrices(M, Nz, 1.0, *_iid(M, Nz, REF_C, 1.0, 0.3,
                                                                           seed=3), om,
                                                         REF_C))).max()
    table = scale * (rng.standard_normal((K, 9, 9)) + 1j * rng.standard_normal((K, 9, 9)))
    table[0] = 0.0
    tidx = rng.integers(0, K, (Nz, M, M))
    plan.set_tgeneral(tidx, table)
    T = table[tidx]
    kh = osl.slab_kernels(M, Nz, 1.0, om, REF_C)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_ref = osl.slab_matvec(x, T, kh, M, Nz)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y, x - y_ref) < 1e-11
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = plan.residual(torch.as_tensor(b).to(hip_device), torch.as_tensor(x).to(hip_device))
    assert _rel(r.cpu().numpy(), b - y_ref) < 1e-11
    y64 = plan.matvec(torch.as_tensor(x).to(hip_device, torch.complex64)).cpu().numpy()
    tol = 2e-5 * np.abs(x - y_ref).max() + 1e-7 * np.abs(x).max()
    assert np.abs((x - y64) - (x - y_ref)).max() < tol
    # source sums: tau = T psi per layer, and the phased sums of the far field
    kP, kS, rh = om / REF_C.alpha, om / REF_C.beta, np.array([-0.8, -0.36, -0.48])
    ls, ph = plan.source_sums(torch.as_tensor(x).to(hip_device), kP, kS, rh)
    tau = np.einsum("lmnab,lmnb->lmna", T, x.reshape(Nz, M, M, 9))
    assert _rel(ls, tau.sum(axis=(1, 2))) < 1e-12
    c = osl.centres(M, Nz, 1.0)
    eP = np.exp(-1j * kP * np.einsum("k,lmnk->lmn", rh, c))
    assert _rel(ph[:, 0], np.einsum("lmn,lmna->la", eP, tau)) < 1e-12
    # an index outside the table is rejected before any device work reads it
    bad = tidx.copy()
    bad[1, 2, 3] = K
    with pytest.raises(ValueError, match="out of range"):
        plan.set_tgeneral(bad, table)
    # back to the cubic T0 of the material
    _, g, r_, mat = _plan(M, Nz, 1.0, REF_C, om, *_iid(M, Nz, REF_C, 0.5, 0.3, seed=2))
    plan.set_material(mat, om)
    T0 = osl.slab_tmatrices(M, Nz, 1.0, mat.Dlambda, mat.Dmu, mat.Drho, om, REF_C)
    y = plan.matvec(torch.as_tensor(x).to(hip_device)).cpu().numpy()
    assert _rel(x - y, x - osl.slab_matvec(x, T0, kh, M, Nz)) < 1e-11


def test_general_tmatrix_solve_and_network_input(hip_device):
    """GMRES with a general T converges to the dense solution; the learned
    preconditioner's input (s / t_max) T r uses the general T, with t_max the
    largest entry modulus of the table."""
    from multiplescatteringcalculations_amd import learned_precond as lp

    M, Nz = 4, 2
    om = 692.8
    plan, *_ = _plan(M, Nz, 1.0, REF_C, om, *_iid(M, Nz, REF_C, 0.5, 0.3, seed=5))
    rng = np.random.default_rng(8)
    K = 3
    T0 = osl.slab_tmatrices(M, Nz, 1.0, *_iid(M, Nz, REF_C, 1.0, 0.4, seed=6), om, REF_C)
    table = np.stack([T0[0, 0, 0] * (1 + 0.3 * rng.standard_normal((9, 9))) for _ in range(K)])
    tidx = rng.integers(0, K, (Nz, M, M))
    plan.set_tgeneral(tidx, table)
    T = table[tidx]
    A = osl.dense_operator(T, M, Nz, 1.0, om, REF_C)
    psi0 = osl.incident_field(M, Nz, 1.0, om, REF_C, [0.893, 0.45, 0.0], "P").ravel()
    x, st = plan.gmres(torch.as_tensor(psi0), "born", tol=1e-13, max_iter=300, restart=30)
    assert st["converged"]
    assert _rel(x.cpu().numpy(), np.linalg.solve(A, psi0)) < 1e-10
    xm, stm = plan.gmres(torch.as_tensor(psi0), "born", tol=1e-8, max_iter=300, restart=30,
                         precision="mixed")
    assert stm["converged"] and _rel(xm.cpu().numpy(), x.cpu().numpy()) < 1e-6
    r = torch.as_tensor(rng.standard_normal(plan.n) + 1j * rng.standard_normal(plan.n))
    inp = lp.network_input(plan, r.to(hip_device)).float().cpu().numpy()
    s = np.sqrt(plan.n) / float(torch.linalg.vector_norm(r))
    tau = np.einsum("lmnab,lmnb->lmna", T, r.numpy().reshape(Nz, M, M, 9)) * s / np.abs(table).max()
    ref = np.stack([tau.real, tau.imag], axis=-1).reshape(Nz, M, M, 18)
    assert np.abs(inp[..., :18] - ref).max() <= 1e-2 * np.abs(ref).max()
    assert np.all(inp[..., 18:] == 0)
    # the learned preconditioner runs on the general T (flexible GMRES reaches 1e-10)
    net = lp.LearnedPreconditioner(seed=0, arch="unet")
    xf, stf = plan.gmres(torch.as_tensor(psi0), "born", tol=1e-10, max_iter=300, restart=30,
                         precision="mixed", precond=net)
    assert stf["converged"] and stf["precond_applies"] > 0
    assert _rel(xf.cpu().numpy(), x.cpu().numpy()) < 1e-8
