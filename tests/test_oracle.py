"""Pins of the complex128 oracle (CPU): each test would fail for a plausible
mistake in the restatement (a dropped term, a wrong sign, index or factor).

Sources of truth: the Mathematica golden values held by the reference's tests,
closed forms re-derived by quadrature, finite differences, physical invariants
(reciprocity, lattice covariance, Born linearity, Kennett layer reflection),
brute-force sums on tiny slabs, and the golden vectors written from the
reference (tests/golden/make_golden.py).
"""

from pathlib import Path

import numpy as np
import pytest
from numpy.polynomial.legendre import leggauss
from scipy import integrate

from oracle import propagator as pr
from oracle import slab as sl
from oracle import tmatrix as tm

GOLD = np.load(Path(__file__).resolve().parent / "golden" / "slab_golden.npz")
REF_T = tm.Medium(5000.0, 3000.0, 2500.0)
OMEGA_T = 150.0
REF_C = tm.Medium(3000.0, 1732.0, 2500.0)


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


# ── T0 ───────────────────────────────────────────────────────────────────────


def test_mathematica_golden_ABC_T123():
    """cubic_scattering/tests/test_cubic_tmatrix.py:463-485 (rel 1e-8)."""
    r = tm.cube_tmatrix(2 * np.pi * 10, 0.01, tm.Medium(5.0, 3.0, 2.5), 1.25, 0.75, 0.0)
    gold = dict(
        Ac=-0.012201107458741135 - 0.00008607635386320063j,
        Bc=0.0026137073560736812 + 0.00003973958017869591j,
        Cc=-0.003587055298886922 - 1.0444731153752018e-10j,
        T1c=-0.002746105632556146 + 0.00015071169827338333j,
        T2c=-0.007190550077000591 - 0.00003475258026337854j,
        T3c=-0.005380582948330384 - 1.566709673062803e-10j,
    )
    for k, v in gold.items():
        assert abs(getattr(r, k) - v) / abs(v) < 1e-8, k


def test_geometric_constants_by_quadrature():
    """g0 = int 1/|x| over [-1,1]^3 = 3 int_{[-1,1]^2} (1+u^2+v^2)^(-1/2) (face pyramids);
    j1, j2, k1 as the surface integrals of effective_contrasts.py:248-269."""
    f = lambda p: integrate.dblquad(lambda v, u: p(u, v), -1, 1, -1, 1, epsabs=1e-13, epsrel=1e-13)[0]
    g0 = 3.0 * f(lambda u, v: (1 + u * u + v * v) ** -0.5)
    j1 = f(lambda u, v: (1 + u * u + v * v) ** -1.5)
    j2 = f(lambda u, v: u * u * (1 + u * u + v * v) ** -2.5)
    k1 = f(lambda u, v: (1 + u * u + v * v) ** -2.5)
    assert abs(g0 - tm.G0_CUBE) < 1e-11
    J1, J2, K1 = tm.eshelby_static_geom()
    assert abs(j1 - J1) < 1e-12 and abs(j2 - J2) < 1e-12 and abs(k1 - K1) < 1e-12


def test_cube_moments_by_gauss_legendre():
    """S0, S1, S2, S11 = int u^m {1, x1^2, x1^4, x1^2 x2^2} over [-a,a]^3 (exact quadrature)."""
    a = 0.7
    x, w = leggauss(16)
    x = a * x
    w = a * w
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    W = w[:, None, None] * w[None, :, None] * w[None, None, :]
    U = X**2 + Y**2 + Z**2
    S0, S1, S2, S11 = tm.cube_moments(a, 6)
    for m in range(7):
        assert np.isclose(np.sum(W * U**m), S0[m], rtol=1e-13)
        assert np.isclose(np.sum(W * X**2 * U**m), S1[m], rtol=1e-13)
        assert np.isclose(np.sum(W * X**4 * U**m), S2[m], rtol=1e-13)
        assert np.isclose(np.sum(W * X**2 * Y**2 * U**m), S11[m], rtol=1e-13)


def _duffy_cube_integral(fun, a, n=24):
    """int over [-a,a]^3 of fun(x) with a 1/r singularity at 0: six face pyramids,
    x = t a (1, u, v) (rotated), Jacobian a^3 t^2, Gauss-Legendre in t, u, v."""
    g, w = leggauss(n)
    t = 0.5 * (g + 1.0)
    wt = 0.5 * w
    total = 0.0
    for axis in range(3):
        for sgn in (1.0, -1.0):
            for it, ti in enumerate(t):
                for iu, ui in enumerate(g):
                    for iv, vi in enumerate(g):
                        p = np.empty(3)
                        p[axis] = sgn
                        p[(axis + 1) % 3] = ui
                        p[(axis + 2) % 3] = vi
                        total = total + wt[it] * w[iu] * w[iv] * a**3 * ti**2 * fun(a * ti * p)
    return total


def test_gamma0_static_limit_by_duffy_cubature():
    """Gamma0 -> int_cube G_11 of the exact Kupradze tensor as ka -> 0.

    The reference's closed form (effective_contrasts.py:112-152) keeps the
    static Kelvin term plus Taylor terms that are all purely imaginary (phi_n,
    psi_n carry odd powers of i w), so it is exact only in the static limit:
    the real dynamic corrections and the P-wave part of Im G(0) are absent
    (1.5% at ka = 0.05, 13% at ka = 0.4).  The cubature pins the static term and
    the normalisation; the digits at finite ka are pinned by the golden
    vectors (test_golden_cube_tmatrix_scalars).
    """
    a = 1.0
    om = 1e-3 * REF_C.beta / a
    num = _duffy_cube_integral(lambda x: pr.greens(x, om, REF_C.alpha, REF_C.beta, REF_C.rho)[1, 1],
                               a, n=12)
    G0 = tm.gamma0(om, a, REF_C.alpha, REF_C.beta, REF_C.rho)
    assert abs(num - G0) / abs(num) < 1e-3
    assert abs(num.real - G0.real) / abs(num.real) < 1e-5


def test_golden_cube_tmatrix_scalars():
    """compute_cube_tmatrix at the configurations' frequencies (golden vectors)."""
    ref = tm.Medium(*GOLD["ref"])
    fields = ["Gamma0", "Ac", "Bc", "Cc", "T1c", "T2c", "T3c", "amp_u", "amp_theta", "amp_e_off",
              "amp_e_diag", "Drho_star", "Dlambda_star", "Dmu_star_off", "Dmu_star_diag"]
    for (om, a, dl, dm, dr), vals in zip(GOLD["cube_inputs"], GOLD["cube_values"]):
        r = tm.cube_tmatrix(om, a, ref, dl, dm, dr)
        for f, v in zip(fields, vals):
            assert abs(getattr(r, f) - v) <= 1e-12 * abs(v), f


def test_born_limit_amplifications():
    """Weak contrast: amplification factors -> 1, effective -> bare (test_cubic_tmatrix.py:42-95)."""
    ref = REF_T
    c = ref.mu * 1e-4
    r = tm.cube_tmatrix(2 * np.pi * 10, 10.0, ref, c, c, ref.rho * 1e-4)
    for amp in (r.amp_u, r.amp_theta, r.amp_e_off, r.amp_e_diag):
        assert abs(amp - 1) < 1e-4
    assert abs(r.Dlambda_star - c) / c < 1e-4 and abs(r.Dmu_star_off - c) / c < 1e-4


def test_t123_closed_form_equals_contraction():
    """The device kernel's T1 = Dl(A+4B+C)+2B Dm, T2 = (A+B)Dm, T3 = 2C Dm equal the
    4-index contraction for random complex A, B, C."""
    rng = np.random.default_rng(0)
    for _ in range(5):
        A, B, C = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        dl, dm = rng.standard_normal(2)
        T1, T2, T3 = tm.t123(A, B, C, dl, dm)
        assert np.isclose(T1, dl * (A + 4 * B + C) + 2 * B * dm, rtol=1e-13)
        assert np.isclose(T2, (A + B) * dm, rtol=1e-13)
        assert np.isclose(T3, 2 * C * dm, rtol=1e-13)


def test_zero_contrast_zero_T_and_cache():
    z = np.zeros((2, 3, 3))
    assert np.all(sl.slab_tmatrices(3, 2, 1.0, z, z, z, OMEGA_T, REF_T) == 0)
    Dl = np.where(np.arange(18).reshape(2, 3, 3) % 2 == 0, 2e9, 0.0)
    Dm = np.where(Dl > 0, 1e9, 0.0)
    Dr = np.where(Dl > 0, 100.0, 0.0)
    T = sl.slab_tmatrices(3, 2, 1.0, Dl, Dm, Dr, OMEGA_T, REF_T)
    T1 = tm.tmatrix_9x9(tm.cube_tmatrix(OMEGA_T, 1.0, REF_T, 2e9, 1e9, 100.0), OMEGA_T, 1.0)
    assert np.allclose(T[0, 0, 0], T1, rtol=1e-13) and np.all(T[0, 0, 1] == 0)


# ── propagator ───────────────────────────────────────────────────────────────


def test_greens_derivatives_by_finite_differences():
    """Gd, Gdd against central differences (test_cubic_tmatrix.py:504-543)."""
    ref = REF_T
    om = 2 * np.pi * 50
    r = np.array([0.3, -0.2, 0.5])
    G, Gd, Gdd = pr.greens_deriv(r, om, ref.alpha, ref.beta, ref.rho)
    assert np.allclose(G, pr.greens(r, om, ref.alpha, ref.beta, ref.rho), rtol=1e-12)
    h = 1e-6
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        fd = (pr.greens(r + e, om, ref.alpha, ref.beta, ref.rho)
              - pr.greens(r - e, om, ref.alpha, ref.beta, ref.rho)) / (2 * h)
        assert _rel(fd, Gd[:, :, k]) < 1e-7
        fd2 = (pr.greens_deriv(r + e, om, ref.alpha, ref.beta, ref.rho)[1]
               - pr.greens_deriv(r - e, om, ref.alpha, ref.beta, ref.rho)[1]) / (2 * h)
        assert _rel(fd2, Gdd[:, :, :, k]) < 1e-6


def test_voigt_blocks_are_strains_of_the_fields():
    """C, H, S by finite differences of G: C[i,(p,q)] = engineering-halved
    sym. derivative of G_i., H = engineering strain of the force field G_.j, and
    S_raw (before the halving of columns >= 3) = engineering strain of C_raw."""
    om = 2 * np.pi * 40
    ref = REF_C
    r = np.array([0.45, -0.3, 0.25])
    P = pr.propagator_9x9(r, om, ref.alpha, ref.beta, ref.rho)
    Cb, Hb, Sb = P[:3, 3:], P[3:, :3], P[3:, 3:]
    h = 1e-6

    def grad(f):
        out = []
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            out.append((f(r + e) - f(r - e)) / (2 * h))
        return np.stack(out, axis=-1)  # last index = derivative direction

    G = lambda x: pr.greens(x, om, ref.alpha, ref.beta, ref.rho)
    dG = grad(G)  # dG[i, j, k] = d_k G_ij
    for al, (p, q) in enumerate(pr.VOIGT_PAIRS):
        for i in range(3):
            c_raw = dG[i, p, p] if p == q else dG[i, p, q] + dG[i, q, p]
            assert abs(Cb[i, al] - (c_raw if al < 3 else 0.5 * c_raw)) < 1e-7 * np.abs(Cb).max()
            h_val = dG[p, i, p] if p == q else dG[p, i, q] + dG[q, i, p]
            assert abs(Hb[al, i] - h_val) < 1e-7 * np.abs(Hb).max()

    def c_raw_field(x, be):
        _, Gd, _ = pr.greens_deriv(x, om, ref.alpha, ref.beta, ref.rho)
        m, n = pr.VOIGT_PAIRS[be]
        return np.array([Gd[i, m, m] if m == n else Gd[i, m, n] + Gd[i, n, m] for i in range(3)])

    for be in range(6):
        dC = grad(lambda x: c_raw_field(x, be))  # dC[u, v] = d_v C_raw[u, be]
        for al, (p, q) in enumerate(pr.VOIGT_PAIRS):
            s_raw = dC[p, p] if p == q else dC[p, q] + dC[q, p]
            want = s_raw if be < 3 else 0.5 * s_raw
            assert abs(Sb[al, be] - want) < 1e-6 * np.abs(Sb).max()


def test_propagator_reciprocity_structure():
    """Reciprocity: G = G^T, H = W^-1 C^T, S W^-1 symmetric (W = diag(1,1,1,1/2,1/2,1/2)).
    The device table stores only these 45 independent entries."""
    W = np.array([1, 1, 1, 0.5, 0.5, 0.5])
    rng = np.random.default_rng(3)
    for _ in range(5):
        r = rng.standard_normal(3)
        P = pr.propagator_9x9(r, 700.0, REF_C.alpha, REF_C.beta, REF_C.rho)
        sc = np.abs(P).max()
        assert np.abs(P[:3, :3] - P[:3, :3].T).max() < 1e-14 * sc
        assert np.abs(P[3:, :3] - P[:3, 3:].T / W[:, None]).max() < 1e-14 * sc
        Sp = P[3:, 3:] / W[None, :]
        assert np.abs(Sp - Sp.T).max() < 1e-13 * sc


def test_lattice_covariance_of_the_propagator():
    """K(sigma r) = R K(r) R for the x, y, z reflections and the x<->y swap
    (lattice_greens.py:522-566 sign tables; z reflection used by the device table)."""
    Rz = np.diag([-1.0, 1, 1, 1, 1, 1, 1, -1, -1])
    r = np.array([0.7, -0.4, 0.9])
    K = lambda x: pr.propagator_9x9(x, 650.0, REF_C.alpha, REF_C.beta, REF_C.rho)
    P = K(r)
    sc = np.abs(P).max()
    assert np.abs(K(r * [1, -1, 1]) - pr.refl_x(P)).max() < 1e-14 * sc
    assert np.abs(K(r * [1, 1, -1]) - pr.refl_y(P)).max() < 1e-14 * sc
    assert np.abs(K(r * [-1, 1, 1]) - Rz @ P @ Rz).max() < 1e-14 * sc
    assert np.abs(K(r[[0, 2, 1]]) - pr.swap_xy(P)).max() < 1e-13 * sc


# ── slab kernel and matvec ───────────────────────────────────────────────────


def _iid(M, Nz, seed, phi=0.5, c=0.4, ref=REF_C):
    rng = np.random.default_rng(seed)
    shape = (Nz, M, M)
    mask = rng.random(shape) < phi
    return (np.where(mask, rng.uniform(-c, c, shape) * ref.lam, 0.0),
            np.where(mask, rng.uniform(-c, c, shape) * ref.mu, 0.0),
            np.where(mask, rng.uniform(-c, c, shape) * ref.rho, 0.0))


def test_kernel_self_term_and_spot_check():
    """Self term zero, offset (dz=d, dx=d) equals the direct propagator (test_slab_scattering.py:180-200)."""
    kh = sl.slab_kernels(3, 2, 1.0, OMEGA_T, REF_T)
    assert kh.shape == (3, 5, 5, 9, 9)
    sp = np.fft.ifft2(kh, axes=(1, 2))
    assert np.abs(sp[1, 2, 2]).max() < 1e-20
    Gd = pr.propagator_9x9([2.0, 2.0, 0.0], OMEGA_T, REF_T.alpha, REF_T.beta, REF_T.rho)
    assert np.abs(sp[2, 3, 2] - Gd).max() < 1e-10 * np.abs(Gd).max()


@pytest.mark.parametrize("M,Nz,uniform", [(2, 2, True), (3, 2, False), (2, 3, False)])
def test_fft_matvec_equals_brute_force(M, Nz, uniform):
    """test_slab_scattering.py:246-273: FFT matvec = dense O(N^2) sum, rtol 1e-10."""
    if uniform:
        Dl, Dm, Dr = (np.full((Nz, M, M), v) for v in (2e9, 1e9, 100.0))
        ref, om = REF_T, OMEGA_T
    else:
        ref, om = REF_C, 692.8
        Dl, Dm, Dr = _iid(M, Nz, seed=M + 7 * Nz)
    T = sl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, ref)
    kh = sl.slab_kernels(M, Nz, 1.0, om, ref)
    rng = np.random.default_rng(123)
    n = Nz * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y_dir = sl.direct_matvec(x, T, M, Nz, 1.0, om, ref)
    assert _rel(sl.slab_matvec(x, T, kh, M, Nz), y_dir) < 1e-12
    pz = 1 << (2 * Nz - 2).bit_length()
    for pad in [None, (pz, 8, 8), (16, 16, 16)]:
        k3 = sl.kernel_hat_3d(kh, M, Nz, pad)
        assert _rel(x - sl.slab_matvec_3d(x, T, k3, M, Nz), x - y_dir) < 1e-11


def test_periodic_matvec_equals_circular_sum():
    """test_slab_scattering.py:663-718: periodic FFT matvec = explicit circular sum."""
    M, Nz = 3, 2
    Dl, Dm, Dr = _iid(M, Nz, seed=9)
    om = 692.8
    T = sl.slab_tmatrices(M, Nz, 1.0, Dl, Dm, Dr, om, REF_C)
    khp = sl.slab_kernels(M, Nz, 1.0, om, REF_C, periodic=True)
    kc = np.fft.ifft2(khp, axes=(1, 2))
    rng = np.random.default_rng(42)
    x = rng.standard_normal(Nz * M * M * 9) + 1j * rng.standard_normal(Nz * M * M * 9)
    psi = x.reshape(Nz, M, M, 9)
    tau = np.einsum("lmnab,lmnb->lmna", T, psi)
    acc = np.zeros_like(psi)
    for i1 in np.ndindex(Nz, M, M):
        for i2 in np.ndindex(Nz, M, M):
            G = kc[i1[0] - i2[0] + Nz - 1, (i1[1] - i2[1]) % M, (i1[2] - i2[2]) % M]
            acc[i1] += G @ tau[i2]
    assert _rel(sl.slab_matvec(x, T, khp, M, Nz, periodic=True), (psi - acc).ravel()) < 1e-12
    # the folded kernel is the aperiodic one summed mod M
    sp = np.fft.ifft2(sl.slab_kernels(M, Nz, 1.0, om, REF_C), axes=(1, 2))
    fold = np.zeros_like(kc)
    for ix in range(2 * M - 1):
        for iy in range(2 * M - 1):
            fold[:, (ix - M + 1) % M, (iy - M + 1) % M] += sp[:, ix, iy]
    assert np.abs(fold - kc).max() < 1e-12 * np.abs(kc).max()


def test_identity_when_T_zero():
    kh = sl.slab_kernels(3, 2, 1.0, OMEGA_T, REF_T)
    x = np.random.default_rng(1).standard_normal(2 * 9 * 9) + 0j
    assert np.allclose(sl.slab_matvec(x, np.zeros((2, 3, 3, 9, 9), complex), kh, 3, 2), x, atol=1e-14)


# ── incident field ───────────────────────────────────────────────────────────


def test_incident_field_closed_form():
    """P: pol = k_hat, phase exp(i kP k.x); strain = (ik/2)(k pol + pol k) with
    engineering shear; S: pol perpendicular to k (test_slab_scattering.py:280-317)."""
    kh = np.array([0.8, 0.36, 0.48])
    psi0 = sl.incident_field(3, 2, 1.0, 500.0, REF_C, kh, "P")
    kP = 500.0 / REF_C.alpha
    c = sl.centres(3, 2, 1.0)
    ph = np.exp(1j * kP * np.einsum("k,lmnk->lmn", kh, c))
    assert np.allclose(psi0[..., :3], kh * ph[..., None], rtol=1e-14)
    eps = 1j * kP * np.array([kh[0] ** 2, kh[1] ** 2, kh[2] ** 2, 2 * kh[1] * kh[2],
                              2 * kh[0] * kh[2], 2 * kh[0] * kh[1]])
    assert np.allclose(psi0[..., 3:], eps * ph[..., None], rtol=1e-14)
    s0 = sl.incident_field(2, 1, 1.0, 500.0, REF_C, kh, "S")
    kS = 500.0 / REF_C.beta
    pol = s0[0, 0, 0, :3] / np.exp(1j * kS * np.dot(kh, sl.centres(2, 1, 1.0)[0, 0, 0]))
    assert abs(np.dot(pol.real, kh)) < 1e-12 and abs(np.linalg.norm(pol) - 1) < 1e-12
    with pytest.raises(ValueError, match="wave_type"):
        sl.incident_field(2, 1, 1.0, 500.0, REF_C, kh, "X")


# ── solve and outputs ────────────────────────────────────────────────────────


def test_solve_equals_dense_direct_solve():
    M, Nz = 3, 2
    Dl, Dm, Dr = _iid(M, Nz, seed=21, phi=0.7)
    x, psi0, T, _, res = sl.solve(M, Nz, 1.0, Dl, Dm, Dr, 692.8, REF_C, [0.893, 0.45, 0.0], "P",
                                  tol=1e-13)
    A = sl.dense_operator(T, M, Nz, 1.0, 692.8, REF_C)
    assert res < 1e-12
    assert _rel(x.ravel(), np.linalg.solve(A, psi0.ravel())) < 1e-11


def test_zero_contrast_and_born_scaling_of_rpp():
    """test_slab_convergence.py:33-68."""
    M, Nz = 4, 1
    z = np.zeros((Nz, M, M))
    x, _, T, _, _ = sl.solve(M, Nz, 1.0, z, z, z, OMEGA_T, REF_T, [1, 0, 0])
    assert abs(sl.rpp_periodic(x, T, M, Nz, 1.0, OMEGA_T, REF_T)) < 1e-20
    R = []
    for s in (1.0, 2.0):
        Dl, Dm, Dr = (np.full((Nz, M, M), s * v) for v in (2e5, 1e5, 0.01))
        x, _, T, _, _ = sl.solve(M, Nz, 1.0, Dl, Dm, Dr, OMEGA_T, REF_T, [1, 0, 0], tol=1e-12)
        R.append(sl.rpp_periodic(x, T, M, Nz, 1.0, OMEGA_T, REF_T))
    assert abs(abs(R[1]) / abs(R[0]) - 2.0) < 0.05 * 2.0


@pytest.mark.parametrize("M,Nz,weak,bound", [(4, 2, True, 0.05), (8, 1, False, 0.15)])
def test_rpp_matches_kennett_layer(M, Nz, weak, bound):
    """test_slab_convergence.py:74-133: periodic-formula R_PP vs the Kennett
    reflection of a uniform layer of thickness N_z d (5% weak, 15% M=8)."""
    s = 1e-4 if weak else 1.0
    c = (2e9 * s, 1e9 * s, 100.0 * s)
    Dl, Dm, Dr = (np.full((Nz, M, M), v) for v in c)
    x, _, T, _, _ = sl.solve(M, Nz, 1.0, Dl, Dm, Dr, OMEGA_T, REF_T, [1, 0, 0], tol=1e-8)
    R_FL = sl.rpp_periodic(x, T, M, Nz, 1.0, OMEGA_T, REF_T)
    R_K = sl.kennett_layer_rpp_normal(REF_T, *c, H=2.0 * Nz, omega=OMEGA_T)
    assert abs(R_FL - R_K) / abs(R_K) < bound


def test_kennett_closed_form_limits():
    assert sl.kennett_layer_rpp_normal(REF_T, 0.0, 0.0, 0.0, 2.0, OMEGA_T) == 0
    # thin-layer limit: R -> 0 as H -> 0, and Born-linear in the contrast
    r1 = sl.kennett_layer_rpp_normal(REF_T, 2e5, 1e5, 0.01, 2.0, OMEGA_T)
    r2 = sl.kennett_layer_rpp_normal(REF_T, 4e5, 2e5, 0.02, 2.0, OMEGA_T)
    assert abs(sl.kennett_layer_rpp_normal(REF_T, 2e9, 1e9, 100.0, 1e-9, OMEGA_T)) < 1e-9
    assert abs(abs(r2 / r1) - 2.0) < 1e-3


# ── golden vectors written from the reference ───────────────────────────────


@pytest.mark.parametrize("ci", [0, 1, 2])
def test_golden_outputs(ci):
    k = f"out{ci}"
    M, Nz = GOLD[k + "_dims"]
    a, om, p = GOLD[k + "_scal"]
    Dl, Dm, Dr = GOLD[k + "_mat"]
    ref = tm.Medium(*GOLD["ref"])
    T = sl.slab_tmatrices(M, Nz, a, Dl, Dm, Dr, om, ref)
    psi = GOLD[k + "_psi"]
    r = sl.rpp_periodic(psi, T, M, Nz, a, om, ref, p=p)
    assert abs(r - GOLD[k + "_rpp"][0]) <= 1e-13 * abs(GOLD[k + "_rpp"][0])
    far = np.array(sl.reflected_field(psi, T, M, Nz, a, om, ref, GOLD[k + "_khat"]))
    assert _rel(far, GOLD[k + "_far"]) < 1e-13


@pytest.mark.parametrize("ci", [0, 1, 2])
def test_golden_swave_polarisation(ci):
    ref = tm.Medium(*GOLD["ref"])
    got = sl.incident_field(2, 2, 1.0, 500.0, ref, GOLD[f"swave{ci}_khat"], "S")
    assert _rel(got, GOLD[f"swave{ci}_psi0"]) < 1e-14


@pytest.mark.parametrize("ci", [0, 1])
def test_golden_converged_solution(ci):
    k = f"solve{ci}"
    M, Nz, per = GOLD[k + "_dims"]
    a, om, p = GOLD[k + "_scal"]
    Dl, Dm, Dr = GOLD[k + "_mat"]
    ref = tm.Medium(*GOLD["ref"])
    x, *_ = sl.solve(M, Nz, a, Dl, Dm, Dr, om, ref, GOLD[k + "_khat"], "P", tol=1e-13,
                     periodic=bool(per))
    assert _rel(x, GOLD[k + "_psi"]) < 1e-10
