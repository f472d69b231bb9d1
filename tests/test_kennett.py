"""CPU tests of the Kennett layer reflectivity (multiplescatteringcalculations_amd.
kennett_layers, boundary-condition derivation) against the oracle's restatement
of the reference's closed forms (oracle/kennett.py) and the physics that pins
both: unitarity and reciprocity of the flux-normalised coefficients, identical
media, the normal-incidence impedance formula, the Airy single-layer formula
and the homogeneous stack (cubic_scattering/tests/test_kennett_layers.py:
151-610 hold the same checks)."""

import numpy as np
import pytest

import multiplescatteringcalculations_amd.kennett_layers  # noqa: F401  (the package root exports the function of that name)
import sys

kl = sys.modules["multiplescatteringcalculations_amd.kennett_layers"]
from oracle import kennett as ok
from oracle import slab as osl
from oracle import tmatrix as otm

MEDIA = [(3000.0, 1732.0, 2500.0), (4500.0, 2600.0, 2700.0), (2000.0, 800.0, 1800.0),
         (6000.0, 3500.0, 2800.0)]


def _solid(al, be, p, Q=np.inf):
    sp, ss = kl.complex_slowness(al, Q), kl.complex_slowness(be, Q)
    return kl.vertical_slowness(sp, p), kl.vertical_slowness(ss, p), 1.0 / ss


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("i,j", [(0, 1), (1, 0), (2, 3), (3, 2), (0, 2)])
@pytest.mark.parametrize("p,Q", [(0.0, np.inf), (5e-5, np.inf), (2.5e-4, np.inf),
                                 (2.5e-4, 80.0), (4.5e-4, np.inf)])
def test_solid_solid_matches_closed_forms(i, j, p, Q):
    (a1, b1, r1), (a2, b2, r2) = MEDIA[i], MEDIA[j]
    e1, n1, c1 = _solid(a1, b1, p, Q)
    e2, n2, c2 = _solid(a2, b2, p, Q)
    got = kl.psv_solid_solid(p, e1, n1, r1, c1, e2, n2, r2, c2)
    want = ok.solid_solid(p, e1, n1, r1, c1, e2, n2, r2, c2)
    for g, w in zip((got.Rd, got.Ru, got.Td, got.Tu), want):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-12 * max(1.0, np.abs(w).max()))


@pytest.mark.parametrize("i", [0, 1, 2])
@pytest.mark.parametrize("p", [0.0, 1e-4, 3e-4, 5.5e-4])
def test_fluid_solid_matches_closed_forms(i, p):
    al, be, rho = MEDIA[i]
    ew = kl.vertical_slowness(kl.complex_slowness(1500.0, np.inf), p)
    e2, n2, c2 = _solid(al, be, p)
    got = kl.psv_fluid_solid(p, ew, 1025.0, e2, n2, rho, c2)
    want = ok.fluid_solid(p, ew, 1025.0, e2, n2, rho, c2)
    for g, w in zip((got.Rd, got.Ru, got.Td, got.Tu), want):
        np.testing.assert_allclose(g, w, rtol=0, atol=1e-12 * max(1.0, np.abs(w).max()))


def _scattering_matrix(c, m1=2):
    S = np.zeros((4, 4), complex)
    S[:2, :2], S[:2, 2:], S[2:, :2], S[2:, 2:] = c.Rd, c.Tu, c.Td, c.Ru
    keep = [k for k in range(4) if not (m1 == 1 and k == 1)]
    return S[np.ix_(keep, keep)]


@pytest.mark.parametrize("p", [0.0, 1e-4, 2.0e-4])
def test_coefficients_are_unitary_and_reciprocal(p):
    """Propagating modes on both sides: [[Rd, Tu], [Td, Ru]] is unitary, Tu = Td^T."""
    (a1, b1, r1), (a2, b2, r2) = MEDIA[0], MEDIA[1]
    e1, n1, c1 = _solid(a1, b1, p)
    e2, n2, c2 = _solid(a2, b2, p)
    c = kl.psv_solid_solid(p, e1, n1, r1, c1, e2, n2, r2, c2)
    S = _scattering_matrix(c)
    np.testing.assert_allclose(S.conj().T @ S, np.eye(4), atol=1e-12)
    np.testing.assert_allclose(c.Tu, c.Td.T, atol=1e-13)
    ew = kl.vertical_slowness(kl.complex_slowness(1500.0, np.inf), p)
    f = kl.psv_fluid_solid(p, ew, 1025.0, e1, n1, r1, c1)
    Sf = _scattering_matrix(f, m1=1)
    np.testing.assert_allclose(Sf.conj().T @ Sf, np.eye(3), atol=1e-12)
    np.testing.assert_allclose(f.Tu[0], f.Td[:, 0], atol=1e-13)


def test_identical_media_and_normal_incidence_impedance():
    al, be, rho = MEDIA[0]
    e, n, c = _solid(al, be, 2e-4)
    same = kl.psv_solid_solid(2e-4, e, n, rho, c, e, n, rho, c)
    np.testing.assert_allclose(same.Rd, 0, atol=1e-14)
    np.testing.assert_allclose(same.Td, np.eye(2), atol=1e-14)
    (a1, b1, r1), (a2, b2, r2) = MEDIA[0], MEDIA[1]
    e1, n1, c1 = _solid(a1, b1, 0.0)
    e2, n2, c2 = _solid(a2, b2, 0.0)
    c = kl.psv_solid_solid(0.0, e1, n1, r1, c1, e2, n2, r2, c2)
    z1, z2 = r1 * a1, r2 * a2
    assert c.Rd[0, 0] == pytest.approx((z2 - z1) / (z2 + z1), rel=1e-13)
    assert abs(c.Rd[0, 1]) < 1e-15 and abs(c.Rd[1, 0]) < 1e-15
    s1, s2 = r1 * b1, r2 * b2
    # SV at normal incidence sees the shear impedance with the opposite sign convention
    assert abs(abs(c.Rd[1, 1]) - abs((s2 - s1) / (s2 + s1))) < 1e-13
    sh = kl.sh_solid_solid(n1, r1, c1, n2, r2, c2)
    assert sh.Rd == pytest.approx((s1 - s2) / (s1 + s2), rel=1e-13)
    assert sh.Ru == pytest.approx(-sh.Rd, rel=1e-13) and sh.Tu == pytest.approx(sh.Td, rel=1e-13)
    assert abs(sh.Rd) ** 2 + abs(sh.Td) ** 2 == pytest.approx(1.0, rel=1e-13)
    want = ok.sh(n1, r1, c1, n2, r2, c2)
    np.testing.assert_allclose([sh.Rd, sh.Ru, sh.Td, sh.Tu], want, rtol=1e-13)


def _stack(spec):
    lays = []
    for al, be, rho, h in spec:
        lays.append(kl.FluidLayer(al, rho, h) if be == 0.0 else kl.IsotropicLayer(al, be, rho, h))
    return kl.LayerStack(lays)


@pytest.mark.parametrize("p", [0.0, 1.5e-4, 3e-4])
def test_recursion_matches_oracle(p):
    spec = [(1500.0, 0.0, 1025.0, 70.0), (2000.0, 800.0, 1800.0, 12.0),
            (3000.0, 1700.0, 2200.0, 30.0), (4500.0, 2600.0, 2700.0, np.inf)]
    omega = 2 * np.pi * np.linspace(5.0, 80.0, 16) + 1j * 2 * np.pi
    res = kl.kennett_layers(_stack(spec), p, omega)
    R, Rsh = ok.reflectivity(spec, p, omega)
    assert _rel(res.RD_psv, R) < 1e-12
    assert _rel(res.RSH, Rsh) < 1e-12 or np.abs(Rsh).max() == 0
    solid = spec[1:]
    res = kl.kennett_layers(_stack(solid), p, omega.real)
    R, Rsh = ok.reflectivity(solid, p, omega.real)
    assert _rel(res.RD_psv, R) < 1e-12 and _rel(res.RSH, Rsh) < 1e-12


def test_single_layer_is_the_airy_formula():
    """[bg | layer(H) | bg] at normal incidence: r (1 - E) / (1 - r^2 E)."""
    ref = otm.Medium(3000.0, 1732.0, 2500.0)
    dl, dm, dr = 2e9, 1e9, 150.0
    lam, mu, rho = ref.lam + dl, ref.mu + dm, ref.rho + dr
    al, be = np.sqrt((lam + 2 * mu) / rho), np.sqrt(mu / rho)
    H = 40.0
    spec = [(ref.alpha, ref.beta, ref.rho, 1.0), (al, be, rho, H),
            (ref.alpha, ref.beta, ref.rho, np.inf)]
    omega = 2 * np.pi * np.array([3.0, 11.0, 27.0, 64.0])
    res = kl.kennett_layers(_stack(spec), 0.0, omega)
    for w, r in zip(omega, res.RPP):
        assert r == pytest.approx(osl.kennett_layer_rpp_normal(ref, dl, dm, dr, H, w), rel=1e-12)


def test_homogeneous_stack_reflects_nothing_and_phase_is_two_way():
    st = kl.LayerStack.homogeneous(3000.0, 1732.0, 2500.0, n_layers=4, thickness=50.0)
    res = kl.kennett_layers(st, 2e-4, 2 * np.pi * np.array([5.0, 40.0]))
    np.testing.assert_allclose(res.RD_psv, 0, atol=1e-14)
    np.testing.assert_allclose(res.RSH, 0, atol=1e-14)
    # a rigid-like half-space under one layer: R = r E^2 with E the one-way phase
    spec = [(3000.0, 1732.0, 2500.0, 1.0), (3000.0, 1732.0, 2500.0, 25.0),
            (3000.0, 1732.0, 25000.0, np.inf)]
    w = 2 * np.pi * np.array([7.0, 19.0])
    res = kl.kennett_layers(_stack(spec), 0.0, w)
    z1, z2 = 2500.0 * 3000.0, 25000.0 * 3000.0
    np.testing.assert_allclose(res.RPP, (z2 - z1) / (z2 + z1) * np.exp(2j * w * 25.0 / 3000.0),
                               rtol=1e-13)


def test_layer_validation_messages():
    with pytest.raises(ValueError, match="alpha must be positive"):
        kl.IsotropicLayer(-1.0, 1.0, 1.0, 1.0)
    with pytest.raises(ValueError, match="beta must be positive"):
        kl.IsotropicLayer(1.0, 0.0, 1.0, 1.0)
    with pytest.raises(ValueError, match="FluidLayer beta must be 0"):
        kl.FluidLayer(1500.0, 1000.0, 10.0, beta=1.0)
    with pytest.raises(ValueError, match="Need >= 2 layers"):
        kl.LayerStack([kl.IsotropicLayer(1.0, 1.0, 1.0, np.inf)])
    with pytest.raises(ValueError, match="half-space"):
        kl.LayerStack([kl.IsotropicLayer(1.0, 0.5, 1.0, 1.0), kl.IsotropicLayer(1.0, 0.5, 1.0, 2.0)])
    st = kl.LayerStack([kl.IsotropicLayer(3000.0, 1700.0, 2000.0, 5.0),
                        kl.FluidLayer(1500.0, 1000.0, np.inf)])
    with pytest.raises(ValueError, match="solid-fluid not supported"):
        kl.kennett_layers(st, 0.0, np.array([10.0]))


def test_planted_sign_slip_in_oracle_is_caught():
    """A flipped sign in the oracle's closed form breaks the agreement."""
    (a1, b1, r1), (a2, b2, r2) = MEDIA[0], MEDIA[1]
    p = 2.5e-4
    e1, n1, c1 = _solid(a1, b1, p)
    e2, n2, c2 = _solid(a2, b2, p)
    Rd, Ru, Td, Tu = ok.solid_solid(p, e1, n1, r1, c1, e2, n2, r2, c2)
    got = kl.psv_solid_solid(p, e1, n1, r1, c1, e2, n2, r2, c2)
    assert np.abs(got.Rd - Rd).max() < 1e-12
    assert np.abs(got.Rd - np.array([[Rd[0, 0], -Rd[0, 1]], [-Rd[1, 0], Rd[1, 1]]])).max() > 1e-3


@pytest.mark.parametrize("H,omega,scale", [(8.0, 346.4, 0.1), (2.0, 692.8, -0.2), (32.0, 100.0, 0.4),
                                           (0.5, 1500.0, 0.05)])
def test_kennett_reference_rpp_matches_airy_layer(H, omega, scale):
    """kennett_reference_rpp (slab_scattering.py:738-783) equals the oracle's
    closed-form normal-incidence reflection of one uniform layer."""
    import multiplescatteringcalculations_amd as msc
    from oracle import slab as osl
    from oracle import tmatrix as otm

    r = msc.ReferenceMedium(3000.0, 1732.0, 2500.0)
    c = msc.MaterialContrast(scale * r.lam, -0.5 * scale * r.mu, 0.7 * scale * r.rho)
    got = msc.kennett_reference_rpp(r, c, H, omega)
    ref = osl.kennett_layer_rpp_normal(otm.Medium(3000.0, 1732.0, 2500.0), c.Dlambda, c.Dmu,
                                       c.Drho, H, omega)
    assert abs(got - ref) <= 1e-12 * max(abs(ref), 1e-300)
    assert msc.kennett_reference_rpp(r, msc.MaterialContrast(0.0, 0.0, 0.0), H, omega) == 0


def test_top_level_names_of_the_reference_are_exported():
    """The reference package's top-level names on this path
    (cubic_scattering/__init__.py:138-307) import from the package root."""
    import multiplescatteringcalculations_amd as msc

    for name in ("ScatteringConfig", "validate_config", "load_config", "run_from_config",
                 "kennett_layers", "kennett_reference_rpp", "cpa_stack_from_phases",
                 "cubic_to_isotropic_layer", "psv_fluid_solid", "voigt_tmatrix_from_resonance_result",
                 "validate_rayleigh_limit", "scattering_order_decomposition", "sub_cell_centres",
                 "phases_from_two_phase", "voigt_average", "random_heterogeneous_stack",
                 "random_velocity_stack", "PSVCoefficients", "SHCoefficients"):
        assert hasattr(msc, name) and name in msc.__all__, name
