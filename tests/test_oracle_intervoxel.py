"""CPU pins of the volume-averaged nearest-neighbour propagator oracle
(oracle/intervoxel.py), against the reference's own known answers
(cubic_scattering/tests/test_inter_voxel_propagator.py): the Mathematica
static propagator values (282-339, rel 1e-12 / 1e-11), the Laplacian
identities that tie every tabulated constant family to the next (165-215,
434-510, 1024-1045 and the order-3 / third-derivative ones), the 3-D Fourier
(WS3) values of the dynamic correction (644-800, 2% / 10%), the G-block trace
identity (872-897) and the block symmetries (916-1020).  A planted transcription
slip in a constant fails the identity that contains it."""

import math

import numpy as np
import pytest

from oracle import intervoxel as iv

MU, NU = 1.0, 0.25
ALPHA, BETA, RHO = 5000.0, 3000.0, 2500.0

MATHEMATICA = [  # (type, index, value, rel)
    ("face", (0, 0, 0, 0), +0.06489609963957935, 1e-12),
    ("face", (0, 0, 1, 1), -0.00994518639570713, 1e-12),
    ("face", (1, 1, 1, 1), -0.02599699603810159, 1e-12),
    ("face", (1, 1, 2, 2), +0.01343931900972618, 1e-12),
    ("face", (0, 1, 0, 1), -0.04369948153183095, 1e-12),
    ("face", (1, 2, 1, 2), -0.02031497612639764, 1e-12),
    ("edge", (0, 0, 0, 0), -0.00080672935720783, 1e-11),
    ("edge", (2, 2, 2, 2), -0.00711840486058437, 1e-11),
    ("edge", (0, 0, 0, 1), +0.03454486653300503, 1e-11),
    ("edge", (0, 1, 2, 2), -0.00833663621412387, 1e-11),
    ("corner", (0, 0, 0, 0), -0.00417024398837613, 1e-11),
    ("corner", (0, 0, 1, 1), +0.00208512199418806, 1e-11),
    ("corner", (0, 0, 0, 1), +0.00988469241293594, 1e-11),
    ("corner", (0, 0, 1, 2), +0.00164678558813909, 1e-11),
]


@pytest.mark.parametrize("t,idx,val,rel", MATHEMATICA)
def test_static_propagator_matches_mathematica(t, idx, val, rel):
    P = iv.p_canonical_static_unit(t, MU, NU)
    assert P[idx] == pytest.approx(val, rel=rel)


def _laplacian_ratio(t, order):
    """sum_k B_ijkk / A_ij over the components where A_ij != 0 (raw values)."""
    A = iv.a_matrix(t, iv.A_RAW[(t, order)])
    B = iv.fill_symmetric(4, iv.B_TUPLES[t], iv.B_RAW[(t, order)])
    tr = np.einsum("ijkk->ij", B)
    return A, tr


@pytest.mark.parametrize("t", ["face", "edge", "corner"])
@pytest.mark.parametrize("order,factor", [(0, 1.0), (1, 12.0), (2, 30.0), (3, 56.0)])
def test_laplacian_identity_ties_b_to_a(t, order, factor):
    """sum_k B^(n)_ijkk = A^(n)_ij for the normalised constants of every order
    (exact in Fourier space); on the raw tabulated values the factor is the
    ratio of the normalisations, 1, 12, 30, 56 for n = 0..3."""
    A = iv.a_matrix(t, iv.A_RAW[(t, order)])
    B = iv.fill_symmetric(4, iv.B_TUPLES[t], iv.B_RAW[(t, order)])
    tr = np.einsum("ijkk->ij", B)
    mask = np.ones((3, 3), bool)
    if t == "corner":
        # the reference's corner fill leaves the {0112} class at zero, which
        # breaks exactly the (0,2) / (2,0) components of this identity
        mask[0, 2] = mask[2, 0] = False
    np.testing.assert_allclose(tr[mask], factor * A[mask], rtol=1e-11, atol=1e-13)
    if order > 0:
        np.testing.assert_allclose(
            np.einsum("ijkk->ij", iv.b_tensor(t, order))[mask],
            iv.a_order(t, order)[mask], rtol=1e-11, atol=1e-20)


@pytest.mark.parametrize("t", ["face", "edge", "corner"])
@pytest.mark.parametrize("order,factor", [(0, 2.0), (1, 12.0), (2, 30.0), (3, 56.0)])
def test_third_derivative_laplacian(t, order, factor):
    """sum_j d3W_jjk = c dW_k (c = 2, 12, 30, 56)."""
    D = iv.d3_tensor(t, order)
    dW = iv.dw_vector(t, order)
    np.testing.assert_allclose(np.einsum("jjk->k", D), factor * dW, rtol=1e-10, atol=1e-12)


def test_a_planted_constant_slip_is_caught():
    saved = iv.B_RAW[("edge", 2)]
    try:
        v = list(saved)
        v[3] *= 1.001  # a digit slip in one dynamic edge constant
        iv.B_RAW[("edge", 2)] = tuple(v)
        with pytest.raises(AssertionError):
            test_laplacian_identity_ties_b_to_a("edge", 2, 30.0)
    finally:
        iv.B_RAW[("edge", 2)] = saved


def test_corner_fill_keeps_the_reference_gap():
    """Documented reference behaviour: B_{0112} (and permutations) stays zero in
    the corner tensors while B_{0012} and B_{0122} carry the tabulated value."""
    B = iv.b_tensor("corner", 0)
    assert B[0, 1, 1, 2] == 0.0 and B[0, 0, 1, 2] == B[0, 1, 2, 2] != 0.0


WS3 = [  # (R, ka_P, {index: Re dP}, rel) -- slab_scattering's WS3 3-D Fourier values
    ((1, 0, 0), 0.05, {(0, 0, 0, 0): -2.5051013578654155e-15, (0, 0, 1, 1): +1.0050372914007568e-15,
                       (1, 1, 1, 1): -4.3123096932947566e-15, (1, 1, 2, 2): +1.9796055507553097e-15,
                       (0, 1, 0, 1): -4.112350835371632e-15, (1, 2, 1, 2): -3.1385435529468117e-15}, 0.02),
    ((1, 0, 0), 0.1, {(0, 0, 0, 0): -9.638611522792203e-15, (0, 0, 1, 1): +3.868208724558371e-15,
                      (1, 1, 1, 1): -1.6891488262665625e-14, (1, 1, 2, 2): +7.799543051112971e-15,
                      (0, 1, 0, 1): -1.6160967846816168e-14, (1, 2, 1, 2): -1.2295026396595533e-14}, 0.02),
    ((1, 0, 0), 0.3, {(0, 0, 0, 0): -6.472738870029643e-14, (0, 0, 1, 1): +2.399726517350551e-14,
                      (1, 1, 1, 1): -1.3182390520301285e-13, (1, 1, 2, 2): +6.274046605242954e-14,
                      (0, 1, 0, 1): -1.358800753743591e-13, (1, 2, 1, 2): -1.0437569909823117e-13}, 0.10),
    ((1, 1, 0), 0.1, {(0, 0, 0, 0): -1.0479240153348353e-14, (0, 0, 1, 1): +3.994235725162921e-15,
                      (0, 0, 2, 2): +4.08158027233474e-15, (2, 2, 2, 2): -1.2226386071473104e-14,
                      (0, 0, 0, 1): +8.722877774926246e-15, (0, 1, 2, 2): -2.2487148661478562e-15}, 0.02),
    ((1, 1, 1), 0.1, {(0, 0, 0, 0): -9.558327302446959e-15, (0, 0, 1, 1): +3.574217220960363e-15,
                      (0, 0, 0, 1): +4.766679766551694e-15, (0, 0, 1, 2): -3.3006205421545777e-16}, 0.02),
]


@pytest.mark.parametrize("R,ka,ref,rel", WS3)
def test_dynamic_correction_matches_fourier_reference(R, ka, ref, rel):
    omega = ka * ALPHA
    dP = (iv.strain_propagator(R, ALPHA, BETA, RHO, omega, 2)
          - iv.strain_propagator(R, ALPHA, BETA, RHO, 0.0, 0))
    for idx, val in ref.items():
        assert dP[idx] == pytest.approx(val, rel=rel), idx


def test_g_block_trace_identity_and_structure():
    mu = RHO * BETA ** 2
    nu = (ALPHA ** 2 - 2 * BETA ** 2) / (2 * (ALPHA ** 2 - BETA ** 2))
    raw11, raw22, _ = iv.A_RAW[("face", 1)]
    Phi = 0.5 * (raw11 + 2 * raw22)
    eta_s = 1.0 / (4.0 * (1.0 - nu))
    P9 = iv.propagator_9x9((1, 0, 0), ALPHA, BETA, RHO, 0.0, 0)
    G = P9[:3, :3].real
    assert np.trace(G) == pytest.approx(Phi * (3 - 2 * eta_s) / (4 * math.pi * mu), rel=1e-12)
    np.testing.assert_allclose(G, G.T, atol=1e-30)
    assert G[1, 1] == pytest.approx(G[2, 2], rel=1e-14) and G[0, 1] == 0.0
    Gc = iv.propagator_9x9((1, 1, 1), ALPHA, BETA, RHO, 0.0, 0)[:3, :3].real
    assert Gc[0, 0] == pytest.approx(Gc[2, 2], rel=1e-12) and Gc[0, 1] == pytest.approx(Gc[1, 2], rel=1e-12)


def test_block_symmetries_for_all_neighbours():
    """H = C^T for all 26 neighbours; G, S even and C, H odd under R -> -R;
    the w^2 series vanishes at w = 0; finite everywhere."""
    omega = 2 * math.pi * 50
    for R in [(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1) if (i, j, k) != (0, 0, 0)]:
        P = iv.propagator_9x9(R, ALPHA, BETA, RHO, omega, 2)
        Pm = iv.propagator_9x9(tuple(-v for v in R), ALPHA, BETA, RHO, omega, 2)
        assert np.isfinite(P).all()
        np.testing.assert_allclose(P[3:, :3], P[:3, 3:].T, atol=1e-30)
        np.testing.assert_allclose(P[:3, :3], Pm[:3, :3], rtol=1e-13, atol=1e-28)
        np.testing.assert_allclose(P[3:, 3:], Pm[3:, 3:], rtol=1e-13, atol=1e-28)
        np.testing.assert_allclose(P[:3, 3:], -Pm[:3, 3:], rtol=1e-13, atol=1e-28)
    P0 = iv.propagator_9x9((1, 0, 0), ALPHA, BETA, RHO, 0.0, 1)
    Ps = iv.propagator_9x9((1, 0, 0), ALPHA, BETA, RHO, 0.0, 0)
    np.testing.assert_allclose(P0, Ps, atol=1e-30)


def test_face_c_block_sparsity_and_rotation():
    """R = (1,0,0): C[0,0], C[0,1] = C[0,2], C[1,5] = C[2,4] and nothing else;
    the R = (0,1,0) block is its x<->z-permuted image."""
    C = iv.propagator_9x9((1, 0, 0), ALPHA, BETA, RHO, 100.0, 2)[:3, 3:].real
    nz = np.argwhere(np.abs(C) > 0)
    assert {tuple(x) for x in nz} == {(0, 0), (0, 1), (0, 2), (1, 5), (2, 4)}
    assert C[0, 1] == C[0, 2] and C[1, 5] == C[2, 4]
    Cy = iv.propagator_9x9((0, 1, 0), ALPHA, BETA, RHO, 100.0, 2)[:3, 3:].real
    assert C[0, 0] == pytest.approx(Cy[1, 1], rel=1e-13)
