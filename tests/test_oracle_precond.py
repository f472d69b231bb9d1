"""CPU pins of the learned-preconditioner oracle (oracle/precond.py).

The oracle is the CPU statement of the network csrc/msc_net.hip runs; before
the GPU kernels are compared with it, its pieces are fixed here against
independent computations: torch.nn.Conv3d for the forward layout,
torch.autograd for the explicit back-propagation, the oracle's own 9x9 T0
matrices (oracle/slab.py) for the T0 gather.
"""

import numpy as np
import pytest
import torch

from oracle import precond as op


def _params(hidden, seed, scale=1.0):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(op.num_params(hidden), generator=g) * 0.2 * scale


def test_parameter_count_matches_formula_and_host_module():
    from multiplescatteringcalculations_amd import learned_precond as lp

    for h in (16, 32, 48, 64):
        expect = 27 * (18 * h + 2 * h * h + 18 * h) + 3 * h + 18
        assert op.num_params(h) == expect == lp.num_params(h)
    assert op.num_params(48) == 171234  # "about 200k parameters"
    assert lp.init_params(48, 0).shape == (171234,)


def test_forward_equals_torch_conv3d_stack():
    """Flat layout [out][in][3][3][3] + bias per layer, channels-last grid,
    zero padding 1, leaky-ReLU 0.1 between layers: the same as nn.Conv3d."""
    h = 16
    p = _params(h, 0)
    x = torch.randn(3, 5, 4, 18, generator=torch.Generator().manual_seed(1))
    out, _ = op.forward(p, h, x, bf16=False)
    chans = [18, h, h, h, 18]
    mods = []
    for l, (W, b) in enumerate(op.unflatten(p, h)):
        conv = torch.nn.Conv3d(chans[l], chans[l + 1], 3, padding=1)
        with torch.no_grad():
            conv.weight.copy_(W)
            conv.bias.copy_(b)
        mods.append(conv)
        if l < 3:
            mods.append(torch.nn.LeakyReLU(0.1))
    net = torch.nn.Sequential(*mods)
    with torch.no_grad():
        ref = net(x.permute(3, 0, 1, 2).unsqueeze(0)).squeeze(0).permute(1, 2, 3, 0)
    np.testing.assert_allclose(out.numpy(), ref.numpy(), rtol=1e-5, atol=1e-6)


def test_explicit_backprop_equals_autograd():
    """Without BF16 rounding the oracle's explicit back-propagation (mirrored,
    transposed data-gradient convolutions, weight-gradient correlations,
    leaky-ReLU derivative taken from the stored activation) is autograd's."""
    h = 16
    p = _params(h, 2)
    gen = torch.Generator().manual_seed(3)
    xs = torch.randn(2, 3, 4, 5, 18, generator=gen)
    ts = torch.randn(2, 3, 4, 5, 18, generator=gen)
    g1, l1 = op.gradients(p, h, xs, ts, bf16=False)
    g2, l2 = op.autograd_gradients(p, h, xs, ts)
    assert l1 == pytest.approx(l2, rel=1e-6)
    np.testing.assert_allclose(g1.numpy(), g2.numpy(), rtol=1e-4, atol=1e-7)


def test_a_wrong_derivative_is_caught():
    """A planted mistake (leaky-ReLU derivative taken as 1) breaks agreement
    with autograd, so the check above has teeth."""
    h = 16
    p = _params(h, 4)
    gen = torch.Generator().manual_seed(5)
    xs = torch.randn(1, 3, 4, 4, 18, generator=gen)
    ts = torch.randn(1, 3, 4, 4, 18, generator=gen)
    g_ref, _ = op.autograd_gradients(p, h, xs, ts)
    g_bad, _ = op.gradients(p, h, xs, ts, bf16=False, slope_bwd=1.0)
    assert not np.allclose(g_bad.numpy(), g_ref.numpy(), rtol=1e-3, atol=1e-6)


def test_tparams_apply_equals_dense_t0():
    """T0 r from the 4 per-voxel parameters equals the oracle's dense 9x9 T0."""
    from oracle import slab as osl
    from oracle import tmatrix as tm

    ref = tm.Medium(3000.0, 1732.0, 2500.0)
    rng = np.random.default_rng(0)
    M, Nz, a, om = 3, 2, 1.0, 0.4 * 1732.0
    shape = (Nz, M, M)
    Dl = rng.uniform(-0.2, 0.2, shape) * ref.lam
    Dm = rng.uniform(-0.2, 0.2, shape) * ref.mu
    Dr = rng.uniform(-0.2, 0.2, shape) * ref.rho
    T = osl.slab_tmatrices(M, Nz, a, Dl, Dm, Dr, om, ref)
    tp = osl.tparams_from_T(T)
    r = rng.standard_normal((Nz, M, M, 9)) + 1j * rng.standard_normal((Nz, M, M, 9))
    got = op.tparams_apply(torch.as_tensor(r), torch.as_tensor(tp))
    ref_v = np.einsum("...ab,...b->...a", T, r)
    np.testing.assert_allclose(got.numpy(), ref_v, rtol=1e-12, atol=1e-12 * np.abs(ref_v).max())


def test_zero_network_is_identity():
    h = 16
    p = torch.zeros(op.num_params(h))
    r = torch.randn(2, 3, 3, 9, dtype=torch.complex128)
    tp = torch.randn(2, 3, 3, 4, dtype=torch.complex128)
    z = op.apply(p, h, r, tp)
    assert torch.equal(z, r)


# ── the four-level network ───────────────────────────────────────────────────


def test_unet_parameter_count():
    from multiplescatteringcalculations_amd import learned_precond as lp

    w = (16, 16, 32, 48)
    c0, c1, c2, c3 = w
    conv = 27 * (18 * c0 + c0 * c0 + c1 * c1 + c2 * c2 + c3 * c3 + c2 * c2 + c1 * c1 + c0 * c0
                 + c0 * 18)
    resample = 8 * 2 * (c0 * c1 + c1 * c2 + c2 * c3)
    bias = (c0 + c0 + c1 + c1 + c2 + c2 + c3 + c3 + c2 + c2 + c1 + c1 + c0 + c0 + 18)
    assert op.unet_num_params(w) == conv + resample + bias == 197938
    assert lp.num_params(arch="unet", widths=w) == 197938
    assert lp.init_params(seed=0, arch="unet", widths=w).shape == (197938,)
    for w2 in ((16, 32, 32, 48), (16, 16, 16, 16)):
        assert op.unet_num_params(w2) == lp.num_params(arch="unet", widths=w2)


@pytest.mark.parametrize("shape", [(8, 8, 8), (5, 7, 6), (3, 4, 9), (1, 2, 1)])
def test_unet_forward_equals_functional_reference(shape):
    """The oracle's space-to-depth down and depth-to-space up layers equal
    conv3d with stride 2 on the zero-extended grid and conv_transpose3d
    cropped to the fine grid."""
    p = torch.randn(op.unet_num_params(), generator=torch.Generator().manual_seed(3)) * 0.15
    x = torch.randn(*shape, 18, generator=torch.Generator().manual_seed(4))
    out, t = op.unet_forward(p, x, bf16=False)
    ref = op.unet_reference_forward(p, x)
    assert float((out - ref).abs().max()) <= 1e-5 * float(ref.abs().max())  # FP32 sums
    # level sizes: ceil halving
    assert t[3].shape[2:] == tuple((s + 1) // 2 for s in shape)


@pytest.mark.parametrize("shape", [(6, 8, 8), (5, 7, 3)])
def test_unet_explicit_backprop_equals_autograd(shape):
    p = torch.randn(op.unet_num_params(), generator=torch.Generator().manual_seed(5)) * 0.12
    g = torch.Generator().manual_seed(6)
    xs = torch.randn(2, *shape, 18, generator=g)
    ts = torch.randn(2, *shape, 18, generator=g)
    g1, l1 = op.unet_gradients(p, xs, ts, bf16=False)
    g2, l2 = op.unet_autograd_gradients(p, xs, ts)
    assert l1 == pytest.approx(l2, rel=1e-6)
    assert float((g1 - g2).norm() / g2.norm()) < 1e-5
    # every layer's weights and bias receive a gradient of their own
    off = 0
    for kind, ci, co, _ in op.unet_layers():
        n = (27 if kind == "conv" else 8) * ci * co
        for m in (n, co):
            a, b = g1[off:off + m], g2[off:off + m]
            assert float((a - b).norm()) <= 1e-4 * float(b.norm()) + 1e-9
            off += m
    # a planted wrong leaky-ReLU derivative is caught
    g3, _ = op.unet_gradients(p, xs, ts, bf16=False, slope_bwd=0.2)
    assert float((g3 - g2).norm() / g2.norm()) > 1e-2
