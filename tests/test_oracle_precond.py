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
