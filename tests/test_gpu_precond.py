"""GPU tests of the learned preconditioner (include/msc_precond.h).

* The BF16-MFMA kernels against the CPU oracle oracle/precond.py, which rounds
  at the same points (input, weights, hidden activations and back-propagated
  gradients to BF16; FP32 accumulation).  Remaining differences come from the
  order of FP32 sums and the BF16 roundings they move, so the tolerances are
  norm-wise: forward 5e-3, gradients 2e-2, loss 1e-4 relative.
* Flexible GMRES: with the zero network P = I it reproduces plain GMRES to
  1e-10; with any network it converges to the same FP64 residual target; a
  network that makes no progress hands over to plain GMRES.
* The ablation: a network trained on pairs harvested from plain solves on
  seeded slabs, scored on a held-out slab against the same solve without it.
"""

import numpy as np
import pytest
import torch

from oracle import precond as op
from oracle import slab as osl
from oracle import tmatrix as otm

pytestmark = pytest.mark.gpu

REF_C = otm.Medium(3000.0, 1732.0, 2500.0)


def _nrel(a, b):
    a, b = torch.as_tensor(a).cpu(), torch.as_tensor(b).cpu()
    dt = torch.complex128 if (a.is_complex() or b.is_complex()) else torch.float64
    a, b = a.to(dt), b.to(dt)
    return float(torch.linalg.vector_norm(a - b) / max(float(torch.linalg.vector_norm(b)), 1e-300))


def _net(hidden, seed=0, scale=1.0):
    from multiplescatteringcalculations_amd.learned_precond import LearnedPreconditioner

    net = LearnedPreconditioner(hidden, seed=None)
    g = torch.Generator().manual_seed(seed)
    p = torch.randn(net.num_params, generator=g) * 0.15 * scale
    net.set_params(p)
    return net, p


def _bf16_input(shape, seed):
    g = torch.Generator().manual_seed(seed)
    x = torch.zeros(*shape, 32)
    x[..., :18] = torch.randn(*shape, 18, generator=g)
    return x.to(torch.bfloat16)


@pytest.mark.parametrize("hidden,grid", [(16, (3, 5, 7)), (48, (4, 9, 16)), (32, (2, 3, 70))])
def test_forward_matches_oracle(hip_device, hidden, grid):
    net, p = _net(hidden, seed=hidden)
    x = _bf16_input(grid, seed=1)
    out = net.forward(x.to(hip_device))
    ref, _ = op.forward(p, hidden, x[..., :18].float(), bf16=True)
    assert out.shape == (*grid, 18)
    assert _nrel(out, ref) < 5e-3, _nrel(out, ref)


@pytest.mark.parametrize("hidden", [16, 32, 48])
def test_gradients_match_oracle(hip_device, hidden):
    net, p = _net(hidden, seed=7)
    grid = (3, 5, 6)
    xs = torch.stack([_bf16_input(grid, seed=s) for s in (2, 3)])
    g = torch.Generator().manual_seed(4)
    ts = torch.randn(2, *grid, 18, generator=g)
    gg, loss = net.gradients(xs.to(hip_device), ts.to(hip_device))
    gr, lr = op.gradients(p, hidden, xs[..., :18].float(), ts, bf16=True)
    assert loss == pytest.approx(lr, rel=1e-4)
    # per layer (weights and bias separately)
    c = op.channels(hidden)
    off = 0
    for l in range(4):
        for n in (27 * c[l] * c[l + 1], c[l + 1]):
            e = _nrel(gg[off:off + n], gr[off:off + n])
            assert e < 2e-2, (l, n, e)
            off += n
    assert off == net.num_params


def test_untested_width_is_refused(hip_device):
    """Only the widths whose kernels the parity tests above pin are built."""
    from multiplescatteringcalculations_amd.learned_precond import LearnedPreconditioner

    with pytest.raises(ValueError, match="hidden width"):
        LearnedPreconditioner(64, seed=None)


def _small_plan(M=6, Nz=4, phi=0.5, c=0.4, ka=0.75, seed=0):
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import slab_scattering as ss

    g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
    ref = msc.ReferenceMedium(REF_C.alpha, REF_C.beta, REF_C.rho)
    mat = msc.random_iid_slab_material(g, ref, phi, c, seed=seed)
    omega = ka * ref.beta / g.a
    plan = ss.SlabPlan(g)
    plan.build_g0(omega, ref)
    plan.set_material(mat, omega)
    khat = np.array([np.sqrt(1 - (1.5e-4 * ref.alpha) ** 2), 1.5e-4 * ref.alpha, 0.0])
    psi0 = plan.incident_field(omega, ref, khat, "P").reshape(-1)
    return plan, psi0, mat, omega, ref


def test_network_input_and_apply_match_oracle(hip_device):
    from multiplescatteringcalculations_amd import learned_precond as lp

    plan, psi0, mat, omega, ref = _small_plan()
    g = plan.geometry
    net, p = _net(16, seed=11, scale=0.5)
    tp = plan.tparams().reshape(g.N_z, g.M, g.M, 4).cpu()
    rng = np.random.default_rng(5)
    r = torch.as_tensor(rng.standard_normal(plan.n) + 1j * rng.standard_normal(plan.n))
    r9 = r.reshape(g.N_z, g.M, g.M, 9)
    # input: (s / t_max) T0 r in bf16, s = sqrt(n)/||r||
    x = lp.network_input(plan, r.to(hip_device))
    s = plan.n ** 0.5 / float(torch.linalg.vector_norm(r))
    xr = op.network_input(r9, tp, s)
    assert 0.05 < float(xr.abs().max()) < 20.0  # order one
    assert torch.all(x[..., 18:] == 0)
    assert _nrel(x[..., :18].float(), xr) < 4e-3
    for dt, tol in ((torch.complex128, 1e-2), (torch.complex64, 1e-2)):
        z = net.apply(plan, r.to(hip_device, dt))
        zr = op.apply(p, 16, r9, tp).reshape(-1)
        corr, corr_r = (z.cpu().to(torch.complex128) - r), (zr - r)
        assert float(torch.linalg.vector_norm(corr_r)) > 1e-3 * float(torch.linalg.vector_norm(r))
        assert _nrel(corr, corr_r) < tol, (dt, _nrel(corr, corr_r))


def test_training_reduces_loss(hip_device):
    """Adam on a fixed set of (input, target) pairs lowers the loss, and the
    first step's loss is the loss of the initial parameters."""
    net, p = _net(16, seed=3, scale=0.5)
    grid = (2, 6, 8)
    xs = torch.stack([_bf16_input(grid, seed=s) for s in range(4)]).to(hip_device)
    # a learnable target: a fixed linear map of the input channels
    A = torch.randn(18, 18, generator=torch.Generator().manual_seed(9)).to(hip_device) * 0.3
    ts = (xs[..., :18].float() @ A).contiguous()
    _, l0 = net.gradients(xs, ts)
    losses = [net.train_step(xs, ts, lr=3e-3) for _ in range(60)]
    assert losses[0] == pytest.approx(l0, rel=1e-6)
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])


def test_zero_network_fgmres_equals_gmres(hip_device):
    """P = I (all parameters zero): flexible GMRES is plain GMRES."""
    plan, psi0, *_ = _small_plan(seed=1)
    from multiplescatteringcalculations_amd.learned_precond import LearnedPreconditioner

    net = LearnedPreconditioner(16, seed=None)  # zero parameters
    x0, s0 = plan.gmres(psi0, "born", tol=1e-11, max_iter=200, restart=20, precision="fp64")
    x1, s1 = plan.gmres(psi0, "born", tol=1e-11, max_iter=200, restart=20, precision="fp64",
                        precond=net)
    assert s1["converged"] and s1["precond_applies"] == s1["iterations"] > 0
    assert abs(s1["iterations"] - s0["iterations"]) <= 1
    assert _nrel(x1, x0) < 1e-10


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_random_network_fgmres_reaches_same_solution(hip_device, precision):
    plan, psi0, *_ = _small_plan(seed=2)
    net, _ = _net(16, seed=5, scale=0.3)
    x0, s0 = plan.gmres(psi0, "born", tol=1e-11, max_iter=300, restart=20, precision="fp64")
    x1, s1 = plan.gmres(psi0, "born", tol=1e-11, max_iter=300, restart=20, precision=precision,
                        precond=net)
    assert s1["converged"] and s1["rel_residual"] <= 1e-11
    assert s1["precond_applies"] > 0
    assert _nrel(x1, x0) < 1e-9
    # the same solve again (cached HIP graphs of the flexible steps) is identical
    x2, s2 = plan.gmres(psi0, "born", tol=1e-11, max_iter=300, restart=20, precision=precision,
                        precond=net)
    assert s2["iterations"] == s1["iterations"]
    assert torch.equal(x2, x1)
    # and a plain solve after it is unaffected by the flexible one
    x3, s3 = plan.gmres(psi0, "born", tol=1e-11, max_iter=300, restart=20, precision="fp64")
    assert torch.equal(x3, x0) and s3["iterations"] == s0["iterations"]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 1e-5)])
def test_fgmres_steps_match_host_fgmres(hip_device, precision, tol):
    """Three flexible GMRES steps on the device (where the network's input
    scale is passed as sqrt(n) for the unit-norm v_j) against the same three
    steps on the host, with z_j = net.apply(v_j) measuring the scale of v_j.
    The device builds its basis from w = v_j - A z_j, so its v_{j+1} is the
    textbook one times -1; the network is not odd (biases, leaky ReLU), so the
    host follows the same basis: H[:, j] = e_j - V^H w, H[j+1, j] = -||w_perp||,
    x_3 = x_0 + Z y, y = argmin ||beta e_1 - H y||.  A wrong input or output
    scale on either path, or the other sign, changes x_3 at order one."""
    import warnings

    plan, psi0, *_ = _small_plan(seed=4)
    net, _ = _net(16, seed=6, scale=0.4)
    k = 3
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")
        xd, sd = plan.gmres(psi0, "born", tol=1e-14, max_iter=k, restart=20, precision=precision,
                            precond=net)
    assert sd["iterations"] == k and sd["precond_applies"] == k and not sd["fell_back"]
    b = psi0.to(hip_device, torch.complex128)
    x0 = b.clone()
    r0 = b - plan.matvec(x0)
    beta = float(torch.linalg.vector_norm(r0))
    V, Z = [r0 / beta], []
    H = torch.zeros(k + 1, k, dtype=torch.complex128)
    for j in range(k):
        z = net.apply(plan, V[j])
        Z.append(z)
        w = V[j] - plan.matvec(z)
        h = torch.zeros(j + 1, dtype=torch.complex128)
        for _ in range(2):  # classical Gram-Schmidt, twice
            for i in range(j + 1):
                c = torch.vdot(V[i], w)
                h[i] += c.cpu()
                w = w - c * V[i]
        H[: j + 1, j] = -h
        H[j, j] += 1.0
        hn = float(torch.linalg.vector_norm(w))
        H[j + 1, j] = -hn
        V.append(w / hn)
    e1 = torch.zeros(k + 1, 1, dtype=torch.complex128)
    e1[0, 0] = beta
    y = torch.linalg.lstsq(H, e1).solution.reshape(-1)
    xh = x0 + sum(y[j].item() * Z[j] for j in range(k))
    xd = xd.reshape(-1).to(hip_device)
    assert _nrel(xh - x0, x0) > 1e-3  # the steps move the iterate
    assert _nrel(xd - x0, xh - x0) < tol, _nrel(xd - x0, xh - x0)


def test_useless_preconditioner_hands_over_to_plain_gmres(hip_device):
    """A network with large random weights maps each v_j to a large, unrelated
    vector, so z_j spans directions that barely reduce the residual: flexible
    cycles stall and plain cycles finish the solve."""
    plan, psi0, *_ = _small_plan(seed=3)
    net, _ = _net(16, seed=13, scale=4.0)
    x0, _ = plan.gmres(psi0, "born", tol=1e-10, max_iter=400, restart=10, precision="fp64")
    x1, s1 = plan.gmres(psi0, "born", tol=1e-10, max_iter=400, restart=10, precision="fp64",
                        precond=net)
    assert s1["converged"]
    assert s1["fell_back"] & 2
    assert _nrel(x1, x0) < 1e-8


def test_ablation_trained_network_vs_plain(hip_device):
    """Train on pairs harvested from plain GMRES on seeded strong-scattering
    slabs (phi 0.5, contrasts +-0.4, ka 0.75), then solve a held-out slab with
    and without the network.  Reports iterations and time; checks that the
    network learned (held-out correction error falls) and that the learned
    solve reaches the same FP64 residual."""
    import time

    from multiplescatteringcalculations_amd import learned_precond as lp

    M, Nz = 16, 8
    sets = []
    for seed in range(1, 7):
        plan, psi0, *_ = _small_plan(M=M, Nz=Nz, seed=seed)
        sets.append(lp.harvest_pairs(plan, psi0, restart=40, tol=1e-8))
        del plan
    pairs = lp.merge_pairs(sets)
    net = lp.LearnedPreconditioner(16, seed=0)
    losses, before, after = lp.train(net, pairs, steps=300, lr=2e-3, batch=8, holdout=8)
    print(f"\npairs {len(pairs)}  train loss {losses[0]:.3e} -> {losses[-1]:.3e}  "
          f"held-out {before:.3e} -> {after:.3e}")
    assert after < 0.7 * before

    plan, psi0, *_ = _small_plan(M=M, Nz=Nz, seed=100)
    res = {}
    for label, pc in (("plain", None), ("learned", net)):
        for prec in ("fp64", "mixed"):
            plan.gmres(psi0, "born", tol=1e-8, max_iter=300, restart=30, precision=prec,
                       precond=pc)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x, st = plan.gmres(psi0, "born", tol=1e-8, max_iter=300, restart=30, precision=prec,
                               precond=pc)
            torch.cuda.synchronize()
            res[(label, prec)] = (x, st, time.perf_counter() - t0)
            print(f"{label:8s} {prec:6s} iters {st['iterations']:3d}  "
                  f"rel {st['rel_residual']:.2e}  {1e3 * res[(label, prec)][2]:.2f} ms  "
                  f"fell_back {st['fell_back']}")
    for key, (x, st, _) in res.items():
        assert st["converged"], key
        assert _nrel(x, res[("plain", "fp64")][0]) < 1e-6, key
    # what the trained network gains: fewer Arnoldi steps than the same solve without it
    for prec in ("fp64", "mixed"):
        assert res[("learned", prec)][1]["fell_back"] == 0
        assert res[("learned", prec)][1]["iterations"] < res[("plain", prec)][1]["iterations"]


# ── the four-level network (msc_net_create_unet) ─────────────────────────────


def _unet(seed=0, widths=(16, 16, 32, 48), jitter=0.02):
    """He-initialised weights plus a small random perturbation of every
    parameter (biases included), so that activations stay of order one."""
    from multiplescatteringcalculations_amd import learned_precond as lp

    net = lp.LearnedPreconditioner(seed=None, arch="unet", widths=widths)
    g = torch.Generator().manual_seed(seed)
    p = torch.as_tensor(lp.init_params(seed=seed, arch="unet", widths=widths, last_scale=1.0))
    p = p + jitter * torch.randn(p.numel(), generator=g)
    net.set_params(p)
    return net, p


@pytest.mark.parametrize("widths,grid", [((16, 16, 32, 48), (8, 8, 8)),
                                         ((16, 16, 32, 48), (5, 7, 13)),
                                         ((16, 16, 32, 48), (4, 24, 72)),
                                         ((16, 16, 32, 48), (1, 2, 3)),
                                         ((16, 32, 32, 48), (6, 10, 12)),
                                         ((16, 16, 16, 16), (3, 9, 66)),
                                         ((16, 16, 32, 48), (24, 96, 96))])
def test_unet_forward_matches_oracle(hip_device, widths, grid):
    net, p = _unet(seed=sum(grid), widths=widths)
    x = _bf16_input(grid, seed=2)
    out = net.forward(x.to(hip_device))
    ref, _ = op.unet_forward(p, x[..., :18].float(), bf16=True, widths=widths)
    assert out.shape == (*grid, 18)
    assert float(ref.abs().max()) > 0.1
    assert _nrel(out, ref) < 1e-2, _nrel(out, ref)


@pytest.mark.parametrize("grid", [(4, 6, 8), (3, 5, 7), (12, 40, 40)])
def test_unet_gradients_match_oracle(hip_device, grid):
    widths = (16, 16, 32, 48)
    net, p = _unet(seed=9)
    xs = torch.stack([_bf16_input(grid, seed=s) for s in (3, 4)])
    ts = torch.randn(2, *grid, 18, generator=torch.Generator().manual_seed(5))
    gg, loss = net.gradients(xs.to(hip_device), ts.to(hip_device))
    gr, lr = op.unet_gradients(p, xs[..., :18].float(), ts, bf16=True, widths=widths)
    assert loss == pytest.approx(lr, rel=1e-3)
    off = 0
    for l, (kind, ci, co, _) in enumerate(op.unet_layers(widths)):
        for n in ((27 if kind == "conv" else 8) * ci * co, co):
            e = _nrel(gg[off:off + n], gr[off:off + n])
            assert e < 3e-2, (l, kind, n, e)
            off += n
    assert off == net.num_params


def test_unet_apply_matches_oracle(hip_device):
    plan, psi0, mat, omega, ref = _small_plan(M=10, Nz=6)
    g = plan.geometry
    net, p = _unet(seed=12)
    tp = plan.tparams().reshape(g.N_z, g.M, g.M, 4).cpu()
    rng = np.random.default_rng(8)
    r = torch.as_tensor(rng.standard_normal(plan.n) + 1j * rng.standard_normal(plan.n))
    r9 = r.reshape(g.N_z, g.M, g.M, 9)
    zr = op.unet_apply(p, r9, tp).reshape(-1)
    for dt in (torch.complex128, torch.complex64):
        z = net.apply(plan, r.to(hip_device, dt))
        corr, corr_r = (z.cpu().to(torch.complex128) - r), (zr - r)
        assert float(torch.linalg.vector_norm(corr_r)) > 1e-2 * float(torch.linalg.vector_norm(r))
        assert _nrel(corr, corr_r) < 2e-2, (dt, _nrel(corr, corr_r))


def test_unet_training_reduces_loss_and_zero_net_is_plain_gmres(hip_device):
    from multiplescatteringcalculations_amd.learned_precond import LearnedPreconditioner

    net, p = _unet(seed=3)
    grid = (4, 8, 8)
    xs = torch.stack([_bf16_input(grid, seed=s) for s in range(4)]).to(hip_device)
    A = torch.randn(18, 18, generator=torch.Generator().manual_seed(9)).to(hip_device) * 0.3
    ts = (xs[..., :18].float() @ A).contiguous()
    _, l0 = net.gradients(xs, ts)
    losses = [net.train_step(xs, ts, lr=2e-3) for _ in range(80)]
    assert losses[0] == pytest.approx(l0, rel=1e-6)
    assert losses[-1] < 0.5 * losses[0], (losses[0], losses[-1])
    plan, psi0, *_ = _small_plan(seed=1)
    zero = LearnedPreconditioner(seed=None, arch="unet")
    x0, s0 = plan.gmres(psi0, "born", tol=1e-11, max_iter=200, restart=20, precision="fp64")
    x1, s1 = plan.gmres(psi0, "born", tol=1e-11, max_iter=200, restart=20, precision="fp64",
                        precond=zero)
    assert s1["converged"] and s1["precond_applies"] == s1["iterations"] > 0
    assert abs(s1["iterations"] - s0["iterations"]) <= 1
    assert _nrel(x1, x0) < 1e-10
    x2, s2 = plan.gmres(psi0, "born", tol=1e-11, max_iter=300, restart=20, precision="mixed",
                        precond=net)
    assert s2["converged"] and _nrel(x2, x0) < 1e-9
