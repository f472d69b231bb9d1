"""CPU oracle of the learned preconditioner network -- TEST INFRASTRUCTURE ONLY.

The reference (tmnestor/MultipleScatteringCalculations) has no learned
preconditioner; its solves are SciPy gmres (cubic_scattering/slab_scattering.py:
532-552) and torch_gmres (cubic_scattering/torch_gmres.py:83-205).  This module
states the network that multiplescatteringcalculations_amd/csrc/msc_net.hip runs
on the GPU, in PyTorch float32 on the CPU, so that the BF16-MFMA forward and
backward kernels can be checked:

* ``forward``/``gradients`` with ``bf16=True`` round at exactly the points the
  GPU kernels do (input, weights, hidden activations, back-propagated
  gradients; FP32 accumulation), so they agree with the kernels up to the
  order of FP32 summation;
* with ``bf16=False`` the manual back-propagation of ``gradients`` is pinned
  against ``torch.autograd`` (tests/test_oracle_precond.py), which fixes the
  transposed/mirrored convolutions and the leaky-ReLU derivative
  independently of the GPU code.

Layout: parameters are one flat vector, per layer the weight in Conv3d layout
[out][in][3][3][3] then the bias [out]; grids are channels-last
[D][H][W][channels] with 18 channels (Re/Im of 9 components interleaved).
"""

from __future__ import annotations

import torch
import torch.nn.functional as F

SLOPE = 0.1
IO = 18
LAYERS = 4


def channels(hidden: int) -> list[int]:
    return [IO, hidden, hidden, hidden, IO]


def num_params(hidden: int) -> int:
    c = channels(hidden)
    return sum(27 * c[l] * c[l + 1] + c[l + 1] for l in range(LAYERS))


def unflatten(params: torch.Tensor, hidden: int):
    """Flat vector -> [(W [out][in][3][3][3], b [out])] per layer."""
    c = channels(hidden)
    out, off = [], 0
    for l in range(LAYERS):
        nw = 27 * c[l] * c[l + 1]
        W = params[off:off + nw].reshape(c[l + 1], c[l], 3, 3, 3)
        off += nw
        b = params[off:off + c[l + 1]]
        off += c[l + 1]
        out.append((W, b))
    assert off == params.numel()
    return out


def _r(x: torch.Tensor, on: bool) -> torch.Tensor:
    return x.to(torch.bfloat16).to(torch.float32) if on else x


def _grid_to_nc(x: torch.Tensor) -> torch.Tensor:  # [D,H,W,C] -> [1,C,D,H,W]
    return x.permute(3, 0, 1, 2).unsqueeze(0)


def _nc_to_grid(x: torch.Tensor) -> torch.Tensor:
    return x.squeeze(0).permute(1, 2, 3, 0)


def forward(params: torch.Tensor, hidden: int, x: torch.Tensor, *, bf16: bool = True):
    """Network output [D,H,W,18] (float32) and the layer inputs a_0..a_3."""
    layers = unflatten(params.float(), hidden)
    a = _r(_grid_to_nc(x.float()), bf16)
    acts = [a]
    for l, (W, b) in enumerate(layers):
        z = F.conv3d(a, _r(W, bf16), b, padding=1)
        if l < LAYERS - 1:
            a = _r(F.leaky_relu(z, SLOPE), bf16)
            acts.append(a)
        else:
            a = z
    return _nc_to_grid(a), acts


def gradients(params: torch.Tensor, hidden: int, xs: torch.Tensor, targets: torch.Tensor, *,
              bf16: bool = True, slope_bwd: float = SLOPE):
    """Gradient (flat, float32) and value of loss = mean((N(x) - t)^2) over the
    samples xs [B,D,H,W,18], targets [B,D,H,W,18], by explicit back-propagation.
    ``slope_bwd`` exists only so that a test can plant a wrong derivative."""
    layers = unflatten(params.float(), hidden)
    count = float(targets.numel())
    grads = [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in layers]
    loss = 0.0
    for x, t in zip(xs, targets):
        out, acts = forward(params, hidden, x, bf16=bf16)
        d = out - t.float()
        loss += float((d.double() ** 2).sum())
        g = _r(_grid_to_nc(2.0 / count * d), bf16)
        for l in range(LAYERS - 1, -1, -1):
            W, _ = layers[l]
            a = acts[l]
            gW = torch.nn.grad.conv3d_weight(a, W.shape, g, padding=1)
            grads[l] = (grads[l][0] + gW, grads[l][1] + g.sum(dim=(0, 2, 3, 4)))
            if l > 0:
                ga = torch.nn.grad.conv3d_input(a.shape, _r(W, bf16), g, padding=1)
                g = _r(torch.where(a > 0, ga, slope_bwd * ga), bf16)
    flat = torch.cat([torch.cat([gW.reshape(-1), gb]) for gW, gb in grads])
    return flat, loss / count


def autograd_gradients(params: torch.Tensor, hidden: int, xs: torch.Tensor,
                       targets: torch.Tensor):
    """The same loss differentiated by torch.autograd (float32, no rounding)."""
    p = params.detach().float().clone().requires_grad_(True)
    layers = unflatten(p, hidden)
    total = 0.0
    for x, t in zip(xs, targets):
        a = _grid_to_nc(x.float())
        for l, (W, b) in enumerate(layers):
            a = F.conv3d(a, W, b, padding=1)
            if l < LAYERS - 1:
                a = F.leaky_relu(a, SLOPE)
        total = total + ((_nc_to_grid(a) - t.float()) ** 2).sum()
    loss = total / targets.numel()
    loss.backward()
    return p.grad.detach(), float(loss)


def tparams_apply(r9: torch.Tensor, tp: torch.Tensor) -> torch.Tensor:
    """T0 r per voxel from the 4 parameters (t_u, t_d, t_o, t_s): r9 [...,9],
    tp [...,4] complex (the block-diagonal T0 of slab_scattering.py:169-205)."""
    tu, td, to, ts = (tp[..., i:i + 1] for i in range(4))
    s = r9[..., 3:6].sum(-1, keepdim=True)
    return torch.cat([tu * r9[..., 0:3], (td - to) * r9[..., 3:6] + to * s, ts * r9[..., 6:9]],
                     dim=-1)


def complex_to_channels(v: torch.Tensor) -> torch.Tensor:
    """[...,9] complex -> [...,18] real, channel 2c = Re, 2c+1 = Im."""
    return torch.view_as_real(v).reshape(*v.shape[:-1], 18)


def network_input(r9: torch.Tensor, tp: torch.Tensor, s: float) -> torch.Tensor:
    """(s / t_max) T0 r as 18 real channels, t_max = max |t| over the grid."""
    tmax = float(tp.abs().max())
    return complex_to_channels(tparams_apply(r9, tp) * (s / tmax)).float()


def apply(params: torch.Tensor, hidden: int, r9: torch.Tensor, tp: torch.Tensor) -> torch.Tensor:
    """z = r + N((s / t_max) T0 r) / s, s = sqrt(n) / ||r||, on a [D,H,W,9] complex grid."""
    n = r9.numel()
    s = (n ** 0.5) / float(torch.linalg.vector_norm(r9))
    x = network_input(r9, tp, s)
    out, _ = forward(params, hidden, x, bf16=True)
    corr = torch.view_as_complex(out.double().reshape(*out.shape[:-1], 9, 2).contiguous())
    return r9 + corr / s
