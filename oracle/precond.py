"""CPU oracle of the learned preconditioner networks -- TEST INFRASTRUCTURE ONLY.

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


# ── the four-level network (msc_net_create_unet) ─────────────────────────────
#
# Layer list in forward order (kind, in, out, skip): tensors are t0 = input,
# then one per layer (t[k + 1] = output of layer k); "skip" is the id of the
# tensor an up layer's output is added to.
#   0 conv io->c0   1 conv c0->c0 (x0)   2 down c0->c1   3 conv c1->c1 (x1)
#   4 down c1->c2   5 conv c2->c2 (x2)   6 down c2->c3   7 conv c3->c3 (x3)
#   8 up c3->c2 +x2  9 conv c2->c2       10 up c2->c1 +x1  11 conv c1->c1
#   12 up c1->c0 +x0 13 conv c0->c0      14 conv c0->io (linear)
# Leaky ReLU after every conv and down layer except the last; up layers are
# linear.  Down: out[c] = sum_s W[:, :, s] in[2c + off(s)], off(s) = (s >> 2,
# (s >> 1) & 1, s & 1) along (D, H, W), fine voxels beyond the grid zero.  Up:
# out[2c + off(s)] = W[:, :, s] in[c] (+ bias + skip), outputs beyond the fine
# grid dropped.  Weights are stored [out][in][2][2][2] for both.

UNET_WIDTHS = (16, 16, 32, 48)


def unet_layers(widths=UNET_WIDTHS):
    c0, c1, c2, c3 = widths
    return [("conv", IO, c0, None), ("conv", c0, c0, None), ("down", c0, c1, None),
            ("conv", c1, c1, None), ("down", c1, c2, None), ("conv", c2, c2, None),
            ("down", c2, c3, None), ("conv", c3, c3, None), ("up", c3, c2, 6),
            ("conv", c2, c2, None), ("up", c2, c1, 4), ("conv", c1, c1, None),
            ("up", c1, c0, 2), ("conv", c0, c0, None), ("conv", c0, IO, None)]
    # the fourth field of an up layer is the id of its skip tensor (t[k + 1] is
    # the output of layer k: x0 = t2, x1 = t4, x2 = t6)


def unet_num_params(widths=UNET_WIDTHS) -> int:
    return sum((27 if k == "conv" else 8) * ci * co + co for k, ci, co, _ in unet_layers(widths))


def unet_unflatten(params: torch.Tensor, widths=UNET_WIDTHS):
    out, off = [], 0
    for kind, ci, co, _ in unet_layers(widths):
        k = 3 if kind == "conv" else 2
        nw = k ** 3 * ci * co
        W = params[off:off + nw].reshape(co, ci, k, k, k)
        off += nw
        b = params[off:off + co]
        off += co
        out.append((W, b))
    assert off == params.numel()
    return out


def _s2d(a: torch.Tensor) -> torch.Tensor:
    """[1,C,D,H,W] -> [1, 8C, ceil(D/2), ceil(H/2), ceil(W/2)], channel s*C + c
    holding a[c] at 2x + off(s); the grid is zero-extended to even sizes."""
    _, C, D, H, W = a.shape
    a = F.pad(a, (0, W % 2, 0, H % 2, 0, D % 2))
    Dc, Hc, Wc = (D + 1) // 2, (H + 1) // 2, (W + 1) // 2
    a = a.reshape(1, C, Dc, 2, Hc, 2, Wc, 2).permute(0, 3, 5, 7, 1, 2, 4, 6)
    return a.reshape(1, 8 * C, Dc, Hc, Wc)


def _d2s(a: torch.Tensor, C: int, fine) -> torch.Tensor:
    """Inverse of _s2d for [1, 8C, Dc, Hc, Wc], cropped to the fine grid."""
    _, _, Dc, Hc, Wc = a.shape
    a = a.reshape(1, 2, 2, 2, C, Dc, Hc, Wc).permute(0, 4, 5, 1, 6, 2, 7, 3)
    a = a.reshape(1, C, 2 * Dc, 2 * Hc, 2 * Wc)
    return a[:, :, :fine[0], :fine[1], :fine[2]]


def _down(a, W, b):
    """out[o] = b[o] + sum_{s,i} W[o, i, s] a_s2d[s*C + i]."""
    co, ci = W.shape[:2]
    Wm = W.reshape(co, ci, 8).permute(0, 2, 1).reshape(co, 8 * ci)
    x = _s2d(a)
    return torch.einsum("ok,nkdhw->nodhw", Wm, x) + b.view(1, -1, 1, 1, 1)


def _up(a, W, b, fine):
    co, ci = W.shape[:2]
    Wm = W.reshape(co, ci, 8).permute(2, 0, 1).reshape(8 * co, ci)
    y = torch.einsum("ki,nidhw->nkdhw", Wm, a)
    return _d2s(y, co, fine) + b.view(1, -1, 1, 1, 1)


def unet_forward(params: torch.Tensor, x: torch.Tensor, *, bf16: bool = True,
                 widths=UNET_WIDTHS):
    """Output [D,H,W,18] (float32) and every tensor t0..t14 (channels first)."""
    lay = unet_layers(widths)
    Ws = unet_unflatten(params.float(), widths)
    t = [_r(_grid_to_nc(x.float()), bf16)]
    for l, ((kind, ci, co, skip), (W, b)) in enumerate(zip(lay, Ws)):
        a = t[-1]
        Wr = _r(W, bf16)
        if kind == "conv":
            z = F.conv3d(a, Wr, b, padding=1)
            if l == len(lay) - 1:
                t.append(z)
                break
            t.append(_r(F.leaky_relu(z, SLOPE), bf16))
        elif kind == "down":
            t.append(_r(F.leaky_relu(_down(a, Wr, b), SLOPE), bf16))
        else:
            fine = t[skip].shape[2:]
            t.append(_r(_up(a, Wr, b, fine) + t[skip], bf16))
    return _nc_to_grid(t[-1]), t


def unet_gradients(params: torch.Tensor, xs: torch.Tensor, targets: torch.Tensor, *,
                   bf16: bool = True, widths=UNET_WIDTHS, slope_bwd: float = SLOPE):
    """Gradient (flat) and value of the mean-square loss by explicit
    back-propagation through the layer graph, with BF16 rounding of the
    back-propagated gradients where the GPU kernels round them.  A tensor
    used as an up layer's skip receives that layer's output gradient; for the
    leaky-ReLU tensors the gradient is taken with respect to the
    pre-activation."""
    lay = unet_layers(widths)
    Ws = unet_unflatten(params.float(), widths)
    count = float(targets.numel())
    grads = [(torch.zeros_like(W), torch.zeros_like(b)) for W, b in Ws]
    loss = 0.0
    L = len(lay)
    skip_user = {skip: l for l, (_, _, _, skip) in enumerate(lay) if skip is not None}

    def mask(g, a):
        return torch.where(a > 0, g, slope_bwd * g)

    for x, tgt in zip(xs, targets):
        out, t = unet_forward(params, x, bf16=bf16, widths=widths)
        d = out - tgt.float()
        loss += float((d.double() ** 2).sum())
        G = {L: _r(_grid_to_nc(2.0 / count * d), bf16)}  # gradient of tensor t[L]
        for l in range(L - 1, -1, -1):
            kind, ci, co, skip = lay[l]
            W, _ = Ws[l]
            Wr = _r(W, bf16)
            g = G[l + 1]
            a = t[l]  # the layer's input tensor
            if kind == "conv":
                gW = torch.nn.grad.conv3d_weight(a, W.shape, g, padding=1)
            elif kind == "down":
                gm = torch.einsum("nodhw,nkdhw->ok", g, _s2d(a))  # [co][8 ci], k = s ci + i
                gW = gm.reshape(co, 8, ci).permute(0, 2, 1).reshape(W.shape)
            else:
                gs = _s2d(g)  # [1, 8 co, ...] at the coarse grid of a
                gm = torch.einsum("nkdhw,nidhw->ki", gs, a)  # [8 co][ci], k = s co + o
                gW = gm.reshape(8, co, ci).permute(1, 2, 0).reshape(W.shape)
            grads[l] = (grads[l][0] + gW, grads[l][1] + g.sum(dim=(0, 2, 3, 4)))
            if l == 0:
                continue
            if kind == "conv":
                ga = torch.nn.grad.conv3d_input(a.shape, Wr, g, padding=1)
                is_act = lay[l - 1][0] != "up"  # up layers are linear
                G[l] = _r(mask(ga, a) if is_act else ga, bf16)
            elif kind == "down":
                Wm = Wr.reshape(co, ci, 8).permute(2, 1, 0).reshape(8 * ci, co)
                ga = _d2s(torch.einsum("ko,nodhw->nkdhw", Wm, g), ci, a.shape[2:])
                if l in skip_user:  # a is also the skip of a later up layer
                    ga = ga + G[skip_user[l] + 1]
                G[l] = _r(mask(ga, a), bf16)
            else:
                Wm = Wr.reshape(co, ci, 8).permute(1, 2, 0).reshape(ci, 8 * co)
                ga = torch.einsum("ik,nkdhw->nidhw", Wm, _s2d(g))
                G[l] = _r(mask(ga, a), bf16)
    flat = torch.cat([torch.cat([gW.reshape(-1), gb]) for gW, gb in grads])
    return flat, loss / count


def unet_autograd_gradients(params: torch.Tensor, xs: torch.Tensor, targets: torch.Tensor,
                            widths=UNET_WIDTHS):
    """The same loss through torch.nn.functional (conv3d with stride 2 on the
    zero-extended grid for down layers, conv_transpose3d cropped to the fine
    grid for up layers), differentiated by torch.autograd; no rounding."""
    p = params.detach().float().clone().requires_grad_(True)
    lay = unet_layers(widths)
    Ws = unet_unflatten(p, widths)
    total = 0.0
    for x, tgt in zip(xs, targets):
        t = [_grid_to_nc(x.float())]
        for l, ((kind, ci, co, skip), (W, b)) in enumerate(zip(lay, Ws)):
            a = t[-1]
            if kind == "conv":
                z = F.conv3d(a, W, b, padding=1)
                t.append(z if l == len(lay) - 1 else F.leaky_relu(z, SLOPE))
            elif kind == "down":
                D, H, Wd = a.shape[2:]
                ap = F.pad(a, (0, Wd % 2, 0, H % 2, 0, D % 2))
                t.append(F.leaky_relu(F.conv3d(ap, W, b, stride=2), SLOPE))
            else:
                fine = t[skip].shape[2:]
                y = F.conv_transpose3d(a, W.transpose(0, 1), b, stride=2)
                t.append(y[:, :, :fine[0], :fine[1], :fine[2]] + t[skip])
        total = total + ((_nc_to_grid(t[-1]) - tgt.float()) ** 2).sum()
    loss = total / targets.numel()
    loss.backward()
    return p.grad.detach(), float(loss.detach())


def unet_reference_forward(params: torch.Tensor, x: torch.Tensor, widths=UNET_WIDTHS):
    """Forward pass through torch.nn.functional only (no rounding), as in
    unet_autograd_gradients: the pin of unet_forward's layer layout."""
    lay = unet_layers(widths)
    Ws = unet_unflatten(params.float(), widths)
    t = [_grid_to_nc(x.float())]
    with torch.no_grad():
        for l, ((kind, ci, co, skip), (W, b)) in enumerate(zip(lay, Ws)):
            a = t[-1]
            if kind == "conv":
                z = F.conv3d(a, W, b, padding=1)
                t.append(z if l == len(lay) - 1 else F.leaky_relu(z, SLOPE))
            elif kind == "down":
                D, H, Wd = a.shape[2:]
                ap = F.pad(a, (0, Wd % 2, 0, H % 2, 0, D % 2))
                t.append(F.leaky_relu(F.conv3d(ap, W, b, stride=2), SLOPE))
            else:
                fine = t[skip].shape[2:]
                y = F.conv_transpose3d(a, W.transpose(0, 1), b, stride=2)
                t.append(y[:, :, :fine[0], :fine[1], :fine[2]] + t[skip])
    return _nc_to_grid(t[-1])


def unet_apply(params: torch.Tensor, r9: torch.Tensor, tp: torch.Tensor,
               widths=UNET_WIDTHS) -> torch.Tensor:
    """z = r + N((s / t_max) T0 r) / s with the four-level network."""
    n = r9.numel()
    s = (n ** 0.5) / float(torch.linalg.vector_norm(r9))
    x = network_input(r9, tp, s)
    out, _ = unet_forward(params, x, bf16=True, widths=widths)
    corr = torch.view_as_complex(out.double().reshape(*out.shape[:-1], 9, 2).contiguous())
    return r9 + corr / s
