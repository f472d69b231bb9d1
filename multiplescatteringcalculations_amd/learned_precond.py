"""Learned 3-D convolutional preconditioner for the slab Foldy-Lax solve.

The reference solves (I - G0 T0) psi = psi0 with unpreconditioned GMRES
(cubic_scattering/slab_scattering.py:532-552, cubic_scattering/torch_gmres.py:
83-205).  This module adds a right preconditioner for flexible GMRES

    P(r) = r + N((s / t_max) T0 r) / s,    s = sqrt(n) / ||r||,

where N is a four-layer 3x3x3 convolutional network 18 -> h -> h -> h -> 18
over the voxel grid (channels: real and imaginary parts of the 9 components)
and t_max the largest modulus of the slab's T0 parameters.
Inference and training run on the GPU in libmsc_slab.so (BF16 MFMA with FP32
accumulation, Adam); the network is trained on residual -> correction pairs
harvested from the plain solver's own Krylov bases: for a basis vector v of a
plain GMRES cycle the pair is (input (I - G0 T0) v, target v).  Plain GMRES
stays the fallback: a flexible cycle that stalls hands over to plain cycles,
and every solve is judged by the same FP64 true residual.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

IO_CHANNELS = 18
PAD_CHANNELS = 32
UNET_WIDTHS = (16, 16, 32, 48)


def layers(arch: str = "local", hidden: int = 16, widths=UNET_WIDTHS) -> list[tuple[str, int, int]]:
    """The layer list (kind, in channels, out channels) in forward order:
    kind "conv" (3x3x3), "down" (2x2x2 stride 2) or "up" (2x2x2 stride-2
    transposed, added to the skip of the finer level)."""
    io = IO_CHANNELS
    if arch == "local":
        return [("conv", io, hidden), ("conv", hidden, hidden), ("conv", hidden, hidden),
                ("conv", hidden, io)]
    if arch != "unet":
        raise ValueError(f"arch must be 'local' or 'unet', got '{arch}'")
    c0, c1, c2, c3 = widths
    return [("conv", io, c0), ("conv", c0, c0), ("down", c0, c1), ("conv", c1, c1),
            ("down", c1, c2), ("conv", c2, c2), ("down", c2, c3), ("conv", c3, c3),
            ("up", c3, c2), ("conv", c2, c2), ("up", c2, c1), ("conv", c1, c1),
            ("up", c1, c0), ("conv", c0, c0), ("conv", c0, io)]


def _taps(kind: str) -> int:
    return 27 if kind == "conv" else 8


def num_params(hidden: int = 16, arch: str = "local", widths=UNET_WIDTHS) -> int:
    return sum(_taps(k) * ci * co + co for k, ci, co in layers(arch, hidden, widths))


def init_params(hidden: int = 16, seed: int = 0, last_scale: float = 0.1, arch: str = "local",
                widths=UNET_WIDTHS) -> np.ndarray:
    """He-normal weights (fan-in taps x in channels; linear up layers use the
    Xavier-like fan-in of their inputs), zero biases, last layer scaled so
    that the untrained preconditioner starts close to the identity."""
    rng = np.random.default_rng(seed)
    lay = layers(arch, hidden, widths)
    parts = []
    for l, (kind, ci, co) in enumerate(lay):
        fan = (_taps(kind) if kind != "up" else 1) * ci
        std = math.sqrt((2.0 if kind != "up" else 1.0) / fan)
        if l == len(lay) - 1:
            std *= last_scale
        parts.append(rng.standard_normal(_taps(kind) * ci * co) * std)
        parts.append(np.zeros(co))
    return np.concatenate(parts).astype(np.float32)


class LearnedPreconditioner:
    """Handle of a device-resident network (msc_net_* of include/msc_precond.h)."""

    def __init__(self, hidden: int = 48, device: torch.device | str | None = None,
                 seed: int | None = 0, arch: str = "local", widths=UNET_WIDTHS):
        self.device = torch.device(device) if device is not None else _lib.require_device()
        if self.device.type != "cuda":
            raise ValueError("the learned preconditioner runs on a HIP device only")
        if arch not in ("local", "unet"):
            raise ValueError(f"arch must be 'local' or 'unet', got '{arch}'")
        self.lib = _lib.load()
        self.arch = arch
        self.widths = tuple(int(c) for c in widths)
        self.hidden = int(hidden) if arch == "local" else self.widths[0]
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            if arch == "local":
                _lib.check(self.lib.msc_net_create(ctypes.byref(h), self.hidden))
            else:
                c4 = (ctypes.c_int * 4)(*self.widths)
                _lib.check(self.lib.msc_net_create_unet(ctypes.byref(h), c4))
        self._h = h
        n = ctypes.c_longlong()
        _lib.check(self.lib.msc_net_num_params(self._h, ctypes.byref(n)))
        self.num_params = int(n.value)
        self.gate_open = True  # see validate()
        if seed is not None:
            self.set_params(init_params(self.hidden, seed, arch=arch, widths=self.widths))

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def _stream(self):
        return _lib.stream_ptr(self.device)

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.msc_net_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ── parameters ──────────────────────────────────────────────────────────
    def params(self) -> torch.Tensor:
        out = torch.empty(self.num_params, dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_net_get_params(self._h, _lib.ptr(out), self._stream()))
        return out

    def set_params(self, params) -> None:
        p = torch.as_tensor(np.asarray(params, dtype=np.float32) if not isinstance(params, torch.Tensor)
                            else params, dtype=torch.float32).to(self.device).contiguous()
        if p.numel() != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.numel()}")
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_net_set_params(self._h, _lib.ptr(p), self._stream()))
            torch.cuda.current_stream(self.device).synchronize()

    def save(self, path) -> None:
        np.savez(path, hidden=self.hidden, arch=self.arch, widths=np.array(self.widths),
                 params=self.params().cpu().numpy())

    @classmethod
    def load(cls, path, device=None) -> "LearnedPreconditioner":
        with np.load(path, allow_pickle=False) as f:
            arch = str(f["arch"]) if "arch" in f else "local"
            widths = tuple(int(c) for c in f["widths"]) if "widths" in f else UNET_WIDTHS
            net = cls(int(f["hidden"]), device=device, seed=None, arch=arch, widths=widths)
            net.set_params(f["params"])
        return net

    # ── network on raw grids ────────────────────────────────────────────────
    def forward(self, x: torch.Tensor) -> torch.Tensor:
        """N(x) for a BF16 input [D, H, W, 32] -> float32 [D, H, W, 18]."""
        D, H, W, C = x.shape
        if C != PAD_CHANNELS or x.dtype != torch.bfloat16:
            raise ValueError("input must be bfloat16 [D, H, W, 32]")
        x = x.to(self.device).contiguous()
        out = torch.empty((D, H, W, IO_CHANNELS), dtype=torch.float32, device=self.device)
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_net_forward(self._h, D, H, W, _lib.ptr(x), _lib.ptr(out),
                                                self._stream()))
        return out

    def gradients(self, x: torch.Tensor, target: torch.Tensor) -> tuple[torch.Tensor, float]:
        """Gradient of mean((N(x) - target)^2) over samples x [B, D, H, W, 32] (bf16),
        target [B, D, H, W, 18] (float32)."""
        x, target, dims = self._batch(x, target)
        g = torch.empty(self.num_params, dtype=torch.float32, device=self.device)
        loss = ctypes.c_double()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_net_gradients(self._h, *dims, _lib.ptr(x), _lib.ptr(target),
                                                  _lib.ptr(g), ctypes.byref(loss), self._stream()))
        return g, loss.value

    def train_step(self, x: torch.Tensor, target: torch.Tensor, lr: float = 1e-3) -> float:
        """One Adam step; returns the loss before the step."""
        x, target, dims = self._batch(x, target)
        loss = ctypes.c_double()
        with torch.cuda.device(self.device):
            _lib.check(self.lib.msc_net_train_step(self._h, *dims, _lib.ptr(x), _lib.ptr(target),
                                                   float(lr), ctypes.byref(loss), self._stream()))
        return loss.value

    def validate(self, plan, b: torch.Tensor, *, repeats: int = 3, **solve_kw) -> dict:
        """Decide whether this network pays for itself on the plan's workload:
        solve ``plan`` x = b with the solver options ``solve_kw`` (default
        GMRES(30), mixed precision, tol 1e-8) with and without the network,
        best of ``repeats`` timed solves each, and open the gate
        (``self.gate_open``) only when flexible GMRES with the network is faster
        than the same solve without it.  The slab should be one held out from
        training.  Returns the measurements."""
        import time

        kw = dict(tol=1e-8, max_iter=3000, restart=30, precision="mixed")
        kw.update(solve_kw)
        out = {}
        for label, pc in (("without", None), ("with", self)):
            best, st = None, None
            for _ in range(repeats + 1):  # the first solve warms caches and graphs
                torch.cuda.synchronize(plan.device)
                t0 = time.perf_counter()
                _, st = plan.gmres(b, "born", precond=pc, **kw)
                torch.cuda.synchronize(plan.device)
                dt = time.perf_counter() - t0
                best = dt if best is None or dt < best else best
            out[label + "_ms"] = best * 1e3
            out[label + "_iterations"] = int(st["iterations"])
        self.gate_open = out["with_ms"] < out["without_ms"]
        out["gate_open"] = self.gate_open
        return out

    def _batch(self, x, target):
        if x.dim() == 4:
            x = x.unsqueeze(0)
        if target.dim() == 4:
            target = target.unsqueeze(0)
        B, D, H, W, C = x.shape
        if C != PAD_CHANNELS or x.dtype != torch.bfloat16:
            raise ValueError("inputs must be bfloat16 [B, D, H, W, 32]")
        if tuple(target.shape) != (B, D, H, W, IO_CHANNELS) or target.dtype != torch.float32:
            raise ValueError("targets must be float32 [B, D, H, W, 18]")
        return (x.to(self.device).contiguous(), target.to(self.device).contiguous(),
                (D, H, W, B))

    # ── as a preconditioner of a slab plan ──────────────────────────────────
    def apply(self, plan, r: torch.Tensor) -> torch.Tensor:
        """z = r + N((s / t_max) T0 r) / s on the plan's grid (complex128 or complex64)."""
        prec = {torch.complex128: _lib.MSC_C128, torch.complex64: _lib.MSC_C64}.get(r.dtype)
        if prec is None:
            raise ValueError("r must be complex128 or complex64")
        r = r.reshape(-1).to(plan.device).contiguous()
        if r.numel() != plan.n:
            raise ValueError(f"r must have {plan.n} entries")
        z = torch.empty_like(r)
        with torch.cuda.device(plan.device):
            _lib.check(self.lib.msc_net_apply(self._h, plan._h, _lib.ptr(r), _lib.ptr(z), prec,
                                              plan.stream))
        return z


def network_input(plan, r: torch.Tensor, scale: float | None = None) -> torch.Tensor:
    """(s / t_max) T0 r as the network's BF16 [N_z, M, M, 32] input (s = sqrt(n)/||r|| by
    default)."""
    prec = {torch.complex128: _lib.MSC_C128, torch.complex64: _lib.MSC_C64}[r.dtype]
    g = plan.geometry
    out = torch.empty((g.N_z, g.M, g.M, PAD_CHANNELS), dtype=torch.bfloat16, device=plan.device)
    r = r.reshape(-1).contiguous()
    with torch.cuda.device(plan.device):
        _lib.check(plan.lib.msc_net_input(plan._h, _lib.ptr(r), prec,
                                          float(scale) if scale else -1.0, _lib.ptr(out),
                                          plan.stream))
    return out


@dataclass
class PairSet:
    """Training pairs on one grid: inputs bf16 [P, N_z, M, M, 32], targets
    float32 [P, N_z, M, M, 18] (the scaled correction s (v - A v))."""
    inputs: torch.Tensor
    targets: torch.Tensor
    sources: list = field(default_factory=list)

    def __len__(self) -> int:
        return int(self.inputs.shape[0])


def harvest_pairs(plan, b: torch.Tensor, *, restart: int = 30, tol: float = 1e-8,
                  max_pairs: int | None = None, precond: "LearnedPreconditioner | None" = None
                  ) -> PairSet:
    """Run one FP64 GMRES(restart) cycle on (I - G0 T0) x = b -- plain, or
    flexible with the right preconditioner ``precond`` -- and turn its
    orthonormal Krylov basis v_0..v_k into pairs.  With z = v (plain) or
    z = P(v) (flexible) and w = A z, the pair is (input (s / t_max) T0 w,
    target s (z - w)), s = sqrt(n)/||w||: the network learns the correction
    that maps the residual w back to the step z that produces it.  Pairs from
    flexible cycles follow the distribution of vectors the preconditioner
    meets inside flexible GMRES."""
    _, st = plan.gmres(b, "born", tol=tol, max_iter=max(restart, 1), restart=restart,
                       precision="fp64", precond=precond)
    k = min(st["iterations"], restart)
    g = plan.geometry
    shape = (g.N_z, g.M, g.M)
    ins, tgts = [], []
    v = torch.empty(plan.n, dtype=torch.complex128, device=plan.device)
    for j in range(k + 1):
        if max_pairs is not None and len(ins) >= max_pairs:
            break
        with torch.cuda.device(plan.device):
            _lib.check(plan.lib.msc_gmres_basis(plan._h, _lib.MSC_C128, j, _lib.ptr(v),
                                                plan.stream))
        z = v if precond is None else precond.apply(plan, v)
        w = plan.matvec(z)
        nw = float(torch.linalg.vector_norm(w))
        if not nw > 0.0:
            continue
        s = math.sqrt(plan.n) / nw
        ins.append(network_input(plan, w, s))
        corr = (z - w) * s
        tgts.append(torch.view_as_real(corr).reshape(*shape, IO_CHANNELS).float())
    if not ins:
        raise RuntimeError("no Krylov vectors harvested")
    return PairSet(torch.stack(ins), torch.stack(tgts), [st])


def train_on_slabs(net: "LearnedPreconditioner", plan, draw, omega: float, ref, k_hat,
                   wave_type: str = "P", *, target_pairs: int = 20000, seeds_per_round: int = 32,
                   first_seed: int = 1, batch: int = 4, lr: float = 2e-3, lr_min: float = 1e-4,
                   flexible_rounds: int = 2, restart: int = 30, log=None) -> dict:
    """Train ``net`` on residual -> correction pairs streamed from solves on
    seeded slabs of one distribution (``draw(seed)`` returns a SlabMaterial of
    the plan's geometry; the plan's G0 is that of ``omega``), without holding
    them all: each round harvests one GMRES(restart) cycle per slab of
    ``seeds_per_round`` slabs, takes one pass of Adam steps over those pairs
    (mini-batches of ``batch``, cosine learning-rate decay from ``lr`` to
    ``lr_min`` over the whole run) and drops them.  Rounds continue until
    ``target_pairs`` pairs have been used; the last ``flexible_rounds`` rounds
    harvest from flexible GMRES cycles with the network being trained.  The
    plan's T0 is left at the last slab drawn."""
    import time

    if target_pairs < 1 or seeds_per_round < 1 or batch < 1:
        raise ValueError("target_pairs, seeds_per_round and batch must be >= 1")
    expect_per_round = seeds_per_round * (restart + 1)
    rounds = max(1, -(-target_pairs // expect_per_round))
    total_steps = max(1, -(-target_pairs // batch))
    step = 0
    used = 0
    seed = first_seed
    losses = []
    t0 = time.perf_counter()
    t_harvest = 0.0
    rng = np.random.default_rng(first_seed)
    for rd in range(rounds):
        flexible = rd >= rounds - flexible_rounds and rd > 0
        th = time.perf_counter()
        sets = []
        for _ in range(seeds_per_round):
            plan.set_material(draw(seed), omega)
            seed += 1
            b = plan.incident_field(omega, ref, k_hat, wave_type).reshape(-1)
            sets.append(harvest_pairs(plan, b, restart=restart, tol=1e-14,
                                      precond=net if flexible else None))
        pairs = merge_pairs(sets)
        del sets
        torch.cuda.synchronize(plan.device)
        t_harvest += time.perf_counter() - th
        n = len(pairs)
        order = torch.as_tensor(rng.permutation(n), device=pairs.inputs.device)
        for i0 in range(0, n - batch + 1, batch):
            frac = min(step / total_steps, 1.0)
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * frac))
            idx = order[i0:i0 + batch]
            losses.append(net.train_step(pairs.inputs[idx], pairs.targets[idx], lr_t))
            step += 1
        used += n
        if log is not None:
            log(f"  round {rd + 1}/{rounds} ({'flexible' if flexible else 'plain'}): {n} pairs "
                f"({used} so far), loss {np.mean(losses[-max(1, n // batch):]):.4e}")
        del pairs
        if used >= target_pairs:
            break
    torch.cuda.synchronize(plan.device)
    return {"pairs": used, "rounds": rd + 1, "steps": step, "slabs": seed - first_seed,
            "seconds": time.perf_counter() - t0, "harvest_seconds": t_harvest,
            "loss_first": float(np.mean(losses[:20])) if losses else None,
            "loss_last": float(np.mean(losses[-20:])) if losses else None}


def merge_pairs(sets: list[PairSet]) -> PairSet:
    return PairSet(torch.cat([s.inputs for s in sets]), torch.cat([s.targets for s in sets]),
                   [x for s in sets for x in s.sources])


def train(net: LearnedPreconditioner, pairs: PairSet, *, steps: int = 200, lr: float = 1e-3,
          batch: int = 4, seed: int = 0, holdout: int = 0, log_every: int = 0):
    """Adam on mean-square correction error over random mini-batches of the
    pairs (the last `holdout` pairs are kept out and scored at the end).
    Returns (train losses, holdout loss before, holdout loss after)."""
    n = len(pairs)
    ntr = n - holdout
    if ntr < 1:
        raise ValueError("no training pairs left after the holdout")
    rng = np.random.default_rng(seed)

    def hold_loss():
        if holdout == 0:
            return None
        _, l = net.gradients(pairs.inputs[ntr:], pairs.targets[ntr:])
        return l

    before = hold_loss()
    losses = []
    for it in range(steps):
        idx = torch.as_tensor(rng.choice(ntr, size=min(batch, ntr), replace=False),
                              device=pairs.inputs.device)
        losses.append(net.train_step(pairs.inputs[idx], pairs.targets[idx], lr))
        if log_every and (it % log_every == 0 or it == steps - 1):
            print(f"  step {it:5d}  loss {losses[-1]:.4e}", flush=True)
    return losses, before, hold_loss()
