"""Benchmark: seconds per solve to a 1e-8 FP64 relative residual on the
128 x 128 x 32 slab (padded FFT grid 64 x 256 x 256), one MI355X.

Workload (synthetic, seeded): background Vp 3000 m/s, Vs 1732 m/s, rho 2500
kg/m^3; cube half-width a = 1 m; Bernoulli(0.3) inclusions with i.i.d.
relative contrasts U[-0.2, 0.2] in lambda, mu, rho; ka = k_S a = 0.4; oblique
P plane wave with horizontal slowness p = 0.15 s/km.  One "step" is one full
GMRES(30) solve of (I - G0 T0) psi = psi0 from the Born guess to
||b - A psi|| / ||b|| <= 1e-8 in FP64, with the operator (G0 table, T0) and
psi0 already resident in HBM.

The headline (``value``) is the learned + mixed solver: GMRES(30) with
complex64 cycles inside FP64 iterative refinement, and the learned four-level
convolutional preconditioner in flexible GMRES when its validation gate is
open.  The network is trained on the GPU before the timed region, on pairs
streamed from GMRES cycles on other seeded slabs of the same distribution, and
validated on a held-out slab: the gate opens only if the solve with the network
beats the same solve without it.  The plain FP64 GMRES(30) solve, the mixed
solve and the solve with the network forced on are timed the same way and
reported beside it; the strong-contrast 96 x 96 x 24 leg
(``detail.other_paths.learned_strong_contrast``) is where the network pays.

Prints one JSON line on stdout; progress goes to stderr.
"""

from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np


def log(*a):
    print(*a, file=sys.stderr, flush=True)


M, NZ, A = 128, 32, 1.0
VP, VS, RHO = 3000.0, 1732.0, 2500.0
PHI, CONTRAST, KA, P_SLOW = 0.3, 0.2, 0.4, 1.5e-4
TOL, RESTART = 1e-8, 30
UNET_WIDTHS, TRAIN_PAIRS, TRAIN_BATCH, TRAIN_LR = (16, 16, 32, 48), 4000, 4, 1e-3


def setup(device):
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd.slab_scattering import SlabPlan

    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=M, N_z=NZ, a=A)
    mat = msc.random_iid_slab_material(g, ref, PHI, CONTRAST, seed=0)
    omega = KA * VS / A
    khat = np.array([np.sqrt(1.0 - (P_SLOW * VP) ** 2), P_SLOW * VP, 0.0])
    plan = SlabPlan(g, device=device)
    plan.build_g0(omega, ref)
    plan.set_material(mat, omega)
    psi0 = plan.incident_field(omega, ref, khat, "P").reshape(-1)
    return plan, psi0, mat, omega


def train_preconditioner(plan, mat0, omega, device):
    """The learned part of the headline solver: the four-level network trained on
    the GPU on residual -> correction pairs streamed from GMRES(30) cycles on
    seeded slabs 1, 2, ... of the benchmark distribution (the timed slab is
    seed 0), then validated on a held-out slab (seed 10000): the gate opens only
    if flexible GMRES with the network beats the same mixed solve without it.
    The plan's G0 is shared (same geometry and frequency); the timed slab's T0
    is restored afterwards."""
    import torch

    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import learned_precond as lp

    ref = msc.ReferenceMedium(VP, VS, RHO)
    khat = np.array([np.sqrt(1.0 - (P_SLOW * VP) ** 2), P_SLOW * VP, 0.0])

    def draw(seed):
        return msc.random_iid_slab_material(plan.geometry, ref, PHI, CONTRAST, seed=seed)

    net = lp.LearnedPreconditioner(device=device, seed=0, arch="unet", widths=UNET_WIDTHS)
    info = lp.train_on_slabs(net, plan, draw, omega, ref, khat, target_pairs=TRAIN_PAIRS,
                             seeds_per_round=16, batch=TRAIN_BATCH, lr=TRAIN_LR, log=log)
    plan.set_material(draw(10000), omega)
    val = net.validate(plan, plan.incident_field(omega, ref, khat, "P").reshape(-1), tol=TOL,
                       restart=RESTART)
    plan.set_material(mat0, omega)
    torch.cuda.synchronize()
    info = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in info.items()}
    info.update({"network": f"unet {UNET_WIDTHS}", "params": net.num_params,
                 "validation_held_out_seed_10000": {k: (round(v, 3) if isinstance(v, float) else v)
                                                    for k, v in val.items()}})
    log(f"learned preconditioner: {info}")
    return net, info


def time_solves(plan, psi0, precision, steps, warmup, precond=None, label=None):
    """K timed solves, after W untimed ones."""
    import torch

    label = label or precision
    st = None
    for _ in range(warmup):
        _, st = plan.gmres(psi0, "born", tol=TOL, max_iter=600, restart=RESTART, precision=precision,
                           precond=precond)
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, st = plan.gmres(psi0, "born", tol=TOL, max_iter=600, restart=RESTART, precision=precision,
                           precond=precond)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        log(f"  {label}: {times[-1] * 1e3:.2f} ms  iters={st['iterations']} "
            f"cycles={st['cycles']} rel_res={st['rel_residual']:.2e} fell_back={st['fell_back']}")
    return times, st, x


def rpp_of(plan, x, mat, omega):
    """Specular R_PP (slab_rpp_periodic) of a device solution of the timed slab."""
    from multiplescatteringcalculations_amd.slab_scattering import SlabResult, slab_rpp_periodic

    g = plan.geometry
    res = SlabResult(psi=None, psi0=None, geometry=g, material=mat, omega=omega,
                     k_hat=np.array([np.sqrt(1.0 - (P_SLOW * VP) ** 2), P_SLOW * VP, 0.0]),
                     wave_type="P", n_gmres_iter=0, gmres_residual=0.0)
    layer_sum, _ = plan.source_sums(x, 0.0, 0.0, np.zeros(3))
    res._sums = (layer_sum, None, (0.0, 0.0), np.zeros(3))
    return slab_rpp_periodic(res, p=P_SLOW)


def working_launches(durations_ms):
    """(count, mean ms) of the launches that did the pass's work: a launch
    queued after its cycle converged finds the done flag raised and returns at
    once (a few microseconds), so launches shorter than a tenth of the longest
    are left out (they move no data)."""
    if not durations_ms:
        return 0, 0.0
    cut = 0.1 * max(durations_ms)
    work = [d for d in durations_ms if d >= cut]
    return len(work), float(np.mean(work))


def time_pass3(plan, psi0, precision, steps):
    """Per-launch duration of the block-multiply pass from HIP events recorded
    around each launch on its stream, over `steps` solves:
    {"c64"|"c128": {"launches", "working", "no_op", "avg_ms"}}, the mean over
    the working launches (working_launches)."""
    import torch

    plan.timing(True)
    for _ in range(steps):
        plan.gmres(psi0, "born", tol=TOL, max_iter=600, restart=RESTART, precision=precision)
    torch.cuda.synchronize()
    out = {}
    for key in ("c64", "c128"):
        d = plan.timing_durations(key)
        nw, avg = working_launches(d)
        out[key] = {"launches": len(d), "working": nw, "no_op": len(d) - nw, "avg_ms": avg}
    plan.timing(False)
    return out


def sweep_cases():
    """The batched-sweep workload: 64 frequencies ka = 0.1..1.0 x 4 slownesses
    p = 0, 0.1, 0.2, 0.3 s/km, each on its own seeded slab of uneven size
    (lateral 32-128, depth 4-32 voxels), phi 0.3, contrasts +-0.2."""
    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(VP, VS, RHO)
    rng = np.random.default_rng(0)
    cases = []
    for i, ka in enumerate(np.linspace(0.1, 1.0, 64)):
        for j, p in enumerate((0.0, 1e-4, 2e-4, 3e-4)):
            g = msc.SlabGeometry(M=int(rng.integers(32, 129)), N_z=int(rng.integers(4, 33)), a=A)
            mat = msc.random_iid_slab_material(g, ref, PHI, CONTRAST, seed=1000 * i + j)
            khat = np.array([np.sqrt(1.0 - (p * VP) ** 2), p * VP, 0.0])
            cases.append(msc.SweepCase(g, mat, float(ka) * VS / A, khat, "P"))
    return cases


def time_sweep(mode: str = "batched", workers: int = 4):
    """Two timed passes of the batched sweep (G0 tables, T0, incident fields
    and solves inside the timed region; inputs already on the host).  The
    first pass also creates the plans of the grids the 8-case warm-up did not
    meet; the second ("repeat") finds the batched mode's plan pools cached.
    mode "batched": the slabs of one padded grid go through every GMRES pass
    together (msc_sweep_batched); "workers": one solve at a time per worker
    thread, concurrently (msc_sweep)."""
    import torch

    import multiplescatteringcalculations_amd as msc

    cases = sweep_cases()
    msc.compute_slab_sweep(cases[:8], gmres_tol=TOL, precision="mixed", workers=workers,
                           keep_on_device=True, mode=mode)  # warm-up (library, allocator)
    torch.cuda.synchronize()
    times = []
    for _ in range(2):  # the second pass finds the batched plan pools cached
        t0 = time.perf_counter()
        res = msc.compute_slab_sweep(cases, gmres_tol=TOL, precision="mixed", workers=workers,
                                     keep_on_device=True, mode=mode)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    dt = times[0]
    its = [r.n_gmres_iter for r in res]
    out = {"mode": mode, "solves": len(res), "seconds": round(dt, 3),
           "solves_per_s": round(len(res) / dt, 1), "seconds_repeat": round(times[1], 3),
           "solves_per_s_repeat": round(len(res) / times[1], 1),
           "workers": workers if mode == "workers" else None,
           "voxels": int(sum(c.geometry.n_cubes for c in cases)),
           "mean_iterations": round(float(np.mean(its)), 2), "max_iterations": int(max(its)),
           "all_converged": bool(all(r.stats["converged"] for r in res)),
           "max_rel_residual": float(max(r.gmres_residual for r in res))}
    log(f"sweep: {out}")
    return out


def z_kernel_name(code: int, real: str = "float") -> str:
    """Kernel of pass 3 that a z tile code selects (Q, + 256 persistent
    direct-to-LDS form, + 512 pruned form)."""
    q = code & 255
    form = "k_z_mult_pr" if code & 512 else ("k_z_mult_dma" if code & 256 else "k_z_mult")
    return f"{form}<{real},64,{q}>"


def measured_traffic(kernel: str):
    """HBM bytes per launch of the roofline kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (scaled by the same counters of a copy of
    known size); None when no summary for this kernel and workload exists."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "k_z_mult_c64_128x128x32_traffic.txt")
    if not os.path.exists(path):
        return None
    cur, vals = None, {}
    with open(path) as f:
        for line in f:
            if line.startswith("kernel ="):
                cur = line.split("=", 1)[1].strip()
            elif line.startswith("traffic_bytes_per_launch") and cur is not None:
                vals[cur] = int(line.split("=")[1])
    return vals.get(kernel)


def time_resonance(n_sub: int = 4, count: int = 1024):
    """Resonance-regime cube T-matrices (internal Foldy-Lax over n_sub^3 sub-cells)
    of `count` random contrasts (+-0.3) at ka = 1.5 in one device call."""
    import torch

    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(VP, VS, RHO)
    rng = np.random.default_rng(5)
    lam, mu = ref.rho * (VP ** 2 - 2 * VS ** 2), ref.rho * VS ** 2
    Dl, Dm, Dr = (rng.uniform(-0.3, 0.3, count) * s for s in (lam, mu, ref.rho))
    om = 1.5 * VS / A
    msc.compute_resonance_tmatrices(om, A, ref, Dl[:4], Dm[:4], Dr[:4], n_sub=n_sub)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    T, ratio = msc.compute_resonance_tmatrices(om, A, ref, Dl, Dm, Dr, n_sub=n_sub,
                                               keep_on_device=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n = 9 * n_sub ** 3
    flops = 8.0 / 3.0 * n ** 3
    out = {"n_sub": n_sub, "cubes": count, "seconds": round(dt, 4),
           "us_per_cube": round(dt / count * 1e6, 2),
           "fp64_tflops": round(flops * count / dt / 1e12, 2), "fp64_peak_tflops": 78.6,
           "max_pivot_ratio": float(np.max(ratio))}
    log(f"resonance: {out}")
    return out


STRONG = {"M": 96, "N_z": 24, "phi": 0.5, "contrast": 0.4, "ka": 0.75, "pairs": 20000,
          "seeds_per_round": 32, "batch": 4, "lr": 1e-3, "widths": (16, 16, 32, 48)}


def time_strong_contrast(pairs: int | None = None):
    """The learned preconditioner where plain GMRES needs many iterations:
    96x96x24 voxels, Bernoulli(0.5) inclusions with contrasts U[-0.4, 0.4],
    ka 0.75, p 0.15 s/km.  The four-level network (about 200k parameters) is
    trained on the GPU on residual -> correction pairs streamed from GMRES(30)
    cycles on seeded slabs 1, 2, ... of the same distribution (the last two
    rounds from flexible cycles with the network), then plain FP64, mixed and
    learned + mixed GMRES(30) solve the held-out slabs (seeds 0 and 100000) to
    the FP64 residual 1e-8 (best of 3 timed solves each)."""
    import torch

    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import learned_precond as lp
    from multiplescatteringcalculations_amd.slab_scattering import SlabPlan

    cfg = dict(STRONG)
    if pairs is not None:
        cfg["pairs"] = pairs
    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=cfg["M"], N_z=cfg["N_z"], a=A)
    om = cfg["ka"] * VS / A
    khat = np.array([np.sqrt(1.0 - (P_SLOW * VP) ** 2), P_SLOW * VP, 0.0])
    plan = SlabPlan(g)
    plan.build_g0(om, ref)

    def draw(seed):
        return msc.random_iid_slab_material(g, ref, cfg["phi"], cfg["contrast"], seed=seed)

    net = lp.LearnedPreconditioner(seed=0, arch="unet", widths=cfg["widths"])
    info = lp.train_on_slabs(net, plan, draw, om, ref, khat, target_pairs=cfg["pairs"],
                             seeds_per_round=cfg["seeds_per_round"], batch=cfg["batch"],
                             lr=cfg["lr"], log=log)
    plan.set_material(draw(200000), om)
    val = net.validate(plan, plan.incident_field(om, ref, khat, "P").reshape(-1), tol=TOL,
                       restart=RESTART)
    out = {"slab": "96x96x24, phi 0.5, contrasts +-0.4, ka 0.75, p 0.15 s/km",
           "network": f"unet {cfg['widths']}", "params": net.num_params,
           "training": {k: (round(v, 4) if isinstance(v, float) else v) for k, v in info.items()},
           "validation_held_out_seed_200000": {k: (round(v, 3) if isinstance(v, float) else v)
                                              for k, v in val.items()}}
    # the network's own cost per application (complex64, as in the solve)
    r = torch.randn(plan.n, dtype=torch.complex64, device=plan.device)
    for _ in range(5):
        net.apply(plan, r)
    torch.cuda.synchronize()
    for seed in (0, 100000):
        plan.set_material(draw(seed), om)
        b = plan.incident_field(om, ref, khat, "P").reshape(-1)
        row = {}
        xs = {}
        for label, kw in (("fp64", dict(precision="fp64")), ("mixed", dict(precision="mixed")),
                          ("learned_mixed", dict(precision="mixed", precond=net))):
            best = None
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                x, st = plan.gmres(b, "born", tol=TOL, max_iter=3000, restart=RESTART, **kw)
                torch.cuda.synchronize()
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            xs[label] = x
            row[label + "_ms"] = round(best * 1e3, 2)
            row[label + "_iterations"] = int(st["iterations"])
            row[label + "_rel_residual"] = float(st["rel_residual"])
            if label == "learned_mixed":
                row["learned_fell_back"] = int(st["fell_back"])
        row["learned_vs_fp64_max_rel_diff"] = float((xs["learned_mixed"] - xs["fp64"]).abs().max()
                                                    / xs["fp64"].abs().max())
        row["learned_speedup_vs_mixed"] = round(row["mixed_ms"] / row["learned_mixed_ms"], 3)
        row["learned_speedup_vs_fp64"] = round(row["fp64_ms"] / row["learned_mixed_ms"], 3)
        out[f"seed_{seed}"] = row
    log(f"strong contrast: {out}")
    return out


RESONANCE_SLAB = {"M": 64, "N_z": 16, "n_sub": 4, "phi": 0.3, "contrast": 0.3,
                  "ka": (1.0, 2.0), "frequencies": 16, "max_iter": 1000}


def time_resonance_slab():
    """The resonance-regime ocean bottom: a 64x64x16 slab (Bernoulli(0.3)
    inclusions, contrasts U[-0.3, 0.3]) whose every cube carries its full-wave
    composite T (4^3 sub-cells), between water (1500 m/s, 1000 kg/m^3) and an
    elastic half-space (4000 / 2300 m/s, 2600 kg/m^3), normal incidence, 16
    frequencies ka 1-2.  Per frequency: the composite T-matrices (one batched
    device solve), the mixed GMRES(30) slab solve (stopped at 1000
    iterations), the same medium resolved into 256x256x64 sub-cells (Rayleigh
    T0 at ka/4) as the cross-check, the embedding, and the Kennett R_PP of the
    CPA medium of the contrast distribution (27-point Gauss-Legendre phases)."""
    import torch

    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd import resonance_slab as rs

    cfg = RESONANCE_SLAB
    ref = msc.ReferenceMedium(VP, VS, RHO)
    g = msc.SlabGeometry(M=cfg["M"], N_z=cfg["N_z"], a=A)
    mat = msc.random_iid_slab_material(g, ref, cfg["phi"], cfg["contrast"], seed=0)
    oms = np.linspace(*cfg["ka"], cfg["frequencies"]) * VS / A
    phases = rs.uniform_contrast_phases(ref, cfg["phi"], cfg["contrast"])
    # warm-up: kernels of both grids and the resonance solver
    rs.resonance_ocean_bottom(g, mat, oms[:1], n_sub=cfg["n_sub"], max_iter=30, subcell=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = rs.resonance_ocean_bottom(g, mat, oms, n_sub=cfg["n_sub"], cpa_phases=phases,
                                  gmres_tol=TOL, precision="mixed", restart=RESTART,
                                  max_iter=cfg["max_iter"], subcell=True)
    total = time.perf_counter() - t0
    both = [i for i in range(len(oms)) if r.converged[i] and r.subcell_converged[i]]
    rel = [abs(r.R_slab[i] - r.R_slab_subcell[i]) / abs(r.R_slab_subcell[i]) for i in both]
    out = {"slab": "64x64x16, n_sub 4, phi 0.3, contrasts +-0.3, ka 1-2 (16), p 0, "
                   "water 1500/1000 over 4000/2300/2600",
           "total_s": round(total, 3),
           "tmatrix_s": round(float(np.sum(r.tmatrix_seconds)), 3),
           "distinct_tmatrices_per_frequency": int(np.mean(r.distinct_tmatrices)),
           "composite_solve_s": round(float(np.sum(r.solve_seconds) - np.sum(r.tmatrix_seconds)), 3),
           "subcell_solve_s": round(float(np.sum(r.subcell_seconds)), 3),
           "ka": [round(float(k), 4) for k in r.ka],
           "composite_iterations": r.iterations, "composite_converged": r.converged,
           "subcell_iterations": r.subcell_iterations, "subcell_converged": r.subcell_converged,
           "R_total_composite": [[round(v.real, 6), round(v.imag, 6)] for v in r.R_total],
           "R_total_subcell": [[round(v.real, 6), round(v.imag, 6)] for v in r.R_total_subcell],
           "R_cpa_kennett": [[round(v.real, 6), round(v.imag, 6)] for v in r.R_cpa],
           "R_bg_kennett": [[round(v.real, 6), round(v.imag, 6)] for v in r.R_bg],
           "R_slab_composite_vs_subcell_rel": [round(float(v), 4) for v in rel],
           "max_abs_R_total_subcell_minus_cpa": round(float(np.max(np.abs(r.R_total_subcell
                                                                           - r.R_cpa))), 5),
           "max_abs_R_total_subcell_minus_bg": round(float(np.max(np.abs(r.R_total_subcell
                                                                          - r.R_bg))), 5)}
    log(f"resonance slab: {out}")
    return out


def time_sphere(n_sub: int = 32):
    """Voxelised-sphere Foldy-Lax T-matrix (9 right-hand sides, FP64 GMRES to 1e-8)
    with n_sub cells per edge, ka_sphere = 2: the 9 right-hand sides as one
    batch (default) and as 3 concurrent worker solves; median of 3, set-up
    (plans, G0) included."""
    import torch

    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(VP, VS, RHO)
    R = n_sub * A
    om = 2.0 * VS / R
    c = msc.MaterialContrast(0.2 * ref.rho * (VP ** 2 - 2 * VS ** 2), 0.2 * ref.rho * VS ** 2,
                             0.2 * ref.rho)
    out = {"n_sub": n_sub}
    for label, workers in (("seconds_9_rhs", 0), ("seconds_9_rhs_3_workers", 3)):
        msc.compute_sphere_foldy_lax_fft_gpu(om, R, ref, c, 4, gmres_tol=TOL,
                                             workers=workers)  # warm-up
        ts = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res = msc.compute_sphere_foldy_lax_fft_gpu(om, R, ref, c, n_sub, gmres_tol=TOL,
                                                       restart=30, workers=workers)
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t0)
        out[label] = round(float(np.median(ts)), 4)
        out["cells"] = int(res.n_cells)
    log(f"sphere: {out}")
    return out


def host_info():
    """CPU model, core counts and thread settings of the host the CPU legs run on."""
    import torch

    model = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OPENBLAS_NUM_THREADS": os.environ.get("OPENBLAS_NUM_THREADS"),
            "torch_threads": torch.get_num_threads()}


def _median_time(fn, repeats=3):
    fn()  # warm-up
    ts = []
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), [round(t, 4) for t in ts]


def time_reference_gpu_formula(plan, n_matvec: int):
    """The reference's GPU operator (slab_scattering_gpu.py:95-155: T-apply,
    zero-padded 9-component grid, fftn, 9x9 spectral block product, ifftn,
    extract, psi - acc) restated in PyTorch and run on this GPU in complex128
    on the same padded 64x256x256 grid, 1 warm-up + 3 timed repeats (median),
    scaled by the plain FP64 solve's operator count like the CPU legs.  The
    spectral table is synthetic (the timing does not depend on its values).
    The spectral 9x9 product in two torch formulations, the faster one
    reported: batched matmuls on contiguous operands, and one broadcast
    multiply and sum over the (9, 9, points) table."""
    import torch

    from multiplescatteringcalculations_amd.slab_scattering import tparams_to_tmatrices

    dev = torch.device("cuda", torch.cuda.current_device())
    Pz, Pxy = plan.Pz, plan.Pxy
    npts = Pz * Pxy * Pxy
    khat = torch.full((npts, 9, 9), 1e-12 + 1e-13j, dtype=torch.complex128, device=dev)
    Tt = torch.as_tensor(tparams_to_tmatrices(plan.tparams().cpu().numpy())).to(dev).contiguous()
    gen = torch.Generator(device=dev).manual_seed(2)
    x = torch.randn(NZ * M * M * 9, dtype=torch.complex128, device=dev, generator=gen)

    def mv():
        psi = x.reshape(NZ, M, M, 9)
        tau = torch.matmul(Tt, psi.unsqueeze(-1)).squeeze(-1)
        pad = torch.zeros((9, Pz, Pxy, Pxy), dtype=torch.complex128, device=dev)
        pad[:, :NZ, :M, :M] = tau.permute(3, 0, 1, 2)
        f = torch.fft.fftn(pad, dim=(1, 2, 3)).reshape(9, npts).t().contiguous()
        acc = torch.bmm(khat, f.unsqueeze(-1)).squeeze(-1).t().reshape(9, Pz, Pxy, Pxy)
        acc = torch.fft.ifftn(acc, dim=(1, 2, 3))
        y = psi - acc[:, :NZ, :M, :M].permute(1, 2, 3, 0)
        torch.cuda.synchronize(dev)
        return y

    t_bmm, reps = _median_time(mv)
    del khat
    torch.cuda.empty_cache()
    khat9 = torch.full((9, 9, npts), 1e-12 + 1e-13j, dtype=torch.complex128, device=dev)

    def mv_bcast():
        psi = x.reshape(NZ, M, M, 9)
        tau = torch.matmul(Tt, psi.unsqueeze(-1)).squeeze(-1)
        pad = torch.zeros((9, Pz, Pxy, Pxy), dtype=torch.complex128, device=dev)
        pad[:, :NZ, :M, :M] = tau.permute(3, 0, 1, 2)
        f = torch.fft.fftn(pad, dim=(1, 2, 3)).reshape(1, 9, npts)
        acc = (khat9 * f).sum(dim=1).reshape(9, Pz, Pxy, Pxy)
        acc = torch.fft.ifftn(acc, dim=(1, 2, 3))
        y = psi - acc[:, :NZ, :M, :M].permute(1, 2, 3, 0)
        torch.cuda.synchronize(dev)
        return y

    t_bc, reps_bc = _median_time(mv_bcast)
    del khat9, Tt
    torch.cuda.empty_cache()
    t = min(t_bmm, t_bc)
    out = {"matvec_ms": round(t * 1e3, 3), "matvec_ms_batched_matmul": round(t_bmm * 1e3, 3),
           "matvec_ms_broadcast": round(t_bc * 1e3, 3), "repeats_s": reps + reps_bc,
           "n_matvec": int(n_matvec), "s_per_solve_extrapolated": round(t * n_matvec, 4),
           "precision": "complex128",
           "formula": "PyTorch fftn / spectral 9x9 product / ifftn on the padded grid "
                      "(slab_scattering_gpu.py:95-155), this GPU; the faster of two formulations"}
    log(f"reference GPU formula (torch): {out}")
    return out


def cpu_baseline(plan, n_matvec: int):
    """CPU legs, all on this host's cores, 1 warm-up + 3 timed repeats (median):

    * the oracle's restatement of the reference CPU matvec (slab_scattering.py:
      359-418: per-layer 2-D FFTs and the N_z^2 layer-pair 9x9 loop, NumPy,
      one core) on the 128 x 128 x 32 slab with its real per-voxel T0 (the
      device's, expanded to 9x9) and a spectral table of the reference's shape
      (2N_z-1, 2M-1, 2M-1, 9, 9) holding constant values (the loop's time does
      not depend on them; the oracle builds the real one in minutes of Python
      loops), scaled by the plain FP64 GPU solve's operator count: the
      reported s/solve (setup excluded, labelled extrapolated);
    * the full 32 x 32 x 8 solve of the oracle (ka 0.2, +-0.1, normal
      incidence; SciPy GMRES(30) to 1e-8 from the Born guess) with its real
      T0 and table: setup once, then the solve timed;
    * the torch-CPU 3-D-FFT formula of slab_scattering_gpu.py:95-155 (fftn,
      81-entry einsum, ifftn over the padded (9, 64, 256, 256) grid) at
      torch.get_num_threads() threads, per matvec, scaled the same way.
    """
    import torch

    from multiplescatteringcalculations_amd.slab_scattering import tparams_to_tmatrices
    from oracle import slab as osl
    from oracle import tmatrix as otm

    info = host_info()
    rng = np.random.default_rng(1)
    n = NZ * M * M * 9
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    H = 2 * M - 1
    kh = np.full((2 * NZ - 1, H, H, 9, 9), 1e-12 + 1e-13j, dtype=np.complex128)
    T = tparams_to_tmatrices(plan.tparams().cpu().numpy())
    t_np, reps_np = _median_time(lambda: osl.slab_matvec(x, T, kh, M, NZ))
    del kh
    log(f"  numpy 2-D matvec (128x128x32): median {t_np:.2f} s of {reps_np}")

    # torch-CPU 3-D FFT formula (slab_scattering_gpu.py:95-155) on the padded grid
    Pz, Pxy = plan.Pz, plan.Pxy
    khat3 = torch.full((9, 9, Pz, Pxy, Pxy), 1e-12 + 1e-13j, dtype=torch.complex128)
    Tt = torch.as_tensor(T)
    xt = torch.as_tensor(x)

    def torch_mv():
        psi = xt.reshape(NZ, M, M, 9)
        tau = torch.einsum("lmnab,lmnb->lmna", Tt, psi).permute(3, 0, 1, 2)
        pad = torch.zeros((9, Pz, Pxy, Pxy), dtype=torch.complex128)
        pad[:, :NZ, :M, :M] = tau
        acc = torch.fft.ifftn(torch.einsum("ijzxy,jzxy->izxy", khat3,
                                           torch.fft.fftn(pad, dim=(1, 2, 3))), dim=(1, 2, 3))
        return psi - acc[:, :NZ, :M, :M].permute(1, 2, 3, 0)

    t_th, reps_th = _median_time(torch_mv)
    del khat3, Tt
    log(f"  torch-CPU 3-D matvec ({torch.get_num_threads()} threads): median {t_th:.3f} s "
        f"of {reps_th}")

    # config-1 full solve of the oracle
    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(VP, VS, RHO)
    g1 = msc.SlabGeometry(M=32, N_z=8, a=A)
    m1 = msc.random_iid_slab_material(g1, ref, 0.3, 0.1, seed=0)
    om1 = 0.2 * VS / A
    ref_o = otm.Medium(VP, VS, RHO)
    t0 = time.perf_counter()
    T1 = osl.slab_tmatrices(32, 8, A, m1.Dlambda, m1.Dmu, m1.Drho, om1, ref_o)
    kh1 = osl.slab_kernels(32, 8, A, om1, ref_o)
    psi01 = osl.incident_field(32, 8, A, om1, ref_o, [1.0, 0.0, 0.0], "P").ravel()
    t_setup1 = time.perf_counter() - t0
    from scipy.sparse.linalg import LinearOperator, gmres

    n1 = psi01.size
    count = [0]

    def mv1(v):
        count[0] += 1
        return osl.slab_matvec(v, T1, kh1, 32, 8)

    def solve1():
        count[0] = 0
        gmres(LinearOperator((n1, n1), matvec=mv1, dtype=complex), psi01, x0=psi01.copy(),
              rtol=TOL, atol=0.0, restart=RESTART, maxiter=50)

    t_s1, reps_s1 = _median_time(solve1)
    log(f"  oracle 32x32x8 solve: setup {t_setup1:.2f} s, solve median {t_s1:.3f} s "
        f"({count[0]} matvecs)")
    return {
        "value": round(t_np * n_matvec, 2), "unit": "s/solve", "cores": 1, "kind": "port",
        "sample": (f"1 warm-up + 3 timed CPU matvecs of the 128x128x32 slab (oracle restatement "
                   f"of slab_scattering.py:359-418, NumPy, 1 core; median {t_np:.2f} s) x {n_matvec} "
                   f"operator applications of the plain FP64 GPU solve = extrapolated s/solve, "
                   f"setup excluded"),
        "extrapolated": True,
        "host": info,
        "numpy_matvec_s": round(t_np, 3), "numpy_matvec_repeats_s": reps_np,
        "torch_cpu_3d_matvec_s": round(t_th, 4), "torch_cpu_3d_repeats_s": reps_th,
        "torch_cpu_3d_s_per_solve": round(t_th * n_matvec, 3),
        "torch_cpu_3d_threads": info["torch_threads"],
        "config1_32x32x8": {"setup_s": round(t_setup1, 3), "solve_s": round(t_s1, 4),
                            "solve_repeats_s": reps_s1, "matvecs": count[0],
                            "s_per_solve_incl_setup": round(t_setup1 + t_s1, 3)},
    }


def time_matvecs(plan, reps: int = 20):
    """Per-operator-application time of the whole five-pass matvec (both
    precisions) from torch events on the plan's stream."""
    import torch

    out = {}
    for dt in (torch.complex64, torch.complex128):
        x = torch.randn(plan.n, dtype=dt, device=plan.device)
        y = torch.empty_like(x)
        for _ in range(3):
            plan.matvec(x, y)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            plan.matvec(x, y)
        e1.record()
        torch.cuda.synchronize()
        out["c64" if dt == torch.complex64 else "c128"] = e0.elapsed_time(e1) / reps
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the resonance-batch, sphere and strong-contrast legs")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: independent (frequency, slowness, slab) solves shard "
              "across GPUs with no communication, so run one process per GPU; --gpus > 1 is "
              "not measured by this build", file=sys.stderr)
        sys.exit(2)

    import torch

    if not torch.cuda.is_available():
        log("no HIP device")
        sys.exit(1)
    dev = torch.device("cuda", 0)
    t0 = time.perf_counter()
    plan, psi0, mat, omega = setup(dev)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    log(f"setup (plan, G0 table, T0, psi0): {t_setup:.2f} s")

    net, learn_info = train_preconditioner(plan, mat, omega, dev)
    log("plain FP64 GMRES(30):")
    t64, st64, x64 = time_solves(plan, psi0, "fp64", args.steps, args.warmup)
    log("mixed (complex64 cycles in FP64 refinement):")
    tmx, stmx, xmx = time_solves(plan, psi0, "mixed", args.steps, args.warmup)
    log("learned + mixed (the validated network in flexible GMRES when its gate is open):")
    tln, stln, xln = time_solves(plan, psi0, "mixed", args.steps, args.warmup,
                                 precond=net if net.gate_open else None, label="learned+mixed")
    log("flexible GMRES with the network forced on (reported beside the headline):")
    ton, ston, _ = time_solves(plan, psi0, "mixed", args.steps, args.warmup, precond=net,
                               label="network on")
    diff = float((xmx - x64).abs().max() / x64.abs().max())
    diff_ln = float((xln - x64).abs().max() / x64.abs().max())
    # R_PP of each timed solve against an FP64 solve to 1e-12
    xt, stt = plan.gmres(psi0, "born", tol=1e-12, max_iter=600, restart=RESTART)
    rpp_t = rpp_of(plan, xt, mat, omega)
    rpp_err = {k: abs(rpp_of(plan, xx, mat, omega) - rpp_t) / abs(rpp_t)
               for k, xx in (("fp64", x64), ("mixed", xmx), ("learned_mixed", xln))}
    log(f"R_PP {rpp_t:.6e}; relative errors vs FP64 at 1e-12: {rpp_err}")
    log(f"tiles (y-fwd, x-fwd, z, x-inv, y-inv): {plan.tiles()}")
    timmx = time_pass3(plan, psi0, "mixed", args.steps)
    tim64 = time_pass3(plan, psi0, "fp64", args.steps)

    # roofline of the block-multiply pass (pass 3) in the headline (complex64) solves
    n64, n64_noop = timmx["c64"]["working"], timmx["c64"]["no_op"]
    nuv, nc = plan.NUV, plan.NC
    s = 8  # bytes per complex64
    spectra = NZ * plan.Pxy * plan.Pxy * 9 * s  # read and written once
    table = nuv * 45 * nc * s
    bytes_pass3 = 2 * spectra + table
    avg_ms = timmx["c64"]["avg_ms"]
    achieved = bytes_pass3 / (avg_ms * 1e-3) / 1e9
    kname = z_kernel_name(plan.tiles()["c64"][2])
    # whole five-pass operator application against the per-matvec bytes of a
    # dense-table formulation, B_mv = s (E P + 2 n + 4 N) with E = 81 or 45
    # table entries per padded point, and against its compulsory bytes (psi
    # in, y out, T0 and the compressed table once each)
    mv = time_matvecs(plan)
    P = plan.Pz * plan.Pxy * plan.Pxy
    nvox = NZ * M * M
    whole = {}
    for key, sb in (("c64", 8), ("c128", 16)):
        t = mv[key] * 1e-3
        b81 = sb * (81 * P + 2 * 9 * nvox + 4 * nvox)
        b45 = sb * (45 * P + 2 * 9 * nvox + 4 * nvox)
        comp = sb * (2 * 9 * nvox + 4 * nvox) + nuv * 45 * nc * sb
        whole[key] = {"ms": round(mv[key], 4), "B_mv_81_GB": round(b81 / 1e9, 3),
                      "frac_B_mv_81": round(b81 / t / 8e12, 3),
                      "B_mv_45_GB": round(b45 / 1e9, 3), "frac_B_mv_45": round(b45 / t / 8e12, 3),
                      "compulsory_GB": round(comp / 1e9, 4),
                      "frac_compulsory": round(comp / t / 8e12, 3)}
    log(f"whole matvec: {whole}")
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": 8000.0, "unit": "GB/s",
                "frac": round(achieved / 8000.0, 4), "traffic": measured_traffic(kname),
                "kernel": f"{kname} (z FFT + 9x9 block multiply "
                          "+ inverse z FFT), complex64 launches of the mixed solves",
                "launches": n64, "no_op_launches_excluded": n64_noop,
                "avg_us": round(avg_ms * 1e3, 2), "bytes_per_launch": bytes_pass3,
                "whole_matvec": whole}

    mean_mx = float(np.mean(tmx))
    mean_64 = float(np.mean(t64))
    mean_ln = float(np.mean(tln))
    sweep = None
    if not args.no_sweep:
        sweep = time_sweep("batched")
        sweep["workers_mode"] = time_sweep("workers")
    extras = None if args.no_extras else {"reference_formula_torch_gpu":
                                          time_reference_gpu_formula(plan, st64["matvecs_c128"]),
                                          "resonance_batch": time_resonance(),
                                          "sphere_fft": time_sphere(),
                                          "resonance_slab_ocean_bottom": time_resonance_slab(),
                                          "learned_strong_contrast": time_strong_contrast()}
    cpu = None
    if not args.no_cpu_baseline:
        log("CPU baseline (oracle restatement of the reference CPU path) ...")
        cpu = cpu_baseline(plan, st64["matvecs_c128"])
        log(f"  -> {cpu['value']} s/solve (extrapolated)")

    out = {
        "metric": "s/solve to 1e-8 residual, 128\u00d7128\u00d732 slab: learned+mixed vs plain FP64 GMRES",
        "value": round(mean_ln, 6),
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_ln * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+c128",
        "data": "synthetic (seeded random slab, Bernoulli(0.3) inclusions, U[-0.2,0.2] contrasts)",
        "config": {"workload": "slab 128x128x32, pad 64x256x256, ka=0.4, p=0.15 s/km, tol 1e-8 "
                               "FP64; GMRES(30) with complex64 cycles in FP64 refinement and the "
                               "learned four-level conv preconditioner in flexible GMRES when its "
                               "validation gate is open",
                   "M": M, "N_z": NZ, "restart": RESTART, "tol": TOL,
                   "precision": "mixed", "preconditioner": f"learned unet {UNET_WIDTHS}",
                   "gate_open": bool(net.gate_open)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "detail": {
            "fp64_s_per_solve": round(mean_64, 6),
            "mixed_s_per_solve": round(mean_mx, 6),
            "speedup_vs_fp64": round(mean_64 / mean_ln, 3),
            "mixed_speedup_vs_fp64": round(mean_64 / mean_mx, 3),
            "network_on_s_per_solve": round(float(np.mean(ton)), 6),
            "network_on_iterations": ston["iterations"],
            "network_on_precond_applies": ston["precond_applies"],
            "learned_iterations": stln["iterations"], "learned_fell_back": stln["fell_back"],
            "learned_precond_applies": stln["precond_applies"],
            "learned_vs_fp64_max_rel_diff": diff_ln, "learned": learn_info,
            "fp64_iterations": st64["iterations"], "mixed_iterations": stmx["iterations"],
            "mixed_matvecs_c64": stmx["matvecs_c64"], "mixed_matvecs_c128": stmx["matvecs_c128"],
            "fp64_rel_residual": st64["rel_residual"], "mixed_rel_residual": stmx["rel_residual"],
            "mixed_vs_fp64_max_rel_diff": diff,
            "rpp_tight_fp64": [rpp_t.real, rpp_t.imag], "rpp_rel_err": rpp_err,
            "psi_rel_err_vs_tight": {k: float((xx - xt).abs().max() / xt.abs().max())
                                     for k, xx in (("fp64", x64), ("mixed", xmx),
                                                   ("learned_mixed", xln))}, "setup_s": round(t_setup, 3),
            "pass3_c128_avg_us": round(tim64["c128"]["avg_ms"] * 1e3, 2),
            "sweep_256_uneven_slabs_mixed": sweep,
            "other_paths": extras,
        },
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
