"""GPU tests of the sweep (compute_slab_sweep in both modes: batches of one
padded grid through msc_sweep_batched, and worker threads through msc_sweep)
and of plan re-shaping: every case of a sweep of uneven slabs, frequencies and
incidences equals the same case solved alone (FP64 solves agree to 1e-9
relative at tol 1e-11; R_PP and the far field to 1e-9), whatever the number of
workers or the batch size."""

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _cases(seed=0):
    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(3000.0, 1732.0, 2500.0)
    rng = np.random.default_rng(seed)
    cases = []
    # periodic (odd i): M = 8 on the circular 8-point grid, M = 12 and M = 3 on
    # zero-padded grids with the folded kernel (M = 3 shares P_z = P_xy = 8
    # with the circular M = 8 case and must not share its plan)
    shapes = [(6, 2), (8, 3), (5, 3), (12, 4), (7, 2), (8, 4), (16, 3), (3, 3), (6, 2)]
    for i, (M, Nz) in enumerate(shapes):
        g = msc.SlabGeometry(M=M, N_z=Nz, a=1.0)
        mat = msc.random_iid_slab_material(g, ref, 0.3, 0.2, seed=100 + i)
        ka = float(rng.uniform(0.1, 1.0))
        p = float(rng.choice([0.0, 1e-4, 2e-4, 3e-4]))
        khat = np.array([np.sqrt(1 - (p * ref.alpha) ** 2), p * ref.alpha, 0.0])
        wave = "S" if i == 2 else "P"
        periodic = i % 2 == 1
        cases.append(msc.SweepCase(g, mat, ka * ref.beta / g.a, khat, wave, periodic))
    return cases


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.parametrize("mode,workers,max_batch",
                         [("workers", 1, 1), ("workers", 3, 1), ("batched", 1, 64),
                          ("batched", 1, 2)])
def test_sweep_equals_individual_solves(hip_device, mode, workers, max_batch):
    import multiplescatteringcalculations_amd as msc

    cases = _cases()
    res = msc.compute_slab_sweep(cases, gmres_tol=1e-11, restart=30, precision="fp64",
                                 workers=workers, mode=mode, max_batch=max_batch)
    assert len(res) == len(cases)
    for c, r in zip(cases, res):
        one = msc.compute_slab_scattering_gpu(c.geometry, c.material, c.omega, c.k_hat,
                                              c.wave_type, gmres_tol=1e-11, restart=30,
                                              periodic=c.periodic)
        assert r.stats["converged"] and r.gmres_residual <= 1e-11
        assert _rel(r.psi0, one.psi0) < 1e-14
        assert _rel(r.psi, one.psi) < 1e-9, (c.geometry, c.omega)
        if c.periodic:
            assert abs(msc.slab_rpp_periodic(r) - msc.slab_rpp_periodic(one)) <= \
                1e-9 * abs(msc.slab_rpp_periodic(one))
        ff = np.array(msc.slab_reflected_field(r))
        ff1 = np.array(msc.slab_reflected_field(one))
        assert _rel(ff, ff1) < 1e-9


@pytest.mark.parametrize("mode", ["batched", "workers"])
def test_mixed_sweep_reaches_fp64_tolerance(hip_device, mode):
    import multiplescatteringcalculations_amd as msc

    cases = _cases(seed=1)
    res = msc.compute_slab_sweep(cases, gmres_tol=1e-8, precision="mixed", workers=4, mode=mode)
    ref = msc.compute_slab_sweep(cases, gmres_tol=1e-11, precision="fp64", workers=2,
                                 mode="workers")
    for r, r0 in zip(res, ref):
        assert r.stats["converged"] and r.gmres_residual <= 1e-8
        assert _rel(r.psi, r0.psi) < 1e-6


def test_plan_reshape_reuses_buffers(hip_device):
    """A plan created at the largest slab of its grid, re-shaped to a smaller
    slab, gives the smaller slab's operator."""
    import multiplescatteringcalculations_amd as msc
    from multiplescatteringcalculations_amd.slab_scattering import SlabPlan

    ref = msc.ReferenceMedium(3000.0, 1732.0, 2500.0)
    big = msc.SlabGeometry(M=8, N_z=4, a=1.0)
    small = msc.SlabGeometry(M=6, N_z=3, a=1.0)
    om = 0.4 * ref.beta
    mat = msc.random_iid_slab_material(small, ref, 0.4, 0.3, seed=3)
    plan = SlabPlan(big, pad=(8, 16))
    plan.build_g0(om, ref)
    plan.reshape(small)
    plan.build_g0(om, ref)
    plan.set_material(mat, om)
    alone = SlabPlan(small, pad=(8, 16))
    alone.build_g0(om, ref)
    alone.set_material(mat, om)
    x = torch.randn(plan.n, dtype=torch.complex128, device=hip_device)
    y1, y2 = plan.matvec(x), alone.matvec(x)
    assert float((y1 - y2).abs().max() / y2.abs().max()) < 1e-14
    # growing beyond the allocation is refused
    with pytest.raises(ValueError, match="capacity"):
        plan.reshape(msc.SlabGeometry(M=9, N_z=4, a=1.0))


@pytest.mark.parametrize("mode", ["batched", "workers"])
def test_volume_averaged_sweep_equals_individual_solves(hip_device, mode):
    """Volume-averaged cases in a sweep, mixed with point-propagator cases of
    the same grid (so the table key must switch): each equals its lone solve."""
    import multiplescatteringcalculations_amd as msc

    ref = msc.ReferenceMedium(3000.0, 1732.0, 2500.0)
    g = msc.SlabGeometry(M=8, N_z=3, a=1.0)
    mat = msc.random_iid_slab_material(g, ref, 0.3, 0.2, seed=7)
    kh = np.array([1.0, 0.0, 0.0])
    om = 0.3 * ref.beta
    cases = [msc.SweepCase(g, mat, om, kh, "P", periodic=True, volume_averaged=True, n_orders=2),
             msc.SweepCase(g, mat, om, kh, "P", periodic=True),
             msc.SweepCase(g, mat, om, kh, "P", periodic=True, volume_averaged=True, n_orders=1),
             msc.SweepCase(g, mat, om, kh, "P", periodic=False, volume_averaged=True)]
    res = msc.compute_slab_sweep(cases, gmres_tol=1e-11, precision="fp64", workers=1, mode=mode)
    for c, r in zip(cases, res):
        one = msc.compute_slab_scattering_gpu(g, mat, om, kh, "P", gmres_tol=1e-11, restart=30,
                                              periodic=c.periodic,
                                              volume_averaged=c.volume_averaged,
                                              n_orders=c.n_orders)
        assert _rel(r.psi, one.psi) < 1e-9, (c.periodic, c.volume_averaged, c.n_orders)
    assert _rel(res[0].psi, res[1].psi) > 1e-6  # the averaging changed the field


def test_sweep_modes_reject_bad_arguments():
    import multiplescatteringcalculations_amd as msc

    with pytest.raises(ValueError, match="mode"):
        msc.compute_slab_sweep([], mode="threads", device=torch.device("cuda", 0))
    with pytest.raises(ValueError, match="max_batch"):
        msc.compute_slab_sweep([], max_batch=0, device=torch.device("cuda", 0))
