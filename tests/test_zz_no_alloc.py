"""GPU test of msc_resonance_tmatrices_noalloc (the resonance T-matrix batch
with its workspace in the library's fixed device scratch instead of device
allocations), through the C ABI, against msc_resonance_tmatrices on the same
inputs: two calls in a row, the second smaller (fewer cubes, smaller n_sub,
the Neumann series), so that nothing of the first call's scratch contents may
leak into the second.  Tolerance 1e-12 relative (the incident patterns' phases
come from the device's sincos instead of the host's)."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

VP, VS, RHO = 3000.0, 1732.0, 2500.0


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _call(name, n_sub, count, wave_s, neumann, seed, device):
    from multiplescatteringcalculations_amd import _lib

    lib = _lib.load()
    rng = np.random.default_rng(seed)
    lam, mu = RHO * (VP**2 - 2 * VS**2), RHO * VS**2
    dl, dm, dr = (torch.as_tensor(rng.uniform(-0.3, 0.3, count) * s).to(device)
                  for s in (lam, mu, RHO))
    n9 = 9 * n_sub**3
    tc = torch.full((count, 9, 9), np.nan, dtype=torch.complex128, device=device)
    px = torch.full((count, n9, 9), np.nan, dtype=torch.complex128, device=device)
    ratio = np.full(count, -1.0)
    iters = np.full(count, -7, dtype=np.int32)
    kh = (ctypes.c_double * 3)(0.8, 0.36, 0.48)
    om = 1.2 * VS
    torch.cuda.synchronize()
    rc = getattr(lib, name)(om, 1.0, VP, VS, RHO, n_sub, ctypes.cast(kh, ctypes.c_void_p), wave_s,
                            count, dl.data_ptr(), dm.data_ptr(), dr.data_ptr(),
                            neumann[0], neumann[1], tc.data_ptr(), px.data_ptr(),
                            ratio.ctypes.data, iters.ctypes.data,
                            _lib.stream_ptr(device))
    _lib.check(rc)
    return tc.cpu().numpy(), px.cpu().numpy(), ratio, iters


@pytest.mark.parametrize("first,second", [
    # (n_sub, count, wave_s, (neumann_order, tol), seed): 20 cubes at n_sub 4 take
    # two scratch chunks; then fewer, smaller cubes with the Neumann series
    ((4, 20, 0, (0, 0.0), 1), (3, 5, 1, (60, 1e-13), 2)),
    ((5, 7, 1, (0, 0.0), 3), (2, 3, 0, (0, 0.0), 4)),
])
def test_noalloc_matches_allocating_entry_point(hip_device, first, second):
    for n_sub, count, wave_s, neu, seed in (first, second):
        ref = _call("msc_resonance_tmatrices", n_sub, count, wave_s, neu, seed, hip_device)
        got = _call("msc_resonance_tmatrices_noalloc", n_sub, count, wave_s, neu, seed, hip_device)
        assert np.isfinite(got[0]).all() and np.isfinite(got[1]).all()
        assert _rel(got[0], ref[0]) < 1e-12, (n_sub, count)
        assert _rel(got[1], ref[1]) < 1e-12, (n_sub, count)
        if neu[0] == 0:
            assert np.allclose(got[2], ref[2], rtol=1e-10, atol=0)
            assert (got[3] == -1).all()
        else:
            assert np.isnan(got[2]).all()
            assert (got[3] == ref[3]).all()


def test_noalloc_rejects_sizes_beyond_its_scratch(hip_device):
    with pytest.raises(ValueError, match="no-allocation scratch"):
        _call("msc_resonance_tmatrices_noalloc", 7, 1, 0, (0, 0.0), 5, hip_device)
