"""CPU tests of the host layer: the C ABI library loads and exports every symbol
include/*.h declares, its host-only entry points and argument checks,
data types, material generators and the YAML configuration (mirroring
cubic_scattering/tests/test_slab_scattering.py:50-130 and
test_solver_config.py:96-230).  No compute call needs a GPU here."""

import ctypes
import re
from pathlib import Path
from textwrap import dedent

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    from multiplescatteringcalculations_amd import _lib

    header = "\n".join(h.read_text() for h in sorted((ROOT / "include").glob("*.h")))
    declared = set(re.findall(r"^(?:int|const char\*)\s+(msc_\w+)\(", header, re.M))
    assert declared, "no declarations parsed"
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    lib = _lib.load()
    for name in declared:
        assert hasattr(lib, name), name
    assert b"gfx950" in lib.msc_version()


def test_host_entry_points_validate_arguments():
    from multiplescatteringcalculations_amd import _lib

    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.msc_plan_create(ctypes.byref(h), 0, 4, 1.0, 0, 0, 0) == _lib.MSC_ERR_INVALID
    assert b"M must be >= 1" in lib.msc_last_error()
    assert lib.msc_plan_create(ctypes.byref(h), 4, 0, 1.0, 0, 0, 0) == _lib.MSC_ERR_INVALID
    assert lib.msc_plan_create(ctypes.byref(h), 4, 4, -1.0, 0, 0, 0) == _lib.MSC_ERR_INVALID
    # padding below 2N-1 is rejected before any device work
    assert lib.msc_plan_create(ctypes.byref(h), 8, 4, 1.0, 0, 4, 16) == _lib.MSC_ERR_INVALID
    # periodic on the M-point lateral grid needs M a power of two (any M runs on
    # the padded grid, the default for such M)
    assert lib.msc_plan_create(ctypes.byref(h), 6, 4, 1.0, 1, 0, 6) == _lib.MSC_ERR_UNSUPPORTED
    assert b"power of two" in lib.msc_last_error()
    assert lib.msc_plan_create(ctypes.byref(h), 6, 4, 1.0, 1, 0, 8) == _lib.MSC_ERR_INVALID
    out = (ctypes.c_double * 8)()
    assert lib.msc_cube_coefficients(-1.0, 1.0, 3000.0, 1732.0, 2500.0, out) == _lib.MSC_ERR_INVALID
    with pytest.raises(ValueError, match="positive"):
        _lib.check(lib.msc_cube_coefficients(1.0, 0.0, 3000.0, 1732.0, 2500.0, out))


def test_cube_coefficients_match_oracle():
    """The library's host FP64 Gamma0, A, B, C equal the oracle's (the device T0
    kernel consumes exactly these)."""
    from multiplescatteringcalculations_amd import ReferenceMedium, MaterialContrast, compute_cube_tmatrix
    from oracle import tmatrix as tm

    for (al, be, rho), om, a in [((5.0, 3.0, 2.5), 2 * np.pi * 10, 0.01),
                                 ((3000.0, 1732.0, 2500.0), 692.8, 1.0),
                                 ((5000.0, 3000.0, 2500.0), 150.0, 1.0)]:
        ref = ReferenceMedium(al, be, rho)
        c = MaterialContrast(0.1 * ref.lam, -0.2 * ref.mu, 0.05 * rho)
        got = compute_cube_tmatrix(om, a, ref, c)
        want = tm.cube_tmatrix(om, a, tm.Medium(al, be, rho), c.Dlambda, c.Dmu, c.Drho)
        for k in ("Gamma0", "Ac", "Bc", "Cc", "T1c", "T2c", "T3c", "Drho_star", "Dlambda_star",
                  "Dmu_star_off", "Dmu_star_diag"):
            assert abs(getattr(got, k) - getattr(want, k)) <= 1e-13 * abs(getattr(want, k)), k


def test_solver_entry_points_fail_loudly_without_gpu():
    import torch

    import multiplescatteringcalculations_amd as msc

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    g = msc.SlabGeometry(M=2, N_z=2, a=1.0)
    ref = msc.ReferenceMedium(5000.0, 3000.0, 2500.0)
    mat = msc.uniform_slab_material(g, ref, msc.MaterialContrast(2e9, 1e9, 100.0))
    with pytest.raises(RuntimeError, match="no HIP device"):
        msc.compute_slab_scattering(g, mat, 150.0, np.array([1.0, 0.0, 0.0]))
    with pytest.raises(ValueError, match="HIP device"):
        msc.compute_slab_scattering_gpu(g, mat, 150.0, np.array([1.0, 0.0, 0.0]),
                                        device=torch.device("cpu"))


# ── data types (test_slab_scattering.py:50-130) ──────────────────────────────


def test_geometry():
    from multiplescatteringcalculations_amd import SlabGeometry

    g = SlabGeometry(M=3, N_z=2, a=1.0)
    assert g.d == 2.0 and g.n_cubes == 18
    assert np.allclose(g.cube_centre(0, 1, 1), [1.0, 0.0, 0.0])
    assert np.allclose(g.cube_centre(1, 0, 0), [3.0, -2.0, -2.0])
    c = g.all_centres()
    assert c.shape == (2, 3, 3, 3)
    for idx in np.ndindex(2, 3, 3):
        assert np.allclose(c[idx], g.cube_centre(*idx))
    with pytest.raises(ValueError, match="M must be >= 1"):
        SlabGeometry(M=0, N_z=1, a=1.0)
    with pytest.raises(ValueError, match="N_z must be >= 1"):
        SlabGeometry(M=1, N_z=0, a=1.0)
    with pytest.raises(ValueError, match="a must be > 0"):
        SlabGeometry(M=1, N_z=1, a=-1.0)


def test_material_validation_and_generators():
    from multiplescatteringcalculations_amd import (
        MaterialContrast, ReferenceMedium, SlabGeometry, SlabMaterial, random_iid_slab_material,
        random_slab_material, uniform_slab_material)

    ref = ReferenceMedium(5000.0, 3000.0, 2500.0)
    with pytest.raises(ValueError, match="shapes must match"):
        SlabMaterial(np.zeros((2, 3, 3)), np.zeros((2, 3, 4)), np.zeros((2, 3, 3)), ref)
    with pytest.raises(ValueError, match="3D"):
        SlabMaterial(np.zeros((2, 3)), np.zeros((2, 3)), np.zeros((2, 3)), ref)
    g = SlabGeometry(M=4, N_z=2, a=1.0)
    c = MaterialContrast(2e9, 1e9, 100.0)
    m1 = random_slab_material(g, ref, c, phi=0.3, seed=99)
    m2 = random_slab_material(g, ref, c, phi=0.3, seed=99)
    assert np.array_equal(m1.Dlambda, m2.Dlambda)
    assert not np.array_equal(m1.Dlambda, random_slab_material(g, ref, c, 0.3, seed=2).Dlambda)
    # reference draw order: one rng.random((M, M)) per layer
    rng = np.random.default_rng(99)
    for lz in range(2):
        assert np.array_equal(m1.Dlambda[lz] > 0, rng.random((4, 4)) < 0.3)
    gd = SlabGeometry(M=8, N_z=3, a=1.0)
    md = random_slab_material(gd, ref, c, phi=lambda lz: 0.1 * (lz + 1), seed=42)
    n_inc = [(md.Dlambda[lz] > 0).sum() for lz in range(3)]
    assert n_inc[0] < n_inc[2]
    u = uniform_slab_material(g, ref, c)
    assert np.all(u.Dmu == 1e9)
    iid = random_iid_slab_material(SlabGeometry(M=32, N_z=8, a=1.0), ref, 0.3, 0.1, seed=0)
    frac = np.mean(iid.Drho != 0)
    assert 0.27 < frac < 0.33
    rel = iid.Dlambda[iid.Dlambda != 0] / ref.lam
    assert rel.min() >= -0.1 and rel.max() <= 0.1


def test_tparams_roundtrip():
    from multiplescatteringcalculations_amd.slab_scattering import (
        tmatrices_to_tparams, tparams_to_tmatrices)

    rng = np.random.default_rng(0)
    tp = rng.standard_normal((2, 3, 3, 4)) + 1j * rng.standard_normal((2, 3, 3, 4))
    T = tparams_to_tmatrices(tp)
    assert T.shape == (2, 3, 3, 9, 9)
    assert np.array_equal(tmatrices_to_tparams(T), tp)
    T[0, 0, 0, 0, 1] = 1.0
    with pytest.raises(ValueError, match="block-diagonal"):
        tmatrices_to_tparams(T)


# ── YAML configuration (test_solver_config.py) ──────────────────────────────

VALID_SLAB = """\
device:
  backend: cpu
  dtype: complex128
  memory_limit_gb: 4.0
solver:
  gmres_tol: 1.0e-6
  max_iter: 500
  initial_guess: born
problem:
  type: slab
  geometry: {M: 4, N_z: 2, a: 1.0}
  reference: {alpha: 5000.0, beta: 3000.0, rho: 2500.0}
  contrast: {Dlambda: 2.0e9, Dmu: 1.0e9, Drho: 100.0}
  frequency: {omega: 150.0}
  incident:
    k_hat: [1.0, 0.0, 0.0]
    wave_type: P
  slab:
    material_type: uniform
"""


@pytest.fixture
def cfg_file(tmp_path):
    def _write(text):
        p = tmp_path / "config.yml"
        p.write_text(dedent(text))
        return p

    return _write


def test_load_valid_and_defaults(cfg_file):
    from multiplescatteringcalculations_amd.solver_config import load_config

    c = load_config(cfg_file(VALID_SLAB))
    assert c.problem.type == "slab" and c.device.backend == "cpu"
    assert c.solver.gmres_tol == 1e-6 and c.solver.max_iter == 500
    assert c.solver.restart is None and c.solver.precision is None
    g = c.problem.to_slab_geometry()
    assert (g.M, g.N_z, g.a) == (4, 2, 1.0)
    assert c.problem.to_reference_medium().alpha == 5000.0
    assert c.problem.to_material_contrast().Dmu == 1e9
    minimal = VALID_SLAB.split("problem:")[1]
    c2 = load_config(cfg_file("problem:" + minimal))
    assert c2.device.backend == "auto" and c2.solver.initial_guess == "born"


def test_reference_example_config_loads(cfg_file):
    """configs/example_slab.yml of the reference, verbatim schema."""
    from multiplescatteringcalculations_amd.solver_config import load_config

    text = """\
    device:
      backend: auto
      dtype: auto
      memory_limit_gb: 8.0
    solver:
      gmres_tol: 1.0e-6
      max_iter: 500
      initial_guess: born
    problem:
      type: slab
      geometry:
        M: 32
        N_z: 10
        a: 1.0
      reference:
        alpha: 5000.0
        beta: 3000.0
        rho: 2500.0
      contrast:
        Dlambda: 2.0e9
        Dmu: 1.0e9
        Drho: 100.0
      frequency:
        omega: 150.0
      incident:
        k_hat: [1.0, 0.0, 0.0]
        wave_type: P
      slab:
        material_type: random
        phi: 0.1
        seed: 42
    """
    c = load_config(cfg_file(text))
    assert c.problem.slab["seed"] == 42


@pytest.mark.parametrize("old,new,match", [
    ("type: slab", "type: cube", "Invalid problem type"),
    ("{alpha: 5000.0, beta", "{beta", "problem.reference.alpha"),
    ("omega: 150.0", "omega: -1.0", "omega must be positive"),
    ("wave_type: P", "wave_type: L", "Invalid wave_type"),
    ("backend: cpu", "backend: tpu", "Invalid backend"),
    ("dtype: complex128", "dtype: float32", "Invalid dtype"),
    ("initial_guess: born", "initial_guess: random", "Invalid initial_guess"),
    ("initial_guess: born", "initial_guess: born\n  precision: half", "Invalid precision"),
    ("initial_guess: born", "initial_guess: born\n  restart: 0", "restart must be >= 1"),
    ("material_type: uniform", "material_type: marble", "material_type"),
])
def test_validation_errors(cfg_file, old, new, match):
    from multiplescatteringcalculations_amd.solver_config import load_config

    assert old in VALID_SLAB
    with pytest.raises(ValueError, match=match):
        load_config(cfg_file(VALID_SLAB.replace(old, new)))


def test_missing_file_and_problem(cfg_file, tmp_path):
    from multiplescatteringcalculations_amd.solver_config import load_config

    with pytest.raises(FileNotFoundError, match="Config file not found"):
        load_config(tmp_path / "nope.yml")
    with pytest.raises(ValueError, match="missing required 'problem'"):
        load_config(cfg_file("device:\n  backend: cpu\n"))


def test_random_iid_schema_and_cpu_backend_refused(cfg_file):
    from multiplescatteringcalculations_amd.solver_config import build_material, load_config, run_from_config

    text = VALID_SLAB.replace("material_type: uniform",
                              "material_type: random_iid\n    phi: 0.3\n    rel_contrast: 0.2\n    seed: 7")
    c = load_config(cfg_file(text))
    g, mat = build_material(c)
    assert mat.Dlambda.shape == (2, 4, 4)
    with pytest.raises(ValueError, match="HIP devices only"):
        run_from_config(c)


def test_sphere_config_validated_and_cpu_backend_refused(cfg_file):
    from multiplescatteringcalculations_amd.solver_config import load_config, run_from_config

    text = VALID_SLAB.replace("type: slab", "type: sphere").replace(
        "  slab:\n    material_type: uniform\n", "  sphere: {radius: 10.0, n_sub: 4}\n")
    c = load_config(cfg_file(text))
    assert c.problem.type == "sphere" and c.problem.sphere == {"radius": 10.0, "n_sub": 4}
    with pytest.raises(ValueError, match="HIP devices only"):
        run_from_config(c)
    with pytest.raises(ValueError, match="radius"):
        load_config(cfg_file(text.replace("{radius: 10.0, n_sub: 4}", "{n_sub: 4}")))


def test_copy_g0_validates_arguments():
    """msc_plan_copy_g0 rejects NULL plans before any device work."""
    from multiplescatteringcalculations_amd import _lib

    lib = _lib.load()
    assert lib.msc_plan_copy_g0(None, None, None) == _lib.MSC_ERR_INVALID
    assert b"NULL" in lib.msc_last_error()


def test_sweep_and_reshape_validate_arguments():
    """msc_sweep and msc_plan_reshape reject bad arguments before any device work."""
    from multiplescatteringcalculations_amd import _lib

    lib = _lib.load()
    opts = _lib.GmresOpts(30, 100, 1e-8, 1, 1, 1e-4, 0, None)
    inst = (_lib.Instance * 1)()
    assert lib.msc_sweep(None, 1, ctypes.byref(opts), 1) == _lib.MSC_ERR_INVALID
    assert lib.msc_sweep(ctypes.cast(inst, ctypes.c_void_p), 1, None, 1) == _lib.MSC_ERR_INVALID
    assert lib.msc_sweep(ctypes.cast(inst, ctypes.c_void_p), 1, ctypes.byref(opts), 0) == _lib.MSC_ERR_INVALID
    # an instance without arrays is invalid and says which one
    assert lib.msc_sweep(ctypes.cast(inst, ctypes.c_void_p), 1, ctypes.byref(opts), 1) == _lib.MSC_ERR_INVALID
    assert inst[0].status == _lib.MSC_ERR_INVALID and b"instance 0" in inst[0].error
    # a preconditioner is not allowed in a sweep
    opts.precond = ctypes.c_void_p(1)
    assert lib.msc_sweep(ctypes.cast(inst, ctypes.c_void_p), 1, ctypes.byref(opts), 1) == _lib.MSC_ERR_INVALID
    assert lib.msc_plan_reshape(None, 4, 4, 1.0) == _lib.MSC_ERR_INVALID


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Every ctypes mirror of a struct in include/msc_slab.h has the C size and
    field offsets (a C program built with gcc prints them)."""
    import shutil
    import subprocess

    from multiplescatteringcalculations_amd import _lib

    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    pairs = [("msc_gmres_opts", _lib.GmresOpts), ("msc_gmres_stats", _lib.GmresStats),
             ("msc_instance", _lib.Instance)]
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "msc_slab.h"', "int main(void){"]
    for cname, cls in pairs:
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf("{cname} {fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parents[1] / "include"
    subprocess.run(["gcc", "-std=c99", f"-I{inc}", str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {}
    for ln in out.splitlines():
        parts = ln.split()
        got[(parts[0], parts[1])] = int(parts[2])
    for cname, cls in pairs:
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert got[(cname, fname)] == getattr(cls, fname).offset, (cname, fname)
