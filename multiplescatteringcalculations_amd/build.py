"""Build libmsc_slab.so (HIP, gfx950) in-tree.

Usage: ``python -m multiplescatteringcalculations_amd.build`` (or ``build()``).
Each ``csrc/*.hip`` file is compiled with hipcc for gfx950 in parallel and the
objects are linked into ``multiplescatteringcalculations_amd/libmsc_slab.so``
with a run path to PyTorch's ``lib`` directory, so that the library binds to the
HIP runtime PyTorch loads (one HIP runtime per process).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
INCLUDE = PKG.parent / "include"
LIB = PKG / "libmsc_slab.so"
OBJ = PKG / "_objs"

ARCH = os.environ.get("MSC_OFFLOAD_ARCH", "gfx950")
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-fPIC",
    "-std=c++17",
    "-Wall",
    "-Wno-unused-function",
    # never emit packed-FP32 instructions (v_pk_add/mul/fma_f32)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the HIP toolchain is required to build libmsc_slab.so")


def _torch_lib() -> str:
    import importlib.util

    spec = importlib.util.find_spec("torch")
    if spec is None or spec.origin is None:
        raise RuntimeError("PyTorch is required (its HIP runtime is the one the library binds to)")
    return str(Path(spec.origin).parent / "lib")


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    hipcc = _hipcc()
    OBJ.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")

    def compile_one(src: Path) -> Path:
        obj = OBJ / (src.stem + ".o")
        cmd = [hipcc, *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
        res = subprocess.run(cmd, capture_output=True, text=True)
        out = (res.stdout + res.stderr).strip()
        # the host pass does not know the packed-fp32 target feature: harmless
        lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
        if res.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n" + "\n".join(lines))
        if verbose and lines:
            print("\n".join(lines), file=sys.stderr)
        return obj

    jobs = jobs or min(len(sources), int(os.environ.get("MAX_JOBS", "8")), 8)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, sources))
    tlib = _torch_lib()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
           f"-L{tlib}", f"-Wl,-rpath,{tlib}", "-lamdhip64"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("link failed:\n" + res.stdout + res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
