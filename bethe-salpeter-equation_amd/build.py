"""Build ``libbse_mi355x.so`` (all HIP sources, gfx950) in-tree.

``python build.py`` or ``build()``: every ``csrc/*.hip`` is compiled with
hipcc for gfx950 (packed-FP32 instruction selection disabled) into
``_build/*.o`` in parallel, then linked into ``libbse_mi355x.so`` next to
this file.  The library has no torch dependency; Python loads it through
ctypes.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = pathlib.Path(os.environ.get("BSE_BUILD_DIR", PKG / "_build"))
LIB = pathlib.Path(os.environ.get("BSE_MI355X_LIB", PKG / "libbse_mi355x.so"))
ARCH = os.environ.get("BSE_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-result", "-munsafe-fp-atomics",
          *os.environ.get("BSE_HIPCC_FLAGS", "").split()]


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> pathlib.Path:
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    jobs = min(len(sources), max(1, min(16, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(objs)} sources", file=sys.stderr)
    return LIB


if __name__ == "__main__":
    build()
