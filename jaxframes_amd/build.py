"""Build libjfx.so (the HIP kernels + C ABI) in-tree for gfx950.

``python -m jaxframes_amd.build`` compiles every ``csrc/*.hip`` with hipcc into
``jaxframes_amd/_lib/libjfx.so``. Objects are compiled in parallel and only
rebuilt when a source or header is newer than its object, unless force=True.
"""

from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
OBJDIR = LIBDIR / "obj"
LIB = LIBDIR / "libjfx.so"
# the benchmark's vendor-library baseline (rocPRIM's radix sort), kept out of
# libjfx.so
CMP_SRC = PKG / "cmp" / "rocprim_sort.hip"
CMP_LIB = LIBDIR / "libjfx_cmp.so"
INCLUDE = PKG.parent / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-function", "-Wno-unused-variable",
    # packed FP32 VALU forms are kept out of every HIP code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def _sources():
    return sorted(CSRC.glob("*.hip"))


def _headers():
    return sorted(CSRC.glob("*.hpp")) + sorted(INCLUDE.glob("*.h"))


def _compile(src: Path, force: bool) -> Path:
    obj = OBJDIR / (src.stem + ".o")
    newest_dep = max([src.stat().st_mtime] + [h.stat().st_mtime for h in _headers()])
    if not force and obj.exists() and obj.stat().st_mtime >= newest_dep:
        return obj
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stderr[-6000:]}")
    return obj


def _build_cmp(force: bool) -> Path:
    """libjfx_cmp.so: rocPRIM's radix_sort_pairs behind one C entry point."""
    if not force and CMP_LIB.exists() and CMP_LIB.stat().st_mtime >= CMP_SRC.stat().st_mtime:
        return CMP_LIB
    tmp = CMP_LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, *FLAGS, "-shared", str(CMP_SRC), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {CMP_SRC.name}:\n{r.stderr[-6000:]}")
    os.replace(tmp, CMP_LIB)
    return CMP_LIB


def build(force: bool = True, jobs: int | None = None) -> Path:
    """Compile all HIP sources and link libjfx.so (and the benchmark's
    libjfx_cmp.so); returns the path of libjfx.so."""
    OBJDIR.mkdir(parents=True, exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs) + 1, max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        cmp_job = ex.submit(_build_cmp, force)
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
        cmp_job.result()
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    force = "--force" in sys.argv
    print(build(force=force))
