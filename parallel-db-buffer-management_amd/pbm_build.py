"""Build recipe of libpbm.so: every csrc/*.hip compiled by hipcc for gfx950.

Run as a script or call build_library().  Objects go to _build/ next to this
file, the shared library to libpbm.so in this directory (written to a
temporary name first, then renamed, so a concurrent loader never sees half
a file).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
INCLUDE = ROOT / "include"
CSRC = PKG / "csrc"
OBJDIR = PKG / "_build"
LIB = PKG / "libpbm.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops: keep the compiler from emitting packed fp32 VALU ops.
HIP_FLAGS = [
    "-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", str(INCLUDE), "-I", str(CSRC),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n"
                           f"{proc.stdout}\n{proc.stderr}")


def build_library(jobs: int = 8) -> Path:
    OBJDIR.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    objs = [OBJDIR / (s.stem + ".o") for s in sources]
    with cf.ThreadPoolExecutor(max_workers=max(1, min(jobs, len(sources)))) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)])
                for s, o in zip(sources, objs)]
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    return LIB


COMPAT_SRC = ROOT / "tests" / "compat" / "compat_check.hip"
COMPAT_STRATEGIES = {"rw_bm": "PBM_RW_BM", "collab": "PBM_COLLAB_RW_BM", "crw": "PBM_CRW_BM",
                     "rw_malloc": "PBM_RW_MALLOC", "rw_basic": "PBM_RW_BASIC",
                     "crw_basic": "PBM_CRW_BASIC", "collab_basic": "PBM_COLLAB_RW_BASIC"}


def compat_binary(name: str) -> Path:
    return OBJDIR / f"compat_check_{name}"


COMPAT_KMALLOC_SRC = ROOT / "tests" / "compat" / "compat_kmalloc.hip"
COMPAT_OUROBOROS_SRC = ROOT / "tests" / "compat" / "compat_ouroboros.hip"
COMPAT_LEGACY_SRC = ROOT / "tests" / "compat" / "compat_legacy.hip"


def build_compat_checks(jobs: int = 4) -> list[Path]:
    """The reference-API drivers against pbm_compat.hpp and libpbm.so:
    tests/compat/compat_check.hip once per policy, the kernel-malloc driver
    tests/compat/compat_kmalloc.hip, the Ouroboros-wrapper driver
    tests/compat/compat_ouroboros.hip and the legacy-API driver
    tests/compat/compat_legacy.hip."""
    OBJDIR.mkdir(exist_ok=True)
    jobs_list = [([f"-DPBM_COMPAT_STRATEGY={m}", str(COMPAT_SRC)], compat_binary(n))
                 for n, m in COMPAT_STRATEGIES.items()]
    jobs_list.append(([str(COMPAT_KMALLOC_SRC)], compat_binary("kmalloc")))
    jobs_list.append((["-DPBM_KMALLOC_LEGACY", str(COMPAT_KMALLOC_SRC)],
                      compat_binary("kmalloc_legacy")))
    jobs_list.append(([str(COMPAT_OUROBOROS_SRC)], compat_binary("ouroboros")))
    jobs_list.append(([str(COMPAT_LEGACY_SRC)], compat_binary("legacy")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, *src, "-o", str(o), "-L", str(PKG), "-lpbm",
                                 "-Wl,-rpath,$ORIGIN/.."])
                for src, o in jobs_list]
        for f in futs:
            f.result()
    return [o for _, o in jobs_list]


if __name__ == "__main__":
    print(build_library(), file=sys.stderr)
    print(build_compat_checks(), file=sys.stderr)
