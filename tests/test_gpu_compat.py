"""The reference's own API names (pbm_compat.hpp) driving the library from
a HIP program built the way a reference user builds theirs: pre-fill,
getPage per thread, uniqueness and occupancy checks, freePage, and for
RWMalloc mallocRW/freeRW.  One program per policy."""
import subprocess

import pytest

import pbm_build
from conftest import log_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", sorted(pbm_build.COMPAT_STRATEGIES))
def test_reference_api_driver(gpu, name):
    exe = pbm_build.compat_binary(name)
    assert exe.exists(), f"{exe} is missing: run __graft_entry__.build()"
    log_case(f"compat driver {name}")
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "compat ok" in proc.stdout
    if name == "collab":
        # collectWarpData prints one line per step: the free bits found by
        # the warp's 32 lanes (at most 32 words of 64 bits)
        lines = proc.stdout.split("collect begin\n", 1)[1].split("collect end", 1)[0].split()
        assert lines and all(0 <= int(v) <= 32 * 64 for v in lines)


@pytest.mark.parametrize("name,ok", [("kmalloc", "compat kmalloc ok"),
                                     ("kmalloc_legacy", "compat kmalloc legacy ok")])
def test_reference_kmalloc_driver(gpu, name, ok):
    # malloc.cuh's classes with the current names and with the legacy header
    exe = pbm_build.compat_binary(name)
    assert exe.exists(), f"{exe} is missing: run __graft_entry__.build()"
    log_case(f"compat kernel malloc driver {name}")
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert ok in proc.stdout


def test_reference_ouroboros_driver(gpu):
    exe = pbm_build.compat_binary("ouroboros")
    assert exe.exists(), f"{exe} is missing: run __graft_entry__.build()"
    log_case("compat Ouroboros driver")
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "compat ouroboros ok" in proc.stdout


def test_reference_legacy_driver(gpu):
    exe = pbm_build.compat_binary("legacy")
    assert exe.exists(), f"{exe} is missing: run __graft_entry__.build()"
    log_case("compat legacy driver")
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "compat legacy ok" in proc.stdout
    # printNumPagesLeft of each family after its gets: 262,144 - 50,000
    assert proc.stdout.count("Number of free pages: 212144 ") == 9
    mon = proc.stdout.split("monitor begin\n", 1)[1].split("monitor end", 1)[0]
    assert mon.count("[RW info] Number of free pages: 262144, percentage free: 100.0 %") == 3
