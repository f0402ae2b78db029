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


def test_reference_kmalloc_driver(gpu):
    exe = pbm_build.compat_binary("kmalloc")
    assert exe.exists(), f"{exe} is missing: run __graft_entry__.build()"
    log_case("compat kernel malloc driver")
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "compat kmalloc ok" in proc.stdout


def test_reference_ouroboros_driver(gpu):
    exe = pbm_build.compat_binary("ouroboros")
    assert exe.exists(), f"{exe} is missing: run __graft_entry__.build()"
    log_case("compat Ouroboros driver")
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    assert "compat ouroboros ok" in proc.stdout
