import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "parallel-db-buffer-management_amd"
for p in (PKG, ROOT / "oracle", ROOT / "tests"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line(
        "markers", "gpu: needs an MI355X (gfx950) GPU and the built libpbm.so")


@pytest.fixture(scope="session")
def gpu():
    """torch on a GPU; GPU tests fail (not skip) without one, so a silent
    host fallback cannot pass them."""
    import torch
    assert torch.cuda.is_available(), "GPU tests need a visible AMD GPU"
    import pbm
    pbm.lib()  # raises when libpbm.so is missing
    return torch


def log_case(msg: str) -> None:
    """Names the case on stdout before the GPU runs it."""
    print(f"[case] {msg}", flush=True)
    sys.stdout.flush()
