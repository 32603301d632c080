import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs on the GPU box)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def ctx():
    from narvalengine_amd import Context
    c = Context(0)
    yield c
    c.close()


def print_case(msg: str):
    """Names the GPU case before it runs (flushed, so a crash still shows it)."""
    sys.stdout.write(f"[case] {msg}\n")
    sys.stdout.flush()
