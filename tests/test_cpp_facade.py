"""The C++ facade (include/ne_gpu.hpp): OfflineEngine tiles vs whole frame,
tile recomputation, untouched remainder pixels, Integrator::Li and
SceneReader on the reference's surfaceAndLight test scene, driven from a
g++-built program over the C ABI (tests/cpp/facade_check.cpp).  Without a
device the reader part still runs (exit 3 after it, 4 if it fails)."""
import json
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "facade_check"
RESOURCES = Path(__file__).resolve().parent / "golden" / "resources"


def _run():
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built (narvalengine_amd.build builds it)")
    print(f"[case] C++ facade check {BIN}", flush=True)
    return subprocess.run([str(BIN), str(RESOURCES)], capture_output=True, text=True, timeout=120)


def test_facade_fails_cleanly_without_a_device():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    r = _run()
    assert r.returncode == 3 and "no HIP device" in r.stderr


@pytest.mark.gpu
def test_facade_tiles_frame_and_li():
    r = _run()
    assert r.returncode == 0, r.stdout + r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["mismatches"] == 0 and out["covered"] == 80 * 40 * 3
