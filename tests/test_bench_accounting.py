"""The bench's algorithmic byte accounting (the numerator of every roofline
fraction it reports), checked on the CPU by relations between the stages
rather than by restating each stage's formula."""
import re
from pathlib import Path

import numpy as np
import pytest

import bench

ROOT = Path(__file__).resolve().parents[1]
EVENTS = ("density_lookups", "majorant_cells", "bvh_nodes", "prim_tests")


def random_counters(seed):
    rng = np.random.default_rng(seed)
    c = {"paths": int(rng.integers(1, 10**6))}
    c["ext_rays"] = c["paths"] + int(rng.integers(0, 10**7))
    c["shadow_rays"] = int(rng.integers(0, 10**7))
    for k in EVENTS:
        c["shadow_" + k] = int(rng.integers(0, 10**8))
        c[k] = c["shadow_" + k] + int(rng.integers(0, 10**8))
    return c


@pytest.mark.parametrize("seed", range(5))
def test_stages_split_the_megakernels_event_bytes(seed):
    """trace + medium + shadow read every density lookup, majorant cell, node
    and primitive the megakernel reads, each once: a stage that drops or
    double-counts an event kind breaks the sum."""
    c = random_counters(seed)
    film = 24 * 1000
    stages = sum(bench.stage_bytes(s, c, film) for s in ("trace", "medium", "shadow"))
    rays = c["ext_rays"] * (bench.RAY_BYTES["trace"] + bench.RAY_BYTES["medium"]) + \
        c["shadow_rays"] * bench.RAY_BYTES["shadow"]
    assert stages - rays == bench.stage_bytes("megakernel", c, film) - film


@pytest.mark.parametrize("seed", range(5))
def test_shading_that_tracks_its_shadow_rays_moves_the_shadow_stage_in(seed):
    """The fused shading stage reads the shadow rays' lookups and cells (the
    shadow stage's event bytes) and writes no shadow queue: no queue record
    out of shade and none read back in."""
    c = random_counters(seed)
    plain = bench.stage_bytes("shade", c, 0)
    fused = bench.stage_bytes("shade", c, 0, shadow_in_shade=True)
    shadow_events = bench.stage_bytes("shadow", c, 0) - c["shadow_rays"] * bench.RAY_BYTES["shadow"]
    assert fused == plain - c["shadow_rays"] * bench.SHADE_OUT + shadow_events
    # no shadow rays: both forms move the same bytes
    c0 = dict(c, shadow_rays=0, **{"shadow_" + k: 0 for k in EVENTS})
    assert bench.stage_bytes("shade", c0, 0, shadow_in_shade=True) == bench.stage_bytes("shade", c0, 0)


def test_regen_bytes_are_the_stores_of_a_new_path():
    """A new path costs the medium stage one record write: REGEN_BYTES is
    4 B per record array that regen_path stores (wavefront.hip), so a field
    added to the new-path record without the accounting fails here."""
    src = (ROOT / "narvalengine_amd" / "csrc" / "wavefront.hip").read_text()
    body = src[src.index("void regen_path("):]
    body = body[:body.index("\n}\n")]
    stores = set(re.findall(r"\bA\.(\w+)\[slot\]\s*=", body))
    assert len(stores) * 4 == bench.REGEN_BYTES, sorted(stores)
    c = random_counters(7)
    assert bench.stage_bytes("medium", c, 0, regen_in_medium=True) - bench.stage_bytes("medium", c, 0) == \
        c["paths"] * bench.REGEN_BYTES


def test_unknown_stage_moves_nothing():
    assert bench.stage_bytes("plan", random_counters(1), 100) == 0
