"""ne_render_noalloc against ne_render: the allocation-free entry point keeps
the megakernel's chunk partial sums in the static device scratch, and its film
sums equal ne_render's (megakernel, same params) bit for bit, on two calls in a
row, the second on a smaller rectangle, and across precisions."""
import numpy as np
import pytest

import scenes
from conftest import print_case
from narvalengine_amd.abi import NE_FP32, NE_FP64, NE_MODE_MEGAKERNEL

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [NE_FP32, NE_FP64])
def test_noalloc_render_equals_render(ctx, precision):
    print_case(f"noalloc vs render, precision {precision}")
    g = ctx.upload(scenes.small_volume(res=32, maj_res=4))
    cam = scenes.VOLUME_CAMERA
    # first call: 64 x 48 pixels, 4 chunks of 8 samples; second call: a smaller rectangle, 3 chunks
    for rect, spp, spt in [((0, 0, 64, 48), 32, 8), ((5, 7, 29, 20), 12, 4)]:
        ref, got = ctx.film(64, 48), ctx.film(64, 48)
        kw = dict(spp=spp, max_bounces=16, seed=3, rect=rect, precision=precision, samples_per_thread=spt)
        ctx.render(g, cam, ref, mode=NE_MODE_MEGAKERNEL, **kw)
        ctx.render(g, cam, got, noalloc=True, **kw)
        a, b = ref.read(), got.read()
        assert a.any()
        np.testing.assert_array_equal(b, a)
        ref.close()
        got.close()


def test_noalloc_render_rejects_what_it_cannot_hold(ctx):
    print_case("noalloc limits")
    from narvalengine_amd import NarvalError
    from narvalengine_amd.abi import NE_MODE_WAVEFRONT
    g = ctx.upload(scenes.small_volume(res=8, maj_res=2))
    f = ctx.film(2048, 2048)
    # 2048^2 pixels x 2 chunks x 24 B = 192 MiB of partial sums > 32 MiB
    with pytest.raises(NarvalError, match="static scratch"):
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=2, max_bounces=2, samples_per_thread=1, noalloc=True)
    with pytest.raises(NarvalError, match="megakernel"):
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=1, max_bounces=2, mode=NE_MODE_WAVEFRONT, noalloc=True)
    with pytest.raises(NarvalError, match="count"):
        ctx.render(g, scenes.VOLUME_CAMERA, f, spp=1, max_bounces=2, count=True, noalloc=True)
    assert not f.read().any()
    f.close()
