"""hana_render_frame_noalloc (the frame entry point whose workspace is a
fixed device array, for callers that cannot allocate) against
hana_render_frame: the same bits, on two calls in a row (the second frame
smaller than the first, so the array keeps the first call's contents), under
the default and the persistent-raster tunings; a frame too large for the
array is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from hana_amd import scenegen as sg

pytestmark = pytest.mark.gpu


def noalloc(ctx, dev):
    d = dev.desc()
    ctx.check(ctx.lib.hana_render_frame_noalloc(ctx.h, C.byref(d)), "hana_render_frame_noalloc")


@pytest.mark.parametrize("tile,wgs,waves", [(16, 0, 4), (32, 4, 2)])
def test_noalloc_matches_render_frame(tile, wgs, waves):
    from hana_amd.gpu import Context, SceneOnDevice
    big = sg.small_scene(sg.NORMALMAP, shadow=True, seed=31, width=640, height=480, segments=96, rings=48,
                         cam=(0.4, 0.3, 1.1))
    small = sg.small_scene(sg.BLINN, shadow=True, seed=32, width=160, height=90)
    with Context(0) as ctx:
        ctx.set_tuning(tile, wgs, raster_waves=waves)
        na = [SceneOnDevice(ctx, s) for s in (big, small)]
        for dev in na:                       # two calls in a row, no wait between them
            noalloc(ctx, dev)
        ctx.sync()
        got = [dev.read() for dev in na]
        ref = []
        for s in (big, small):
            dev = SceneOnDevice(ctx, s)
            dev.render()
            ref.append(dev.read())
        for s, (c0, d0), (c1, d1) in zip((big, small), got, ref):
            assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32)), s.name
            assert np.array_equal(c0, c1), s.name
        # the post-frame shadow-map clear and the next frame: render both again
        for dev in na:
            dev.clear()
            noalloc(ctx, dev)
        ctx.sync()
        for dev, (c1, d1) in zip(na, ref):
            c0, d0 = dev.read()
            assert np.array_equal(d0.view(np.uint32), d1.view(np.uint32))
            assert np.array_equal(c0, c1)


def test_noalloc_refuses_what_does_not_fit():
    """A frame that needs more than the scratch array is refused with
    HANA_E_NOMEM, and the context renders on; a call while per-call draws are
    pending is refused with HANA_E_INVALID."""
    from hana_amd.gpu import HANA_E_INVALID, HANA_E_NOMEM, Context, HanaError, SceneOnDevice
    huge = sg.config2_scene()
    small = sg.small_scene(sg.BLINN, shadow=True, seed=33, width=96, height=64)
    with Context(0) as ctx:
        dev = SceneOnDevice(ctx, huge)
        with pytest.raises(HanaError) as err:
            noalloc(ctx, dev)
        assert err.value.code == HANA_E_NOMEM and "scratch" in str(err.value)
        s = SceneOnDevice(ctx, small)
        s.render()                           # a per-call draw, not yet waited for
        with pytest.raises(HanaError) as err:
            noalloc(ctx, s)
        assert err.value.code == HANA_E_INVALID
        ctx.sync()
        ref = s.read()
        noalloc(ctx, s)
        c, d = s.read()
        assert np.array_equal(d.view(np.uint32), ref[1].view(np.uint32)) and np.array_equal(c, ref[0])
