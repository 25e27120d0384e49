"""Batched rendering: many scenes of uneven size in one launch sequence,
captured in a hipGraph and replayed.  Each scene must equal the oracle (same
bar as the per-frame path) and the per-frame GPU path bit for bit.

Finalisation renders every frame once (a trial run that sizes the
workspace), so before each checked run every target and shadow map is filled
with a sentinel (colour bytes 0x5A, NaN depth): what is compared was written
by that run."""
import numpy as np
import pytest

from edge_scenes import empty_mesh
from hana_amd import scenegen as sg
from parity import assert_color_ok, assert_depth_exact
from pyoracle import CPURenderer

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from hana_amd.gpu import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def scenes():
    # a reduced config-3 mix: 1-8 displaced spheres/boxes, uneven sizes,
    # Blinn/NormalMap x shadow on/off, random aspect and cameras
    return sg.config3_scenes(count=12, tri_lo=200, tri_hi=20_000, px_lo=48 * 48, px_hi=320 * 320)


@pytest.fixture(scope="module")
def oracle_frames(scenes):
    o = CPURenderer("oracle")
    return [o.render_frame(s) for s in scenes]


def read_all(ctx, devs):
    return [d.read() for d in devs]


def poison(ctx, devs):
    """Fills every target and shadow map of the frames with the sentinel."""
    for d in devs:
        maps = [(d.target, d.scene.width, d.scene.height)]
        if d.shadow_map is not None:
            maps.append((d.shadow_map,) + tuple(d.scene.shadow_wh))
        for t, w, h in maps:
            ctx.upload(t, np.full((h, w, 4), 0x5A, np.uint8), np.full((h, w), np.nan, np.float32))


def test_config3_full_batch(ctx):
    """The whole batched workload: 256 scenes of 1-8 displaced spheres/boxes
    (1k-200k triangles, 128^2-1024^2 pixels, random aspect, Blinn/NormalMap x
    shadow on/off) in one captured graph, against the oracle frame by frame:
    depth bit-exact, colour within the bar, on two runs."""
    from hana_amd.gpu import Batch, SceneOnDevice
    scenes = sg.config3_scenes()
    assert len(scenes) == 256
    o = CPURenderer("oracle")
    refs = [o.render_frame(s)[0] for s in scenes]
    devs = [SceneOnDevice(ctx, s) for s in scenes]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.finalize(use_graph=True)
    for run in range(2):
        poison(ctx, devs)
        batch.run()
        for s, d, ref in zip(scenes, devs, refs):
            col, dep = d.read()
            assert_depth_exact(dep, ref.depth, f"{s.name} run {run}")
            assert_color_ok(col, ref.color, f"{s.name} run {run}")
    batch.close()


@pytest.mark.parametrize("tile", [16, 32])
def test_many_tiny_frames(ctx, tile):
    """Forty frames of 8^2-40^2 pixels (1-9 tiles each) in one batch part: the
    raster's runs of 64 workgroups each span many passes, so every workgroup
    steps to its pass from the run's first.  Against the oracle, twice."""
    from hana_amd.gpu import Batch, SceneOnDevice
    tiny = sg.config3_scenes(count=40, tri_lo=50, tri_hi=2_000, px_lo=8 * 8, px_hi=40 * 40)
    o = CPURenderer("oracle")
    refs = [o.render_frame(sc)[0] for sc in tiny]
    devs = [SceneOnDevice(ctx, sc) for sc in tiny]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.set_tuning(tile, 0, 1, 4)
    batch.finalize(use_graph=True)
    for run in range(2):
        poison(ctx, devs)
        batch.run()
        for sc, (col, dep), ref in zip(tiny, read_all(ctx, devs), refs):
            assert_depth_exact(dep, ref.depth, f"{sc.name} run {run}")
            assert_color_ok(col, ref.color, f"{sc.name} run {run}")
    batch.close()


def test_batch_matches_oracle_and_per_frame(ctx, scenes, oracle_frames):
    from hana_amd.gpu import Batch, SceneOnDevice
    devs = [SceneOnDevice(ctx, s) for s in scenes]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.finalize(use_graph=True)
    poison(ctx, devs)
    batch.run()
    out_b = read_all(ctx, devs)
    for s, (col, dep), (tgt, _, _) in zip(scenes, out_b, oracle_frames):
        assert_depth_exact(dep, tgt.depth, s.name)
        assert_color_ok(col, tgt.color, s.name)
    # per-frame path on fresh targets gives the same bits
    for s, d, (col, dep) in zip(scenes, devs, out_b):
        d2 = SceneOnDevice(ctx, s)
        d2.render()
        c2, z2 = d2.read()
        assert np.array_equal(c2, col) and np.array_equal(z2.view(np.uint32), dep.view(np.uint32)), s.name
    batch.close()


def test_graph_replays(ctx, scenes, oracle_frames):
    """Replays of the captured graph, with eager work in between, all equal."""
    from hana_amd.gpu import Batch, SceneOnDevice
    devs = [SceneOnDevice(ctx, s) for s in scenes[:6]]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.finalize(use_graph=True)
    first = None
    for rep in range(4):
        poison(ctx, devs)
        batch.run()
        # eager work between replays: a per-frame render on another target
        extra = SceneOnDevice(ctx, scenes[6 + rep % 4])
        extra.render()
        ctx.sync()
        out = read_all(ctx, devs)
        if first is None:
            first = out
            for s, (col, dep), (tgt, _, _) in zip(scenes[:6], out, oracle_frames[:6]):
                assert_depth_exact(dep, tgt.depth, s.name)
                assert_color_ok(col, tgt.color, s.name)
        else:
            for (c0, z0), (c1, z1) in zip(first, out):
                assert np.array_equal(c0, c1) and np.array_equal(z0.view(np.uint32), z1.view(np.uint32)), rep
    batch.close()


def test_batch_without_graph(ctx, scenes, oracle_frames):
    from hana_amd.gpu import Batch, SceneOnDevice
    devs = [SceneOnDevice(ctx, s) for s in scenes[:5]]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.finalize(use_graph=False)
    poison(ctx, devs)
    batch.run()
    for s, (col, dep), (tgt, _, _) in zip(scenes[:5], read_all(ctx, devs), oracle_frames[:5]):
        assert_depth_exact(dep, tgt.depth, s.name)
        assert_color_ok(col, tgt.color, s.name)
    batch.close()


@pytest.mark.parametrize("pick", [[3], [0, 7, 11]])
@pytest.mark.parametrize("use_graph", [True, False])
def test_small_batches(ctx, scenes, oracle_frames, pick, use_graph):
    """One frame (a single part) and three frames (two uneven parts that run
    concurrently), captured and eager, replayed twice."""
    from hana_amd.gpu import Batch, SceneOnDevice
    devs = [SceneOnDevice(ctx, scenes[i]) for i in pick]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.finalize(use_graph=use_graph)
    for _ in range(2):
        poison(ctx, devs)
        batch.run()
        for i, (col, dep) in zip(pick, read_all(ctx, devs)):
            tgt = oracle_frames[i][0]
            assert_depth_exact(dep, tgt.depth, scenes[i].name)
            assert_color_ok(col, tgt.color, scenes[i].name)
    batch.close()


def test_batch_rejects_shared_targets(ctx, scenes):
    from hana_amd.gpu import Batch, HanaError, SceneOnDevice
    d = SceneOnDevice(ctx, scenes[0])
    batch = Batch(ctx)
    batch.add(d.desc())
    batch.add(d.desc())
    with pytest.raises(HanaError):
        batch.finalize()
    batch.close()


@pytest.mark.parametrize("empty_first", [False, True])
def test_empty_mesh_frames(ctx, empty_first):
    """A batch whose parts include one with only an empty mesh (no vertex, no
    face), in a graph and eagerly, each run twice: that frame's target keeps
    its contents and the other frame equals the oracle."""
    from hana_amd.gpu import Batch, SceneOnDevice
    o = CPURenderer("oracle")
    full = sg.small_scene(sg.BLINN, shadow=True, seed=77, width=120, height=90)
    empty = sg.Scene(empty_mesh(), 96, 64, sg.NORMALMAP, True, full.cam_pos, full.light_pos, full.diffuse,
                     full.normal, name="empty")
    scenes = [empty, full] if empty_first else [full, empty]
    refs = [o.render_frame(s)[0] for s in scenes]
    for graph in (True, False):
        devs = [SceneOnDevice(ctx, s) for s in scenes]
        batch = Batch(ctx)
        for d in devs:
            batch.add(d.desc())
        batch.finalize(use_graph=graph)
        for _ in range(2):
            poison(ctx, devs)
            batch.run()
            for s, d, ref in zip(scenes, devs, refs):
                col, dep = d.read()
                assert_depth_exact(dep, ref.depth, s.name)
                assert_color_ok(col, ref.color, s.name)
        batch.close()


@pytest.mark.parametrize("tile,wgs,parts,waves", [(32, 0, 4, 4), (16, 4, 2, 4), (32, 8, 1, 2), (16, 2, 4, 4),
                                                  (16, 0, 4, 2), (16, 4, 4, 2)])
def test_batch_tunings(ctx, scenes, oracle_frames, tile, wgs, parts, waves):
    """Every batch tuning (tile size, persistent raster, raster waves, parts)
    renders the oracle's frames, on two graph runs."""
    from hana_amd.gpu import Batch, SceneOnDevice
    devs = [SceneOnDevice(ctx, s) for s in scenes]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.set_tuning(tile, wgs, parts, waves)
    batch.finalize(use_graph=True)
    t = batch.tuning()
    assert (t["tile_size"], t["raster_wgs_per_cu"], t["parts"], t["raster_waves"]) == (tile, wgs, parts, waves)
    for run in range(2):
        poison(ctx, devs)
        batch.run()
        for s, (col, dep), (tgt, _, _) in zip(scenes, read_all(ctx, devs), oracle_frames):
            assert_depth_exact(dep, tgt.depth, f"{s.name} run {run}")
            assert_color_ok(col, tgt.color, f"{s.name} run {run}")
    batch.close()


def test_autotune_picks_a_tuning_and_renders_exactly(ctx, scenes, oracle_frames):
    """hana_batch_autotune times the candidates, finalizes with the fastest,
    and the tuned batch still renders the oracle's frames."""
    from hana_amd.gpu import Batch, SceneOnDevice
    devs = [SceneOnDevice(ctx, s) for s in scenes]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    chosen = batch.autotune(use_graph=True, reps=2)
    assert chosen["tile_size"] in (16, 32) and chosen["raster_wgs_per_cu"] in (0, 2, 4, 8)
    assert chosen["parts"] in (1, 2, 4) and chosen["raster_waves"] in (2, 4) and chosen["ms"] > 0
    assert batch.tuning() == chosen
    poison(ctx, devs)
    batch.run()
    for s, (col, dep), (tgt, _, _) in zip(scenes, read_all(ctx, devs), oracle_frames):
        assert_depth_exact(dep, tgt.depth, s.name)
        assert_color_ok(col, tgt.color, s.name)
    batch.close()
