"""Consecutive draws with no read in between, on small (150x110) and large
(2048x2048) targets.  The per-call path rotates over streams with their
own workspaces, so a draw's geometry stages overlap the rasters before it;
the results must still be those of the calls run one after another:
depth tests across draws, shadow maps written by one call and sampled by the
next, targets read, uploaded or presented between calls, timing switched
on in the middle of a chain, chains that mix calls of at least 2^19 faces
(which rotate over two slots) with smaller ones (all four), and calls on
several targets interleaved, which wait only for the calls that share a
target or shadow map with them."""
import numpy as np
import pytest

from hana_amd import scenegen as sg
from parity import assert_color_ok, assert_depth_exact
from pyoracle import CPURenderer, HostTarget

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", params=[(150, 110), (2048, 2048)], ids=["small", "large"])
def wh(request):
    return request.param


@pytest.fixture(scope="module")
def ctx():
    from hana_amd.gpu import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    return CPURenderer("oracle")


def chain_scenes(n, seed0, wh):
    kinds = [sg.NORMALMAP, sg.BLINN, sg.TEXTURE_LIGHT, sg.TOON, sg.GROUND, sg.TEXTURE]
    cams = [(0.0, 0.0, 2.0), (0.4, 0.3, 1.8), (-0.5, 0.1, 1.6), (0.2, -0.6, 2.2)]
    return [sg.small_scene(kinds[i % len(kinds)], seed=seed0 + i, width=wh[0], height=wh[1], cam=cams[i % len(cams)])
            for i in range(n)]


def oracle_chain(oracle, scenes, target=None):
    t = target or HostTarget(scenes[0].width, scenes[0].height)
    for sc in scenes:
        oracle.draw(sc.mesh, sc.shader, oracle.build_uniforms(sc), t, sc.diffuse, sc.normal)
    return t


@pytest.mark.parametrize("n", [2, 3, 7])
def test_draw_chain_into_one_target(ctx, oracle, wh, n):
    """n draws of different meshes into one target, depth-tested against each other."""
    from hana_amd.gpu import build_uniforms
    W, H = wh
    scenes = chain_scenes(n, 300 + n, wh)
    tgt = ctx.target(W, H)
    keep = []
    for sc in scenes:
        u = build_uniforms(sc)
        keep.append(u)
        ctx.draw(ctx.mesh(sc.mesh), sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
    col, dep = ctx.read(tgt, W, H)
    ref = oracle_chain(oracle, scenes)
    assert_depth_exact(dep, ref.depth, f"chain{n}")
    assert_color_ok(col, ref.color, f"chain{n}")


def test_alternating_frames_with_shadows(ctx, oracle, wh):
    """Two shadowed scenes rendered in turn, each frame drawing its shadow map
    and sampling it in the next call; read only at the end."""
    from hana_amd.gpu import SceneOnDevice
    W, H = wh
    scenes = [sg.small_scene(k, shadow=True, seed=310 + k, width=W, height=H) for k in (sg.NORMALMAP, sg.BLINN)]
    devs = [SceneOnDevice(ctx, sc) for sc in scenes]
    for _ in range(3):
        for d in devs:
            d.clear()
            d.render()
    for sc, d in zip(scenes, devs):
        col, dep = d.read()
        ref, _, _ = oracle.render_frame(sc)
        assert_depth_exact(dep, ref.depth, sc.name)
        assert_color_ok(col, ref.color, sc.name)


def test_reads_uploads_and_presents_between_draws(ctx, oracle, wh):
    """Target operations on the context stream are ordered after draws of
    either slot, and later draws after them."""
    from hana_amd.gpu import build_uniforms
    W, H = wh
    scenes = chain_scenes(5, 320, wh)
    tgt = ctx.target(W, H)
    ref = HostTarget(W, H)
    keep = []

    def draw(i):
        sc = scenes[i]
        u = build_uniforms(sc)
        keep.append(u)
        ctx.draw(ctx.mesh(sc.mesh), sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
        oracle_chain(oracle, [sc], ref)

    draw(0)
    draw(1)
    col, dep = ctx.read(tgt, W, H)   # after a draw on the second slot
    assert_depth_exact(dep, ref.depth, "read1")
    draw(2)
    out = ctx.present(tgt, W, H)
    assert_color_ok(out[::-1][..., [2, 1, 0, 3]], ref.color, "present")
    # replace the target's contents, then keep drawing over them
    rng = np.random.default_rng(5)
    up_c = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    up_d = rng.uniform(0.2, 1.0, (H, W)).astype(np.float32)
    ctx.upload(tgt, up_c, up_d)
    ref.color[...] = up_c
    ref.depth[...] = up_d
    draw(3)
    ctx.set_timing(True)   # switches to one stream mid-chain
    draw(4)
    ctx.set_timing(False)
    ctx.read_timing()
    col, dep = ctx.read(tgt, W, H)
    assert_depth_exact(dep, ref.depth, "after upload")
    assert_color_ok(col, ref.color, "after upload")


def test_large_and_small_calls_mixed(ctx, oracle, wh):
    """Calls of 588 k faces (two slots) between calls of a few thousand
    (all four slots), into one target: every raster waits for the call before
    it whatever slots the two use."""
    from hana_amd.gpu import build_uniforms
    W, H = wh
    big = sg.small_scene(sg.BLINN, seed=330, width=W, height=H, segments=768, rings=384, cam=(0.3, 0.2, 1.9))
    small = chain_scenes(4, 340, wh)
    order = [small[0], big, small[1], small[2], big, small[3], big]
    tgt = ctx.target(W, H)
    keep, meshes = [], {}
    for sc in order:
        u = build_uniforms(sc)
        keep.append(u)
        m = meshes.setdefault(id(sc), ctx.mesh(sc.mesh))
        ctx.draw(m, sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
    col, dep = ctx.read(tgt, W, H)
    ref = oracle_chain(oracle, order)
    assert_depth_exact(dep, ref.depth, "mixed")
    assert_color_ok(col, ref.color, "mixed")


def test_interleaved_targets(ctx, oracle, wh):
    """28 calls in seeded random order over four targets: plain draws, and
    shadow passes into one target sampled by a main pass into another.  Calls
    on disjoint targets may run concurrently; every target still equals the
    oracle running the calls one after another, and so does a read in the
    middle."""
    from hana_amd.gpu import SHADOW, build_uniforms
    W, H = wh
    rng = np.random.default_rng(77)
    kinds = [sg.BLINN, sg.NORMALMAP, sg.TOON, sg.TEXTURE]
    scenes = [sg.small_scene(kinds[i % 4], shadow=True, seed=350 + i, width=W, height=H,
                             cam=[(0.0, 0.0, 2.0), (0.4, 0.3, 1.8), (-0.5, 0.1, 1.6)][i % 3]) for i in range(6)]
    devm = [ctx.mesh(sc.mesh) for sc in scenes]
    devt = [(ctx.texture(sc.diffuse), ctx.texture(sc.normal)) for sc in scenes]
    tg = [ctx.target(W, H) for _ in range(4)]
    ref = [HostTarget(W, H) for _ in range(4)]
    keep = []
    for step in range(28):
        i = int(rng.integers(6))
        sc = scenes[i]
        u = build_uniforms(sc)
        ou = oracle.build_uniforms(sc)
        keep.append(u)
        j = int(rng.integers(4))
        if rng.random() < 0.35:
            m = int((j + 1 + rng.integers(3)) % 4)        # main pass into another target, sampling j
            ctx.draw(devm[i], SHADOW, u, tg[j], *devt[i])
            ctx.draw(devm[i], sc.shader, u, tg[m], *devt[i], shadow_map=tg[j])
            oracle.draw(sc.mesh, sg.SHADOW, ou, ref[j], sc.diffuse, sc.normal)
            oracle.draw(sc.mesh, sc.shader, ou, ref[m], sc.diffuse, sc.normal, ref[j])
        else:
            ctx.draw(devm[i], sc.shader, u, tg[j], *devt[i])
            oracle.draw(sc.mesh, sc.shader, ou, ref[j], sc.diffuse, sc.normal)
        if step == 13:
            col, dep = ctx.read(tg[2], W, H)
            assert_depth_exact(dep, ref[2].depth, "middle")
    for k in range(4):
        col, dep = ctx.read(tg[k], W, H)
        assert_depth_exact(dep, ref[k].depth, f"target {k}")
        assert_color_ok(col, ref[k].color, f"target {k}")
