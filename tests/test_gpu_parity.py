"""GPU pipeline against the oracle on the same seeded inputs (through the C ABI).

Bar: depth and the winning primitive of every pixel bit-exact; colour within
1 LSB on >= 99.9 % of pixels and 2 LSB everywhere; the shadow map (R8 of the
interpolated light-space depth) bit-exact.
"""
import numpy as np
import pytest

from edge_scenes import (IDENTITY_VP, empty_mesh, faceless, nan_depth_scene, screen_inputs, sliver_scene,
                         tiny_input_scene, with_camera_vp)
from hana_amd import scenegen as sg
from parity import assert_color_ok, assert_depth_exact
from pyoracle import CPURenderer, HostTarget

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from hana_amd.gpu import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    return CPURenderer("oracle")


def gpu_passes(ctx, scene, winner=True):
    """Shadow pass then main pass as two hana_draw calls; returns target and shadow map."""
    from hana_amd.gpu import SHADOW, build_uniforms
    u = build_uniforms(scene)
    mesh = ctx.mesh(scene.mesh)
    dif = ctx.texture(scene.diffuse)
    nrm = ctx.texture(scene.normal)
    tgt = ctx.target(scene.width, scene.height, winner=winner)
    sm = None
    smap = None
    if scene.shadow:
        sw, sh = scene.shadow_wh
        sm = ctx.target(sw, sh, winner=winner)
        ctx.draw(mesh, SHADOW, u, sm, dif, nrm)
    ctx.draw(mesh, scene.shader, u, tgt, dif, nrm, sm)
    col, dep = ctx.read(tgt, scene.width, scene.height)
    win = ctx.read_winner(tgt, scene.width, scene.height) if winner else None
    if sm is not None:
        sw, sh = scene.shadow_wh
        sc, sd = ctx.read(sm, sw, sh)
        swin = ctx.read_winner(sm, sw, sh) if winner else None
        smap = (sc, sd, swin)
    return col, dep, win, smap


def oracle_passes(oracle, scene):
    u = oracle.build_uniforms(scene)
    tgt = HostTarget(scene.width, scene.height)
    sm = None
    st_s = None
    if scene.shadow:
        sw, sh = scene.shadow_wh
        sm = HostTarget(sw, sh)
        st_s, swin = oracle.draw(scene.mesh, sg.SHADOW, u, sm, scene.diffuse, scene.normal, winner=True)
        sm.winner = swin
    st, win = oracle.draw(scene.mesh, scene.shader, u, tgt, scene.diffuse, scene.normal, sm, winner=True)
    return tgt, win, sm, st, st_s


def as_u32_ids(owin):
    """Oracle winner map (int32, -1 = none) in the GPU encoding (uint32, 0xFFFFFFFF = none)."""
    w = owin.astype(np.int64)
    w[w < 0] = 0xFFFFFFFF
    return w


def check_scene(ctx, oracle, scene):
    col, dep, win, smap = gpu_passes(ctx, scene)
    tgt, owin, osm, st, st_s = oracle_passes(oracle, scene)
    if scene.shadow:
        sc, sd, swin = smap
        assert_depth_exact(sd, osm.depth, scene.name + " shadow depth")
        sw_o = as_u32_ids(osm.winner)
        assert np.array_equal(swin.astype(np.int64), sw_o), \
            f"{scene.name}: shadow coverage differs at {int((swin.astype(np.int64) != sw_o).sum())} pixels"
        assert np.array_equal(sc[..., :3], osm.color[..., :3]), scene.name + " shadow colour"
    owin_u = as_u32_ids(owin)
    assert np.array_equal(win.astype(np.int64), owin_u), \
        f"{scene.name}: coverage differs at {int((win.astype(np.int64) != owin_u).sum())} pixels"
    assert_depth_exact(dep, tgt.depth, scene.name)
    rep = assert_color_ok(col, tgt.color, scene.name)
    return rep, st


SHADERS = [sg.GROUND, sg.TOON, sg.TEXTURE, sg.TEXTURE_LIGHT, sg.BLINN, sg.NORMALMAP]


@pytest.mark.parametrize("kind", SHADERS)
def test_small_each_shader(ctx, oracle, kind):
    scene = sg.small_scene(kind, shadow=False, seed=10 + kind)
    check_scene(ctx, oracle, scene)


@pytest.mark.parametrize("kind", [sg.BLINN, sg.NORMALMAP])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_small_shadowed(ctx, oracle, kind, seed):
    scene = sg.small_scene(kind, shadow=True, seed=seed)
    check_scene(ctx, oracle, scene)


@pytest.mark.parametrize("cam", [(0.3, 0.2, 1.2), (0.0, 0.0, 1.05), (0.9, -0.4, 0.6), (0.0, 1.1, 0.2)])
def test_small_clipping(ctx, oracle, cam):
    """Cameras close to or inside the mesh: near-plane and screen-edge clipping."""
    scene = sg.small_scene(sg.NORMALMAP, shadow=True, seed=7, cam=cam, width=128, height=96)
    check_scene(ctx, oracle, scene)


@pytest.mark.parametrize("wh", [(1, 1), (7, 5), (33, 31), (65, 17), (200, 37)])
def test_ragged_sizes(ctx, oracle, wh):
    scene = sg.small_scene(sg.BLINN, shadow=True, seed=5, width=wh[0], height=wh[1])
    check_scene(ctx, oracle, scene)


def test_config1(ctx, oracle):
    for kind in (sg.NORMALMAP, sg.BLINN):
        scene = sg.config1_scene(kind)
        check_scene(ctx, oracle, scene)


def test_render_frame_semantics(ctx, oracle):
    """DrawModel::draw: shadow pass, main pass, shadow map cleared afterwards;
    two frames in a row give the same image."""
    from hana_amd.gpu import SceneOnDevice
    scene = sg.small_scene(sg.NORMALMAP, shadow=True, seed=21)
    dev = SceneOnDevice(ctx, scene)
    dev.render()
    col1, dep1 = dev.read()
    sw, sh = scene.shadow_wh
    sc, sd = ctx.read(dev.shadow_map, sw, sh)
    assert (sc[..., :3] == 0).all() and (sd == np.float32(3.4028234663852886e38)).all()
    dev.clear()
    dev.render()
    col2, dep2 = dev.read()
    assert np.array_equal(col1, col2) and np.array_equal(dep1.view(np.uint32), dep2.view(np.uint32))
    tgt, st_s, st_m = oracle.render_frame(scene)
    assert_depth_exact(dep1, tgt.depth, "frame")
    assert_color_ok(col1, tgt.color, "frame")


@pytest.mark.parametrize("wh", [(4096, 4096), (4112, 4096)])
def test_tile_scan_large_targets(ctx, oracle, wh):
    """Large targets: 65,536 tiles (16 scan chunks) and 65,792 (a partial
    17th chunk) give the oracle's frame."""
    scene = sg.small_scene(sg.BLINN, seed=77, width=wh[0], height=wh[1], segments=48, rings=24)
    check_scene(ctx, oracle, scene)


def soup_scene(n, width, height, px_lo, px_hi, seed):
    """n front-facing triangles of px_lo..px_hi pixels across, scattered in
    random order over the view of the z = 0 plane from (0, 0, 2)."""
    rng = np.random.default_rng(seed)
    half = 2.0 * np.tan(np.radians(30.0))                  # visible half-height at z = 0
    aspect = width / height
    centre = rng.uniform(-0.95, 0.95, (n, 2)) * (half * aspect, half)
    size = half * 2 / height * rng.uniform(px_lo, px_hi, n) / 2
    ang = rng.uniform(0, 2 * np.pi, n)
    corners = np.stack([ang, ang + 2.1, ang + 4.2], 1)     # counter-clockwise: front-facing
    xy = centre[:, None, :] + size[:, None, None] * np.stack([np.cos(corners), np.sin(corners)], 2)
    z = np.repeat(rng.uniform(-0.3, 0.3, (n, 1, 1)), 3, 1)
    pos = np.concatenate([xy, z], 2).reshape(-1, 3).astype(np.float32)
    nrm = sg.fixed_point_normals(np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (3 * n, 1)))
    uv = rng.uniform(0, 1, (3 * n, 2)).astype(np.float32)
    mesh = sg.Mesh(pos, nrm, uv, np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    base = sg.small_scene(sg.BLINN, width=width, height=height, seed=seed)
    return sg.Scene(mesh, width, height, sg.BLINN, False, (0.0, 0.0, 2.0), base.light_pos, base.diffuse,
                    base.normal, name=f"soup{n}")


@pytest.mark.parametrize("n,wh,px", [(70_000, (512, 512), (1.5, 4.0)), (60_000, (2048, 2048), (8.0, 40.0))],
                         ids=["small-tris", "wide"])
def test_scattered_triangle_soup(ctx, oracle, n, wh, px):
    """Triangles in random order, so each segment of 1024 faces touches more
    distinct tiles than the tile tables of set-up and binning hold (512 slots
    for sequences of small triangles, at most 4 target pixels per face; 1024
    and 2048 otherwise): those tiles fall back to the global counters and
    cursors, and the frame still equals the oracle's."""
    scene = soup_scene(n, wh[0], wh[1], px[0], px[1], seed=2024 + n)
    # the premise: most segments list more distinct tiles than the tables hold
    w, h = wh
    p = scene.mesh.pos[scene.mesh.idx]                     # (F, 3, 3); eye at z = 2, 60 degree fovy
    k = 1.0 / (np.tan(np.radians(30.0)) * (2.0 - p[..., 2]))
    sx = (p[..., 0] * k / (w / h) + 1) * w / 2
    sy = (p[..., 1] * k + 1) * h / 2
    tx0, tx1 = (np.clip([sx.min(1), sx.max(1)], 0, w - 1) // 16).astype(int)
    ty0, ty1 = (np.clip([sy.min(1), sy.max(1)], 0, h - 1) // 16).astype(int)
    slots = 512 if w * h <= 4 * n else 2048
    over = 0
    for s0 in range(0, n - 1023, 1024):
        tiles = {(y, x) for f in range(s0, s0 + 1024)
                 for y in range(ty0[f], ty1[f] + 1) for x in range(tx0[f], tx1[f] + 1)}
        over += len(tiles) > slots
    assert over > n // 1024 // 2
    check_scene(ctx, oracle, scene)


def test_empty_meshes(ctx, oracle):
    """Draws of a mesh with no vertices, and of one with vertices but no
    faces, right after draws that left tile counts and bins in both draw
    slots: the target keeps its contents, or takes a pending clear."""
    from hana_amd.gpu import build_uniforms
    W, H = 150, 110
    scenes = [sg.small_scene(k, seed=400 + i, width=W, height=H) for i, k in enumerate((sg.BLINN, sg.TEXTURE))]
    tgt = ctx.target(W, H)
    ref = HostTarget(W, H)
    keep = []

    def draw(mesh, sc):
        u = build_uniforms(sc)
        keep.append(u)
        ctx.draw(ctx.mesh(mesh), sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
        oracle.draw(mesh, sc.shader, oracle.build_uniforms(sc), ref, sc.diffuse, sc.normal)

    for sc in scenes:
        draw(sc.mesh, sc)
    draw(empty_mesh(), scenes[0])
    draw(empty_mesh(), scenes[1])
    draw(faceless(scenes[0].mesh), scenes[0])
    col, dep = ctx.read(tgt, W, H)
    assert (dep < 1e30).sum() > 1000
    assert_depth_exact(dep, ref.depth, "empty draws")
    assert_color_ok(col, ref.color, "empty draws")
    ctx.clear(tgt, (10, 20, 30, 255), 0.5)
    draw(empty_mesh(), scenes[1])
    col, dep = ctx.read(tgt, W, H)
    assert (col == np.array([10, 20, 30, 255], np.uint8)).all() and (dep == np.float32(0.5)).all()


def test_widest_target(ctx, oracle):
    """A target 2^20 pixels wide (65,536 tiles in a row, the most the 16-bit
    tile coordinates hold) renders like the oracle; one pixel wider is
    refused."""
    from hana_amd.gpu import HANA_E_INVALID, HanaError, build_uniforms
    with pytest.raises(HanaError) as err:
        ctx.target((1 << 20) + 1, 4)
    assert err.value.code == HANA_E_INVALID
    W, H = 1 << 20, 4
    pos = np.array([[-0.99, -0.9, 0.2], [1.0, -0.95, 0.1], [1.0, 0.99, 0.3], [-0.98, 1.0, 0.4],
                    [0.999, -1.0, -0.5], [1.0, -1.0, -0.5], [1.0, 1.0, -0.5]], np.float32)
    idx = np.array([[0, 1, 2], [0, 2, 3], [4, 5, 6]], np.int32)
    rng = np.random.default_rng(8)
    mesh = sg.Mesh(pos, np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (7, 1)),
                   rng.uniform(0, 1, (7, 2)).astype(np.float32), idx)
    base = sg.small_scene(sg.TEXTURE, width=64, height=64, seed=8)
    sc = sg.Scene(mesh, W, H, sg.TEXTURE, False, (0.0, 0.0, 2.0), base.light_pos, base.diffuse, base.normal,
                  name="widest")
    tgt = ctx.target(W, H, winner=True)
    u = with_camera_vp(build_uniforms(sc), IDENTITY_VP)
    ctx.draw(ctx.mesh(mesh), sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
    col, dep = ctx.read(tgt, W, H)
    win = ctx.read_winner(tgt, W, H)
    ref = HostTarget(W, H)
    _, owin = oracle.draw(mesh, sc.shader, with_camera_vp(oracle.build_uniforms(sc), IDENTITY_VP), ref, sc.diffuse,
                          sc.normal, winner=True)
    assert (owin[:, -16:] >= 0).any()                       # the last tile column has fragments
    assert np.array_equal(win.astype(np.int64), as_u32_ids(owin))
    assert_depth_exact(dep, ref.depth, "widest")
    assert_color_ok(col, ref.color, "widest")


def test_tiny_screen_inputs(ctx, oracle):
    """Coverage of set-ups with tiny inputs: corners within 2^-8 px of the
    target's edge and edges a few ulps from vertical or horizontal (their
    triangles take the raster's division path), and corners exactly 2^-8 px
    from the edge (the shared-reciprocal path at its boundary).  Winners,
    depth and colour equal the oracle's."""
    from hana_amd.gpu import build_uniforms
    sc = tiny_input_scene()
    a = np.abs(screen_inputs(sc))
    assert ((a > 0) & (a < 2.0 ** -8)).any(1).sum() > 1000 and (a == 2.0 ** -8).any(1).sum() > 100
    W, H = sc.width, sc.height
    tgt = ctx.target(W, H, winner=True)
    u = with_camera_vp(build_uniforms(sc), IDENTITY_VP)
    ctx.draw(ctx.mesh(sc.mesh), sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
    col, dep = ctx.read(tgt, W, H)
    win = ctx.read_winner(tgt, W, H)
    ref = HostTarget(W, H)
    st, owin = oracle.draw(sc.mesh, sc.shader, with_camera_vp(oracle.build_uniforms(sc), IDENTITY_VP), ref,
                           sc.diffuse, sc.normal, winner=True)
    assert st["fragments"] > W * H
    assert np.array_equal(win.astype(np.int64), as_u32_ids(owin)), "winners"
    assert_depth_exact(dep, ref.depth, "tiny inputs")
    assert_color_ok(col, ref.color, "tiny inputs")


def test_pass_statistics(ctx, oracle):
    """hana_gpu_last_stats of a shadow pass and a main pass with clipping:
    faces and faces failing the trivial accept equal the oracle's counts,
    pixels written equal the pixels the oracle's fragments reached, and every
    winning primitive has a bin entry."""
    from hana_amd.gpu import SHADOW, build_uniforms
    sc = sg.small_scene(sg.BLINN, shadow=True, seed=7, cam=(0.3, 0.2, 1.2))
    tgt_o, owin, osm, st, st_s = oracle_passes(oracle, sc)
    u = build_uniforms(sc)
    mesh, dif, nrm = ctx.mesh(sc.mesh), ctx.texture(sc.diffuse), ctx.texture(sc.normal)
    sw, sh = sc.shadow_wh
    sm, tgt = ctx.target(sw, sh), ctx.target(sc.width, sc.height)
    ctx.set_stats(True)
    try:
        ctx.draw(mesh, SHADOW, u, sm, dif, nrm)
        (gs,) = ctx.last_stats()
        ctx.draw(mesh, sc.shader, u, tgt, dif, nrm, sm)
        (gm,) = ctx.last_stats()
    finally:
        ctx.set_stats(False)
    for g, o, win in ((gs, st_s, osm.winner), (gm, st, owin)):
        assert g["faces"] == o["faces"] and g["clipped_faces"] == o["clipped_faces"], (g, o)
        assert g["polygons"] <= g["clipped_faces"]
        assert g["pixels_written"] == int((win >= 0).sum())
        assert g["tile_entries"] >= len(np.unique(win[win >= 0]))
    assert st["clipped_faces"] > 0 and gm["polygons"] > 0


def test_nan_depth_fragments(ctx, oracle):
    """Fragments of NaN depth (a vertex whose clip z and w overflow to inf):
    they replace any stored depth, every later fragment at their pixels
    passes, and a second draw into the same target starts from the stored
    NaNs.  Depth (its NaN bit pattern included), winners and colour equal
    the oracle's."""
    from hana_amd.gpu import build_uniforms
    W, H = 96, 72
    tgt = ctx.target(W, H, winner=True)
    ref = HostTarget(W, H)
    keep = []
    for shift in (0.0, 0.3):
        sc = nan_depth_scene(W, H, shift=shift)
        u = with_camera_vp(build_uniforms(sc))
        keep.append(u)
        ctx.draw(ctx.mesh(sc.mesh), sc.shader, u, tgt, ctx.texture(sc.diffuse), ctx.texture(sc.normal))
        _, owin = oracle.draw(sc.mesh, sc.shader, with_camera_vp(oracle.build_uniforms(sc)), ref, sc.diffuse,
                              sc.normal, winner=True)
    col, dep = ctx.read(tgt, W, H)
    win = ctx.read_winner(tgt, W, H)
    assert np.isnan(ref.depth).sum() > 300
    assert_depth_exact(dep, ref.depth, "nan depth")
    assert np.array_equal(win.astype(np.int64), as_u32_ids(owin)), "winners"
    assert_color_ok(col, ref.color, "nan depth")


@pytest.mark.slow
def test_config2_full(ctx, oracle):
    scene = sg.config2_scene()
    rep, st = check_scene(ctx, oracle, scene)
    assert st["fragments"] > 500_000


def test_config5_reduced(ctx, oracle):
    """Micro-triangles (1-4 px) in layer pairs with identical positions: every
    covered pixel of the front pair is an exact depth tie, resolved by
    submission order (LEQUAL, the later face wins), in both passes."""
    scene = sg.config5_scene(layers=8, grid=180, width=640, height=360, shadow_size=512, tex_size=256)
    rep, st = check_scene(ctx, oracle, scene)
    assert st["fragments"] > scene.width * scene.height      # ties pass again


@pytest.mark.slow
def test_config5_full(ctx, oracle):
    """The full micro-triangle stress frame: 20 M triangles, 3840x2160, 4096^2 shadow map."""
    scene = sg.config5_scene()
    rep, st = check_scene(ctx, oracle, scene)
    assert st["faces"] == 19_998_784


def straddling_scene(n=40_000):
    """n copies of one triangle that crosses the left frustum plane: every
    face is clipped into a polygon (two fan triangles each), far more than
    the polygon pool and the fan-triangle list hold at first, and all of
    them tie in depth (the last copy wins every pixel)."""
    tri = np.array([[-3.0, -0.4, 0.0], [0.5, -0.3, 0.0], [0.2, 0.5, 0.0]], np.float32)
    uv = np.array([[0.1, 0.2], [0.8, 0.3], [0.5, 0.9]], np.float32)
    mesh = sg.Mesh(np.tile(tri, (n, 1)), np.tile(np.array([[0, 0, 1]], np.float32), (3 * n, 1)),
                   np.tile(uv, (n, 1)), np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    d, nm = sg.texture_pair(32, 32, 5)
    return sg.Scene(mesh, 48, 32, sg.BLINN, False, (0.0, 0.0, 2.0), (2.0, 2.0, 2.0), d, nm, name="straddling")


def test_clip_pool_overflow_regrows(oracle):
    """The first render overflows the clip pool: the read grows the workspace
    and renders the draw again by itself, and returns the exact image; a
    second render needs no replay.  A context of its own, so that no earlier
    test has grown the pool."""
    from hana_amd.gpu import Context
    with Context(0) as ctx:
        check_regrowth(ctx, oracle)


def check_regrowth(ctx, oracle):
    from hana_amd.gpu import build_uniforms
    scene = straddling_scene()
    u = build_uniforms(scene)
    mesh, dif, nrm = ctx.mesh(scene.mesh), ctx.texture(scene.diffuse), ctx.texture(scene.normal)
    tgt = ctx.target(scene.width, scene.height, winner=True)
    ref, owin, _, st, _ = oracle_passes(oracle, scene)
    assert st["clipped_faces"] == 40_000
    r0 = ctx.replay_count()
    for attempt in range(2):
        ctx.clear(tgt)
        ctx.draw(mesh, scene.shader, u, tgt, dif, nrm)
        col, dep = ctx.read(tgt, scene.width, scene.height)
        win = ctx.read_winner(tgt, scene.width, scene.height)
        assert ctx.replay_count() == r0 + 1, attempt   # the first draw overflowed, the second fits
        assert np.array_equal(win.astype(np.int64), as_u32_ids(owin)), attempt
        assert_depth_exact(dep, ref.depth, "straddling")
        assert_color_ok(col, ref.color, "straddling")


def test_overflow_replays_calls_in_order(oracle):
    """Calls since the last wait are rendered again in call order when one of
    them outgrew the workspace: an overflowing draw into a target, a draw on
    top of it, a frame into a second target whose shadow map a later draw
    samples after its post-frame clear, then the reads.  Every image equals
    the oracle's for the same call sequence."""
    from hana_amd.gpu import SHADOW, Context, build_uniforms
    big = straddling_scene()
    top = sg.small_scene(sg.TEXTURE, seed=21, width=big.width, height=big.height)
    fr = sg.small_scene(sg.BLINN, shadow=True, seed=22, width=80, height=60)
    with Context(0) as ctx:
        ub, ut, uf = build_uniforms(big), build_uniforms(top), build_uniforms(fr)
        t1 = ctx.target(big.width, big.height)
        t2 = ctx.target(fr.width, fr.height)
        sm = ctx.target(fr.width, fr.height)
        r0 = ctx.replay_count()
        ctx.draw(ctx.mesh(big.mesh), big.shader, ub, t1, ctx.texture(big.diffuse), ctx.texture(big.normal))
        ctx.draw(ctx.mesh(top.mesh), top.shader, ut, t1, ctx.texture(top.diffuse), ctx.texture(top.normal))
        fm, fd, fn = ctx.mesh(fr.mesh), ctx.texture(fr.diffuse), ctx.texture(fr.normal)
        ctx.render_frame(fm, fr.shader, True, uf, t2, sm, fd, fn)
        ctx.draw(fm, sg.BLINN, uf, t2, fd, fn, sm)   # samples the cleared shadow map
        c1, d1 = ctx.read(t1, big.width, big.height)
        c2, d2 = ctx.read(t2, fr.width, fr.height)
        assert ctx.replay_count() == r0 + 1
    o1 = HostTarget(big.width, big.height)
    oracle.draw(big.mesh, big.shader, oracle.build_uniforms(big), o1, big.diffuse, big.normal)
    oracle.draw(top.mesh, top.shader, oracle.build_uniforms(top), o1, top.diffuse, top.normal)
    o2, osm = HostTarget(fr.width, fr.height), HostTarget(fr.width, fr.height)
    ou = oracle.build_uniforms(fr)
    oracle.draw(fr.mesh, SHADOW, ou, osm, fr.diffuse, fr.normal)
    oracle.draw(fr.mesh, fr.shader, ou, o2, fr.diffuse, fr.normal, osm)
    osm.clear((0.0, 0.0, 0.0, 1.0), float(np.finfo(np.float32).max))   # DrawModel::draw's post-frame clear
    oracle.draw(fr.mesh, sg.BLINN, ou, o2, fr.diffuse, fr.normal, osm)
    assert_depth_exact(d1, o1.depth, "overflowed target")
    assert_color_ok(c1, o1.color, "overflowed target")
    assert_depth_exact(d2, o2.depth, "second target")
    assert_color_ok(c2, o2.color, "second target")


def test_bin_overflow_regrows(oracle):
    """800 k slivers of six (triangle, tile) pairs each overflow the bins the
    first draw allocates (four pairs per face): the read grows the bins and
    renders the draw again by itself, exactly; later draws on every draw slot
    need no replay.  A context of its own, so that no earlier test has grown
    the bins."""
    from hana_amd.gpu import Context, build_uniforms
    scene = sliver_scene()
    W, H = scene.width, scene.height
    ref = HostTarget(W, H)
    st, owin = oracle.draw(scene.mesh, scene.shader, with_camera_vp(oracle.build_uniforms(scene), IDENTITY_VP),
                           ref, scene.diffuse, scene.normal, winner=True)
    assert st["fragments"] > 1_000_000
    with Context(0) as ctx:
        u = with_camera_vp(build_uniforms(scene), IDENTITY_VP)
        mesh, dif, nrm = ctx.mesh(scene.mesh), ctx.texture(scene.diffuse), ctx.texture(scene.normal)
        tgt = ctx.target(W, H, winner=True)
        r0 = ctx.replay_count()
        for attempt in range(3):                             # every draw slot has grown
            ctx.clear(tgt)
            ctx.draw(mesh, scene.shader, u, tgt, dif, nrm)
            col, dep = ctx.read(tgt, W, H)
            win = ctx.read_winner(tgt, W, H)
            assert ctx.replay_count() == r0 + 1, attempt
            assert np.array_equal(win.astype(np.int64), as_u32_ids(owin)), attempt
            assert_depth_exact(dep, ref.depth, "slivers")
            assert_color_ok(col, ref.color, "slivers")


def test_pair_limit_refused(oracle):
    """A draw whose triangles cover more (triangle, tile) pairs than 32-bit
    bin positions reach (32,768 bands of 196,608 tiles each on a 2^20 x 2^10
    target: 6.4 G pairs) is refused with HANA_E_NOMEM at the next sync and
    bins nothing; the context keeps rendering.  A context of its own, which
    frees the 8 GB target at the end."""
    from hana_amd.gpu import HANA_E_NOMEM, Context, HanaError, build_uniforms
    W, H, n = 1 << 20, 1 << 10, 32768
    rng = np.random.default_rng(12)
    ty = rng.integers(0, H // 16 - 3, n)
    y0, y1 = (ty * 16 + 8) / H * 2 - 1, (ty * 16 + 40) / H * 2 - 1    # 3 tile rows, the full width
    z = rng.uniform(-0.5, 0.5, n)
    tri = np.stack([np.stack([-np.ones(n), y0, z], 1), np.stack([np.ones(n), y0, z], 1),
                    np.stack([np.ones(n), y1, z], 1)], 1)
    mesh = sg.Mesh(tri.reshape(-1, 3).astype(np.float32), np.tile(np.array([[0.0, 0.0, 1.0]], np.float32), (3 * n, 1)),
                   rng.uniform(0, 1, (3 * n, 2)).astype(np.float32), np.arange(3 * n, dtype=np.int32).reshape(n, 3))
    base = sg.small_scene(sg.TEXTURE, width=64, height=64, seed=12)
    bands = sg.Scene(mesh, W, H, sg.TEXTURE, False, (0.0, 0.0, 2.0), base.light_pos, base.diffuse, base.normal)
    with Context(0) as ctx:
        big = ctx.target(W, H)
        u = with_camera_vp(build_uniforms(bands), IDENTITY_VP)
        ctx.draw(ctx.mesh(mesh), bands.shader, u, big, ctx.texture(bands.diffuse), ctx.texture(bands.normal))
        with pytest.raises(HanaError) as err:
            ctx.sync()
        assert err.value.code == HANA_E_NOMEM and "(triangle, tile) pairs" in str(err.value)
        check_scene(ctx, oracle, sg.small_scene(sg.BLINN, shadow=True, seed=6, width=96, height=72))


def test_clip_pool_overflow_in_a_batch(ctx, oracle):
    """Batch finalisation sizes the pool and the lists with its trial run."""
    from hana_amd.gpu import Batch, SceneOnDevice
    scene = straddling_scene()
    dev = SceneOnDevice(ctx, scene)
    batch = Batch(ctx)
    batch.add(dev.desc())
    batch.finalize(use_graph=True)
    batch.run()
    col, dep = dev.read()
    tgt, _, _ = oracle.render_frame(scene)
    assert_depth_exact(dep, tgt.depth, "straddling batch")
    assert_color_ok(col, tgt.color, "straddling batch")
    batch.close()


# Tunings of the per-call path: 32x32 tiles, the persistent raster (both pass
# groups in one launch, main passes waiting for the shadow tiles), and both.
TUNINGS = [(32, 0, 4), (16, 4, 4), (32, 8, 2), (16, 1, 4), (16, 0, 2), (32, 4, 2)]


def frame_scenes():
    return [sg.small_scene(sg.NORMALMAP, shadow=True, seed=1),
            sg.small_scene(sg.BLINN, shadow=True, seed=2, width=65, height=17),
            sg.small_scene(sg.NORMALMAP, shadow=True, seed=7, cam=(0.9, -0.4, 0.6), width=128, height=96),
            sg.small_scene(sg.TEXTURE_LIGHT, shadow=False, seed=8, width=200, height=37),
            sg.config5_scene(layers=8, grid=180, width=640, height=360, shadow_size=512, tex_size=256)]


@pytest.mark.parametrize("tile,wgs,waves", TUNINGS)
def test_tunings_match_oracle(oracle, tile, wgs, waves):
    """Every tuning renders the oracle's bits: separate draws (depth, winners,
    shadow maps) and DrawModel::draw frames, whose main pass waits inside the
    persistent raster for its own shadow pass; 32x32 tiles, 2-wave raster
    workgroups."""
    from hana_amd.gpu import Context, SceneOnDevice
    with Context(0) as ctx:
        ctx.set_tuning(tile, wgs, raster_waves=waves)
        for scene in frame_scenes():
            check_scene(ctx, oracle, scene)
            dev = SceneOnDevice(ctx, scene)
            dev.render()
            col, dep = dev.read()
            tgt, _, _ = oracle.render_frame(scene)
            assert_depth_exact(dep, tgt.depth, f"{scene.name} frame, tiles {tile}, wgs {wgs}")
            assert_color_ok(col, tgt.color, f"{scene.name} frame, tiles {tile}, wgs {wgs}")


def test_tunings_nan_depth(oracle):
    """NaN-depth fragments resolve the same way with 32x32 tiles in the persistent raster."""
    from hana_amd.gpu import Context
    with Context(0) as ctx:
        ctx.set_tuning(32, 4, raster_waves=2)
        check_scene(ctx, oracle, nan_depth_scene())


def test_tuning_rejects_bad_values():
    from hana_amd.gpu import Context, HanaError
    with Context(0) as ctx:
        for bad in [(8, 0, 4), (16, -1, 4), (16, 17, 4), (16, 0, 0), (16, 0, 5), (16, 0, 4, 3), (16, 0, 4, 8)]:
            with pytest.raises(HanaError):
                ctx.set_tuning(*bad)


def fp32(scene):
    scene.shadow_fp32 = True
    scene.name += " fp32 shadow"
    return scene


@pytest.mark.parametrize("tile,wgs", [(16, 0), (32, 4)])
def test_fp32_shadow_extension(oracle, tile, wgs):
    """The fp32 shadow-map extension (depth-only shadow pass, the main pass
    comparing against the map's fp32 depth) against the oracle extended the
    same way: shadow-map depth and winners bit-exact, its colour untouched,
    the main pass at the usual bar; per-call draws and whole frames."""
    from hana_amd.gpu import Context, SceneOnDevice
    scenes = [fp32(sg.small_scene(sg.NORMALMAP, shadow=True, seed=1)),
              fp32(sg.small_scene(sg.BLINN, shadow=True, seed=2, width=97, height=61)),
              fp32(sg.config5_scene(layers=8, grid=180, width=640, height=360, shadow_size=512, tex_size=256))]
    with Context(0) as ctx:
        ctx.set_tuning(tile, wgs)
        for scene in scenes:
            check_scene(ctx, oracle, scene)
            dev = SceneOnDevice(ctx, scene)
            dev.render()
            col, dep = dev.read()
            tgt, _, _ = oracle.render_frame(scene)
            assert_depth_exact(dep, tgt.depth, scene.name + " frame")
            assert_color_ok(col, tgt.color, scene.name + " frame")
