"""On-disk formats through the GPU path: a scene written to OBJ + TGA files
(raw and RLE, 3 and 4 bytes per texel), read back by the library's loaders
and rendered on the device, equals the oracle on the in-memory scene; and
hana_target_present equals the window blit (row flip, R/B swap).

The renderer's two shipped models (african_head/ and diablo3_pose/, OBJ + TGA
as it ships them) are rendered at 1000x600 and compared with the oracle too.
They are read from oracle/_ref/assets/, which the build stages from the
reference tree (HANA_ASSETS_DIR names another directory of the same layout);
the cases skip only where neither exists."""
import os

import numpy as np
import pytest

from hana_amd import scenegen as sg
from parity import assert_color_ok, assert_depth_exact
from pyoracle import CPURenderer, write_obj

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASSETS = os.environ.get("HANA_ASSETS_DIR", "") or os.path.join(REPO, "oracle", "_ref", "assets")
HAVE_ASSETS = all(os.path.exists(os.path.join(ASSETS, m, m + ".obj")) for m in ("african_head", "diablo3_pose"))


@pytest.fixture(scope="module")
def ctx():
    from hana_amd.gpu import Context
    c = Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle():
    return CPURenderer("oracle")


def render_gpu(ctx, scene):
    from hana_amd.gpu import SceneOnDevice
    dev = SceneOnDevice(ctx, scene)
    dev.render()
    return dev, dev.read()


@pytest.mark.parametrize("rle,bpp", [(True, 3), (False, 4), (True, 4)])
def test_files_roundtrip_render(ctx, oracle, tmp_path, rle, bpp):
    from hana_amd.gpu import load_obj, load_tga, save_tga
    sc = sg.small_scene(sg.NORMALMAP, shadow=True, seed=60 + bpp, width=200, height=150, tex=128)
    write_obj(sc.mesh, tmp_path / "m.obj")
    tex = [np.ascontiguousarray(t[..., :bpp]) for t in (sc.diffuse, sc.normal)]
    for name, t in zip(("d", "n"), tex):
        save_tga(tmp_path / f"{name}.tga", t, rle)
    loaded = sg.Scene(load_obj(tmp_path / "m.obj"), sc.width, sc.height, sc.shader, sc.shadow, sc.cam_pos,
                      sc.light_pos, load_tga(tmp_path / "d.tga"), load_tga(tmp_path / "n.tga"), name="files")
    assert all(np.array_equal(a, b) for a, b in zip(tex, (loaded.diffuse, loaded.normal)))
    _, (col, dep) = render_gpu(ctx, loaded)
    ref = sg.Scene(sc.mesh, sc.width, sc.height, sc.shader, sc.shadow, sc.cam_pos, sc.light_pos, tex[0], tex[1])
    tgt, _, _ = oracle.render_frame(ref)
    assert_depth_exact(dep, tgt.depth, "files")
    assert_color_ok(col, tgt.color, "files")


def test_present_is_the_window_blit(ctx):
    sc = sg.small_scene(sg.BLINN, seed=64, width=173, height=91)
    dev, (col, _) = render_gpu(ctx, sc)
    out = ctx.present(dev.target, sc.width, sc.height)
    expect = col[::-1][..., [2, 1, 0, 3]]
    assert np.array_equal(out, expect)


@pytest.mark.skipif(not HAVE_ASSETS, reason="no staged assets: build() stages them from the reference tree")
@pytest.mark.parametrize("name", ["african_head", "diablo3_pose"])
def test_real_assets(ctx, oracle, name):
    from hana_amd.gpu import load_obj, load_tga
    base = os.path.join(ASSETS, name, name)
    sc = sg.Scene(load_obj(base + ".obj"), 1000, 600, sg.NORMALMAP, True, (0.0, 0.0, 1.5), (2.0, 2.0, 2.0),
                  load_tga(base + "_diffuse.tga", True), load_tga(base + "_nm_tangent.tga", True), name=name)
    _, (col, dep) = render_gpu(ctx, sc)
    tgt, _, _ = oracle.render_frame(sc)
    assert_depth_exact(dep, tgt.depth, name)
    assert_color_ok(col, tgt.color, name)
