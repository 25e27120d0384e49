"""The C++ drop-in: the reference's own host code (Model read by its OBJ
parser, its IShader subclasses bound to a ShaderData, its RenderBuffer) with
graphics_draw_triangle taken from hana-softwarerenderer_amd/dropin/
graphics_hana.cpp, i.e. rendered by libhana_gpu.so.

The library under test, oracle/_ref/libhanaref_dropin.so, is compiled from the
reference's sources (oracle/Makefile, `make dropin`) where they are present;
the tests are skipped where it was not built."""
import os

import numpy as np
import pytest

from hana_amd import scenegen as sg
from parity import assert_color_ok, assert_depth_exact
from pyoracle import DROPIN_SO, REF_SO, CPURenderer, HostTarget, write_obj

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(DROPIN_SO),
                                 reason="oracle/_ref/libhanaref_dropin.so not built (needs the reference sources)")]


@pytest.fixture(scope="module")
def dropin():
    return CPURenderer("dropin")


@pytest.fixture(scope="module")
def oracle():
    return CPURenderer("oracle")


def obj_of(scene, tmp_path):
    path = tmp_path / f"{scene.name.replace('/', '_')}.obj"
    write_obj(scene.mesh, path)
    return str(path)


@pytest.mark.parametrize("kind", [sg.GROUND, sg.TOON, sg.TEXTURE, sg.TEXTURE_LIGHT, sg.BLINN, sg.NORMALMAP,
                                  sg.SHADOW])
def test_each_shader_with_clipping(dropin, oracle, kind, tmp_path):
    """A close camera clips part of the mesh at the near plane and the sides."""
    sc = sg.small_scene(kind, shadow=False, seed=40 + kind, width=96, height=72, cam=(0.2, 0.1, 1.2))
    u = oracle.build_uniforms(sc)
    got = HostTarget(sc.width, sc.height)
    nf = dropin.draw_model(obj_of(sc, tmp_path), kind, u, got, sc.diffuse, sc.normal)
    assert nf == sc.mesh.nfaces
    ref = HostTarget(sc.width, sc.height)
    st, _ = oracle.draw(sc.mesh, kind, u, ref, sc.diffuse, sc.normal)
    assert st["clipped_faces"] > 0
    assert_depth_exact(got.depth, ref.depth, f"kind {kind}")
    assert_color_ok(got.color, ref.color, f"kind {kind}")


def test_shadow_then_main_pass(dropin, oracle, tmp_path):
    """DrawModel::draw's two calls: the shadow pass into a host RenderBuffer,
    then the main pass sampling it."""
    sc = sg.small_scene(sg.NORMALMAP, shadow=True, seed=47, width=120, height=90)
    u = oracle.build_uniforms(sc)
    path = obj_of(sc, tmp_path)
    sw, sh = sc.shadow_wh
    sm, tgt = HostTarget(sw, sh), HostTarget(sc.width, sc.height)
    dropin.draw_model(path, sg.SHADOW, u, sm, sc.diffuse, sc.normal)
    dropin.draw_model(path, sc.shader, u, tgt, sc.diffuse, sc.normal, sm)
    ref, ref_sm, _, _ = oracle.render_passes(sc, u)
    assert_depth_exact(sm.depth, ref_sm.depth, "shadow map")
    assert np.array_equal(sm.color[..., :3], ref_sm.color[..., :3])
    assert_depth_exact(tgt.depth, ref.depth, "main pass")
    assert_color_ok(tgt.color, ref.color, "main pass")


def test_matches_unmodified_reference(dropin, tmp_path):
    """Same OBJ, same shader, same RenderBuffer contents: the reference's own
    graphics_draw_triangle on the CPU and the drop-in agree (no clipped
    triangles, where the unmodified CPU function is defined)."""
    if not os.path.exists(REF_SO):
        pytest.skip("oracle/_ref/libhanaref.so not built")
    ref = CPURenderer("reference")
    sc = sg.small_scene(sg.BLINN, shadow=False, seed=5, width=80, height=60, cam=(0.0, 0.3, 3.0))
    u = ref.build_uniforms(sc)
    path = obj_of(sc, tmp_path)
    a, b = HostTarget(sc.width, sc.height), HostTarget(sc.width, sc.height)
    ref.draw_model(path, sg.BLINN, u, a, sc.diffuse, sc.normal)
    dropin.draw_model(path, sg.BLINN, u, b, sc.diffuse, sc.normal)
    assert_depth_exact(b.depth, a.depth, "drop-in vs reference")
    assert_color_ok(b.color, a.color, "drop-in vs reference")


def test_keeps_alpha_and_accumulates(dropin, oracle, tmp_path):
    """set_color writes RGB only, and a second call draws over the first
    call's colour and depth (the RenderBuffer is the state, as on the CPU)."""
    sc1 = sg.small_scene(sg.BLINN, seed=50, width=64, height=48, cam=(0.0, 0.0, 2.5))
    sc2 = sg.small_scene(sg.TOON, seed=51, width=64, height=48, cam=(0.3, 0.2, 2.2))
    got = HostTarget(sc1.width, sc1.height, rgba=(0.2, 0.4, 0.6, 1.0), depth=1.5)
    ref = HostTarget(sc1.width, sc1.height, rgba=(0.2, 0.4, 0.6, 1.0), depth=1.5)
    got.color[..., 3] = ref.color[..., 3] = 7
    for sc in (sc1, sc2):
        u = oracle.build_uniforms(sc)
        dropin.draw_model(obj_of(sc, tmp_path), sc.shader, u, got, sc.diffuse, sc.normal)
        oracle.draw(sc.mesh, sc.shader, u, ref, sc.diffuse, sc.normal)
    assert (got.color[..., 3] == 7).all()
    assert_depth_exact(got.depth, ref.depth, "two calls")
    assert_color_ok(got.color, ref.color, "two calls")
