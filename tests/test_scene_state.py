"""The interactive scene state (hana_scene_*): the orbit camera of
Camera::update_transform (SR/camera.cpp:31-70), the key handlers of Scene and
SingleModelScene (SR/scene.cpp:45-63,127-147) and the camera reset
(SR/camera.cpp:167-169).

Pinned three ways:
  * closed forms and invariants, which hold without the reference: the
    initial state, the target fixed without panning, dolly scaling the distance
    by 0.95^d, a full orbit turn returning, the polar clamp, the pan length
    distance*tan(30 deg)*2*|pan|*sin(view, up) across the view, the light's
    rotation about (0,1,1) preserving its axial component and length, the
    shader cycle and its shadow rule;
  * bit for bit against the reference's own Camera and SingleModelScene,
    compiled into oracle/_ref (skipped where that library or the reference
    assets are absent);
  * hana_scene_uniforms equal to hana_build_uniforms on the state's camera,
    light and shadow switch."""
import math
import os

import numpy as np
import pytest

from hana_amd import gpu
from hana_amd.gpu import KEY_A, KEY_D, KEY_E, KEY_Q, KEY_SPACE, InteractiveScene
from pyoracle import REF_SO, CPURenderer

ASSET = "/root/reference/Hana-SoftwareRenderer/assets/african_head/african_head.obj"
need_ref = pytest.mark.skipif(not os.path.exists(REF_SO), reason="oracle/_ref not built")


def f32(x):
    return np.asarray(x, np.float32)


def test_initial_state():
    s = InteractiveScene(800, 600)
    assert s.camera_position == (0.0, 0.0, 2.0) and s.camera_target == (0.0, 0.0, 0.0)
    assert s.light_position == (2.0, 2.0, 2.0)
    assert s.shadow and s.shader == gpu_kind("NORMALMAP")
    assert np.float32(s.s.aspect) == np.float32(800) / np.float32(600)


def gpu_kind(name):
    from hana_amd import scenegen as sg
    return getattr(sg, name)


def test_dolly_scales_distance_and_keeps_target():
    s = InteractiveScene(640, 480)
    s.move(orbit=(0.1, -0.05))
    r0 = np.linalg.norm(f32(s.camera_position) - f32(s.camera_target))
    t0 = s.camera_target
    for d in (1.0, -2.5, 3.0):
        s.move(dolly=d)
        r1 = np.linalg.norm(f32(s.camera_position) - f32(s.camera_target))
        assert s.camera_target == t0                       # no pan: target untouched, exactly
        assert math.isclose(r1, r0 * 0.95 ** d, rel_tol=2e-6)
        r0 = r1


def test_full_orbit_turn_returns_and_pole_clamps():
    s = InteractiveScene(640, 480)
    s.move(orbit=(0.2, 0.03))
    p0 = f32(s.camera_position)
    s.move(orbit=(1.0, 0.0))                              # theta -= 2 pi
    assert np.allclose(s.camera_position, p0, atol=2e-5)
    s.move(orbit=(0.0, 5.0))                              # far past the pole: phi clamps at EPSILON
    p = f32(s.camera_position) - f32(s.camera_target)
    r = np.linalg.norm(p)
    h = math.hypot(p[0], p[2])                            # r * sin(1e-5): above the pole, not past it
    assert p[1] > 0 and 0 < h < 2e-5 * r


def test_pan_moves_target_across_the_view():
    s = InteractiveScene(640, 480)
    s.move(orbit=(0.07, 0.02))
    pos0, tgt0 = f32(s.camera_position), f32(s.camera_target)
    dist = np.linalg.norm(tgt0 - pos0)
    fwd = (tgt0 - pos0) / dist
    s.move(pan=(0.1, -0.04))
    d = f32(s.camera_target) - tgt0
    # left = cross(UP, forward) is not normalised (SR/camera.cpp:34): its
    # length, and up's, is the sine of the angle between UP and the view
    expect = dist * math.tan(math.radians(30)) * 2 * math.hypot(0.1, -0.04) * math.sqrt(1 - fwd[1] ** 2)
    assert math.isclose(np.linalg.norm(d), expect, rel_tol=1e-5)
    assert abs(float(np.dot(d, fwd))) < 1e-6               # perpendicular to the view
    # the camera keeps its offset from the (moved) target
    assert np.allclose(f32(s.camera_position) - f32(s.camera_target), pos0 - tgt0, atol=2e-5)


def test_keys():
    s = InteractiveScene(320, 240)
    l0 = f32(s.light_position)
    axis = f32((0, 1, 1)) / math.sqrt(2)
    s.key(KEY_A)
    l1 = f32(s.light_position)
    assert not np.allclose(l1, l0)
    assert math.isclose(float(np.dot(l1, axis)), float(np.dot(l0, axis)), rel_tol=1e-6)
    assert math.isclose(np.linalg.norm(l1), np.linalg.norm(l0), rel_tol=1e-6)
    cosang = np.dot(l0 - np.dot(l0, axis) * axis, l1 - np.dot(l1, axis) * axis) / \
        np.linalg.norm(l0 - np.dot(l0, axis) * axis) ** 2
    assert math.isclose(cosang, math.cos(math.radians(5)), rel_tol=1e-5)
    s.key(KEY_D)
    assert np.allclose(s.light_position, l0, atol=1e-6)
    # Q cycles NormalMap, Blinn, Texture+Light, Texture, Toon, Ground; shadow for the first three
    kinds = []
    for _ in range(6):
        s.key(KEY_Q)
        kinds.append((s.shader, s.shadow))
    names = ["BLINN", "TEXTURE_LIGHT", "TEXTURE", "TOON", "GROUND", "NORMALMAP"]
    assert kinds == [(gpu_kind(n), n in ("BLINN", "TEXTURE_LIGHT", "NORMALMAP")) for n in names]
    s.key(KEY_E)
    assert not s.shadow
    s.move(orbit=(0.3, 0.1), pan=(0.2, 0.1), dolly=2)
    s.key(KEY_SPACE)
    assert s.camera_position == (0.0, 0.0, 2.0) and s.camera_target == (0.0, 0.0, 0.0)
    with pytest.raises(gpu.HanaError):
        s.key(99)


def test_uniforms_follow_the_state():
    from hana_amd import scenegen as sg
    s = InteractiveScene(200, 150)
    s.move(orbit=(0.05, 0.02), pan=(0.03, 0.01), dolly=0.5)
    s.key(KEY_A)
    s.key(KEY_E)
    sc = sg.small_scene(sg.BLINN, width=200, height=150)
    fr = s.frame(sc.mesh, sc.diffuse, sc.normal)
    assert bytes(s.uniforms()) == bytes(gpu.build_uniforms(fr))
    assert fr.cam_target == s.camera_target and fr.shadow is False


@need_ref
@pytest.mark.parametrize("seed", range(4))
def test_camera_matches_reference(seed):
    ref = CPURenderer("reference")
    rng = np.random.default_rng(seed)
    s = InteractiveScene(int(rng.integers(64, 2000)), int(rng.integers(64, 2000)))
    for i in range(60):
        m = rng.normal(0, 0.1, 5).astype(np.float32)
        m[4] = np.float32(rng.normal(0, 2))
        if i % 7 == 0:
            m[2:4] = 0                                    # pure orbit/dolly steps too
        pos, tgt = ref.camera_update(s.camera_position, s.camera_target, s.s.aspect, m)
        s.move(orbit=(m[0], m[1]), pan=(m[2], m[3]), dolly=m[4])
        assert np.array_equal(f32(s.camera_position).view(np.uint32), pos.view(np.uint32)), (seed, i)
        assert np.array_equal(f32(s.camera_target).view(np.uint32), tgt.view(np.uint32)), (seed, i)


@need_ref
@pytest.mark.skipif(not os.path.exists(ASSET), reason="reference assets not present")
def test_keys_match_reference():
    ref = CPURenderer("reference")
    rng = np.random.default_rng(7)
    keys = [int(k) for k in rng.choice([KEY_A, KEY_D, KEY_E, KEY_Q], 40)]
    out = ref.scene_keys(ASSET, 64, 48, keys)
    s = InteractiveScene(64, 48)
    for i, k in enumerate(keys):
        s.key(k)
        assert np.array_equal(f32(s.light_position).view(np.uint32), out[i, :3].view(np.uint32)), i
        assert s.shadow == bool(out[i, 3]) and s.s.shader_index == int(out[i, 4]), i
