"""Frame sequences driven by the interactive scene state, as the reference's
main loop drives SingleModelScene::tick: camera motions and key presses
between frames (orbit, pan, dolly, light rotation, shader and shadow
switches, camera reset), each frame rendered with hana_render_frame from
hana_scene_uniforms on one target and one frame-sized shadow map, cleared
between frames (SR/main.cpp:152-153).  Every frame equals the oracle's
render of the same state."""
import numpy as np
import pytest

from hana_amd import scenegen as sg
from hana_amd.gpu import KEY_A, KEY_D, KEY_E, KEY_Q, KEY_SPACE, InteractiveScene
from parity import assert_color_ok, assert_depth_exact
from pyoracle import CPURenderer

pytestmark = pytest.mark.gpu

W, H = 220, 160
STEPS = [dict(), dict(orbit=(0.05, 0.02)), dict(key=KEY_A), dict(pan=(0.04, -0.02)), dict(key=KEY_Q),
         dict(dolly=1.5), dict(key=KEY_D), dict(key=KEY_E), dict(key=KEY_Q), dict(key=KEY_Q),
         dict(orbit=(-0.12, 0.05), dolly=-1.0), dict(key=KEY_Q), dict(key=KEY_Q), dict(key=KEY_Q),
         dict(key=KEY_SPACE)]


def test_interactive_sequence():
    from hana_amd.gpu import Context
    ctx = Context(0)
    oracle = CPURenderer("oracle")
    base = sg.small_scene(sg.NORMALMAP, shadow=True, seed=91, width=W, height=H)
    mesh = ctx.mesh(base.mesh)
    dif, nrm = ctx.texture(base.diffuse), ctx.texture(base.normal)
    tgt = ctx.target(W, H)
    smap = ctx.target(W, H)
    state = InteractiveScene(W, H)
    seen = set()
    for i, step in enumerate(STEPS):
        if "key" in step:
            state.key(step["key"])
        else:
            state.move(**step)
        seen.add((state.shader, state.shadow))
        ctx.clear(tgt)
        ctx.render_frame(mesh, state.shader, state.shadow, state.uniforms(), tgt, smap if state.shadow else None,
                         dif, nrm)
        col, dep = ctx.read(tgt, W, H)
        ref, _, _ = oracle.render_frame(state.frame(base.mesh, base.diffuse, base.normal))
        assert_depth_exact(dep, ref.depth, f"step {i}")
        assert_color_ok(col, ref.color, f"step {i}")
    assert len({k for k, _ in seen}) == 6        # every shader of the scene took a frame
    ctx.close()
