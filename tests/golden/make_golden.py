"""Writes the golden vectors in this directory from the reference's own code.

The reference is compiled from its sources by oracle/Makefile (`make ref`,
oracle/_ref/libhanaref.so); this script calls it through the same entry
points as the oracle and stores inputs and outputs as .npz.  Run it in a
container that has the reference tree:

    python tests/golden/make_golden.py

The fixtures are data only: seeded inputs and what the reference computed.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "oracle"), os.path.join(REPO, "hana-softwarerenderer_amd")]

sys.path.insert(0, os.path.dirname(HERE))

from edge_scenes import NAN_SCENE_VP, nan_depth_scene, with_camera_vp  # noqa: E402
from hana_amd import scenegen as sg  # noqa: E402
from pyoracle import CPURenderer, HostTarget  # noqa: E402

SEED = 0x5EED


def math_kats(ref, rng):
    out = {}
    a = rng.normal(size=(64, 4, 4)).astype(np.float32)
    b = rng.normal(size=(64, 4, 4)).astype(np.float32)
    v = rng.normal(size=(64, 4)).astype(np.float32)
    out["mm_a"], out["mm_b"], out["mv_v"] = a, b, v
    out["mm_out"] = np.stack([ref.mat4_mul(a[i], b[i]) for i in range(64)])
    out["mv_out"] = np.stack([ref.mat4_mul_vec4(a[i], v[i]) for i in range(64)])
    # invertible transforms: model matrices and random well-conditioned ones
    inv_in = np.concatenate([a[:32] + 4 * np.eye(4, dtype=np.float32)[None],
                             np.stack([ref.model_matrix(rng.uniform(-2, 2, 3), rng.uniform(-180, 180, 3),
                                                        rng.uniform(0.2, 3, 3)) for _ in range(32)])]).astype(np.float32)
    out["inv_in"] = inv_in
    out["inv_out"] = np.stack([ref.invert(m) for m in inv_in])
    eyes = rng.normal(size=(32, 3)).astype(np.float32) * 3
    tgts = rng.normal(size=(32, 3)).astype(np.float32) * 0.3
    out["look_eye"], out["look_tgt"] = eyes, tgts
    out["look_out"] = np.stack([ref.lookat(eyes[i], tgts[i], np.array([0, 1, 0], np.float32)) for i in range(32)])
    pp = np.stack([rng.uniform(0.2, 2.5, 32), rng.uniform(0.3, 3, 32), rng.uniform(0.01, 1, 32),
                   rng.uniform(2, 1e4, 32)], 1).astype(np.float32)
    out["persp_in"] = pp
    out["persp_out"] = np.stack([ref.perspective(*p) for p in pp])
    out["ortho_out"] = np.stack([ref.orthographic(p[1], 1.0, 0.0, p[3]) for p in pp])
    ang = rng.uniform(-7, 7, 32).astype(np.float32)
    out["rot_in"] = ang
    out["rotx_out"] = np.stack([ref.rotate_axis("x", t) for t in ang])
    out["roty_out"] = np.stack([ref.rotate_axis("y", t) for t in ang])
    out["rotz_out"] = np.stack([ref.rotate_axis("z", t) for t in ang])
    ax = rng.normal(size=(32, 3)).astype(np.float32)
    out["rot_axis"] = ax
    out["rot_out"] = np.stack([ref.rotate(ang[i], *ax[i]) for i in range(32)])
    mp = rng.uniform(-2, 2, (32, 3)).astype(np.float32)
    mr = rng.uniform(-180, 180, (32, 3)).astype(np.float32)
    ms = rng.uniform(0.2, 3, (32, 3)).astype(np.float32)
    out["model_p"], out["model_r"], out["model_s"] = mp, mr, ms
    out["model_out"] = np.stack([ref.model_matrix(mp[i], mr[i], ms[i]) for i in range(32)])
    ndc = rng.uniform(-1.2, 1.2, (64, 3)).astype(np.float32)
    wh = rng.integers(1, 4096, (64, 2)).astype(np.int32)
    out["vp_ndc"], out["vp_wh"] = ndc, wh
    out["vp_out"] = np.stack([ref.viewport(int(wh[i, 0]), int(wh[i, 1]), ndc[i]) for i in range(64)])
    return out


def uniform_kats(ref):
    scenes = [sg.small_scene(sg.BLINN, seed=s) for s in range(4)] + [sg.config1_scene()]
    rows = []
    for sc in scenes:
        u = ref.build_uniforms(sc)
        rows.append((np.array(sc.cam_pos, np.float32), np.array(sc.light_pos, np.float32), sc.width, sc.height,
                     np.frombuffer(bytes(u), np.uint8).copy()))
    return {"u_cam": np.stack([r[0] for r in rows]), "u_light": np.stack([r[1] for r in rows]),
            "u_wh": np.array([[r[2], r[3]] for r in rows], np.int32), "u_bytes": np.stack([r[4] for r in rows])}


def raster_kats(ref, rng):
    out = {}
    n = 2000
    tri = rng.uniform(0, 64, (n, 3, 2)).astype(np.float32)
    # snap some vertices to integers and some points onto edges/vertices
    tri[: n // 4] = np.round(tri[: n // 4])
    P = rng.integers(0, 64, (n, 2)).astype(np.float32)
    P[n // 2: n // 2 + 200] = np.round(tri[n // 2: n // 2 + 200, 0])
    deg = tri[-100:].copy()
    deg[:, 2] = deg[:, 0] + (deg[:, 1] - deg[:, 0]) * 0.5   # collinear
    tri[-100:] = deg
    out["bary_tri"], out["bary_p"] = tri, P
    out["bary_out"] = np.stack([ref.barycentric(tri[i, 0], tri[i, 1], tri[i, 2], P[i]) for i in range(n)])
    ndc = rng.uniform(-1, 1, (500, 9)).astype(np.float32)
    out["bf_ndc"] = ndc
    out["bf_out"] = np.array([ref.is_back_facing(x) for x in ndc], np.int32)
    # clipping: triangles crossing each plane and several at once
    cases = []
    for k in range(400):
        v = np.zeros((3, 13), np.float32)
        v[:, 3] = rng.uniform(-0.5, 2.0, 3)
        v[:, :3] = rng.uniform(-2.5, 2.5, (3, 3)) * np.abs(v[:, 3:4]).clip(0.2, None)
        v[:, 4:] = rng.normal(size=(3, 9))
        cases.append(v)
    cases = np.stack(cases).astype(np.float32)
    outs = np.zeros((len(cases), 10, 13), np.float32)
    counts = np.zeros(len(cases), np.int32)
    for i, c in enumerate(cases):
        r = ref.clip_triangle(c)
        counts[i] = r.shape[0]
        outs[i, : r.shape[0]] = r
    out["clip_in"], out["clip_n"], out["clip_out"] = cases, counts, outs
    return out


def shader_kats(ref, rng):
    out = {}
    sc = sg.small_scene(sg.NORMALMAP, shadow=True, seed=77, width=48, height=40, tex=32)
    u = ref.build_uniforms(sc)
    out["u_bytes"] = np.frombuffer(bytes(u), np.uint8).copy()
    out["diffuse"], out["normal"] = sc.diffuse, sc.normal
    sm = HostTarget(48, 40)
    sm.color[..., 0] = rng.integers(0, 256, (40, 48))
    sm.color[..., 1] = sm.color[..., 0]
    sm.color[..., 2] = sm.color[..., 0]
    out["shadow_color"] = sm.color.copy()
    a2v = rng.normal(size=(300, 8)).astype(np.float32)
    a2v[:, 6:8] = rng.uniform(0, 1, (300, 2))
    out["a2v"] = a2v
    vs, fs = {}, {}
    v2f = rng.normal(size=(1000, 13)).astype(np.float32)
    v2f[:, 10:12] = rng.uniform(-0.05, 1.05, (1000, 2))     # uv, partly out of range
    v2f[:, 4:7] *= 0.7
    v2f[:, 12] = rng.uniform(-0.2, 1.2, 1000)               # intensity
    v2f[:20, 7] = 0.0
    v2f[:20, 9] = 0.0                                        # normal with x = z = 0
    out["v2f"] = v2f
    for kind in range(7):
        out[f"vert_{kind}"] = np.stack([ref.vertex(kind, u, a) for a in a2v])
        for shadow in (0, 1):
            uu = ref.build_uniforms(sc, enable_shadow=shadow)
            out[f"frag_{kind}_{shadow}"] = np.stack([ref.fragment(kind, uu, v, sc.diffuse, sc.normal, sm) for v in v2f])
    return out


def frame_fixtures(ref):
    out = {}
    cases = [
        ("blinn", sg.small_scene(sg.BLINN, shadow=False, seed=31, width=64, height=48, segments=16, rings=8, tex=32)),
        ("normalmap", sg.small_scene(sg.NORMALMAP, shadow=False, seed=32, width=64, height=48, segments=16, rings=8, tex=32)),
        ("shadow_nm", sg.small_scene(sg.NORMALMAP, shadow=True, seed=33, width=64, height=48, segments=16, rings=8, tex=32)),
        ("edge_clip", sg.small_scene(sg.BLINN, shadow=True, seed=34, width=64, height=48, segments=16, rings=8, tex=32,
                                     cam=(0.4, 0.3, 1.3))),
        ("near_clip", sg.small_scene(sg.NORMALMAP, shadow=True, seed=35, width=64, height=48, segments=16, rings=8, tex=32,
                                     cam=(0.1, 0.05, 1.02))),
        ("toon", sg.small_scene(sg.TOON, shadow=False, seed=36, width=64, height=48, segments=16, rings=8, tex=32)),
        ("texture_light", sg.small_scene(sg.TEXTURE_LIGHT, shadow=False, seed=37, width=64, height=48, segments=16,
                                         rings=8, tex=32)),
    ]
    for name, sc in cases:
        tgt, sm, st_s, st_m = ref.render_passes(sc)
        d = {"pos": sc.mesh.pos, "nrm": sc.mesh.nrm, "uv": sc.mesh.uv, "idx": sc.mesh.idx,
             "diffuse": sc.diffuse, "normal": sc.normal,
             "cam": np.array(sc.cam_pos, np.float32), "light": np.array(sc.light_pos, np.float32),
             "meta": np.array([sc.width, sc.height, sc.shader, int(sc.shadow)], np.int32),
             "color": tgt.color, "depth": tgt.depth, "fragments": np.int64(st_m["fragments"])}
        if sm is not None:
            d["shadow_color"], d["shadow_depth"] = sm.color, sm.depth
            d["shadow_fragments"] = np.int64(st_s["fragments"])
        for k, v in d.items():
            out[f"{name}__{k}"] = v
    return out


def edge_fixtures(ref):
    """Two draws of the NaN-depth scene into one target: the mesh and camera
    transform of each draw, and the target after both."""
    out = {"camera_vp": NAN_SCENE_VP}
    t = HostTarget(96, 72)
    for i, shift in enumerate((0.0, 0.3)):
        sc = nan_depth_scene(shift=shift)
        ref.draw(sc.mesh, sc.shader, with_camera_vp(ref.build_uniforms(sc)), t, sc.diffuse, sc.normal)
        for k, v in {"pos": sc.mesh.pos, "nrm": sc.mesh.nrm, "uv": sc.mesh.uv, "idx": sc.mesh.idx,
                     "diffuse": sc.diffuse, "light": np.array(sc.light_pos, np.float32)}.items():
            out[f"nan{i}__{k}"] = v
    out["nan__color"], out["nan__depth"] = t.color, t.depth
    return out


def main(which=None):
    """Writes every fixture file, or only those named (without .npz)."""
    ref = CPURenderer("reference")
    rng = np.random.default_rng(SEED)
    files = {
        "kat_math": lambda: {**math_kats(ref, rng), **uniform_kats(ref)},
        "kat_raster": lambda: raster_kats(ref, rng),
        "kat_shaders": lambda: shader_kats(ref, rng),
        "frames": lambda: frame_fixtures(ref),
        "edge_frames": lambda: edge_fixtures(ref),
    }
    for name, make in files.items():
        if not which or name in which:
            np.savez_compressed(os.path.join(HERE, name + ".npz"), **make())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main(sys.argv[1:])
