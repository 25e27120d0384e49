"""Serving: continuous batching with HBM-budget admission.

Each request's image must be the same whether it runs alone (the per-frame
path) or inside a batch mixed with other requests of other sizes on
concurrent lanes; a small budget must throttle admission without losing or
corrupting requests."""
import numpy as np
import pytest

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
def workload():
    return sg.config4_workload(n_meshes=6, tri_lo=500, tri_hi=20_000, px_lo=32 * 32, px_hi=256 * 256,
                               seconds=1.0, max_requests=40)


@pytest.fixture(scope="module")
def resident(ctx, workload):
    meshes = [ctx.mesh(m) for m in workload.meshes]
    texs = [(ctx.texture(d), ctx.texture(n)) for d, n in workload.textures]
    return meshes, texs


def submit_all(server, workload, resident, reqs):
    meshes, texs = resident
    ids = {}
    for i, r in enumerate(reqs):
        d, n = texs[r["tex"]]
        rid = server.submit(meshes[r["mesh"]], r["shader"], r["shadow"], r["width"], r["height"], r["cam_pos"],
                            r["light_pos"], d, n, tag=i)
        ids[rid] = i
    return ids


def collect(server, count, timeout_ms=60_000):
    out = {}
    waited = 0
    while len(out) < count and waited < timeout_ms:
        got = server.poll(64, wait_ms=50)
        if not got:
            waited += 50
        for c in got:
            img = server.image(c) if c.status == 0 else None
            dep = server.depth(c) if c.status == 0 else None
            out[int(c.user_tag)] = (c.status, img, c.batch_frames, dep)
            server.release(int(c.id))
    return out


def alone(ctx, workload, resident, r):
    """The same request through the per-frame path."""
    from hana_amd.gpu import build_uniforms
    meshes, texs = resident
    sc = sg.request_scene(workload, r)
    u = build_uniforms(sc)
    tgt = ctx.target(sc.width, sc.height)
    sm = ctx.target(sc.width, sc.height) if sc.shadow else None
    d, n = texs[r["tex"]]
    ctx.render_frame(meshes[r["mesh"]], sc.shader, sc.shadow, u, tgt, sm, d, n)
    col, _ = ctx.read(tgt, sc.width, sc.height)
    return col


def test_same_alone_and_in_the_mix(ctx, workload, resident):
    from hana_amd.gpu import Server
    reqs = workload.requests
    server = Server(ctx, lanes=2, max_batch_frames=8)
    submit_all(server, workload, resident, reqs)
    res = collect(server, len(reqs))
    st = server.stats()
    server.close()
    assert len(res) == len(reqs) and all(v[0] == 0 for v in res.values())
    assert st["batches"] >= len(reqs) // 8 and max(v[2] for v in res.values()) > 1   # really batched
    for i, r in enumerate(reqs):
        ref = alone(ctx, workload, resident, r)
        assert np.array_equal(res[i][1], ref), f"request {i} differs alone vs in the mix"


def test_matches_oracle(ctx, workload, resident):
    from hana_amd.gpu import Server
    reqs = workload.requests[:6]
    server = Server(ctx, lanes=3, max_batch_frames=4, return_depth=True)
    submit_all(server, workload, resident, reqs)
    res = collect(server, len(reqs))
    server.close()
    o = CPURenderer("oracle")
    for i, r in enumerate(reqs):
        tgt, _, _ = o.render_frame(sg.request_scene(workload, r))
        assert_color_ok(res[i][1], tgt.color, f"request {i}")
        assert_depth_exact(res[i][3], tgt.depth, f"request {i}")


def test_config4_full_size_matches_oracle(ctx):
    """The serving workload at its own sizes: the first 64 requests of the
    Poisson stream over the 64 resident meshes (5k-500k triangles), 256^2 to
    1920x1080, Blinn/NormalMap, half shadowed, submitted together and batched
    across lanes; every completion's colour and depth against the oracle."""
    from hana_amd.gpu import Server
    wl = sg.config4_workload(seconds=1.0, max_requests=64)
    assert len(wl.meshes) == 64 and len(wl.requests) == 64
    meshes = [ctx.mesh(m) for m in wl.meshes]
    texs = [(ctx.texture(d), ctx.texture(n)) for d, n in wl.textures]
    server = Server(ctx, lanes=3, max_batch_frames=16, return_depth=True)
    for i, r in enumerate(wl.requests):
        d, n = texs[r["tex"]]
        server.submit(meshes[r["mesh"]], r["shader"], r["shadow"], r["width"], r["height"], r["cam_pos"],
                      r["light_pos"], d, n, tag=i)
    got = {}
    waited = 0
    while len(got) < len(wl.requests) and waited < 120_000:
        cs = server.poll(64, wait_ms=50)
        waited += 0 if cs else 50
        for c in cs:
            got[int(c.user_tag)] = (c.status, server.image(c) if c.status == 0 else None,
                                    server.depth(c) if c.status == 0 else None, c.batch_frames)
            server.release(int(c.id))
    st = server.stats()
    server.close()
    assert len(got) == len(wl.requests) and all(v[0] == 0 for v in got.values())
    assert max(v[3] for v in got.values()) > 1          # batched across requests
    o = CPURenderer("oracle")
    for i, r in enumerate(wl.requests):
        sc = sg.request_scene(wl, r)
        tgt, _, _ = o.render_frame(sc)
        assert_depth_exact(got[i][2], tgt.depth, f"request {i} ({sc.width}x{sc.height})")
        assert_color_ok(got[i][1], tgt.color, f"request {i} ({sc.width}x{sc.height})")
    assert st["completed"] == len(wl.requests)


def test_budget_throttles_admission(ctx, workload, resident):
    """A budget of about two requests keeps the in-flight footprint under it;
    a request larger than the whole budget is rejected, the others complete."""
    from hana_amd.gpu import Server
    reqs = workload.requests[:12]
    sizes = [r["width"] * r["height"] for r in reqs]
    budget = 24 * (1 << 20)
    server = Server(ctx, lanes=3, max_batch_frames=16, hbm_budget_bytes=budget)
    submit_all(server, workload, resident, reqs)
    res = collect(server, len(reqs))
    st = server.stats()
    server.close()
    assert len(res) == len(reqs)
    assert st["peak_inflight_bytes"] <= budget
    ok = [i for i, v in res.items() if v[0] == 0]
    assert len(ok) >= 1
    for i in ok:
        assert np.array_equal(res[i][1], alone(ctx, workload, resident, reqs[i]))
    rejected = [i for i, v in res.items() if v[0] != 0]
    assert st["rejected"] == len(rejected)
    assert all(v[0] in (0, -3) for v in res.values()), sizes
