#!/usr/bin/env python3
"""Benchmark: shaded Mfragments/s and frames/s of the gfx950 pipeline.

Workload (default): the single large frame -- icosphere subdivided 8x
(1,310,720 triangles) with value-noise radial displacement (amp 0.05),
2048^2 diffuse + normal maps, NormalMap shader with a 2048^2 shadow pre-pass,
1920x1080 target; one step = one DrawModel::draw frame (shadow pass + main
pass, target cleared first) with every input resident in HBM.

The shaded-fragment count is the reference's: one fragment() invocation per
covered pixel that passes the depth test at its turn in submission order,
counted by the CPU leg (compiled reference, or the oracle where the reference
is not built) for the shadow and main passes.  The CPU leg also gives the
cpu_baseline: the same frame on one host core.

Prints ONE JSON line (rank 0).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(REPO, "hana-softwarerenderer_amd"), os.path.join(REPO, "oracle")]

METRIC = "shaded Mfrag/s + frames/s (batched vs per-frame, same GPU); p50/p99 request ms"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_leg(scene, seconds):
    """Times whole frames on one host core; returns stats and the last frame."""
    from pyoracle import REF_SO, CPURenderer
    which = "reference" if os.path.exists(REF_SO) else "oracle"
    r = CPURenderer(which)
    u = r.build_uniforms(scene)
    times = []
    out = None
    t_start = time.perf_counter()
    while True:
        t0 = time.perf_counter()
        out = r.render_passes(scene, u)
        times.append(time.perf_counter() - t0)
        log(f"cpu frame {len(times)}: {times[-1]:.2f} s ({r.kind})")
        if time.perf_counter() - t_start >= seconds:
            break
    tgt, sm, st_s, st_m = out
    frags = st_m["fragments"] + (st_s["fragments"] if st_s else 0)
    return {"kind": r.kind, "frames": len(times), "sec_per_frame": float(np.mean(times)), "fragments": frags,
            "fragments_main": st_m["fragments"], "fragments_shadow": st_s["fragments"] if st_s else 0,
            "target": tgt}


def cpu_frame_counts(scenes, label):
    """CPU leg for a multi-scene workload: renders every scene once with the
    oracle on one core; returns per-scene fragment counts and the time."""
    from pyoracle import CPURenderer
    o = CPURenderer("oracle")
    counts = []
    t0 = time.perf_counter()
    for sc in scenes:
        tgt, sm, st_s, st_m = o.render_passes(sc)
        counts.append(st_m["fragments"] + (st_s["fragments"] if st_s else 0))
    dt = time.perf_counter() - t0
    log(f"cpu {label}: {len(scenes)} scenes in {dt:.1f} s, {sum(counts)} fragments")
    return counts, dt


def bench_batched(ctx, args):
    """256 independent scenes (1-8 displaced spheres/boxes, 1k-200k tris,
    128^2-1024^2 px, mixed shaders and shadows): per-frame launches against one
    captured graph for the whole batch."""
    from hana_amd import scenegen as sg
    from hana_amd.gpu import Batch, SceneOnDevice
    t0 = time.perf_counter()
    scenes = sg.config3_scenes()
    log(f"batched: generated {len(scenes)} scenes in {time.perf_counter() - t0:.1f} s")
    frags = None
    cpu = None
    if not args.no_cpu:
        counts, cpu_dt = cpu_frame_counts(scenes, "batched")
        frags = sum(counts)
        cpu = {"value": round(frags / cpu_dt / 1e6, 4), "unit": "Mfrag/s", "cores": 1, "kind": "port",
               "sample": f"all {len(scenes)} scenes once with the oracle, single-threaded, {cpu_dt:.1f} s"}
    devs = [SceneOnDevice(ctx, s) for s in scenes]
    # per-frame: one launch sequence per scene
    for d in devs:
        d.clear()
        d.render()
    ctx.sync()
    reps = args.batch_reps
    t0 = time.perf_counter()
    for _ in range(reps):
        for d in devs:
            d.clear()
            d.render()
    ctx.sync()
    t_frame = (time.perf_counter() - t0) / reps
    per_frame_out = [d.read()[0] for d in devs[:8]]
    batch = Batch(ctx)
    for d in devs:
        batch.add(d.desc())
    batch.finalize(use_graph=True)
    batch.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.run()
    ctx.sync()
    t_batch = (time.perf_counter() - t0) / reps
    same = all(np.array_equal(d.read()[0], c) for d, c in zip(devs[:8], per_frame_out))
    batch.close()
    # per-stage kernel times of the same launch sequence, run eagerly with events
    eager = Batch(ctx)
    for d in devs:
        eager.add(d.desc())
    eager.finalize(use_graph=False)
    eager.run()
    ctx.sync()
    ctx.read_timing()
    ctx.set_timing(True)
    for _ in range(reps):
        eager.run()
    ctx.sync()
    st = ctx.read_timing()
    ctx.set_timing(False)
    eager.close()
    # per batch: the sum over the batch's parts (each part is one launch sequence)
    stage_ms = {k: round(v[0] / reps, 4) for k, v in st.items()}
    out = {"workload": f"{len(scenes)} scenes, {sum(s.mesh.nfaces for s in scenes)} tris, "
                       f"{sum(s.width * s.height for s in scenes)} px, {sum(s.shadow for s in scenes)} with shadow",
           "per_frame_ms_per_batch": round(t_frame * 1e3, 3), "batched_ms_per_batch": round(t_batch * 1e3, 3),
           "per_frame_frames_per_s": round(len(scenes) / t_frame, 1),
           "batched_frames_per_s": round(len(scenes) / t_batch, 1),
           "speedup": round(t_frame / t_batch, 2), "fragments_per_batch": frags,
           "per_frame_Mfrag_s": round(frags / t_frame / 1e6, 2) if frags else None,
           "batched_Mfrag_s": round(frags / t_batch / 1e6, 2) if frags else None,
           "batched_equals_per_frame": same, "stage_ms": stage_ms, "cpu_baseline": cpu}
    log("batched:", json.dumps(out))
    return out


def bench_serving(ctx, args):
    """Poisson arrivals over 64 resident meshes (5k-500k tris) at 256^2 to
    1920x1080, continuous batching within an HBM budget; p50/p99 latency from
    submission to the colour image in host memory."""
    import threading
    from hana_amd import scenegen as sg
    from hana_amd.gpu import Server
    t0 = time.perf_counter()
    wl = sg.config4_workload(rate_per_s=args.serve_rate, seconds=args.serve_seconds)
    log(f"serving: {len(wl.requests)} requests over {args.serve_seconds} s at {args.serve_rate}/s, "
        f"{sum(m.nfaces for m in wl.meshes)} resident tris (generated in {time.perf_counter() - t0:.1f} s)")
    frags_per_req = None
    if not args.no_cpu and args.serve_cpu_sample > 0:
        sample = wl.requests[: args.serve_cpu_sample]
        counts, cpu_dt = cpu_frame_counts([sg.request_scene(wl, r) for r in sample], "serving sample")
        frags_per_req = sum(counts) / len(counts)
        cpu_req_s = len(sample) / cpu_dt
    meshes = [ctx.mesh(m) for m in wl.meshes]
    texs = [(ctx.texture(d), ctx.texture(n)) for d, n in wl.textures]
    server = Server(ctx, lanes=args.serve_lanes, max_batch_frames=args.serve_max_batch,
                    hbm_budget_bytes=int(args.serve_budget_gb * (1 << 30)))
    lat = []
    batch_sizes = []
    errors = [0]
    done = threading.Event()

    def poller():
        got = 0
        while got < len(wl.requests):
            cs = server.poll(256, wait_ms=20)
            for c in cs:
                if c.status != 0:
                    errors[0] += 1
                else:
                    lat.append(c.t_done_ms - c.t_submit_ms)
                    batch_sizes.append(c.batch_frames)
                server.release(int(c.id))
            got += len(cs)
            if done.is_set() and not cs and got >= len(wl.requests):
                break
    th = threading.Thread(target=poller, daemon=True)
    th.start()
    start = time.perf_counter()
    for i, r in enumerate(wl.requests):
        target_t = start + r["t_ms"] / 1000.0
        while True:
            dt = target_t - time.perf_counter()
            if dt <= 0:
                break
            time.sleep(min(dt, 0.0005))
        d, n = texs[r["tex"]]
        server.submit(meshes[r["mesh"]], r["shader"], r["shadow"], r["width"], r["height"], r["cam_pos"],
                      r["light_pos"], d, n, tag=i)
    done.set()
    th.join(timeout=120)
    wall = time.perf_counter() - start
    st = server.stats()
    server.close()
    lat_a = np.array(lat) if lat else np.zeros(1)
    out = {"workload": f"{len(wl.requests)} requests, Poisson {args.serve_rate}/s for {args.serve_seconds} s, "
                       f"64 meshes 5k-500k tris, 256^2-1920x1080, half shadowed, budget {args.serve_budget_gb} GB",
           "completed": len(lat), "errors": errors[0], "requests_per_s": round(len(lat) / wall, 1),
           "p50_ms": round(float(np.percentile(lat_a, 50)), 3), "p99_ms": round(float(np.percentile(lat_a, 99)), 3),
           "max_ms": round(float(lat_a.max()), 3), "mean_batch_frames": round(float(np.mean(batch_sizes)), 2)
           if batch_sizes else None, "gpu_busy_frac": round(st["busy_ms"] / (wall * 1e3), 3),
           "batches": st["batches"], "peak_inflight_GB": round(st["peak_inflight_bytes"] / 2 ** 30, 3)}
    if frags_per_req:
        out["est_Mfrag_s"] = round(frags_per_req * len(lat) / wall / 1e6, 2)
        out["cpu_baseline"] = {"value": round(cpu_req_s, 3), "unit": "requests/s", "cores": 1, "kind": "port",
                               "sample": f"first {args.serve_cpu_sample} requests with the oracle, single-threaded"}
        out["fragments_per_request_sample"] = round(frags_per_req, 1)
    log("serving:", json.dumps(out))
    return out


def bench_micro(ctx, args):
    """Micro-triangle stress: 8 jittered grid layers (20 M triangles of 1-4 px)
    over 3840x2160 with exact depth ties between layer pairs, NormalMap with a
    4096^2 shadow pre-pass; frames/s and shaded Mfrag/s as for the large frame."""
    from hana_amd import scenegen as sg
    from hana_amd.gpu import SceneOnDevice
    t0 = time.perf_counter()
    scene = sg.config5_scene()
    log(f"micro: {scene.mesh.nfaces} tris, {scene.width}x{scene.height}, shadow {scene.shadow_wh} "
        f"({time.perf_counter() - t0:.1f} s to generate)")
    frags = None
    cpu = None
    if not args.no_cpu:
        from pyoracle import CPURenderer
        o = CPURenderer("oracle")
        t0 = time.perf_counter()
        tgt, sm, st_s, st_m = o.render_passes(scene)
        cpu_dt = time.perf_counter() - t0
        frags = st_m["fragments"] + st_s["fragments"]
        log(f"cpu micro: 1 frame in {cpu_dt:.1f} s, {frags} fragments")
        cpu = {"value": round(frags / cpu_dt / 1e6, 4), "unit": "Mfrag/s", "cores": 1, "kind": "port",
               "sample": f"one whole frame with the oracle, single-threaded, {cpu_dt:.1f} s"}
    dev = SceneOnDevice(ctx, scene)
    for _ in range(2):
        dev.clear()
        dev.render()
    ctx.sync()
    steps = args.micro_steps
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.clear()
        dev.render()
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps
    ctx.set_timing(True)
    for _ in range(steps):
        dev.clear()
        dev.render()
    stages = ctx.read_timing()
    ctx.set_timing(False)
    stage_ms = {k: round(v[0] / v[1], 5) if v[1] else 0.0 for k, v in stages.items()}
    parity = None
    if cpu is not None:
        col, dep = dev.read()
        d = np.abs(col[..., :3].astype(np.int16) - tgt.color[..., :3].astype(np.int16)).max(-1)
        parity = {"depth_bit_exact": bool(np.array_equal(dep.view(np.uint32), tgt.depth.view(np.uint32))),
                  "color_max_lsb": int(d.max()), "color_frac_le1": float((d <= 1).mean())}
    _, frame_bytes = alg_bytes(scene)
    out = {"workload": f"{scene.mesh.nfaces} tris in 8 layers, {scene.width}x{scene.height}, "
                       f"shadow {scene.shadow_wh[0]}^2, NormalMap",
           "ms_per_frame": round(dt * 1e3, 3), "frames_per_s": round(1.0 / dt, 2),
           "fragments_per_frame": frags, "Mfrag_s": round(frags / dt / 1e6, 2) if frags else None,
           "frame_alg_bytes": frame_bytes, "frame_achieved_GBps": round(frame_bytes / dt / 1e9, 2),
           "stage_ms": stage_ms, "parity": parity, "cpu_baseline": cpu}
    log("micro:", json.dumps(out))
    del dev
    return out


TRAFFIC_FILE = os.path.join(REPO, "profiles", "traffic_large_frame.json")
TRAFFIC_SOURCE = ("profiles/counters_large_frame.md: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE per dispatch, "
                  "scaled by a plain copy's readings")


def measured_traffic(stage):
    """HBM bytes per launch of a stage from the committed counter summary of
    this workload (None if absent)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    per = json.load(open(TRAFFIC_FILE)).get("per_launch", {})
    for k, v in per.items():
        if k.split(" ")[0] == stage:
            return int(v["read_bytes"] + v["write_bytes"])
    return None


def measured_valu(stage):
    """VALU wave instructions per launch of a stage (SQ_INSTS_VALU) from the
    committed counter summary of this workload (None if absent)."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    per = json.load(open(TRAFFIC_FILE)).get("per_launch", {})
    for k, v in per.items():
        if k.split(" ")[0] == stage and "valu_instr" in v:
            return int(v["valu_instr"])
    return None


# VALU issue peak of MI355X: 256 CUs x 4 SIMDs at 2.4 GHz, one wave64 FP32
# instruction per SIMD every 4 cycles (16 lanes a cycle; packed FP32 is off).
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4


def alg_bytes(scene):
    """Algorithmic HBM bytes of one frame, split by stage:
    a2v 32 B and indices 12 B per vertex/triangle per pass, colour+depth 8 B
    per pixel written once, shadow map written (8 B/texel) and read (4 B/texel),
    each texture read once at 4 B/texel."""
    V, T = scene.mesh.nverts, scene.mesh.nfaces
    passes = 2 if scene.shadow else 1
    W, H = scene.width, scene.height
    S = scene.shadow_wh[0] * scene.shadow_wh[1] if scene.shadow else 0
    tex = scene.diffuse.shape[0] * scene.diffuse.shape[1] * 4 + scene.normal.shape[0] * scene.normal.shape[1] * 4
    b = {"vertex": passes * V * 32, "setup": passes * T * 12, "bin": 0, "clip": 0, "scan": 0, "unused": 0,
         "raster_g0": S * 8, "raster_g1": W * H * 8 + S * 4 + tex}
    if not scene.shadow:
        b["raster_g0"] = W * H * 8 + tex
        b["raster_g1"] = 0
    frame = passes * (V * 32 + T * 12) + W * H * 8 + (S * 8 + S * 4) + tex
    return b, frame


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-batched", action="store_true")
    ap.add_argument("--skip-serving", action="store_true")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--serve-max-batch", type=int, default=64)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-cpu-sample", type=int, default=40)
    ap.add_argument("--skip-micro", action="store_true")
    ap.add_argument("--micro-steps", type=int, default=5)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (scales rocprof FETCH/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 4, "copy_bytes_written": dst.numel() * 4, "launches": 5}))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and not run here.", file=sys.stderr)
        sys.exit(2)

    from hana_amd import scenegen as sg
    from hana_amd.gpu import Context, SceneOnDevice

    t0 = time.perf_counter()
    scene = sg.config2_scene()
    log(f"scene: {scene.mesh.nfaces} tris, {scene.width}x{scene.height}, shadow {scene.shadow_wh} "
        f"({time.perf_counter() - t0:.1f} s to generate)")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_leg(scene, args.cpu_seconds)

    ctx = Context(0)
    dev = SceneOnDevice(ctx, scene)
    for _ in range(args.warmup):
        dev.clear()
        dev.render()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.clear()
        dev.render()
    ctx.sync()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    log(f"gpu: {ms_per_step:.3f} ms/frame")

    # per-stage kernel times (events on the library's stream), a separate loop
    ctx.set_timing(True)
    for _ in range(args.steps):
        dev.clear()
        dev.render()
    stages = ctx.read_timing()
    ctx.set_timing(False)
    stage_ms = {k: (v[0] / v[1] if v[1] else 0.0) for k, v in stages.items()}
    log("stage ms:", json.dumps({k: round(v, 4) for k, v in stage_ms.items()}))

    col, dep = dev.read()
    parity = None
    frags = None
    if cpu is not None:
        ref = cpu["target"]
        d = np.abs(col[..., :3].astype(np.int16) - ref.color[..., :3].astype(np.int16)).max(-1)
        parity = {"depth_bit_exact": bool(np.array_equal(dep.view(np.uint32), ref.depth.view(np.uint32))),
                  "color_max_lsb": int(d.max()), "color_frac_le1": float((d <= 1).mean())}
        frags = cpu["fragments"]

    per_stage_bytes, frame_bytes = alg_bytes(scene)
    dom = max(stage_ms, key=lambda k: stage_ms[k])
    traffic = measured_traffic(dom)
    dom_gbps = per_stage_bytes.get(dom, 0) / (stage_ms[dom] * 1e-3) / 1e9 if stage_ms[dom] > 0 else 0.0
    value = frags * args.steps / elapsed / 1e6 if frags else None
    res = {
        "metric": METRIC,
        "value": round(value, 3) if value is not None else None,
        "unit": "Mfrag/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "single large frame: icosphere 8x subdivided (1,310,720 tris), displacement 0.05, "
                               "2048^2 diffuse + normal map, NormalMap + 2048^2 shadow pre-pass, 1920x1080",
                   "triangles": scene.mesh.nfaces, "vertices": scene.mesh.nverts,
                   "width": scene.width, "height": scene.height, "shadow_map": list(scene.shadow_wh)},
        "frames_per_s": round(args.steps / elapsed, 3),
        "fragments_per_frame": frags,
        "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(dom_gbps, 2), "peak": HBM_PEAK_GBPS,
                     "unit": "GB/s", "frac": round(dom_gbps / HBM_PEAK_GBPS, 5),
                     "bytes_per_launch": per_stage_bytes.get(dom, 0), "traffic": traffic,
                     "traffic_source": TRAFFIC_SOURCE if traffic is not None else None,
                     "launch_us": round(stage_ms[dom] * 1e3, 2),
                     "time_source": f"HIP events around each launch on the library stream, {args.steps} frames "
                                    "with stage timing on (one stream, frames not overlapped)",
                     "frame_alg_bytes": frame_bytes,
                     "frame_achieved": round(frame_bytes / (ms_per_step * 1e-3) / 1e9, 2)},
        "parity": parity,
    }
    valu = measured_valu(dom)
    if valu is not None and stage_ms[dom] > 0:
        rate = valu / (stage_ms[dom] * 1e-3)
        res["valu"] = {"kernel": dom, "wave_instr_per_launch": valu, "achieved": round(rate / 1e9, 2),
                       "peak": round(VALU_PEAK_PER_S / 1e9, 1), "unit": "G wave64 VALU instr/s",
                       "frac": round(rate / VALU_PEAK_PER_S, 4),
                       "source": "SQ_INSTS_VALU per dispatch from " + os.path.relpath(TRAFFIC_FILE, REPO) +
                                 " over the launch time above: the bound of this kernel"}
    if cpu is not None:
        cpu_mfrag = cpu["fragments"] / cpu["sec_per_frame"] / 1e6
        res["cpu_baseline"] = {"value": round(cpu_mfrag, 4), "unit": "Mfrag/s", "cores": 1, "kind": cpu["kind"],
                               "sample": f"{cpu['frames']} whole frame(s) of this workload, single-threaded, "
                                         f"{cpu['sec_per_frame']:.2f} s/frame",
                               "frames_per_s": round(1.0 / cpu["sec_per_frame"], 4)}
        res["gpu_over_cpu"] = round(value / cpu_mfrag, 1) if value else None
    if not args.skip_batched:
        res["batched"] = bench_batched(ctx, args)
    if not args.skip_serving:
        res["serving"] = bench_serving(ctx, args)
    if not args.skip_micro:
        res["micro"] = bench_micro(ctx, args)
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
