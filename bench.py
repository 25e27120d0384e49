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
    b = {"vertex": passes * V * 32, "setup_count": passes * T * 12, "setup_scatter": passes * T * 12,
         "clip_count": 0, "clip_scatter": 0, "scan": 0,
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
                     "bytes_per_launch": per_stage_bytes.get(dom, 0), "traffic": None,
                     "frame_alg_bytes": frame_bytes,
                     "frame_achieved": round(frame_bytes / (ms_per_step * 1e-3) / 1e9, 2)},
        "parity": parity,
    }
    if cpu is not None:
        cpu_mfrag = cpu["fragments"] / cpu["sec_per_frame"] / 1e6
        res["cpu_baseline"] = {"value": round(cpu_mfrag, 4), "unit": "Mfrag/s", "cores": 1, "kind": cpu["kind"],
                               "sample": f"{cpu['frames']} whole frame(s) of this workload, single-threaded, "
                                         f"{cpu['sec_per_frame']:.2f} s/frame",
                               "frames_per_s": round(1.0 / cpu["sec_per_frame"], 4)}
        res["gpu_over_cpu"] = round(value / cpu_mfrag, 1) if value else None
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
