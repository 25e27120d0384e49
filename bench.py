#!/usr/bin/env python3
"""Benchmark: shaded Mfragments/s and frames/s, batched vs per-frame on one
GPU, and p50/p99 request latency, on the five workloads of the hot path.

Headline (`value`): the batched workload -- 256 independent scenes of 1-8
seeded displaced spheres/boxes (1k-200k triangles, 128^2-1024^2 pixels,
random aspect, Blinn/NormalMap x shadow on/off) rendered by one captured
hipGraph whose tuning (tile size, raster form, parts) the library's
autotuner picks; one step = one run of the whole batch, inputs resident in
HBM.  Beside it, the two per-frame forms of the same 256 frames: the C ABI's
resident per-call path, and the plain drop-in's call pattern (every pass
uploads its mesh corners, textures and target and reads them back, as
graphics_draw_triangle's drop-in does).

Sections:
  c1  plain method, small: UV sphere 64x32, 800x600, NormalMap, no shadow
  c2  single large frame: 1.31M-triangle icosphere, 1920x1080, 2048^2 shadow
  c3  the batched workload above (also the headline)
  c4  serving: Poisson arrivals at 2k req/s over 64 resident meshes for 60 s
  c5  micro-triangle stress: 20M triangles, 3840x2160, 4096^2 shadow

The shaded-fragment count is the reference's: one fragment() invocation per
covered pixel that passes the depth test at its turn in submission order,
counted by the CPU leg.  The CPU legs run the reference's own path compiled
from its sources (oracle/_ref; the oracle port where it is not built) on the
host: one core, and one worker process per core of the box's share, over
bounded samples; they run before the GPU is touched.  Every GPU image those
samples cover is compared with the CPU's (depth bit-exact, colour within
1 LSB on 99.9 % and 2 LSB everywhere) and reported as `parity`.

Prints ONE JSON line.
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
# VALU issue peak of MI355X for unpacked FP32 code (packed FP32 is off): 256
# CUs x 4 SIMDs at 2.4 GHz, one wave64 instruction per SIMD every 4 cycles
# (16 lanes a cycle; the published 157.3 TFLOP/s counts packed FP32, two
# operations per lane).  SQ_INSTS_VALU counts an instruction whatever its
# active lanes, and the micro-triangle frame's shadow raster (43 % of lanes
# active) retires 1.2x this rate: instructions with lanes masked off must,
# on average, issue in fewer cycles.  The fraction indicates VALU pressure;
# it is not a bound.
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 4
STAGES = ("vertex", "setup", "clip", "scan", "bin", "unused", "raster_g0", "raster_g1")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# ---------------------------------------------------------------------------
# CPU legs (run before the GPU is touched: worker processes are forked)
# ---------------------------------------------------------------------------

_WORK = {}   # scenes the forked workers render, by key


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_workers():
    """Worker processes: one per core of this process's share (at most 16,
    the host share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_kind():
    from pyoracle import REF_SO
    return "reference" if os.path.exists(REF_SO) else "oracle"


def _frags(out):
    tgt, sm, st_s, st_m = out
    return st_m["fragments"] + (st_s["fragments"] if st_s else 0)


def _render_many(args):
    """Worker: renders scenes _WORK[key][i] for i in idx; returns fragments."""
    key, idx, kind = args
    from pyoracle import CPURenderer
    r = CPURenderer(kind)
    scenes = _WORK[key]
    n = 0
    for i in idx:
        n += _frags(r.render_passes(scenes[i]))
    return n


def cpu_single(scenes, kind, keep=False, min_frames=1, min_seconds=0.0, label=""):
    """One core: renders scenes in order (cycling until min_frames and
    min_seconds are reached); returns stats and, with keep, the images."""
    from pyoracle import CPURenderer
    r = CPURenderer(kind)
    frames, frags, images = 0, 0, []
    t0 = time.perf_counter()
    i = 0
    while True:
        sc = scenes[i % len(scenes)]
        out = r.render_passes(sc)
        frags += _frags(out)
        if keep and i < len(scenes):
            images.append((out[0].color.copy(), out[0].depth.copy(), _frags(out)))
        frames += 1
        i += 1
        dt = time.perf_counter() - t0
        if frames >= max(min_frames, len(scenes) if keep else 1) and dt >= min_seconds:
            break
    log(f"cpu {label}: {frames} frames on 1 core in {dt:.1f} s ({kind})")
    return {"frames": frames, "seconds": dt, "fragments": frags, "images": images}


def cpu_all(key, scenes, kind, repeat=1, label=""):
    """Every worker process renders a share of the scenes (each `repeat`
    times); returns frames, fragments and wall seconds."""
    import multiprocessing as mp
    workers = cpu_workers()
    _WORK[key] = scenes
    jobs = []
    order = list(range(len(scenes))) * repeat
    for w in range(workers):
        idx = order[w::workers]
        if idx:
            jobs.append((key, idx, kind))
    t0 = time.perf_counter()
    with mp.get_context("fork").Pool(len(jobs)) as pool:
        frags = sum(pool.map(_render_many, jobs, chunksize=1))
    dt = time.perf_counter() - t0
    log(f"cpu {label}: {len(order)} frames on {len(jobs)} processes in {dt:.1f} s ({kind})")
    return {"frames": len(order), "seconds": dt, "fragments": frags, "workers": len(jobs)}


def cpu_baseline_block(single, allc, kind, unit_frames, sample_single, sample_all):
    """cpu_baseline: the all-core figure, with the single-core one beside it."""
    def rate(x):
        return x["fragments"] / x["seconds"] / 1e6
    out = {"value": round(rate(allc), 4), "unit": "Mfrag/s", "cores": allc["workers"], "kind": kind,
           "sample": sample_all, "frames_per_s": round(allc["frames"] / allc["seconds"], 4),
           "single_core": {"value": round(rate(single), 4), "unit": "Mfrag/s", "cores": 1, "sample": sample_single,
                           "ms_per_frame": round(single["seconds"] / single["frames"] * 1e3, 3),
                           "frames_per_s": round(single["frames"] / single["seconds"], 4)},
           "cpu_model": cpu_model()}
    if unit_frames:
        out["frames_unit"] = unit_frames
    return out


# ---------------------------------------------------------------------------
# algorithmic bytes (the roofline numerator)
# ---------------------------------------------------------------------------

def tex_bytes(scene):
    """Textures the main pass reads once: the diffuse map for every shader that
    samples it, the normal map for NormalMap only (RGBA8 words on the device)."""
    b = 0
    if scene.shader in (2, 3, 4, 5) and scene.diffuse is not None:   # Texture, TextureLight, Blinn, NormalMap
        b += scene.diffuse.shape[0] * scene.diffuse.shape[1] * 4
    if scene.shader == 5 and scene.normal is not None:
        b += scene.normal.shape[0] * scene.normal.shape[1] * 4
    return b


def alg_bytes(scene):
    """Per stage and per frame: a2v 32 B per vertex and 12 B per face per pass,
    colour+depth 8 B per pixel written once, the shadow map written (8 B per
    texel) and sampled (4 B per texel), each texture read once."""
    V, T = scene.mesh.nverts, scene.mesh.nfaces
    passes = 2 if scene.shadow else 1
    W, H = scene.width, scene.height
    S = scene.shadow_wh[0] * scene.shadow_wh[1] if scene.shadow else 0
    tex = tex_bytes(scene)
    main = W * H * 8 + S * 4 + tex
    # the fp32 extension's shadow pass writes its fp32 depth only
    sw = S * (4 if getattr(scene, "shadow_fp32", False) else 8)
    b = {"vertex": passes * V * 32, "setup": passes * T * 12, "clip": 0, "scan": 0, "bin": 0,
         "raster_g0": sw if scene.shadow else main, "raster_g1": main if scene.shadow else 0,
         "raster_all": sw + main}
    frame = passes * (V * 32 + T * 12) + W * H * 8 + sw + S * 4 + tex
    return b, frame


def batch_alg_bytes(scenes):
    """Per stage for a batch: its shadow passes form raster group 0 and its
    main passes group 1 (every frame's main pass is in group 1)."""
    tot = {"vertex": 0, "setup": 0, "clip": 0, "scan": 0, "bin": 0, "raster_g0": 0, "raster_g1": 0, "raster_all": 0}
    frame = 0
    for s in scenes:
        b, f = alg_bytes(s)
        S = s.shadow_wh[0] * s.shadow_wh[1] if s.shadow else 0
        main = b["raster_all"] - S * 8
        tot["vertex"] += b["vertex"]
        tot["setup"] += b["setup"]
        tot["raster_g0"] += S * 8
        tot["raster_g1"] += main
        tot["raster_all"] += b["raster_all"]
        frame += f
    return tot, frame


TRAFFIC_FILES = {"c2": os.path.join(REPO, "profiles", "traffic_large_frame.json"),
                 "c3": os.path.join(REPO, "profiles", "traffic_batched.json"),
                 "c5": os.path.join(REPO, "profiles", "traffic_micro.json")}


def measured(section, stage, field):
    """A per-launch figure of a stage from the committed counter summary of
    the workload (rocprofv3 --pmc), or None."""
    f = TRAFFIC_FILES.get(section)
    if not f or not os.path.exists(f):
        return None
    per = json.load(open(f)).get("per_launch", {})
    v = per.get(stage)
    if not v:
        return None
    if field == "traffic":
        return int(v["read_bytes"] + v["write_bytes"]) if "read_bytes" in v else None
    return int(v[field]) if field in v else None


def roofline(section, stage_ms, launches, stage_bytes):
    """The dominant stage against the HBM peak: algorithmic bytes per launch
    over its mean launch time (HIP events on the stream it runs on)."""
    stage = max(stage_ms, key=lambda k: stage_ms[k])
    n = max(launches.get(stage, 1), 1)
    per_launch_ms = stage_ms[stage] / n
    nbytes = stage_bytes.get(stage, 0) / n
    gbps = nbytes / (per_launch_ms * 1e-3) / 1e9 if per_launch_ms > 0 else 0.0
    tr = measured(section, stage, "traffic")
    out = {"bound": "hbm", "kernel": stage, "achieved": round(gbps, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": round(gbps / HBM_PEAK_GBPS, 5), "bytes_per_launch": int(nbytes), "traffic": tr,
           "traffic_source": (os.path.relpath(TRAFFIC_FILES[section], REPO) + ": rocprofv3 --pmc FETCH_SIZE and "
                              "WRITE_SIZE per dispatch, scaled by a plain copy's readings") if tr else None,
           "launch_us": round(per_launch_ms * 1e3, 2), "launches_per_step": n,
           "time_source": "HIP events recorded around each launch on the library's stream (stage timing on: "
                          "launches not overlapped)"}
    valu = measured(section, stage, "valu_instr")
    if valu is not None and per_launch_ms > 0:
        rate = valu / (per_launch_ms * 1e-3)
        out["valu"] = {"wave_instr_per_launch": valu, "achieved": round(rate / 1e9, 2),
                       "peak": round(VALU_PEAK_PER_S / 1e9, 1), "unit": "G wave64 VALU instr/s",
                       "frac": round(rate / VALU_PEAK_PER_S, 4)}
    return out


def parity_block(pairs):
    """pairs: (gpu colour, gpu depth, cpu colour, cpu depth) per image."""
    exact, worst, frac = True, 0, 1.0
    for gc, gd, cc, cd in pairs:
        exact = exact and np.array_equal(gd.view(np.uint32), cd.view(np.uint32))
        d = np.abs(gc[..., :3].astype(np.int16) - cc[..., :3].astype(np.int16)).max(-1)
        worst = max(worst, int(d.max()) if d.size else 0)
        frac = min(frac, float((d <= 1).mean()) if d.size else 1.0)
    return {"images": len(pairs), "depth_bit_exact": bool(exact), "color_max_lsb": worst,
            "color_min_frac_le1": round(frac, 6), "ok": bool(exact and worst <= 2 and frac >= 0.999)}


# ---------------------------------------------------------------------------
# GPU legs
# ---------------------------------------------------------------------------

def timed_stages(ctx, fn, reps):
    """Runs fn() reps times with per-launch HIP events; mean ms per rep and
    launches per rep, by stage."""
    ctx.sync()
    ctx.read_timing()
    ctx.set_timing(True)
    for _ in range(reps):
        fn()
    st = ctx.read_timing()
    ctx.set_timing(False)
    ms = {k: v[0] / reps for k, v in st.items() if v[1]}
    n = {k: max(1, round(v[1] / reps)) for k, v in st.items() if v[1]}
    return ms, n


def per_frame_section(ctx, scene, name, steps, warmup, cpu, images):
    """One scene drawn frame after frame (DrawModel::draw per step)."""
    from hana_amd.gpu import SceneOnDevice
    dev = SceneOnDevice(ctx, scene)
    for _ in range(warmup):
        dev.clear()
        dev.render()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        dev.clear()
        dev.render()
    ctx.sync()
    dt = (time.perf_counter() - t0) / steps

    def one():
        dev.clear()
        dev.render()
    stage_ms, launches = timed_stages(ctx, one, min(steps, 20))
    stage_bytes, frame_bytes = alg_bytes(scene)
    out = {"workload": scene.name, "triangles": scene.mesh.nfaces, "width": scene.width, "height": scene.height,
           "shadow_map": list(scene.shadow_wh) if scene.shadow else None,
           "ms_per_frame": round(dt * 1e3, 4), "frames_per_s": round(1.0 / dt, 2), "steps": steps,
           "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
           "frame_alg_bytes": frame_bytes, "frame_achieved_GBps": round(frame_bytes / dt / 1e9, 2),
           "roofline": roofline(name, stage_ms, launches, stage_bytes)}
    if cpu is not None:
        frags = cpu["fragments_per_frame"]
        out["fragments_per_frame"] = frags
        out["Mfrag_s"] = round(frags / dt / 1e6, 2)
        if cpu.get("baseline"):
            out["cpu_baseline"] = cpu["baseline"]
            out["gpu_over_cpu_baseline"] = round(out["Mfrag_s"] / cpu["baseline"]["value"], 1)
    if images:
        col, dep = dev.read()
        out["parity"] = parity_block([(col, dep, images[0][0], images[0][1])])
    return out


def dropin_frame(lib, ctx, scene, u, corners, host_tgt, host_sm):
    """The drop-in's calls for one DrawModel::draw frame (shadow pass, main
    pass): per pass, the mesh corners, both material textures, a sampled
    shadow map and the target go up, one pass runs, the target comes back,
    the uploads are freed (graphics_hana.cpp keeps nothing between calls)."""
    import ctypes as C
    from hana_amd import gpu
    pos, nrm, uv, idx = corners
    F = idx.shape[0]
    passes = ([(gpu.SHADOW, host_sm, None)] if scene.shadow else []) + [(scene.shader, host_tgt, host_sm)]
    for kind, tgt_host, sampled in passes:
        h = ctx.h
        mesh, dif, nm = C.c_void_p(), C.c_void_p(), C.c_void_p()
        ctx.check(lib.hana_mesh_create(h, 3 * F, gpu._fp(pos), gpu._fp(nrm), gpu._fp(uv), F,
                                       idx.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(mesh)), "mesh")
        for img, out in ((scene.diffuse, dif), (scene.normal, nm)):
            ctx.check(lib.hana_texture_create(h, img.shape[1], img.shape[0], img.shape[2],
                                              img.ctypes.data_as(C.POINTER(C.c_uint8)), C.byref(out)), "texture")
        dev_t, (hc, hd) = tgt_host
        smap = None
        if kind != gpu.SHADOW and scene.shadow:
            smap, (sc, sd) = sampled
            ctx.upload(smap, sc, sd)
        ctx.upload(dev_t, hc, hd)
        d = gpu.DrawDesc(mesh, kind, 0, C.pointer(u), dif, nm, smap, dev_t)
        ctx.check(lib.hana_draw(h, C.byref(d)), "draw")
        ctx.check(lib.hana_target_read(h, dev_t, hc.ctypes.data_as(C.POINTER(C.c_uint8)), gpu._fp(hd)), "read")
        lib.hana_texture_destroy(h, dif)
        lib.hana_texture_destroy(h, nm)
        lib.hana_mesh_destroy(h, mesh)


def bench_c3(ctx, args, scenes, cpu):
    """The batched workload: resident per-call frames, the drop-in's call
    pattern, and the captured batch (default tuning, then autotuned)."""
    from hana_amd import gpu
    from hana_amd.gpu import Batch, SceneOnDevice
    lib = ctx.lib
    frags = cpu["fragments_per_batch"] if cpu else None
    devs = [SceneOnDevice(ctx, s) for s in scenes]
    reps = args.batch_reps
    t_call = None
    if not args.c3_batch_only:
        # resident per-call frames (hana_render_frame per scene)
        for d in devs:
            d.clear()
            d.render()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            for d in devs:
                d.clear()
                d.render()
        ctx.sync()
        t_call = (time.perf_counter() - t0) / reps
    # the drop-in's calls, per frame, with host RenderBuffers
    t_drop = None
    if not args.skip_dropin and not args.c3_batch_only:
        hosts = []
        for s, d in zip(scenes, devs):
            m = s.mesh
            ix = np.ascontiguousarray(m.idx, np.int32).reshape(-1)
            corners = (np.ascontiguousarray(m.pos[ix], np.float32), np.ascontiguousarray(m.nrm[ix], np.float32),
                       np.ascontiguousarray(m.uv[ix], np.float32), np.arange(ix.size, dtype=np.int32).reshape(-1, 3))
            tgt = (d.target, (np.zeros((s.height, s.width, 4), np.uint8), np.full((s.height, s.width), np.float32(3.4e38))))
            sm = None
            if s.shadow:
                sw, sh = s.shadow_wh
                sm = (d.shadow_map, (np.zeros((sh, sw, 4), np.uint8), np.full((sh, sw), np.float32(3.4e38))))
            hosts.append((corners, tgt, sm))
        dropin_frame(lib, ctx, scenes[0], devs[0].u, *hosts[0])   # warm
        t0 = time.perf_counter()
        for s, d, h in zip(scenes, devs, hosts):
            dropin_frame(lib, ctx, s, d.u, *h)
        ctx.sync()
        t_drop = time.perf_counter() - t0
        del hosts
    # the captured batch, default tuning
    t_default, tuning_default = None, None
    if not args.c3_batch_only:
        b0 = Batch(ctx)
        for d in devs:
            b0.add(d.desc())
        b0.finalize(use_graph=True)
        for _ in range(args.warmup):
            b0.run()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b0.run()
        ctx.sync()
        t_default = (time.perf_counter() - t0) / args.steps
        tuning_default = b0.tuning()
        b0.close()
    # every candidate of the autotuner's space, timed alone (evidence for the choice)
    sweep = None
    if args.c3_sweep:
        sweep = []
        for tile in (16, 32):
            for wgs in (0, 2, 4, 8):
                for waves in (4, 2):
                    for parts in (1, 2, 4):
                        bs = Batch(ctx)
                        for d in devs:
                            bs.add(d.desc())
                        bs.set_tuning(tile, wgs, parts, waves)
                        bs.finalize(use_graph=True)
                        for _ in range(2):
                            bs.run()
                        ctx.sync()
                        t0 = time.perf_counter()
                        for _ in range(args.steps):
                            bs.run()
                        ctx.sync()
                        ms_ = (time.perf_counter() - t0) / args.steps * 1e3
                        bs.close()
                        sweep.append({"tile_size": tile, "raster_wgs_per_cu": wgs, "raster_waves": waves,
                                      "parts": parts, "ms_per_batch": round(ms_, 4)})
                        log(f"c3 sweep: {sweep[-1]}")
    # autotuned (or the tuning given)
    b = Batch(ctx)
    for d in devs:
        b.add(d.desc())
    t0 = time.perf_counter()
    if args.tuning:
        b.set_tuning(*[int(x) for x in args.tuning.split(",")])
        b.finalize(use_graph=True)
        chosen = b.tuning()
    else:
        chosen = b.autotune(use_graph=True, reps=3)
    t_tune = time.perf_counter() - t0
    for _ in range(args.warmup):
        b.run()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    ctx.sync()
    t_batch = (time.perf_counter() - t0) / args.steps
    def ms(t):
        return round(t * 1e3, 4) if t else None
    log(f"c3: per-call {ms(t_call)} ms, drop-in pattern {ms(t_drop)} ms, batch default {ms(t_default)} ms, "
        f"tuned {ms(t_batch)} ms ({chosen})")
    # parity: every frame of the tuned batch, its targets poisoned first
    parity = None
    if cpu and cpu.get("images"):
        for d in devs:
            maps = [(d.target, d.scene.width, d.scene.height)]
            if d.shadow_map is not None:
                maps.append((d.shadow_map,) + tuple(d.scene.shadow_wh))
            for t, w, h in maps:
                ctx.upload(t, np.full((h, w, 4), 0x5A, np.uint8), np.full((h, w), np.nan, np.float32))
        b.run()
        pairs = []
        for d, (cc, cd, _) in zip(devs, cpu["images"]):
            gc, gd = d.read()
            pairs.append((gc, gd, cc, cd))
        parity = parity_block(pairs)
        log("c3 parity:", parity)
    # stage times of the same launch sequences, eagerly, with events
    eager = Batch(ctx)
    for d in devs:
        eager.add(d.desc())
    eager.set_tuning(chosen["tile_size"], chosen["raster_wgs_per_cu"], chosen["parts"])
    eager.finalize(use_graph=False)
    eager.run()
    stage_ms, launches = timed_stages(ctx, eager.run, reps)
    eager.close()
    b.close()
    stage_bytes, batch_bytes = batch_alg_bytes(scenes)
    if chosen["raster_wgs_per_cu"]:
        # the persistent raster runs both groups in one launch, timed as raster_g0
        stage_bytes = dict(stage_bytes, raster_g0=stage_bytes["raster_all"], raster_g1=0)
    n = len(scenes)
    out = {"workload": f"{n} scenes, {sum(s.mesh.nfaces for s in scenes)} tris, "
                       f"{sum(s.width * s.height for s in scenes)} px, {sum(s.shadow for s in scenes)} with shadow",
           "ms_per_batch": round(t_batch * 1e3, 4), "frames_per_s": round(n / t_batch, 1),
           "tuning": chosen, "tuning_source": "given" if args.tuning else "hana_batch_autotune",
           "tuning_default": tuning_default, "autotune_s": round(t_tune, 2),
           "default_tuning_ms_per_batch": ms(t_default),
           "per_call_ms_per_batch": ms(t_call), "per_call_frames_per_s": round(n / t_call, 1) if t_call else None,
           "dropin_ms_per_batch": ms(t_drop), "dropin_frames_per_s": round(n / t_drop, 1) if t_drop else None,
           "speedup_vs_dropin": round(t_drop / t_batch, 2) if t_drop else None,
           "speedup_vs_per_call": round(t_call / t_batch, 2) if t_call else None,
           "stage_ms": {k: round(v, 5) for k, v in stage_ms.items()},
           "batch_alg_bytes": batch_bytes, "batch_achieved_GBps": round(batch_bytes / t_batch / 1e9, 2),
           "roofline": roofline("c3", stage_ms, launches, stage_bytes), "parity": parity}
    if sweep:
        out["tuning_sweep"] = sweep
    if frags:
        out["fragments_per_batch"] = frags
        out["Mfrag_s"] = round(frags / t_batch / 1e6, 2)
        out["per_call_Mfrag_s"] = round(frags / t_call / 1e6, 2) if t_call else None
        out["dropin_Mfrag_s"] = round(frags / t_drop / 1e6, 2) if t_drop else None
        out["cpu_baseline"] = cpu["baseline"]
        out["gpu_over_cpu_baseline"] = round(out["Mfrag_s"] / cpu["baseline"]["value"], 1)
    return out


def bench_c4(ctx, args, wl, cpu):
    """Poisson arrivals over 64 resident meshes (5k-500k tris) at 256^2 to
    1920x1080, continuous batching within an HBM budget; p50/p99 latency from
    submission to the colour image in host memory."""
    import threading
    from hana_amd.gpu import Server
    meshes = [ctx.mesh(m) for m in wl.meshes]
    texs = [(ctx.texture(d), ctx.texture(n)) for d, n in wl.textures]
    budget = int(args.serve_budget_gb * (1 << 30))
    n_req = len(wl.requests)

    def submit(server, i, r):
        d, n = texs[r["tex"]]
        server.submit(meshes[r["mesh"]], r["shader"], r["shadow"], r["width"], r["height"], r["cam_pos"],
                      r["light_pos"], d, n, tag=i)

    # parity pass: the CPU sample's requests, colour and depth
    parity = None
    if cpu and cpu.get("images"):
        k = len(cpu["images"])
        server = Server(ctx, lanes=args.serve_lanes, max_batch_frames=args.serve_max_batch, hbm_budget_bytes=budget,
                        return_depth=True)
        for i, r in enumerate(wl.requests[:k]):
            submit(server, i, r)
        got = {}
        while len(got) < k:
            for c in server.poll(64, wait_ms=50):
                if c.status == 0:
                    got[int(c.user_tag)] = (server.image(c), server.depth(c))
                else:
                    got[int(c.user_tag)] = None
                server.release(int(c.id))
        server.close()
        pairs = [(got[i][0], got[i][1], cc, cd) for i, (cc, cd, _) in enumerate(cpu["images"]) if got.get(i)]
        parity = parity_block(pairs)
        parity["completed"] = len(pairs)
        log("c4 parity:", parity)
    server = Server(ctx, lanes=args.serve_lanes, max_batch_frames=args.serve_max_batch, hbm_budget_bytes=budget)
    # warm-up: the first second of the stream (fills the target and pinned pools), untimed
    warm = [r for r in wl.requests if r["t_ms"] < 1000.0]
    for i, r in enumerate(warm):
        submit(server, n_req + i, r)
    left = len(warm)
    while left > 0:
        cs = server.poll(256, wait_ms=50)
        for c in cs:
            server.release(int(c.id))
        left -= len(cs)
    lat, batch_sizes, errors = [], [], [0]
    t_lat = []
    done = threading.Event()

    def poller():
        got = 0
        while got < n_req:
            cs = server.poll(256, wait_ms=20)
            for c in cs:
                if c.status != 0:
                    errors[0] += 1
                else:
                    lat.append(c.t_done_ms - c.t_submit_ms)
                    t_lat.append(c.t_submit_ms)
                    batch_sizes.append(c.batch_frames)
                server.release(int(c.id))
            got += len(cs)
            if done.is_set() and not cs and got >= n_req:
                break
    st0 = server.stats()
    th = threading.Thread(target=poller, daemon=True)
    th.start()
    start = time.perf_counter()
    next_log = start + 15.0
    for i, r in enumerate(wl.requests):
        target_t = start + r["t_ms"] / 1000.0
        while True:
            now = time.perf_counter()
            dt = target_t - now
            if dt <= 0:
                break
            time.sleep(min(dt, 0.0005))
        if now >= next_log:
            log(f"c4: {i} requests submitted, {len(lat)} completed")
            next_log += 15.0
        submit(server, i, r)
    done.set()
    th.join(timeout=120)
    wall = time.perf_counter() - start
    st = server.stats()
    server.close()
    lat_a = np.array(lat) if lat else np.zeros(1)
    worst = int(np.argmax(lat_a))
    out = {"workload": f"{n_req} requests, Poisson {args.serve_rate:g}/s for {args.serve_seconds:g} s, "
                       f"{len(wl.meshes)} meshes 5k-500k tris, 256^2-1920x1080, half shadowed, "
                       f"budget {args.serve_budget_gb:g} GB, {args.serve_lanes} lanes",
           "completed": len(lat), "errors": errors[0], "requests_per_s": round(len(lat) / wall, 1),
           "p50_ms": round(float(np.percentile(lat_a, 50)), 3), "p99_ms": round(float(np.percentile(lat_a, 99)), 3),
           "p999_ms": round(float(np.percentile(lat_a, 99.9)), 3), "max_ms": round(float(lat_a.max()), 3),
           "max_at_s": round(float(t_lat[worst]) / 1e3, 2) if t_lat else None,
           "mean_batch_frames": round(float(np.mean(batch_sizes)), 2) if batch_sizes else None,
           "max_batch_frames": int(max(batch_sizes)) if batch_sizes else None,
           "gpu_busy_frac": round((st["busy_ms"] - st0["busy_ms"]) / (wall * 1e3), 3),
           "batches": st["batches"] - st0["batches"], "retries": st["retries"],
           "peak_inflight_GB": round(st["peak_inflight_bytes"] / 2 ** 30, 3), "parity": parity}
    if cpu:
        fpr = cpu["fragments_per_request"]
        out["fragments_per_request_sample"] = round(fpr, 1)
        out["est_Mfrag_s"] = round(fpr * len(lat) / wall / 1e6, 2)
        out["cpu_baseline"] = cpu["baseline"]
    return out


# ---------------------------------------------------------------------------
# profile summaries (rocprofv3 csv output -> the committed figures)
# ---------------------------------------------------------------------------

def stage_of_kernel(name):
    """The timing stage a kernel's dispatches belong to."""
    import re
    base = re.sub(r"^void ", "", name).split("(")[0].replace("hana::", "").strip()
    if base.startswith("k_raster_q"):
        return "raster_g0"
    m = re.match(r"k_raster<\d+, (\d)(, \d+)?>", base)
    if m:
        return "raster_g" + m.group(1)
    for k, st in (("k_vertex", "vertex"), ("k_setup", "setup"), ("k_clip", "clip"), ("k_scan", "scan"),
                  ("k_bin", "bin"), ("k_fill32", "fill"), ("k_present", "present")):
        if base.startswith(k):
            return st
    return base


def _csv_rows(d, suffix):
    import csv
    import glob
    rows = []
    for f in glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True):
        with open(f) as fh:
            rows.extend(csv.DictReader(fh))
    return rows


def _counters(d):
    """{stage: {counter: total}} and {stage: dispatches} over a --pmc run."""
    tot, disp, seen = {}, {}, set()
    for r in _csv_rows(d, "counter_collection.csv"):
        st = stage_of_kernel(r["Kernel_Name"])
        tot.setdefault(st, {})
        c = r["Counter_Name"]
        tot[st][c] = tot[st].get(c, 0.0) + float(r["Counter_Value"])
        key = (r["Dispatch_Id"], st)
        if key not in seen:
            seen.add(key)
            disp[st] = disp.get(st, 0) + 1
    return tot, disp


def summarize_counters(workload, d, out_json):
    """Per-launch HBM bytes (FETCH_SIZE, WRITE_SIZE scaled by a plain copy's
    readings) and SQ counters, by stage; writes OUT_JSON, prints markdown."""
    copy_bytes = 64 * 1024 * 1024 * 4

    def cal(sub, counter):
        # the largest per-dispatch reading of the run is the 256 MiB copy's
        t, n = _counters(os.path.join(d, sub))
        vals = [v[counter] / n[st] for st, v in t.items() if counter in v]
        return copy_bytes / (max(vals) * 1024) if vals else 1.0
    fs, ws_ = cal("cal_fetch", "FETCH_SIZE"), cal("cal_write", "WRITE_SIZE")
    per = {}
    counts = {}
    import glob
    for sub in sorted(glob.glob(os.path.join(d, "pmc_*"))):
        tot, disp = _counters(sub)
        for st, cs in tot.items():
            n = disp[st]
            counts[st] = max(counts.get(st, 0), n)
            e = per.setdefault(st, {})
            for c, v in cs.items():
                e[c] = v / n
    out = {"workload": workload, "fetch_scale": fs, "write_scale": ws_, "per_launch": {}}
    for st, e in per.items():
        o = {}
        if "FETCH_SIZE" in e:
            o["read_bytes"] = int(e["FETCH_SIZE"] * 1024 * fs)
        if "WRITE_SIZE" in e:
            o["write_bytes"] = int(e["WRITE_SIZE"] * 1024 * ws_)
        for c, k in (("SQ_INSTS_VALU", "valu_instr"), ("SQ_INSTS_LDS", "lds_instr"),
                     ("SQ_LDS_BANK_CONFLICT", "lds_bank_conflict_cycles"), ("SQ_WAVES", "waves"),
                     ("SQ_BUSY_CYCLES", "busy_cycles"), ("SQ_INSTS_VALU_TRANS_F32", "valu_trans_f32"),
                     ("SQ_ACTIVE_INST_VALU", "active_inst_valu"), ("SQ_THREAD_CYCLES_VALU", "thread_cycles_valu"),
                     ("SQ_INSTS_VMEM_RD", "vmem_rd"), ("SQ_INSTS_VMEM_WR", "vmem_wr"), ("SQ_WAVE_CYCLES", "wave_cycles"),
                     ("SQ_INST_CYCLES_VMEM", "inst_cycles_vmem"), ("SQ_WAIT_INST_ANY", "wait_inst_any"),
                     ("SQ_ACTIVE_INST_ANY", "active_inst_any"), ("TCC_HIT_sum", "l2_hit"), ("TCC_MISS_sum", "l2_miss")):
            if c in e:
                o[k] = int(e[c])
        out["per_launch"][st] = o
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1)
    print(f"FETCH_SIZE scale {fs:.3f}, WRITE_SIZE scale {ws_:.3f} (a 256 MiB copy)\n")
    print("| stage | dispatches | HBM read MB | HBM write MB | VALU instr | LDS instr | LDS conflict cycles | "
          "VALU thread util | trans F32 |")
    print("|---|---:|---:|---:|---:|---:|---:|---:|---:|")
    for st, o in sorted(out["per_launch"].items()):
        util = (o["thread_cycles_valu"] / (64.0 * o["active_inst_valu"])
                if o.get("active_inst_valu") and "thread_cycles_valu" in o else None)
        print(f"| {st} | {counts.get(st, 0)} | {o.get('read_bytes', 0) / 1e6:.2f} | {o.get('write_bytes', 0) / 1e6:.2f} | "
              f"{o.get('valu_instr', '')} | {o.get('lds_instr', '')} | {o.get('lds_bank_conflict_cycles', '')} | "
              f"{'' if util is None else f'{util:.3f}'} | {o.get('valu_trans_f32', '')} |")


def _kernel_name(raw):
    return raw.replace("void ", "").split("(")[0].replace("hana::", "")


def summarize_trace(d, last=0):
    """The rocprofv3 --stats summary under d as a markdown table; with last > 0,
    also each kernel's mean over its last `last` dispatches in start order."""
    rows = _csv_rows(d, "kernel_stats.csv")
    total = sum(float(r["TotalDurationNs"]) for r in rows) or 1.0
    print("| kernel | calls | average µs | min µs | max µs | share of kernel time |")
    print("|---|---:|---:|---:|---:|---:|")
    for r in sorted(rows, key=lambda r: -float(r["TotalDurationNs"])):
        print(f"| `{_kernel_name(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['MinNs']) / 1e3:.2f} | {float(r['MaxNs']) / 1e3:.2f} | "
              f"{100 * float(r['TotalDurationNs']) / total:.1f} % |")
    if last <= 0:
        return
    by = {}
    for r in sorted(_csv_rows(d, "kernel_trace.csv"), key=lambda r: int(r["Start_Timestamp"])):
        by.setdefault(_kernel_name(r["Kernel_Name"]), []).append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3)
    print(f"\n## The last {last} dispatches of each kernel (kernel_trace.csv)\n")
    print("| kernel | dispatches | average µs |")
    print("|---|---:|---:|")
    for name, ds in sorted(by.items(), key=lambda kv: -sum(kv[1][-last:])):
        if name.startswith("__amd") or len(ds) < last:
            continue
        t = ds[-last:]
        print(f"| `{name}` | {len(t)} | {sum(t) / len(t):.2f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed runs of the batch (the headline)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--sections", default="c1,c2,c3,c4,c5")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tuning", default="", help="batch tuning tile,wgs_per_cu,parts[,waves] instead of the autotuner")
    ap.add_argument("--c3-batch-only", action="store_true", help="the batch alone (profiling runs)")
    ap.add_argument("--c3-sweep", action="store_true", help="time every tuning of the autotuner's space on the batch")
    ap.add_argument("--skip-dropin", action="store_true")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--frame-steps", type=int, default=20)
    ap.add_argument("--c1-steps", type=int, default=200)
    ap.add_argument("--micro-steps", type=int, default=5)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-lanes", type=int, default=3)
    ap.add_argument("--serve-max-batch", type=int, default=64)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--serve-cpu-sample", type=int, default=40)
    ap.add_argument("--serve-cpu-replay", type=int, default=2000)
    ap.add_argument("--c5-one-core", action="store_true",
                    help="time the micro-triangle frame's CPU leg on one core only (not on every core as well)")
    ap.add_argument("--no-fp32", action="store_true", help="skip the fp32 shadow-map variants of c2 and c5")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (scales rocprof FETCH/WRITE_SIZE)")
    ap.add_argument("--summarize-counters", nargs=3, metavar=("WORKLOAD", "DIR", "OUT_JSON"),
                    help="summarise rocprofv3 --pmc runs under DIR (subfolders pmc_* and the copy calibration "
                         "cal_fetch, cal_write) into per-launch figures by stage; prints a markdown table")
    ap.add_argument("--summarize-trace", nargs="+", metavar="DIR [LAST]",
                    help="print the rocprofv3 --stats kernel summary under DIR as a markdown table, and with LAST "
                         "each kernel's mean over its last LAST dispatches")
    args = ap.parse_args()
    if args.summarize_counters:
        summarize_counters(*args.summarize_counters)
        return
    if args.summarize_trace:
        summarize_trace(args.summarize_trace[0], int(args.summarize_trace[1]) if len(args.summarize_trace) > 1 else 0)
        return
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
        print(f"bench.py measures one GPU only (asked for {args.gpus}); run it with --gpus 1. Frames, batch "
              "parts and server requests are independent, so several GPUs would each take their own.", file=sys.stderr)
        sys.exit(2)
    sections = set(args.sections.split(","))

    from hana_amd import scenegen as sg
    t0 = time.perf_counter()
    wl = {}
    if "c1" in sections:
        wl["c1"] = sg.config1_scene()
    if "c2" in sections:
        wl["c2"] = sg.config2_scene()
    if "c3" in sections:
        wl["c3"] = sg.config3_scenes()
    if "c4" in sections:
        wl["c4"] = sg.config4_workload(rate_per_s=args.serve_rate, seconds=args.serve_seconds)
    if "c5" in sections:
        wl["c5"] = sg.config5_scene()
    log(f"workloads generated in {time.perf_counter() - t0:.1f} s")

    # ---- CPU legs: before any HIP call ----
    cpu = {k: None for k in wl}
    kind = cpu_kind()
    if not args.no_cpu:
        W = cpu_workers()
        if "c1" in wl:
            s1 = cpu_single([wl["c1"]], kind, keep=True, min_frames=100, label="c1")
            a1 = cpu_all("c1", [wl["c1"]], kind, repeat=20 * W, label="c1")
            cpu["c1"] = {"fragments_per_frame": s1["fragments"] // s1["frames"], "images": s1["images"][:1],
                         "baseline": cpu_baseline_block(s1, a1, kind, None, f"{s1['frames']} frames on one core",
                                                        f"{a1['frames']} frames over {a1['workers']} processes")}
        if "c2" in wl:
            s2 = cpu_single([wl["c2"]], kind, keep=True, min_frames=3, label="c2")
            a2 = cpu_all("c2", [wl["c2"]], kind, repeat=W, label="c2")
            cpu["c2"] = {"fragments_per_frame": s2["fragments"] // s2["frames"], "images": s2["images"][:1],
                         "baseline": cpu_baseline_block(s2, a2, kind, None, f"{s2['frames']} frames on one core",
                                                        f"{a2['frames']} frames, one per process")}
        if "c3" in wl:
            s3 = cpu_single(wl["c3"], kind, keep=True, label="c3")
            a3 = cpu_all("c3", wl["c3"], kind, label="c3")
            cpu["c3"] = {"fragments_per_batch": s3["fragments"], "images": s3["images"],
                         "baseline": cpu_baseline_block(s3, a3, kind, None, f"all {s3['frames']} scenes on one core",
                                                        f"all {a3['frames']} scenes over {a3['workers']} processes")}
        if "c4" in wl:
            w4 = wl["c4"]
            sample = [sg.request_scene(w4, r) for r in w4.requests[:args.serve_cpu_sample]]
            s4 = cpu_single(sample, kind, keep=True, label="c4 sample")
            replay = [sg.request_scene(w4, r) for r in w4.requests[:args.serve_cpu_replay]]
            a4 = cpu_all("c4", replay, kind, label="c4 replay")
            b4 = cpu_baseline_block(s4, a4, kind, None, f"first {len(sample)} requests on one core",
                                    f"offline replay of the first {len(replay)} requests over {a4['workers']} "
                                    "processes (throughput only)")
            b4["requests_per_s"] = round(a4["frames"] / a4["seconds"], 2)
            cpu["c4"] = {"fragments_per_request": s4["fragments"] / s4["frames"], "images": s4["images"],
                         "baseline": b4}
            del replay
        if "c5" in wl:
            s5 = cpu_single([wl["c5"]], kind, keep=True, label="c5")
            a5 = cpu_all("c5", [wl["c5"]], kind, repeat=W, label="c5") if not args.c5_one_core else None
            cpu["c5"] = {"fragments_per_frame": s5["fragments"], "images": s5["images"][:1],
                         "baseline": cpu_baseline_block(s5, a5 or dict(s5, workers=1), kind, None,
                                                        "one whole frame on one core",
                                                        f"{a5['frames']} frames, one per process" if a5 else
                                                        "one frame on one core (all-core run off)")}
        # the fp32 shadow-map extension of the large and micro-triangle frames:
        # one frame with the oracle extended the same way (no reference mode)
        import dataclasses
        from pyoracle import CPURenderer
        for key in ("c2", "c5"):
            if key in wl and not args.no_fp32:
                sc = dataclasses.replace(wl[key], shadow_fp32=True, name=wl[key].name + " fp32 shadow")
                wl[key + "_fp32"] = sc
                t0 = time.perf_counter()
                o = CPURenderer("oracle").render_passes(sc)
                log(f"cpu {key}_fp32: 1 frame with the oracle in {time.perf_counter() - t0:.1f} s")
                cpu[key + "_fp32"] = {"fragments_per_frame": _frags(o), "baseline": None,
                                      "images": [(o[0].color.copy(), o[0].depth.copy(), _frags(o))]}
        _WORK.clear()

    # ---- GPU ----
    from hana_amd.gpu import Context
    ctx = Context(0)
    res = {}
    if "c1" in wl:
        res["c1"] = per_frame_section(ctx, wl["c1"], "c1", args.c1_steps, args.warmup, cpu["c1"],
                                      cpu["c1"]["images"] if cpu["c1"] else None)
        log("c1:", json.dumps(res["c1"]))
    if "c2" in wl:
        res["c2"] = per_frame_section(ctx, wl["c2"], "c2", args.frame_steps, args.warmup, cpu["c2"],
                                      cpu["c2"]["images"] if cpu["c2"] else None)
        log("c2:", json.dumps(res["c2"]))
    if "c5" in wl:
        res["c5"] = per_frame_section(ctx, wl["c5"], "c5", args.micro_steps, 2, cpu["c5"],
                                      cpu["c5"]["images"] if cpu["c5"] else None)
        log("c5:", json.dumps(res["c5"]))
    for key, steps in (("c2_fp32", args.frame_steps), ("c5_fp32", args.micro_steps)):
        if key in wl:
            res[key] = per_frame_section(ctx, wl[key], key, steps, 2, cpu[key], cpu[key]["images"])
            res[key]["mode"] = "fp32 shadow-map extension (depth-only shadow pass); parity against the oracle"
            log(f"{key}:", json.dumps(res[key]))
    if "c3" in wl:
        res["c3"] = bench_c3(ctx, args, wl["c3"], cpu["c3"])
        log("c3:", json.dumps(res["c3"]))
    if "c4" in wl:
        res["c4"] = bench_c4(ctx, args, wl["c4"], cpu["c4"])
        log("c4:", json.dumps(res["c4"]))
    ctx.close()

    head = res.get("c3")
    out = {"metric": METRIC,
           "value": head.get("Mfrag_s") if head else None,
           "unit": "Mfrag/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": head["ms_per_batch"] if head else None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded generators; the reference publishes no number for this metric)",
           "config": {"workload": "batched: 256 independent scenes, each 1-8 seeded displaced spheres/boxes "
                                  "totalling 1k-200k tris (log-uniform), 128^2-1024^2 px (log-uniform, aspect "
                                  "1/2-2), Blinn/NormalMap x shadow on/off; one captured hipGraph per step, "
                                  "tuning by the library's autotuner",
                      "scenes": len(wl["c3"]) if "c3" in wl else None,
                      "triangles": sum(s.mesh.nfaces for s in wl["c3"]) if "c3" in wl else None,
                      "pixels": sum(s.width * s.height for s in wl["c3"]) if "c3" in wl else None},
           "frames_per_s": head["frames_per_s"] if head else None,
           "roofline": head["roofline"] if head else None,
           "cpu_baseline": head.get("cpu_baseline") if head else None,
           "per_frame": {"resident_per_call_Mfrag_s": head.get("per_call_Mfrag_s"),
                         "dropin_Mfrag_s": head.get("dropin_Mfrag_s"),
                         "batched_over_dropin": head.get("speedup_vs_dropin"),
                         "batched_over_per_call": head.get("speedup_vs_per_call")} if head else None,
           "parity": head.get("parity") if head else None,
           "p50_ms": res["c4"]["p50_ms"] if "c4" in res else None,
           "p99_ms": res["c4"]["p99_ms"] if "c4" in res else None}
    out.update(res)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
