#!/usr/bin/env python3
"""Benchmark of the chunk pipeline (noise -> fill -> cull/emit) on one MI355X.

Workload (the metric's configuration): a uniform batch of 4096 chunks of
32^3 voxels on a 64 x 64 grid at y = 0; 3-D fBm Perlin density (seed 7,
4 octaves, lacunarity 2, gain 0.5, base frequency 1/64), solid where
density - y/128 > 0; every in-grid neighbour present.  Outputs per step:
u8 block ids, packed u32 occupancy rows, boundary planes, and the mesh as
4 vertices (f32 xyz+uv) + 6 u32 indices per visible face.  A step is one
replay of the captured graph fill -> count -> scan -> emit over the batch;
inputs (the request table) are resident in HBM before timing starts.

Prints ONE JSON line on stdout.  Progress goes to stderr.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-plain]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "wgpu-voxel-engine_amd"))

import numpy as np  # noqa: E402

import voxel_amd as V  # noqa: E402

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3    # MI355X published fp32 peak
# VALU issue rate behind that figure: 1024 SIMDs x 32 lanes x 2.4 GHz (one FMA = 2 flops per lane)
VALU_LANE_OPS_T = 78.64
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "traffic_uniform4096x32.txt")
GRID = 64                   # 64 x 64 chunks
EDGE = 32
# FP32 operations of one restated 3-D Improved Noise evaluation inside fBm
# (oracle/voxel_oracle.c noise3f + vxo_fbm3): 3 scale, 3 floor, 3 frac,
# 3 fades x 7, 3 "-1", 8 gradient adds, 7 lerps x 3, 2 accumulate = 64.
K_EVAL = 64


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def workload(ctx):
    t = V.Terrain(kind=V.TERRAIN_DENSITY, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0,
                  gain=0.5, y_scale=1 / 128)
    coords = np.array([(x, 0, z) for x in range(GRID) for z in range(GRID)], np.int32)
    return t, coords, V.grid_neighbours(GRID, GRID)


def cpu_baseline(target_s: float = 12.0):
    """The oracle's multithreaded driver (one chunk per task, all host
    cores) on a bounded sub-grid of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    O.build()
    threads = len(os.sched_getaffinity(0))
    t = O.terrain(kind=3, seed=7, octaves=4, frequency=1 / 64, lacunarity=2.0, gain=0.5,
                  y_scale=1 / 128)
    # warm-up and sizing on an 8x8 sub-grid
    side = 8
    coords = np.array([(x, 0, z) for x in range(side) for z in range(side)], np.int32)
    t0 = time.perf_counter()
    O.batch_run(t, coords, V.grid_neighbours(side, side), EDGE, threads)
    per_chunk = (time.perf_counter() - t0) / len(coords)
    reps = 5
    n_target = max(64, int(target_s / reps / max(per_chunk, 1e-6)))
    side = int(min(GRID, max(8, int(np.sqrt(n_target)))))
    coords = np.array([(x, 0, z) for x in range(side) for z in range(side)], np.int32)
    nb = V.grid_neighbours(side, side)
    rates = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.batch_run(t, coords, nb, EDGE, threads)
        rates.append(len(coords) / (time.perf_counter() - t0))
    cpu_model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(statistics.median(rates), 1), "unit": "chunks/s", "cores": threads,
            "kind": "port",
            "sample": f"{side}x{side} sub-grid ({side * side} chunks) of the same 32^3 workload, "
                      f"fill (AVX2 noise) + scalar cull/emit to vertices+indices, median of "
                      f"{reps} runs on {threads} threads ({cpu_model})"}


def frame_leg(ctx, faces_per_chunk: float, frames: int = 20):
    """Full frame (config 5): 65,536 resident 32^3 chunks (256 x 256, fBm seed
    5) plus, per frame, 512 new chunks (the columns x = 256, 257 beside the
    grid; their -x neighbours resident, +-z in the same batch) meshed and the
    65,536 resident AABBs culled against a moving seeded camera (fov 70 deg,
    near 0.1, far 1024) in one launch."""
    import camera as T  # the package's camera matrices (look_at / perspective)

    t5 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=5, octaves=4, frequency=1 / 64, lacunarity=2.0,
                   gain=0.5, y_scale=1 / 128)
    tid = ctx.terrain(t5)
    side = 256
    rc = np.array([(x, 0, z) for x in range(side) for z in range(side)], np.int32)
    cap = int(len(rc) * faces_per_chunk * 1.15) + 4096
    R = V.Batch(ctx, rc, EDGE, tid, V.grid_neighbours(side, side), face_capacity=cap)
    ctx.synchronize()
    t0 = time.perf_counter()
    R.run()
    resident_faces = R.wait()
    gen_s = time.perf_counter() - t0
    ext = [R.chunk((side - 1) * side + z) for z in range(side)]
    nc = [(side + dx, 0, z) for dx in range(2) for z in range(side)]
    nb = []
    for dx in range(2):
        for z in range(side):
            i = dx * side + z
            nb.append([(-2 - z) if dx == 0 else (i - side), (i + side) if dx == 0 else -1,
                       i - 1 if z > 0 else -1, i + 1 if z < side - 1 else -1])
    B = V.Batch(ctx, nc, EDGE, tid, nb, formats=V.FMT_VERTICES, ext=ext)
    cbuf = V.DeviceBuffer.of(ctx, rc)
    mbuf = V.DeviceBuffer(ctx, len(rc))
    rng = np.random.default_rng(5)
    eye = np.array([4096.0, 60.0, 4096.0])
    heading = 0.0

    def camera():
        nonlocal eye, heading
        heading += rng.normal(0, 0.05)
        d = np.array([np.cos(heading), -0.2, np.sin(heading)])
        eye = eye + 2.0 * d
        m = T.colmajor(T.perspective(np.radians(70), 16 / 9, 0.1, 1024) @ T.look_at(eye, eye + d))
        return V.Frustum(m)

    B.set_cull(camera(), cbuf, len(rc), EDGE, mbuf)
    B.run()
    new_faces = B.wait()
    frame_ms = {}
    for form, mode in (("staged", V.BATCH_STAGED), ("fused", V.BATCH_FUSED),
                       ("pipelined", V.BATCH_PIPELINED)):
        # fused: new chunks and the cull in one persistent launch (config 5's form)
        B.set_mode(mode)
        B.set_cull(camera(), cbuf, len(rc), EDGE, mbuf)
        B.run()
        assert B.wait() == new_faces
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(frames):
            B.set_cull(camera(), cbuf, len(rc), EDGE, mbuf)
            B.run()
        ctx.synchronize()
        frame_ms[form] = (time.perf_counter() - t0) * 1e3 / frames
        assert B.wait() == new_faces
    # the autotuner's pick for this 512-chunk batch (forms, and for the
    # staged form the fill and emit workgroups per chunk and per item)
    B.set_mode(V.BATCH_AUTO)
    tune = B.autotune(3)
    B.set_cull(camera(), cbuf, len(rc), EDGE, mbuf)
    B.run()
    assert B.wait() == new_faces
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        B.set_cull(camera(), cbuf, len(rc), EDGE, mbuf)
        B.run()
    ctx.synchronize()
    frame_ms["autotuned"] = (time.perf_counter() - t0) * 1e3 / frames
    assert B.wait() == new_faces
    visible = int(mbuf.download(np.uint8, len(rc)).sum())
    # the frame figure is the autotuned form's (what VX_BATCH_AUTO runs)
    form = tune["mode"]
    el = frame_ms["autotuned"] * frames / 1e3
    # World::render's loop over the resident meshes as an indirect draw list
    f = camera()
    args, cnt = V.DeviceBuffer(ctx, len(rc) * V.DRAW_DTYPE.itemsize), V.DeviceBuffer(ctx, 8)
    draw = lambda: ctx.check(ctx.L.vx_batch_draw_list(ctx.h, R.h, C.byref(f.c), V.DRAW_EXPANDED,
                                                       args.ptr, cnt.ptr), "vx_batch_draw_list")
    draw()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(frames):
        draw()
    ctx.synchronize()
    draw_ms = (time.perf_counter() - t0) * 1e3 / frames
    draws = int(cnt.download(np.uint32, 2)[0])
    args.free()
    cnt.free()
    B.close()
    R.close()
    for c in ext:
        c.release()
    cbuf.free()
    mbuf.free()
    return {"resident_chunks": len(rc), "resident_faces": resident_faces,
            "resident_generation_ms": round(gen_s * 1e3, 2),
            "resident_chunks_per_s": round(len(rc) / gen_s, 1),
            "new_chunks_per_frame": len(nc), "new_faces_per_frame": new_faces,
            "frame_ms": round(el * 1e3 / frames, 4), "frames": frames,
            "visible_last_frame": visible, "form": form,
            "frame_ms_by_form": {k: round(v, 4) for k, v in frame_ms.items()},
            "autotune": tune,
            "draw_list_ms": round(draw_ms, 4), "draws": draws}


def world_leg(ticks: int = 130):
    """World + ChunkGenerator stepped tick by tick (vx_world, jobs on the GPU):
    a scripted walk (move, stand while the loads drain and rebuilds run, turn,
    jump far, stand), render distance 5, results receivable one tick after
    submission, fBm seed 5."""
    import camera as T  # the package's camera matrices

    with V.Context(0) as c:
        t5 = V.Terrain(kind=V.TERRAIN_DENSITY, seed=5, octaves=4, frequency=1 / 64,
                       lacunarity=2.0, gain=0.5, y_scale=1 / 128)
        start = np.array([16.0, 40.0, 16.0])
        w = V.World(c, c.terrain(t5), start, render_distance=5, result_delay=1)
        pos, look = start.copy(), (7.0, -8.0, 3.5)
        t0 = time.perf_counter()
        for k in range(ticks):
            step = ((7.0, 0.0, 3.0) if k < 25 else (0.0, 0.0, 0.0) if k < 90 else
                    (-4.0, 0.0, 6.0) if k < 100 else (400.0, 0.0, 0.0) if k == 100 else
                    (0.0, 0.0, 0.0))
            if step[0] or step[2]:
                look = (step[0], -8.0, step[2] + 0.5)
            pos = pos + np.array(step)
            m = T.colmajor(T.perspective(np.radians(70), 16 / 9, 0.1, 1024) @
                           T.look_at(pos, pos + np.array(look)))
            w.tick(pos.astype(np.float32), V.Frustum(m))
        c.synchronize()
        el = time.perf_counter() - t0
        st = w.stats()
        w.close()
    return {"ticks": ticks, "ms_per_tick": round(el * 1e3 / ticks, 4),
            "jobs_loaded": st["jobs_loaded"], "jobs_rebuilt": st["jobs_rebuilt"],
            "unloaded": st["unloaded"], "resident_chunks": st["chunks"]}


def config3_batch(ctx):
    """Config 3: 2048 chunks, edges uniform over 16/32/64, classes 30% air /
    10% solid / 60% 4-octave heightmap surface with height = edge, one seeded
    stream (seed 11), no neighbours."""
    rng = np.random.default_rng(11)
    n3 = 2048
    edges = rng.choice([16, 32, 64], size=n3)
    cls = rng.choice(3, size=n3, p=[0.3, 0.1, 0.6])
    tids = {}
    for e, k in zip(edges, cls):
        key = (int(k), int(e))
        if key not in tids:
            tv = (V.Terrain(kind=V.TERRAIN_AIR) if k == 0 else V.Terrain(kind=V.TERRAIN_SOLID)
                  if k == 1 else V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=11, octaves=4,
                                           frequency=1 / 32, height_scale=float(e),
                                           height_max=int(e) - 1))
            tids[key] = ctx.terrain(tv)
    coords3 = [(i % 64, 0, i // 64) for i in range(n3)]
    b3 = V.Batch(ctx, coords3, edges, [tids[(int(k), int(e))] for e, k in zip(edges, cls)],
                 np.full((n3, 4), -1, np.int32), formats=V.FMT_VERTICES)
    b3.edges = edges
    return b3, n3


def configs_leg(ctx, reps: int = 20):
    """The other batch configurations, each one captured graph replayed
    `reps` times after a sizing run: config 1 (4 x 4 chunks of 16^3, 2-D
    heightmap seed 42, 1 octave, f 1/32, heights 0..15, in-grid neighbours)
    and config 3 (2048 chunks, edges uniform over 16/32/64, classes 30% air /
    10% solid / 60% 4-octave heightmap surface with height = edge, one
    seeded stream, seed 11, no neighbours)."""
    out = {}
    c1 = V.Terrain(kind=V.TERRAIN_HEIGHTMAP, seed=42, octaves=1, frequency=1 / 32,
                   height_scale=16.0, height_max=15)
    coords1 = [(x, 0, z) for x in range(4) for z in range(4)]
    b1 = V.Batch(ctx, coords1, 16, ctx.terrain(c1), V.grid_neighbours(4, 4),
                 formats=V.FMT_VERTICES)
    b3, n3 = config3_batch(ctx)
    for name, b, n in (("config1", b1, len(coords1)), ("config3", b3, n3)):
        b.run()
        faces = b.wait()  # sizes the arena (config 3 outgrows the default once)
        # the autotuner over this batch's forms and launch and tile shapes;
        # the batch then runs its choice (VX_BATCH_AUTO)
        tune = b.autotune(3)
        b.run()
        assert b.wait() == faces
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run()
        ctx.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[name] = {"chunks": n, "faces": faces, "ms_per_batch": round(ms, 4),
                     "chunks_per_s": round(n / (ms * 1e-3), 1), "form": tune["mode"],
                     "autotune": tune}
        if name == "config3":
            # stage times (HIP events); the emit's algorithmic bytes: 104 B per
            # face written + every chunk's occupancy rows read (no neighbours)
            p = b.profile(reps)
            rows = int(sum(int(e) * int(e) * ((int(e) + 31) // 32) * 4 for e in b.edges))
            emit_bytes = faces * 104 + rows
            if tune["mode"] == "staged":
                out[name]["stages_ms"] = {k: round(float(v), 4) for k, v in
                                          zip(("fill", "count", "scan", "emit"), p[:4])}
                out[name]["emit_algorithmic_bytes"] = emit_bytes
                out[name]["emit_GB/s"] = round(emit_bytes / (float(p[3]) * 1e-3) / 1e9, 1)
            else:
                out[name]["kernel_ms"] = round(float(p[0]), 4)
        b.close()
    return out


def measured_counter(kernel: str, faces: int, key: str = "bytes"):
    """A per-launch figure of `kernel` from the committed rocprofv3 summary
    (key "bytes": HBM bytes from FETCH_SIZE / WRITE_SIZE; "valu_insts":
    SQ_INSTS_VALU), when it was measured on this same workload (same face
    count); None otherwise."""
    try:
        lines = open(TRAFFIC_FILE).read().splitlines()
    except OSError:
        return None
    if not any(f"{faces} faces" in ln for ln in lines if ln.startswith("#")):
        return None
    for ln in lines:
        parts = ln.split()
        if len(parts) >= 2 and parts[0] == "kernel" and parts[1] == kernel:
            v = dict(p.split("=") for p in parts[2:]).get(key)
            return int(v) if v is not None else None
    return None


def measured_traffic(kernel: str, faces: int):
    return measured_counter(kernel, faces, "bytes")


def two_stream_leg(t, coords, nb, b, ctx, faces: int, steps: int):
    """The same 4096-chunk batch twice, each on its own context (its own HIP
    stream), replayed alternately: one batch's emit (HBM stores) overlaps the
    other's fill (FP32 issue).  Reported beside `value`, which times one
    batch at a time on one stream."""
    ctx2 = V.Context(0)
    try:
        b2 = V.Batch(ctx2, coords, EDGE, ctx2.terrain(t), nb, formats=V.FMT_VERTICES)
        b2.run()
        assert b2.wait() == faces
        b2.autotune(3)
        for _ in range(2):
            b.run()
            b2.run()
        ctx.synchronize()
        ctx2.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
            b2.run()
        ctx.synchronize()
        ctx2.synchronize()
        el = time.perf_counter() - t0
        assert b.wait() == faces and b2.wait() == faces
        b2.close()
    finally:
        ctx2.close()
    batches = 2 * steps
    return {"streams": 2, "batches": batches, "ms_per_batch": round(el * 1e3 / batches, 4),
            "chunks_per_s": round(batches * len(coords) / el, 1)}


def instances_leg(ctx, t, coords, nb, steps: int):
    """The same 4096-chunk batch emitting InstanceRaw records only (68 B per
    face, src/renderer/instance.rs:1-6), the format the Rust renderer
    consumes: graph replays timed like `value`, the emit kernel timed with
    HIP events."""
    b = V.Batch(ctx, coords, EDGE, ctx.terrain(t), nb, formats=V.FMT_INSTANCES)
    b.run()
    faces = b.wait()
    b.set_mode(V.BATCH_STAGED)
    b.run()
    b.wait()
    for _ in range(3):
        b.run()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.run()
    ctx.synchronize()
    el = time.perf_counter() - t0
    assert b.wait() == faces
    ms = b.profile(5)
    n = len(coords)
    N3, N2 = EDGE ** 3, EDGE ** 2
    emit_bytes = n * (N3 // 8 + 4 * N2 // 8 + 4 + 8) + 68 * faces
    b.close()
    return {"format": "InstanceRaw (68 B/face)", "faces": faces,
            "ms_per_step": round(el * 1e3 / steps, 4), "chunks_per_s": round(n * steps / el, 1),
            "emit_ms": round(float(ms[3]), 4),
            "emit_GB/s": round(emit_bytes / (ms[3] * 1e-3) / 1e9, 1),
            "emit_algorithmic_bytes": emit_bytes}


def serving_leg(rates=(2000, 5000, 10000, 20000), duration=1.5,
                pressure=((20000, 0.25), (20000, 0.125))):
    """Streaming (config 4): camera random walk at 20 blocks/s (seed 3),
    render radius 16 chunks of 32^3, Poisson arrivals, sizes 16/32/64
    uniform, 8 GB HBM admission budget with LRU eviction.  In 1.5 s the
    working set (about 1 GB) stays under 8 GB, so the `pressure` runs repeat
    the top rate under budgets of 0.25 and 0.125 GB, where LRU admission
    evicts throughout."""
    out = []
    runs = [(rate, 8.0) for rate in rates] + list(pressure)
    for rate, budget_gb in runs:
        with V.Context(0, hbm_budget=int(budget_gb * 2**30)) as c:
            t = V.Terrain(kind=V.TERRAIN_DENSITY, seed=7, octaves=4, frequency=1 / 64,
                          lacunarity=2.0, gain=0.5, y_scale=1 / 128)
            tid = c.terrain(t)
            r = V.simulate(c, seed=3, rate=rate, duration=duration, speed=20.0, radius=16,
                           classes=((16, 1.0, tid), (32, 1.0, tid), (64, 1.0, tid)))
            out.append({"rate_req_s": rate, "budget_GB": budget_gb,
                        "requests": r.requests, "completed": r.completed,
                        "cache_hits": r.cache_hits, "generated": r.generated,
                        "batches": r.batches, "evictions": r.evictions,
                        "throughput_req_s": round(r.throughput_rps, 1),
                        "generated_chunks_s": round(r.generated_per_s, 1),
                        "p50_ms": round(r.p50_ms, 4), "p99_ms": round(r.p99_ms, 4),
                        "max_ms": round(r.max_ms, 4),
                        "miss_p50_ms": round(r.miss_p50_ms, 4),
                        "miss_p99_ms": round(r.miss_p99_ms, 4),
                        "peak_resident_GB": round(r.peak_resident_bytes / 2**30, 3)})
            log(f"[bench] serving {rate} req/s, {budget_gb} GB: p50 {r.p50_ms:.3f} ms, p99 "
                f"{r.p99_ms:.3f} ms (misses p99 {r.miss_p99_ms:.3f} ms), {r.generated} generated, "
                f"{r.evictions} evicted")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-chunk GPU path")
    ap.add_argument("--profile-reps", type=int, default=5)
    ap.add_argument("--no-frame", action="store_true", help="skip the full-frame leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the streaming leg")
    ap.add_argument("--no-world", action="store_true", help="skip the tick-by-tick world leg")
    ap.add_argument("--no-two", action="store_true",
                    help="skip the two-batches-in-flight leg (two streams)")
    ap.add_argument("--pmc-run", choices=["fused", "staged", "pipelined", "instances", "config3",
                                          "frame", "server"], default=None,
                    help="profiler run of one workload, no timing: the config-2 batch in one "
                         "form (vertices+indices, or InstanceRaw staged), the config-3 ragged "
                         "batch, the config-5 frame, or server flights at 5000 req/s; then a "
                         "256 MiB plain copy kernel (the byte-counter reference)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(chunks shard by spatial tile; see the design notes). Run with --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    ctx = V.Context(0)
    if args.pmc_run:
        t, coords, nb = workload(ctx)
        faces = None
        if args.pmc_run in ("fused", "staged", "pipelined", "instances"):
            fmt = V.FMT_INSTANCES if args.pmc_run == "instances" else V.FMT_VERTICES
            b = V.Batch(ctx, coords, EDGE, ctx.terrain(t), nb, formats=fmt)
            b.set_mode({"fused": V.BATCH_FUSED, "pipelined": V.BATCH_PIPELINED}.get(
                args.pmc_run, V.BATCH_STAGED))
            if args.pmc_run == "pipelined":
                b.set_grid(8 * 256)
        elif args.pmc_run == "config3":
            b, _ = config3_batch(ctx)
            b.set_mode(V.BATCH_STAGED)
        if args.pmc_run == "frame":
            faces = frame_leg(ctx, 2746.6, frames=2)["new_faces_per_frame"]
        elif args.pmc_run == "server":
            tid = ctx.terrain(t)
            r = V.simulate(ctx, seed=3, rate=5000, duration=0.3, speed=20.0, radius=16,
                           classes=((16, 1.0, tid), (32, 1.0, tid), (64, 1.0, tid)))
            faces = int(r.faces)
        else:
            b.run()
            faces = b.wait()  # sizes the arena
            b.run()
            b.wait()
        nbytes = 256 << 20
        src, dst = V.DeviceBuffer(ctx, nbytes), V.DeviceBuffer(ctx, nbytes)
        for _ in range(2):
            ctx.check(ctx.L.vx_buffer_copy(ctx.h, dst.ptr, src.ptr, nbytes), "vx_buffer_copy")
        ctx.synchronize()
        print(json.dumps({"pmc_run": args.pmc_run, "faces": faces, "copy_bytes": nbytes}))
        return
    t, coords, nb = workload(ctx)
    tid = ctx.terrain(t)
    n = len(coords)
    log(f"[bench] creating batch: {n} chunks of {EDGE}^3")
    b = V.Batch(ctx, coords, EDGE, tid, nb, formats=V.FMT_VERTICES)
    b.run()
    total_faces = b.wait()           # sizes the mesh arena (grows once if needed)
    tune = b.autotune(5)             # picks the form and launch shape for this batch
    mode = b.mode
    log(f"[bench] autotune: {tune}; faces per step: {total_faces}")

    def timed(steps, warmup):
        for _ in range(warmup):
            b.run()
        ctx.synchronize()
        # timed region: K graph replays bracketed by device synchronisation
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
        ctx.synchronize()
        el = time.perf_counter() - t0
        assert b.wait() == total_faces
        return el

    elapsed = timed(args.steps, args.warmup)
    ms_per_step = elapsed * 1e3 / args.steps
    value = n * args.steps / elapsed
    log(f"[bench] {ms_per_step:.3f} ms/step, {value:.0f} chunks/s")
    chosen_ms = b.profile(args.profile_reps)

    N3, N2 = EDGE ** 3, EDGE ** 2
    fill_bytes = n * (N3 + N3 // 8 + 4 * N2 // 8 + 4)          # ids + occupancy + planes + meta
    fill_flops = n * N3 * t.octaves * K_EVAL
    emit_bytes = n * (N3 // 8 + 4 * N2 // 8 + 4 + 8) + 104 * total_faces
    fused_bytes = fill_bytes + 104 * total_faces + n * 12       # one pass: writes only

    # every other form, for comparison (same outputs): timed like `value`,
    # kernels timed with HIP events
    names_of = {V.BATCH_STAGED: "staged", V.BATCH_FUSED: "fused", V.BATCH_PIPELINED: "pipelined",
                V.BATCH_SPLIT: "split"}
    form_ms = {mode: chosen_ms}
    other = []
    reps2 = max(5, args.steps // 2)
    for m in (V.BATCH_STAGED, V.BATCH_FUSED, V.BATCH_PIPELINED, V.BATCH_SPLIT):
        if m == mode:
            continue
        b.set_mode(m)
        # a persistent form at the autotuner's fastest grid and roles for it
        shapes = [(c["grid"], c.get("ratio", 0))
                  for c in sorted(tune.get("candidates", []), key=lambda c: c["ms"])
                  if c["mode"] == names_of[m]]
        b.set_grid(shapes[0][0] if shapes else 0)
        b.set_roles(shapes[0][1] if shapes else 0)
        b.run()
        assert b.wait() == total_faces
        oe = timed(reps2, 1)
        form_ms[m] = b.profile(args.profile_reps)
        other.append({"form": names_of[m], "value": round(n * reps2 / oe, 1), "unit": "chunks/s",
                      "ms_per_step": round(oe * 1e3 / reps2, 4),
                      "grid": shapes[0][0] if shapes else 0,
                      "ratio": shapes[0][1] if shapes else 0})
        b.set_grid(0)
        b.set_roles(0)
    b.set_mode(V.BATCH_AUTO)
    stage_ms = form_ms[V.BATCH_STAGED]
    names = ["fill", "count", "scan", "emit"]
    log("[bench] staged ms: " + ", ".join(f"{k}={v:.4f}" for k, v in zip(names, stage_ms[:4])))
    stages = {
        "fill": {"ms": round(float(stage_ms[0]), 4),
                 "GB/s": round(fill_bytes / (stage_ms[0] * 1e-3) / 1e9, 1),
                 "TFLOP/s": round(fill_flops / (stage_ms[0] * 1e-3) / 1e12, 2)},
        "count": {"ms": round(float(stage_ms[1]), 4)},
        "scan": {"ms": round(float(stage_ms[2]), 4)},
        "emit": {"ms": round(float(stage_ms[3]), 4),
                 "GB/s": round(emit_bytes / (stage_ms[3] * 1e-3) / 1e9, 1)},
        "algorithmic_bytes": {"fill": fill_bytes, "emit": emit_bytes, "fused": fused_bytes},
        "algorithmic_flops": fill_flops,
    }
    one_launch = {}
    for m, label in ((V.BATCH_FUSED, "fused_kernel"), (V.BATCH_PIPELINED, "pipe_kernel")):
        k_ms = float(form_ms[m][0])
        one_launch[m] = (k_ms, fill_flops / (k_ms * 1e-3) / 1e12, fused_bytes / (k_ms * 1e-3) / 1e9)
        stages[label] = {"ms": round(k_ms, 4), "TFLOP/s": round(one_launch[m][1], 2),
                         "GB/s": round(one_launch[m][2], 1)}
        log(f"[bench] {label} {k_ms:.4f} ms: {one_launch[m][1]:.1f} TFLOP/s, "
            f"{one_launch[m][2]:.0f} GB/s")
    if mode in one_launch:
        k_ms, f_tf, f_gbs = one_launch[mode]
        kname = ("fused_kernel<32> (noise+fill+cull+scan+emit)" if mode == V.BATCH_FUSED else
                 "pipe_kernel<32> (fill tasks and cull+scan+emit tasks, one launch)")
        kshort = "fused_kernel<32>" if mode == V.BATCH_FUSED else "pipe_kernel<32>"
        roof = {"kernel": kname, "bound": "valu",
                "achieved": round(f_tf, 2), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(f_tf / FP32_PEAK_TFLOPS, 4),
                "traffic": measured_traffic(kshort, total_faces),
                "ms": round(k_ms, 4), "hbm_GB/s": round(f_gbs, 1),
                "hbm_frac": round(f_gbs / HBM_PEAK_GBS, 4)}
    elif stage_ms[0] >= stage_ms[3]:
        ach = fill_flops / (stage_ms[0] * 1e-3) / 1e12
        roof = {"kernel": "fill_kernel<32> (noise+fill)", "bound": "valu", "achieved": round(ach, 2),
                "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(ach / FP32_PEAK_TFLOPS, 4),
                "traffic": measured_traffic("fill_kernel<32>", total_faces),
                "ms": round(float(stage_ms[0]), 4)}
        vi = measured_counter("fill_kernel<32>", total_faces, "valu_insts")
        if vi:  # executed VALU stream against the SIMDs' issue rate
            lane_ops = vi * 64 / (stage_ms[0] * 1e-3) / 1e12
            roof["valu_issue"] = {"insts_per_launch": vi, "lane_ops_T_per_s": round(lane_ops, 2),
                                  "peak": VALU_LANE_OPS_T, "frac": round(lane_ops / VALU_LANE_OPS_T, 4)}
    else:
        ach = emit_bytes / (stage_ms[3] * 1e-3) / 1e9
        roof = {"kernel": "emit_kernel<32> (cull+compact+emit)", "bound": "hbm",
                "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(ach / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic("emit_kernel<32>", total_faces),
                "ms": round(float(stage_ms[3]), 4)}
    if roof.get("traffic") is not None:
        roof["traffic_source"] = ("read from profiles/traffic_uniform4096x32.txt: rocprofv3 "
                                  "FETCH_SIZE + WRITE_SIZE of this kernel on this workload, scaled "
                                  "by a plain copy kernel's counters; not measured in this run")
    stages["fill"]["traffic"] = measured_traffic("fill_kernel<32>", total_faces)
    stages["emit"]["traffic"] = measured_traffic("emit_kernel<32>", total_faces)

    plain = None
    if not args.no_plain:
        b.run_per_chunk()
        ctx.synchronize()
        reps = 2
        t0 = time.perf_counter()
        for _ in range(reps):
            b.run_per_chunk()
        ctx.synchronize()
        pe = time.perf_counter() - t0
        assert b.wait() == total_faces
        plain = {"value": round(n * reps / pe, 1), "unit": "chunks/s",
                 "ms_per_step": round(pe * 1e3 / reps, 3),
                 "launches_per_step": 3 * n + 1}
        log(f"[bench] plain per-chunk path: {plain['value']:.0f} chunks/s")

    inst = instances_leg(ctx, t, coords, nb, args.steps)
    log(f"[bench] InstanceRaw output: {inst['ms_per_step']} ms/step, emit {inst['emit_ms']} ms "
        f"({inst['emit_GB/s']} GB/s)")

    two = None
    if not args.no_two:
        two = two_stream_leg(t, coords, nb, b, ctx, total_faces, args.steps)
        log(f"[bench] two batches in flight: {two['ms_per_batch']} ms per batch, "
            f"{two['chunks_per_s']:.0f} chunks/s")

    frame = None
    if not args.no_frame:
        log("[bench] full frame: 65,536 resident + 512 new chunks per frame + cull...")
        frame = frame_leg(ctx, total_faces / n)
        log(f"[bench] frame: {frame['frame_ms']} ms ({frame['form']}), resident generation "
            f"{frame['resident_chunks_per_s']:.0f} chunks/s")

    configs = configs_leg(ctx)
    log("[bench] other configs: " + ", ".join(
        f"{k} {v['ms_per_batch']} ms ({v['chunks_per_s']:.0f} chunks/s)" for k, v in configs.items()))

    serving = None
    if not args.no_serve:
        serving = serving_leg()

    world = None
    if not args.no_world:
        world = world_leg()
        log(f"[bench] world: {world['ms_per_tick']} ms per tick over {world['ticks']} ticks")

    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (oracle, all host cores)...")
        cpu = cpu_baseline()
        log(f"[bench] CPU: {cpu['value']} chunks/s on {cpu['cores']} threads")

    out = {
        "metric": "chunks/sec (noise+mesh, 32\u00b3, bit-exact) vs CPU & plain GPU; "
                  "p50/p99 chunk latency",
        "value": round(value, 1),
        "unit": "chunks/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+u8",
        "data": "synthetic (seeded fBm terrain)",
        "config": {"workload": "uniform batch: 4096 chunks of 32^3 on a 64x64 grid at y=0, "
                               "3-D fBm density seed 7 (4 octaves, lac 2, gain 0.5, f0 1/64), "
                               "solid iff density - y/128 > 0, in-grid neighbours present; "
                               "outputs u8 ids + u32 occupancy + vertices/indices",
                   "chunks_per_step": n, "edge": EDGE, "faces_per_step": total_faces},
        "roofline": roof,
        "stages": stages,
        "form": names_of.get(mode, str(mode)),
        "autotune": tune,
        "cpu_baseline": cpu,
        "plain_gpu": plain,
        "two_batches_in_flight": two,
        "instances_format": inst,
        "other_forms_gpu": other,
        "full_frame": frame,
        "other_configs": configs,
        "serving": serving,
        "world": world,
    }
    if serving:
        # request-to-mesh latency of the streaming config (8 GB budget) at its
        # highest rate, and the same rate with LRU eviction running
        top = max((r for r in serving if r["budget_GB"] == 8.0), key=lambda r: r["rate_req_s"])
        out["latency_ms"] = {"rate_req_s": top["rate_req_s"], "p50": top["p50_ms"],
                             "p99": top["p99_ms"], "miss_p50": top["miss_p50_ms"],
                             "miss_p99": top["miss_p99_ms"]}
        tight = [r for r in serving if r["budget_GB"] < 8.0]
        if tight:
            tt = min(tight, key=lambda r: r["budget_GB"])
            out["latency_ms"]["under_pressure"] = {
                "budget_GB": tt["budget_GB"], "evictions": tt["evictions"], "p50": tt["p50_ms"],
                "p99": tt["p99_ms"], "miss_p99": tt["miss_p99_ms"]}
    if cpu:
        out["vs_cpu"] = round(value / cpu["value"], 1)
    if plain:
        out["vs_plain_gpu"] = round(value / plain["value"], 2)
    b.close()
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
