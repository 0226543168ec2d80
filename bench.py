#!/usr/bin/env python
"""Benchmark: batched vs plain per-draw execution of the 256-draw workload on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--tile 0|8|16|32] [--no-serve] ...

Workload: one 1920x1080 BGRA8 + D32 target, 256 draws of log-uniform 16..65536
triangles (about 1.85M), per-draw MVP, 4 mipmapped 512x512 textures sampled trilinear repeat;
even draws opaque (LESS), odd draws alpha-blended with depth writes off. Inputs (vertex, index,
texture data) are resident in HBM; the target stays in HBM.

A step is one vkQueueSubmit of the recorded render pass (all 256 draws), timed on the host
between synchronisations. value = covered fragments per second in the batched path. The plain
per-draw path (setup, bin, raster and a host wait per draw, as the reference's DrawTriangles) is
timed beside it, and the CPU baseline is the oracle's reference-structured threaded mode on a
bounded sample of the draws.
"""

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "batched/served Gfragments/s vs plain per-draw on 1 GPU; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0


def raster_bytes(st, width, height, color_bytes, texel_bytes):
    """Algorithmic HBM bytes of one raster launch (per frame): depth test reads of covered
    fragments, depth writes, colour writes (+ reads when blending) and texels of shaded fragments,
    and the one write of the final colour and depth of every pixel (the clear is fused)."""
    return (4 * st["pixels_written"] + 4 * st["shaded_dwrite"] + color_bytes * st["depth_passed"]
            + color_bytes * st["shaded_blend"] + texel_bytes * st["depth_passed"]
            + width * height * (color_bytes + 4))


def profiled_value(tile, key):
    """A per-launch figure of the raster kernel from the committed rocprofv3 summary of this
    workload and tile size (None when no such summary exists)."""
    path = os.path.join(ROOT, "profiles", "va_raster_%d__256draws_1080p.md" % tile)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(float(line.split(":")[1]))
    except OSError:
        pass
    return None


def profiled_traffic(tile):
    """HBM bytes per raster launch (FETCH_SIZE / WRITE_SIZE, calibrated), or None."""
    return profiled_value(tile, "traffic_bytes_per_launch")


# VALU issue: 256 CUs x 4 SIMDs, each issuing one wave64 VALU instruction per 4 cycles (16 lanes)
# at 2.4 GHz; this is the non-packed FP32 rate (78.6 TFLOP/s, half the published packed 157.3).
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4


def serve(ctx, args):
    """64 clients, Poisson arrivals at each rate in turn: served fragments/s and submit -> present
    latency (the client's colour target back in host memory)."""
    from visor_amd import serving
    t0 = time.perf_counter()
    world = serving.ServingWorld(ctx, nclients=args.serve_clients, seed=4)
    srv = serving.Server(ctx, hbm_budget_bytes=16 << 30, slots=2, max_batch=64)
    for c in world.clients:  # warm: every client once
        srv.wait(srv.submit(c.cb))
    print("[bench] serving world: %d clients, %d-%d triangles per request, built in %.1f s" % (
        len(world.clients), min(c.triangles for c in world.clients), max(c.triangles for c in world.clients),
        time.perf_counter() - t0), file=sys.stderr, flush=True)
    out = {"clients": len(world.clients), "slots": 2, "hbm_budget_gib": 16, "seconds_per_rate": args.serve_seconds,
           "workload": "targets 256^2/512^2/1024^2/1920x1080 BGRA8+D32, 1-32 indexed draws of resident UV-sphere "
                       "meshes (log-uniform total triangles, 1k-500k), Lambert x mipmapped texture 256^2-2048^2, "
                       "bilinear clamp, back-face cull", "rates": []}
    for k, rate in enumerate(float(x) for x in args.serve_rates.split(",")):
        st, wall = serving.run_clients(world, srv, rate, args.serve_seconds, seed=k + 1)
        row = {"offered_rps": rate, "completed": st["completed"], "achieved_rps": round(st["completed"] / wall, 1),
               "gfragments_s": round(st["fragments"] / wall / 1e9, 4), "p50_ms": round(st["p50_ms"], 3),
               "p99_ms": round(st["p99_ms"], 3), "mean_batch": round(st["mean_batch"], 2),
               "gpu_busy_frac": round(st["busy_ms"] / (wall * 1e3), 3)}
        out["rates"].append(row)
        print("[bench] served %s" % json.dumps(row), file=sys.stderr, flush=True)
    srv.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=16, help="untimed frames; they also let the autotuner try "
                    "its 7 launch shapes twice")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--tile", type=int, default=0, help="8, 16 or 32 fixes the tile size; 0 = autotuned")
    ap.add_argument("--chunk", type=int, default=0, help="16..64 fixes the LDS chunk (with --tile); 0 = default")
    ap.add_argument("--cpu-draws", type=int, default=128, help="draws in the CPU-baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch the batched pass kernel by kernel")
    ap.add_argument("--no-serve", action="store_true", help="skip the 64-client serving run")
    ap.add_argument("--serve-rates", type=str, default="1000,4000,1000000",
                    help="aggregate Poisson arrival rates (requests/s) of the serving run; a rate far above "
                         "capacity makes every client resubmit as soon as its frame is done")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-clients", type=int, default=64)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.chunk:
        os.environ["VISOR_CHUNK"] = str(args.chunk)  # read when the queue is created
    from visor_amd import render, scenes

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only; the ICD synchronises its own stream
        sync = lambda: None

    scene = scenes.many_draws()
    ctx = render.Context()
    r = render.SceneRenderer(ctx, scene, host_visible_targets=False)
    ctx.set_mode(batched=True, tile=args.tile, graphs=not args.no_graphs)
    cb = r.record()
    r.cb = cb

    print("[bench] many_draws: %d draws, %d triangles; warmup %d" % (len(scene.draws), scene.triangles, args.warmup),
          file=sys.stderr, flush=True)
    # The autotuner tries its launch shapes on the first frames of a new pass shape. Those trial
    # frames run here, as part of setting up (like building the pipelines), so that the warmup and
    # the timed frames all run the chosen shape whatever --warmup is.
    trial_frames = 0
    if args.tile == 0:
        while trial_frames < 64:
            r.render()
            trial_frames += 1
            if not ctx.stats(0)["tuning"]:
                break
    for _ in range(args.warmup):
        r.render()
    stats_one = ctx.stats(0)
    tile = stats_one["tile_size"]
    print("[bench] frame stats: %s" % json.dumps(stats_one), file=sys.stderr, flush=True)

    # timed batched steps
    ctx.icd.vh_reset_stats(ctx.queue)
    lat = []
    sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lat.append(r.render())
    sync()
    elapsed = time.perf_counter() - t0
    acc = ctx.stats(1)
    frags_per_frame = stats_one["pixels_written"]
    value = frags_per_frame * args.steps / elapsed / 1e9
    ms_per_step = elapsed * 1e3 / args.steps
    raster_ms = acc["raster_ms"] / max(acc["raster_launches"], 1)
    print("[bench] batched %.3f ms/frame, raster kernel %.3f ms" % (ms_per_step, raster_ms), file=sys.stderr, flush=True)

    # roofline of the dominant kernel (the raster launch, one per frame)
    b_raster = raster_bytes(stats_one, scene.width, scene.height, 4, 32)
    achieved = b_raster / (raster_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": profiled_traffic(tile)}
    # the bound the raster kernel actually meets: wave-level VALU instruction issue
    # (SQ_INSTS_VALU per launch from the committed summary, over the live kernel time)
    valu = profiled_value(tile, "valu_insts_per_launch")
    valu_roofline = None
    if valu:
        got = valu / (raster_ms * 1e-3) / 1e9
        valu_roofline = {"bound": "valu_issue", "achieved": round(got, 1), "peak": VALU_PEAK_GINST,
                         "unit": "G wave-instructions/s", "frac": round(got / VALU_PEAK_GINST, 4)}

    # plain per-draw path
    plain = None
    if not args.no_plain:
        ctx.set_mode(batched=False)
        r.render()
        ctx.icd.vh_reset_stats(ctx.queue)
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            r.render()
        pe = time.perf_counter() - t0
        plain = frags_per_frame * args.plain_steps / pe / 1e9
        print("[bench] plain %.3f ms/frame" % (pe * 1e3 / args.plain_steps), file=sys.stderr, flush=True)
        ctx.set_mode(batched=True)

    cpu = None
    if not args.no_cpu:
        from oracle import refrast
        k = min(args.cpu_draws, len(scene.draws))
        res = refrast.render(scene, threads=7, draws=scene.draws[:k], time_only=True)
        ser = refrast.render(scene, threads=0, draws=scene.draws[:k], time_only=True)
        cpu = {"value": round(res["counters"]["pixels_written"] / res["seconds"] / 1e9, 6), "unit": "Gfragments/s",
               "cores": 8, "kind": "port",
               "serial_one_core_value": round(ser["counters"]["pixels_written"] / ser["seconds"] / 1e9, 6),
               "sample": "first %d of the 256 draws (%d covered fragments, %.1f s), oracle threaded mode: "
                         "7 workers + work-stealing submitter over 32x32 blocks" % (
                             k, res["counters"]["pixels_written"], res["seconds"])}
        print("[bench] cpu baseline %.4f Gfrag/s" % cpu["value"], file=sys.stderr, flush=True)

    served = None
    if not args.no_serve:
        served = serve(ctx, args)

    lat_ms = sorted(x * 1e3 for x in lat)
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gfragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "256 draws into 1920x1080 BGRA8+D32, %d triangles, trilinear repeat, "
                               "half alpha-blended" % scene.triangles,
                   "mode": "batched", "tile": tile, "chunk": stats_one["chunk"], "persistent": bool(stats_one["persistent"]),
                   "launch_shape": "autotuned" if args.tile == 0 else "fixed", "tuner_trial_frames": trial_frames, "fragments_per_frame": frags_per_frame,
                   "shaded_per_frame": stats_one["depth_passed"]},
        "plain_per_draw_gfrag_s": round(plain, 4) if plain else None,
        "speedup_vs_plain": round(value / plain, 2) if plain else None,
        "frame_latency_ms": {"p50": round(statistics.median(lat_ms), 3),
                             "p99": round(lat_ms[min(len(lat_ms) - 1, int(0.99 * len(lat_ms)))], 3)},
        "kernel_ms": {"raster": round(raster_ms, 4), "setup": round(acc["setup_ms"] / max(args.steps, 1), 4),
                      "bin": round(acc["bin_ms"] / max(args.steps, 1), 4)},
        "graph_replays": acc["graph_launches"],
        "roofline": roofline,
        "valu_roofline": valu_roofline,
        "cpu_baseline": cpu,
        "served": served,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
