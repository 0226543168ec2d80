#!/usr/bin/env python
"""Benchmark: batched vs plain per-draw execution of the 256-draw workload on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--tile 0|8|16|32] [--no-serve] ...

Workload: one 1920x1080 BGRA8 + D32 target, 256 draws of log-uniform 16..65536
triangles (about 1.85M), per-draw MVP, 4 mipmapped 512x512 textures sampled trilinear repeat;
even draws opaque (LESS), odd draws alpha-blended with depth writes off. Inputs (vertex, index,
texture data) are resident in HBM; the target stays in HBM.

A step is one synchronous vkQueueSubmit of the recorded render pass (all 256 draws), timed on the
host between synchronisations. value = covered fragments per second in the batched path; the
p50/p99 frame latency comes from the same steps, and back_to_back_gfrag_s from at least 100 more
frames submitted asynchronously with one wait at the end. The plain
per-draw path (setup, bin, raster and a host wait per draw, as the reference's DrawTriangles) is
timed beside it, and the CPU baseline is the oracle's reference-structured threaded mode on a
bounded sample of the draws. The other single-GPU workloads (1000 random triangles at 256x256,
the 200k-triangle lit sphere at 1920x1080, 4M micro-triangles at 3840x2160 RGBA16F) get the same
figures under "configs", and the 64-client serving run under "served", with the same requests
submitted one at a time through vkQueueSubmit as its baseline.
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
REFERENCE_WORKERS = 7  # the reference's NUM_THREADS (rasterizer.cpp:8) beside its work-stealing submitter


def raster_bytes(st, width, height, color_bytes, texel_bytes):
    """Algorithmic HBM bytes of one raster launch (per frame): depth test reads of covered
    fragments, depth writes, colour writes (+ reads when blending) and texels of shaded fragments,
    and the one write of the final colour and depth of every pixel (the clear is fused)."""
    return (4 * st["pixels_written"] + 4 * st["shaded_dwrite"] + color_bytes * st["depth_passed"]
            + color_bytes * st["shaded_blend"] + texel_bytes * st["depth_passed"]
            + width * height * (color_bytes + 4))


def profiled_value(tile, key):
    """A per-launch figure of the raster kernel from the committed rocprofv3 summary of the
    256-draw workload at this tile size (None when no such summary exists)."""
    path = os.path.join(ROOT, "profiles", "va_raster_%d__256draws_1080p.md" % tile)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return int(float(line.split(":")[1]))
    except OSError:
        pass
    return None


# VALU issue: 256 CUs x 4 SIMDs, each issuing one wave64 VALU instruction per 4 cycles (16 lanes)
# at 2.4 GHz; this is the non-packed FP32 rate (78.6 TFLOP/s, half the published packed 157.3).
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, draws, what, nproc_run):
    """The oracle on the host cores: the reference-structured threaded mode (7 workers plus the
    work-stealing submitter, 32x32 blocks through one locked FIFO, a wake per triangle), the same
    with nproc - 1 workers, and the serial mode on one core, over the sample `draws`."""
    from oracle import refrast
    t = refrast.render(scene, threads=REFERENCE_WORKERS, draws=draws, time_only=True)
    frags = t["counters"]["pixels_written"]
    out = {"value": round(frags / t["seconds"] / 1e9, 6), "unit": "Gfragments/s", "cores": REFERENCE_WORKERS + 1,
           "kind": "port", "ms_per_frame": round(t["seconds"] * 1e3, 2),
           "sample": "%s: %d covered fragments, %.2f s; oracle threaded mode, %d workers + work-stealing submitter "
                     "over 32x32 blocks" % (what, frags, t["seconds"], REFERENCE_WORKERS),
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    workers = min(max((os.cpu_count() or 8) - 1, 1), 63)
    if nproc_run and workers != REFERENCE_WORKERS:
        n = refrast.render(scene, threads=workers, draws=draws, time_only=True)
        out["nproc_minus_1"] = {"workers": workers, "value": round(frags / n["seconds"] / 1e9, 6),
                                "seconds": round(n["seconds"], 2)}
    s = refrast.render(scene, threads=0, draws=draws, time_only=True)
    out["serial_one_core_value"] = round(frags / s["seconds"] / 1e9, 6)
    print("[bench] cpu %s: %.5f Gfrag/s threaded, %.5f serial" % (what, out["value"], out["serial_one_core_value"]),
          file=sys.stderr, flush=True)
    return out


def device_sync():
    """torch's synchronisation around the timed frames (plumbing: every frame already ends with
    vkQueueWaitIdle on the ICD's own stream)."""
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def measure(ctx, scene, texel_bytes, steps, warmup, plain_steps, tile=0, graphs=True):
    """Batched (autotuned unless tile is fixed) and plain per-draw GPU figures of one scene."""
    from visor_amd import render
    r = render.SceneRenderer(ctx, scene, host_visible_targets=False)
    ctx.set_mode(batched=True, tile=tile, graphs=graphs)
    r.cb = r.record()
    # The autotuner tries its launch shapes on the first frames of a new pass shape. Those trial
    # frames run here, as part of setting up (like building the pipelines), so that the warmup and
    # the timed frames all run the chosen shape whatever --warmup is.
    trials = 0
    if tile == 0:
        while trials < 64:
            r.render()
            trials += 1
            if not ctx.stats(0)["tuning"]:
                break
    for _ in range(warmup):
        r.render()
    one = ctx.stats(0)
    ctx.icd.vh_reset_stats(ctx.queue)
    # the timed loop calls the ICD directly with a submit info built once (what render() does,
    # without rebuilding the ctypes arguments every frame)
    import ctypes
    from visor_amd import vk
    cbs = (vk.Handle * 1)(r.cb)
    si = vk.SubmitInfo(commandBufferCount=1, pCommandBuffers=cbs)
    submit, wait_idle, queue, clock = ctx.icd.vkQueueSubmit, ctx.icd.vkQueueWaitIdle, ctx.queue, time.perf_counter
    psi = ctypes.byref(si)
    lat = [0.0] * steps
    bad = 0
    device_sync()
    t0 = clock()
    for k in range(steps):
        t = clock()
        bad |= submit(queue, 1, psi, None) | wait_idle(queue)
        lat[k] = clock() - t
    device_sync()
    elapsed = clock() - t0
    ctx.icd.check(bad, "vkQueueSubmit / vkQueueWaitIdle")
    acc = ctx.stats(1)
    frags = one["pixels_written"]
    raster_ms = acc["raster_ms"] / max(acc["raster_launches"], 1)
    color_bytes = 8 if scene.color_bpp == 8 else 4
    b_raster = raster_bytes(one, scene.width, scene.height, color_bytes, texel_bytes)
    achieved = b_raster / (raster_ms * 1e-3) / 1e9
    out = {"gfragments_s": frags * steps / elapsed / 1e9, "ms_per_frame": elapsed * 1e3 / steps, "raster_ms": raster_ms,
           "setup_ms": acc["setup_ms"] / max(steps, 1), "bin_ms": acc["bin_ms"] / max(steps, 1),
           "fragments": frags, "shaded": one["depth_passed"], "stats": one, "latencies": lat, "trials": trials,
           "graph_replays": acc["graph_launches"], "algorithmic_bytes": b_raster,
           "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": round(achieved / HBM_PEAK_GBS, 4)}}
    # back-to-back submissions (asynchronous submit, one wait at the end): the queue launches each
    # pass behind the one still running. A few such frames first start the queue's worker thread
    # and build the second output slot's graph; the rate is then taken over at least 100 frames,
    # since the first and last frames of a burst run without a pass beside them.
    ctx.set_mode(batched=True, async_submit=True)
    for _ in range(4):
        ctx.icd.check(ctx.submit([r.cb]), "vkQueueSubmit")
    ctx.icd.check(ctx.icd.vkQueueWaitIdle(ctx.queue), "vkQueueWaitIdle")
    n_b2b = max(steps, 100)
    t0 = time.perf_counter()
    for _ in range(n_b2b):
        ctx.icd.check(ctx.submit([r.cb]), "vkQueueSubmit")
    ctx.icd.check(ctx.icd.vkQueueWaitIdle(ctx.queue), "vkQueueWaitIdle")
    out["back_to_back_gfragments_s"] = frags * n_b2b / (time.perf_counter() - t0) / 1e9
    ctx.set_mode(batched=True, async_submit=False)
    if plain_steps:
        ctx.set_mode(batched=False)
        r.render()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            r.render()
        pe = time.perf_counter() - t0
        out["plain_gfragments_s"] = frags * plain_steps / pe / 1e9
        out["plain_ms_per_frame"] = pe * 1e3 / plain_steps
        ctx.set_mode(batched=True)
    return out


def config_line(m, cpu):
    line = {"batched_gfragments_s": round(m["gfragments_s"], 4), "batched_ms_per_frame": round(m["ms_per_frame"], 4),
            "batched_back_to_back_gfragments_s": round(m["back_to_back_gfragments_s"], 4),
            "raster_ms": round(m["raster_ms"], 4), "fragments_per_frame": m["fragments"],
            "shaded_per_frame": m["shaded"], "tile": m["stats"]["tile_size"], "chunk": m["stats"]["chunk"],
            "roofline": m["roofline"], "cpu_baseline": cpu}
    if "plain_gfragments_s" in m:
        line["plain_gfragments_s"] = round(m["plain_gfragments_s"], 4)
        line["plain_ms_per_frame"] = round(m["plain_ms_per_frame"], 3)
        line["batched_vs_plain"] = round(m["gfragments_s"] / m["plain_gfragments_s"], 2)
    return line


def other_configs(ctx, args):
    """The other single-GPU workloads: GPU plain and batched figures, raster roofline and a CPU
    baseline each."""
    from visor_amd import scenes
    out = {}
    specs = [("random_triangles_256", "1000 random non-indexed triangles, 256x256 BGRA8+D32, passthrough colour, "
              "LESS", scenes.random_triangles, 0, 20, None),
             ("lit_sphere_1080p", "200k-triangle indexed UV sphere at 1920x1080, MVP + normal matrix, Lambert x "
              "1024^2 texture bilinear clamp, back-face cull", scenes.lit_sphere, 16, 20, None),
             ("micro_triangles_4k", "4M indexed micro-triangles (~4 px) in 50 depth layers, 3840x2160 RGBA16F+D32, "
              "8 trilinear taps of a 2048^2 x 4-layer array texture, 25% additive", scenes.micro_triangles, 256, 10, 1)]
    for key, what, make, texel_bytes, steps, cpu_draws in specs:
        t0 = time.perf_counter()
        scene = make()
        m = measure(ctx, scene, texel_bytes, steps=steps, warmup=4, plain_steps=0 if cpu_draws else 3)
        print("[bench] %s: batched %.3f ms (%.2f Gfrag/s), raster %.3f ms, frac %.4f (%.1f s)" % (
            key, m["ms_per_frame"], m["gfragments_s"], m["raster_ms"], m["roofline"]["frac"], time.perf_counter() - t0),
            file=sys.stderr, flush=True)
        cpu = None
        if not args.no_cpu:
            draws = scene.draws[:cpu_draws] if cpu_draws else None
            sample = "whole frame" if not cpu_draws else "first %d of %d draws" % (cpu_draws, len(scene.draws))
            cpu = cpu_baseline(scene, draws, sample, nproc_run=key == "random_triangles_256")
        line = config_line(m, cpu)
        line["workload"] = what
        out[key] = line
    return out


def serve(ctx, args):
    """64 clients, Poisson arrivals at each rate in turn: served fragments/s and submit -> present
    latency (the client's colour target back in host memory); then the same requests one at a
    time through vkQueueSubmit."""
    from visor_amd import serving
    t0 = time.perf_counter()
    world = serving.ServingWorld(ctx, nclients=args.serve_clients, seed=4)
    srv = serving.Server(ctx, hbm_budget_bytes=16 << 30, slots=args.serve_slots, max_batch=64)
    for c in world.clients:  # warm: every client once
        srv.wait(srv.submit(c.cb))
    print("[bench] serving world: %d clients, %d-%d triangles per request, built in %.1f s" % (
        len(world.clients), min(c.triangles for c in world.clients), max(c.triangles for c in world.clients),
        time.perf_counter() - t0), file=sys.stderr, flush=True)
    out = {"clients": len(world.clients), "slots": args.serve_slots, "hbm_budget_gib": 16, "seconds_per_rate": args.serve_seconds,
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
    # baseline: the same clients' requests, one synchronous vkQueueSubmit at a time
    ctx.set_mode(batched=True)
    done, frags, wall, lat = serving.run_sequential(world, ctx, args.serve_seconds)
    lat.sort()
    seq = {"completed": done, "achieved_rps": round(done / wall, 1), "gfragments_s": round(frags / wall / 1e9, 4),
           "p50_ms": round(statistics.median(lat), 3), "p99_ms": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 3),
           "how": "one thread, round-robin over the 64 clients, synchronous vkQueueSubmit + vkQueueWaitIdle per "
                  "request (batched execution of each request's own draws)"}
    out["sequential_submit"] = seq
    out["served_vs_sequential"] = round(max(r["gfragments_s"] for r in out["rates"]) / max(seq["gfragments_s"], 1e-9), 2)
    print("[bench] sequential %s" % json.dumps(seq), file=sys.stderr, flush=True)
    return out


def transfer(ctx, gib, reps=3):
    """vkCmdCopyBuffer between two device-local buffers of `gib` GiB (the transfer path,
    cmd_exec.cpp:174-182): one submission per copy; also the known-size copy that calibrates the
    HBM byte counters of the profiles."""
    from visor_amd import vk
    import ctypes as C
    n = int(gib * (1 << 30))
    src, smem = ctx.buffer(n, host_visible=False)
    dst, dmem = ctx.buffer(n, host_visible=False)
    region = vk.BufferCopy(srcOffset=0, dstOffset=0, size=n)
    cb = ctx.begin()
    ctx.icd.vkCmdCopyBuffer(cb, src, dst, 1, C.byref(region))
    ctx.icd.check(ctx.icd.vkEndCommandBuffer(cb), "vkEndCommandBuffer")
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        ctx.icd.check(ctx.submit([cb]), "vkQueueSubmit")
        ctx.icd.check(ctx.icd.vkQueueWaitIdle(ctx.queue), "vkQueueWaitIdle")
        times.append(time.perf_counter() - t0)
    ctx.free_buffer(src, smem)
    ctx.free_buffer(dst, dmem)
    best = min(times)
    return {"bytes_copied": n, "best_ms": round(best * 1e3, 3), "gb_s_read_plus_write": round(2 * n / best / 1e9, 1)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=16, help="untimed frames after the autotuner's trial frames")
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--tile", type=int, default=0, help="8, 16 or 32 fixes the tile size; 0 = autotuned")
    ap.add_argument("--chunk", type=int, default=0, help="16..64 fixes the LDS chunk (with --tile); 0 = default")
    ap.add_argument("--cpu-draws", type=int, default=32, help="draws of the 256-draw workload in its CPU sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-graphs", action="store_true", help="launch the batched pass kernel by kernel")
    ap.add_argument("--no-serve", action="store_true", help="skip the 64-client serving run")
    ap.add_argument("--no-configs", action="store_true", help="skip the other single-GPU workloads")
    ap.add_argument("--serve-rates", type=str, default="1000,4000,1000000",
                    help="aggregate Poisson arrival rates (requests/s) of the serving run; a rate far above "
                         "capacity makes every client resubmit as soon as its frame is done")
    ap.add_argument("--serve-seconds", type=float, default=3.0)
    ap.add_argument("--serve-clients", type=int, default=64)
    ap.add_argument("--serve-slots", type=int, default=2, help="batches on the GPU at once (1..4)")
    ap.add_argument("--transfer-gib", type=float, default=0.0,
                    help="only time a device-local vkCmdCopyBuffer of this many GiB (three copies) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one MI355X; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    if args.chunk:
        os.environ["VISOR_CHUNK"] = str(args.chunk)  # read when the queue is created
    from visor_amd import render, scenes

    if args.transfer_gib > 0:
        print(json.dumps({"transfer": transfer(render.Context(), args.transfer_gib)}))
        return
    scene = scenes.many_draws()
    ctx = render.Context()
    print("[bench] many_draws: %d draws, %d triangles; warmup %d" % (len(scene.draws), scene.triangles, args.warmup),
          file=sys.stderr, flush=True)
    m = measure(ctx, scene, 32, args.steps, args.warmup, 0 if args.no_plain else args.plain_steps, tile=args.tile,
                graphs=not args.no_graphs)
    one = m["stats"]
    tile = one["tile_size"]
    print("[bench] frame stats: %s" % json.dumps(one), file=sys.stderr, flush=True)
    print("[bench] batched %.3f ms/frame, raster kernel %.3f ms" % (m["ms_per_frame"], m["raster_ms"]), file=sys.stderr,
          flush=True)
    roofline = dict(m["roofline"])
    roofline["traffic"] = profiled_value(tile, "traffic_bytes_per_launch")
    # the bound the raster kernel actually meets: wave-level VALU instruction issue
    # (SQ_INSTS_VALU per launch from the committed summary, over the live kernel time)
    valu = profiled_value(tile, "valu_insts_per_launch")
    valu_roofline = None
    if valu:
        got = valu / (m["raster_ms"] * 1e-3) / 1e9
        valu_roofline = {"bound": "valu_issue", "achieved": round(got, 1), "peak": VALU_PEAK_GINST,
                         "unit": "G wave-instructions/s", "frac": round(got / VALU_PEAK_GINST, 4)}
    plain = m.get("plain_gfragments_s")

    cpu = None
    if not args.no_cpu:
        k = min(args.cpu_draws, len(scene.draws))
        cpu = cpu_baseline(scene, scene.draws[:k], "first %d of the 256 draws" % k, nproc_run=True)

    configs = None if args.no_configs else other_configs(ctx, args)
    served = None if args.no_serve else serve(ctx, args)

    lat_ms = sorted(x * 1e3 for x in m["latencies"])
    out = {
        "metric": METRIC,
        "value": round(m["gfragments_s"], 4),
        "unit": "Gfragments/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_frame"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "256 draws into 1920x1080 BGRA8+D32, %d triangles, trilinear repeat, "
                               "half alpha-blended" % scene.triangles,
                   "mode": "batched", "tile": tile, "chunk": one["chunk"], "persistent": bool(one["persistent"]),
                   "launch_shape": "autotuned" if args.tile == 0 else "fixed", "tuner_trial_frames": m["trials"],
                   "fragments_per_frame": m["fragments"], "shaded_per_frame": m["shaded"]},
        "back_to_back_gfrag_s": round(m["back_to_back_gfragments_s"], 4),
        "plain_per_draw_gfrag_s": round(plain, 4) if plain else None,
        "speedup_vs_plain": round(m["gfragments_s"] / plain, 2) if plain else None,
        "frame_latency_ms": {"p50": round(statistics.median(lat_ms), 3),
                             "p99": round(lat_ms[min(len(lat_ms) - 1, int(0.99 * len(lat_ms)))], 3)},
        "kernel_ms": {"raster": round(m["raster_ms"], 4), "setup": round(m["setup_ms"], 4), "bin": round(m["bin_ms"], 4)},
        "graph_replays": m["graph_replays"],
        "roofline": roofline,
        "valu_roofline": valu_roofline,
        "cpu_baseline": cpu,
        "configs": configs,
        "served": served,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
