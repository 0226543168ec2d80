"""Asynchronous submits of batched render passes back to back: the executor launches each pass
before it has waited for the one ahead of it, and completes that one while the next runs. Every
frame equals the oracle's, a pass that continues another's target keeps its order, and a pass
whose bins overflow runs again together with the pass queued behind it, in order, with every
fragment counted once."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from oracle import refrast
from visor_amd import render, scenes, vk
from visor_amd.vk import Handle

pytestmark = pytest.mark.gpu


def _device_bytes(ctx, mem, nbytes):
    """Bytes of device-local memory: a buffer bound to the same memory is copied into a
    host-visible one by a synchronous submit."""
    src = Handle()
    ctx.icd.check(ctx.icd.vkCreateBuffer(ctx.device, C.byref(vk.BufferCreateInfo(size=nbytes)), None, C.byref(src)),
                  "vkCreateBuffer")
    ctx.icd.check(ctx.icd.vkBindBufferMemory(ctx.device, src, mem, 0), "vkBindBufferMemory")
    dst, dmem = ctx.buffer(nbytes, host_visible=True)
    cb = ctx.begin()
    ctx.icd.vkCmdCopyBuffer(cb, src, dst, 1, C.byref(vk.BufferCopy(srcOffset=0, dstOffset=0, size=nbytes)))
    ctx.end_submit(cb)
    out = np.frombuffer((C.c_uint8 * nbytes).from_address(ctx.map(dmem, nbytes)), np.uint8).copy()
    ctx.icd.vkDestroyBuffer(ctx.device, src, None)
    ctx.free_buffer(dst, dmem)
    return out


def _check_target(ctx, r, ref):
    s = r.scene
    col = _device_bytes(ctx, r.cmem, s.width * s.height * 4).reshape(s.height, s.width, 4)
    dep = _device_bytes(ctx, r.dmem, s.width * s.height * 4).view(np.uint32).reshape(s.height, s.width)
    bad = np.argwhere((col != ref["color"]).any(-1))
    assert bad.size == 0, "%s: colour differs at %d pixels, first %s" % (s.name, len(bad), bad[:3].tolist())
    assert np.array_equal(dep, ref["depth"].view(np.uint32)), s.name


def _two_passes(ctx, s):
    """The scene's one draw as two render passes in two command buffers: the first clears and
    draws the first half of the triangles, the second loads colour and depth and draws the rest
    (together the same fragments in the same order as the scene in one pass)."""
    d = s.draws[0]
    half = 3 * (d.count // 6)
    a = dataclasses.replace(d, count=half)
    b = dataclasses.replace(d, count=d.count - half, first=d.first + half)
    r = render.SceneRenderer(ctx, s, host_visible_targets=False)
    return r, r.record(passes=[([a], True)]), r.record(passes=[([b], False)])


def _submit_back_to_back(ctx, cbs):
    ctx.set_mode(batched=True, tile=8, async_submit=True)
    try:
        ctx.icd.vh_reset_stats(ctx.queue)
        for cb in cbs:
            ctx.icd.check(ctx.submit([cb]), "vkQueueSubmit")
        ctx.icd.check(ctx.icd.vkQueueWaitIdle(ctx.queue), "vkQueueWaitIdle")
        return ctx.stats(1), ctx.stats(0)
    finally:
        ctx.set_mode(batched=True, async_submit=False)


def test_back_to_back_passes_equal_the_oracle(ctx):
    """Four scenes of different sizes, each as two passes on one target, submitted as A1 B1 A2
    B2 C1 D1 C2 D2 with no wait in between: each target equals the oracle's whole scene, and
    the counters of the eight submissions add up to the oracle's."""
    scs = [scenes.random_triangles(ntris=400 + 150 * k, width=160 + 16 * k, height=120 - 8 * k, seed=40 + k)
           for k in range(4)]
    refs = [refrast.render(s) for s in scs]
    parts = [_two_passes(ctx, s) for s in scs]
    order = []
    for a, b in ((0, 1), (2, 3)):
        order += [parts[a][1], parts[b][1], parts[a][2], parts[b][2]]
    acc, _ = _submit_back_to_back(ctx, order)
    for (r, _, _), ref in zip(parts, refs):
        _check_target(ctx, r, ref)
    for k in ("pixels_written", "depth_passed", "tris_in", "tris_out", "pixels_tested"):
        assert acc[k] == sum(ref["counters"][k] for ref in refs), k
    assert acc["raster_launches"] == 8


def test_back_to_back_overflow_runs_both_passes_again_in_order():
    """A queue whose bin capacity has not grown (max(2^20, 2 x triangles)): the first pass draws
    1000 large random triangles on 640x480 with 8x8 tiles, about 1.25M (tile, triangle) pairs, so
    its bins overflow; a small pass on another, smaller target is submitted right behind it (its
    pair count is far below the capacity, and the pairs past it are the first pass's), then the
    pass that loads the first target and draws the other 1000 triangles. The first writes nothing,
    the one behind it sees the overflow flag and writes nothing either, and both run again in
    order: both targets equal the oracle's, the first pass's pairs exceed the first capacity, and
    each fragment is counted once."""
    fresh = render.Context()
    s = scenes.random_triangles(ntris=2000, width=640, height=480, seed=9)
    small = scenes.random_triangles(ntris=60, width=96, height=64, seed=10)
    ref, ref_small = refrast.render(s), refrast.render(small)
    r, first, second = _two_passes(fresh, s)
    rs = render.SceneRenderer(fresh, small, host_visible_targets=False)
    acc, last = _submit_back_to_back(fresh, [first, rs.record(), second])
    _check_target(fresh, r, ref)
    _check_target(fresh, rs, ref_small)
    assert acc["bin_pairs"] - last["bin_pairs"] > 1 << 20
    for k in ("pixels_written", "depth_passed", "tris_in", "pixels_tested"):
        assert acc[k] == ref["counters"][k] + ref_small["counters"][k], k
    assert acc["raster_launches"] == 3
