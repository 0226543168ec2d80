"""The API paths around the draw on the GPU, each against the oracle or its specified behaviour:
shaders fed from uniform buffers (reference spirv_compile.cpp:552-558), vkCmdDrawIndexed's
vertexOffset (recorded, never applied: rasterizer.cpp:100-119), vertices far outside the int32
window range and out-of-range clear values (the reference's x64 float->int conversions), buffer
copies with regions outside their buffers (refused), asynchronous submits completed through
fences, and the headless swapchain: image order, presented output and per-frame statistics
(wsi.cpp:185-219, stats.cpp:25-41)."""
import ctypes as C
import dataclasses

import numpy as np
import pytest

from oracle import refrast
from visor_amd import render, scenes, shaders, vk
from visor_amd.vk import CMP, Handle

pytestmark = pytest.mark.gpu


def _render(ctx, s, batched=True, tile=8):
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=batched, tile=tile)
    r.render()
    return r.read_color(), r.read_depth(), ctx.stats(0)


def _check(ctx, s, batched=True, tile=8):
    ref = refrast.render(s)
    col, dep, st = _render(ctx, s, batched, tile)
    bad = np.argwhere((col != ref["color"]).any(-1))
    assert bad.size == 0, "colour differs at %d pixels, first %s gpu %s ref %s" % (
        len(bad), bad[:2].tolist(), col[tuple(bad[0])].tolist(), ref["color"][tuple(bad[0])].tolist())
    assert np.array_equal(dep.view(np.uint32), ref["depth"].view(np.uint32))
    for k in ("pixels_written", "depth_passed", "tris_in", "tris_out", "blocks", "pixels_tested"):
        assert st[k] == ref["counters"][k], (k, st[k], ref["counters"][k])
    return ref


def _ubo_scene(seed):
    s = scenes.random_triangles(ntris=400, width=96, height=80, seed=seed)
    s.pipelines = {"main": scenes.Pipeline("vs_ubo_pos4_col", "fs_ubo_tint")}
    rng = np.random.default_rng(seed)
    n = s.draws[0].count // 3
    s.uniforms, s.draws = [], []
    for k in range(4):
        m = np.eye(4)
        m[:3, :3] = scenes.rotation(0.0, 0.0, rng.uniform(-0.4, 0.4)) * rng.uniform(0.7, 1.2)
        m[0, 3], m[1, 3] = rng.uniform(-0.2, 0.2, 2)
        tint = rng.uniform(0.2, 1.5, 4)
        bias = rng.uniform(-0.2, 0.3, 4)
        s.uniforms.append(shaders.ubo_block_bytes(m, tint, bias))
        lo, hi = 3 * (n * k // 4), 3 * (n * (k + 1) // 4)
        s.draws.append(scenes.Draw("main", hi - lo, lo, False, b"\0" * 128, uniform=k))
    return s


@pytest.mark.parametrize("batched", [True, False])
def test_uniform_buffer_programs(ctx, batched):
    """Four draws, each with its own uniform buffer (MVP in the vertex program, tint and bias in
    the fragment program), bit-exact against the oracle."""
    ref = _check(ctx, _ubo_scene(31), batched=batched)
    assert ref["counters"]["depth_passed"] > 0


@pytest.mark.parametrize("batched", [True, False])
def test_vertex_offset_is_not_applied(ctx, batched):
    rng = scenes.Rng(41)
    nv, nt = 300, 250
    pos = np.zeros((nv, 4), np.float32)
    pos[:, 0] = rng.uniform(nv, -1.0, 1.0)
    pos[:, 1] = rng.uniform(nv, -1.0, 1.0)
    pos[:, 2] = rng.uniform(nv, 0.0, 1.0)
    pos[:, 3] = 1.0
    col = rng.bytes(nv * 4).reshape(nv, 4)
    idx = (rng.raw(3 * nt) % np.uint64(nv // 2)).astype(np.uint32)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color")}
    draws = [scenes.Draw("main", 3 * 120, 0, True, scenes.push_block(scenes.IDENTITY), vertex_offset=nv // 2),
             scenes.Draw("main", 3 * 130, 3 * 120, True, scenes.push_block(scenes.IDENTITY), vertex_offset=-7)]
    s = scenes.Scene("voff", 88, 72, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes, draws,
                     index_data=idx.tobytes(), index_type=vk.INDEX_U32)
    _check(ctx, s, batched=batched)


@pytest.mark.parametrize("batched", [True, False])
def test_far_vertices_and_out_of_range_clear(ctx, batched):
    """Triangles with a vertex at w = 1e-7 (window coordinates ~6e8, or outside int32 where the
    conversion gives 0x80000000) or behind the eye (w < 0): the reference does no clipping, and
    its wrapping int32 edge arithmetic decides coverage; clear values whose x255 is out of range
    store the low byte of 0x80000000."""
    rng = scenes.Rng(57)
    n = 240
    pos = np.zeros((3 * n, 4), np.float32)
    pos[:, 0] = rng.uniform(3 * n, -1.0, 1.0)
    pos[:, 1] = rng.uniform(3 * n, -1.0, 1.0)
    pos[:, 2] = rng.uniform(3 * n, 0.0, 1.0)
    pos[:, 3] = 1.0
    k = np.arange(n)
    near = 3 * k[k % 4 == 0] + 2
    pos[near, 3] = 1e-7                          # in-range huge window coordinates
    out = 3 * k[k % 4 == 1] + 1
    pos[out, 0] *= 400.0                         # x/w up to 4e9: outside int32
    pos[out, 3] = 1e-7
    behind = 3 * k[k % 4 == 2]
    pos[behind, 3] = -0.5
    col = rng.bytes(3 * n * 4).reshape(3 * n, 4)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color")}
    s = scenes.Scene("far", 64, 48, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes,
                     [scenes.Draw("main", 3 * n, 0, False, scenes.push_block(scenes.IDENTITY))],
                     clear_color=(1e10, 2.0, -0.5, -1e10))
    ref = _check(ctx, s, batched=batched)
    assert ref["counters"]["pixels_written"] > 0


def test_buffer_copy_outside_its_buffers_is_refused(ctx):
    src, smem = ctx.buffer(256, host_visible=True)
    dst, dmem = ctx.buffer(256, host_visible=True)
    data = np.arange(256, dtype=np.uint8)
    C.memmove(ctx.map(smem, 256), data.tobytes(), 256)
    C.memset(ctx.map(dmem, 256), 0xEE, 256)
    for srcOffset, dstOffset, size in [(200, 0, 100), (0, 250, 8), (0, 0, 1 << 62)]:
        cb = ctx.begin()
        ctx.icd.vkCmdCopyBuffer(cb, src, dst, 1, C.byref(vk.BufferCopy(srcOffset=srcOffset, dstOffset=dstOffset, size=size)))
        ctx.icd.vkEndCommandBuffer(cb)
        assert ctx.submit([cb]) == -8  # VK_ERROR_FEATURE_NOT_PRESENT
    assert np.all(np.frombuffer(C.string_at(ctx.map(dmem, 256), 256), np.uint8) == 0xEE)
    cb = ctx.begin()
    ctx.icd.vkCmdCopyBuffer(cb, src, dst, 1, C.byref(vk.BufferCopy(srcOffset=100, dstOffset=0, size=156)))
    ctx.end_submit(cb)
    got = np.frombuffer(C.string_at(ctx.map(dmem, 256), 256), np.uint8)
    assert np.array_equal(got[:156], data[100:]) and np.all(got[156:] == 0xEE)


def _fence(ctx, signalled=False):
    f = Handle()
    ctx.icd.check(ctx.icd.vkCreateFence(ctx.device, C.byref(vk.FenceCreateInfo(flags=1 if signalled else 0)), None,
                                        C.byref(f)), "vkCreateFence")
    return f


def test_async_submit_completes_through_fences(ctx):
    """With asynchronous submits vkQueueSubmit returns at once; the fence passed to it signals
    when the results are in host-visible memory, for wait-all and for wait-any (a second,
    never-signalled fence in the set)."""
    s = scenes.random_triangles(ntris=600, width=128, height=96, seed=8)
    ref = refrast.render(s)
    r = render.SceneRenderer(ctx, s)
    r.cb = r.record()
    ctx.set_mode(batched=True, tile=8, async_submit=True)
    try:
        never = _fence(ctx)
        for wait_all in (1, 0):
            done = _fence(ctx)
            C.memset(ctx.map(r.cmem, s.width * s.height * 4), 0, s.width * s.height * 4)
            assert ctx.submit([r.cb], fence=done) == 0
            fences = (Handle * 2)(done, never) if not wait_all else (Handle * 1)(done)
            assert ctx.icd.vkWaitForFences(ctx.device, len(fences), fences, wait_all, 30 * 10 ** 9) == 0
            assert np.array_equal(r.read_color(), ref["color"])
            assert np.array_equal(r.read_depth().view(np.uint32), ref["depth"].view(np.uint32))
            ctx.icd.vkDestroyFence(ctx.device, done, None)
        ctx.icd.vkDestroyFence(ctx.device, never, None)
    finally:
        ctx.set_mode(batched=True, async_submit=False)


def test_headless_swapchain_present_and_frame_stats(ctx):
    """A swapchain of minImageCount images: each acquire returns the next image in turn
    (reference wsi.cpp:185-195, starting after image 0), each present shows the image it names,
    and its bytes equal the oracle's frame; a present ends the frame's statistics and begins the
    next frame's (stats.cpp:25-41)."""
    W, H, n = 64, 48, 3
    fmt = vk.FMT["B8G8R8A8_UNORM"]
    sci = vk.SwapchainCreateInfoKHR(minImageCount=n, imageFormat=fmt, imageExtent=vk.Extent2D(width=W, height=H),
                                    imageArrayLayers=1)
    sw = Handle()
    ctx.icd.check(ctx.icd.vkCreateSwapchainKHR(ctx.device, C.byref(sci), None, C.byref(sw)), "vkCreateSwapchainKHR")
    cnt = vk.u32(0)
    ctx.icd.vkGetSwapchainImagesKHR(ctx.device, sw, C.byref(cnt), None)
    assert cnt.value == n
    imgs = (Handle * n)()
    ctx.icd.vkGetSwapchainImagesKHR(ctx.device, sw, C.byref(cnt), imgs)
    views = [ctx.view(imgs[k], fmt) for k in range(n)]
    ctx.set_mode(batched=True, tile=8)
    pixels = vk.vp()
    idx, presents = vk.u32(), vk.u64()
    assert ctx.icd.vh_swapchain_frame(sw, C.byref(idx), C.byref(presents), C.byref(pixels)) == 0
    assert presents.value == 0 and not pixels.value
    ctx.icd.vh_reset_stats(ctx.queue)  # the frame begins here (earlier tests shared the queue)
    for frame in range(2 * n):
        ctx.icd.check(ctx.icd.vkAcquireNextImageKHR(ctx.device, sw, 2 ** 64 - 1, None, None, C.byref(idx)), "acquire")
        assert idx.value == (frame + 1) % n
        s = scenes.random_triangles(ntris=150 + 40 * frame, width=W, height=H, seed=100 + frame)
        ref = refrast.render(s)
        r = render.SceneRenderer(ctx, s, target=(imgs[idx.value], views[idx.value]))
        r.render()
        # the frame's statistics run until the present
        cur = ctx.stats(3)
        assert cur["pixels_written"] == ref["counters"]["pixels_written"]
        res = (vk.i32 * 1)(99)
        pi = vk.PresentInfoKHR(swapchainCount=1, pSwapchains=(Handle * 1)(sw), pImageIndices=(vk.u32 * 1)(idx.value),
                               pResults=res)
        assert ctx.icd.vkQueuePresentKHR(ctx.queue, C.byref(pi)) == 0 and res[0] == 0
        shown = vk.u32()
        assert ctx.icd.vh_swapchain_frame(sw, C.byref(shown), C.byref(presents), C.byref(pixels)) == 0
        assert shown.value == idx.value and presents.value == frame + 1
        got = np.frombuffer(C.string_at(pixels.value, W * H * 4), np.uint8).reshape(H, W, 4)
        assert np.array_equal(got, ref["color"])
        ended, begun = ctx.stats(2), ctx.stats(3)
        for k in ("pixels_written", "depth_passed", "tris_in", "draws"):
            assert ended[k] == ref["counters"][k], (frame, k)
        assert begun["pixels_written"] == 0 and begun["draws"] == 0
    # a present of an image index the swapchain does not have is refused
    pi = vk.PresentInfoKHR(swapchainCount=1, pSwapchains=(Handle * 1)(sw), pImageIndices=(vk.u32 * 1)(n))
    assert ctx.icd.vkQueuePresentKHR(ctx.queue, C.byref(pi)) == -1000001004
    ctx.icd.vkDestroySwapchainKHR(ctx.device, sw, None)
