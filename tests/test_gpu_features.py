"""Fixed-function and sampler features on the GPU, each against the oracle: every depth compare op
with and without depth writes, every cull mode x front face, all blend-factor pairs the reference
implements, triangle strips, 16-bit indices, a loading (not clearing) second render pass, clear
values outside [0, 1], sampler modes, BC2/BC3 textures and cube maps. Small targets; bit-exact
colour and depth."""
import dataclasses

import numpy as np
import pytest

from oracle import refrast
from visor_amd import render, scenes
from visor_amd.vk import (ADDR_CLAMP, ADDR_REPEAT, BF, CMP, CULL_BACK, CULL_FRONT, CULL_NONE, FILTER_LINEAR,
                          FILTER_NEAREST, FMT, FRONT_CCW, FRONT_CW, INDEX_U16, MIPMAP_LINEAR, MIPMAP_NEAREST,
                          TOPO_STRIP)

pytestmark = pytest.mark.gpu


def _run(ctx, s, passes=None, batched=True):
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=batched, tile=8)
    if passes is not None:
        r.cb = r.record(passes=passes)
    r.render()
    return r.read_color(), (r.read_depth() if r.has_depth else None), ctx.stats(0)


def _check(ctx, s, ref=None, passes=None, batched=True):
    ref = ref if ref is not None else refrast.render(s)
    col, dep, st = _run(ctx, s, passes, batched)
    bad = np.argwhere((col != ref["color"]).any(-1))
    assert bad.size == 0, "colour differs at %d pixels, first %s gpu %s ref %s" % (
        len(bad), bad[:2].tolist(), col[tuple(bad[0])].tolist(), ref["color"][tuple(bad[0])].tolist())
    if dep is not None:
        bad = np.argwhere(dep.view(np.uint32) != ref["depth"].view(np.uint32))
        assert bad.size == 0, "depth differs at %d pixels" % len(bad)
    if passes is None:
        assert st["pixels_written"] == ref["counters"]["pixels_written"]
        assert st["depth_passed"] == ref["counters"]["depth_passed"]
    return ref


def _soup(ntris=300, w=96, h=80, seed=5, **pipe):
    s = scenes.random_triangles(ntris, w, h, seed=seed)
    s.pipelines = {"main": dataclasses.replace(s.pipelines["main"], **pipe)}
    return s


@pytest.mark.parametrize("op", sorted(CMP))
@pytest.mark.parametrize("write", [True, False])
def test_depth_compare_ops(ctx, op, write):
    s = _soup(depth_op=CMP[op], depth_write=write)
    s.clear_depth = 0.5
    ref = _check(ctx, s)
    if op not in ("NEVER", "EQUAL"):
        assert ref["counters"]["depth_passed"] > 0


def test_equal_redraw_passes_where_depth_was_written(ctx):
    """A LESS pass writes depth; a loading pass redraws the same triangles with EQUAL and other
    colours: exactly the frontmost fragments pass again."""
    s = _soup(ntris=200, w=80, h=64, seed=14)
    s.pipelines["eq"] = dataclasses.replace(s.pipelines["main"], depth_op=CMP["EQUAL"], depth_write=False,
                                            blend=True, src=BF["ONE"], dst=BF["ONE"])
    d = s.draws[0]
    first = scenes.Draw("main", d.count, 0, False, d.push)
    again = scenes.Draw("eq", d.count, 0, False, d.push)
    s.draws = [first, again]
    ref = refrast.render(s)
    assert ref["counters"]["depth_passed"] > 0
    _check(ctx, s, ref=ref, passes=[([first], True), ([again], False)])


@pytest.mark.parametrize("cull", [CULL_NONE, CULL_FRONT, CULL_BACK])
@pytest.mark.parametrize("front", [FRONT_CCW, FRONT_CW])
def test_cull_and_front_face(ctx, cull, front):
    _check(ctx, _soup(cull=cull, front=front))


FACTORS = ["ZERO", "ONE", "SRC_ALPHA", "ONE_MINUS_SRC_ALPHA"]


@pytest.mark.parametrize("src", FACTORS)
@pytest.mark.parametrize("dst", FACTORS)
def test_blend_factor_pairs(ctx, src, dst):
    s = _soup(ntris=200, blend=True, src=BF[src], dst=BF[dst], depth_write=False)
    s.clear_color = (0.2, 0.6, 0.9, 0.5)
    _check(ctx, s)


def test_triangle_strip(ctx):
    rng = scenes.Rng(8)
    n = 240
    pos = np.zeros((n, 4), np.float32)
    pos[:, 0] = np.linspace(-1.1, 1.1, n).astype(np.float32) + rng.uniform(n, -0.05, 0.05)
    pos[:, 1] = np.where(np.arange(n) % 2 == 0, -0.6, 0.6).astype(np.float32) + rng.uniform(n, -0.3, 0.3)
    pos[:, 2] = rng.uniform(n, 0.0, 1.0)
    pos[:, 3] = 1.0
    col = rng.bytes(n * 4).reshape(n, 4)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color", topology=TOPO_STRIP, cull=CULL_BACK)}
    s = scenes.Scene("strip", 128, 96, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes,
                     [scenes.Draw("main", n, 0, False, scenes.push_block(scenes.IDENTITY))])
    ref = _check(ctx, s)
    assert ref["counters"]["tris_out"] < ref["counters"]["tris_in"] == n - 2


@pytest.mark.parametrize("batched", [True, False])
def test_indexed_u16(ctx, batched):
    rng = scenes.Rng(9)
    nv, nt = 400, 350
    pos = np.zeros((nv, 4), np.float32)
    pos[:, 0] = rng.uniform(nv, -1.0, 1.0)
    pos[:, 1] = rng.uniform(nv, -1.0, 1.0)
    pos[:, 2] = rng.uniform(nv, 0.0, 1.0)
    pos[:, 3] = 1.0
    col = rng.bytes(nv * 4).reshape(nv, 4)
    idx = (rng.raw(3 * nt) % np.uint64(nv)).astype(np.uint16)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color")}
    draws = [scenes.Draw("main", 3 * 200, 0, True, scenes.push_block(scenes.IDENTITY)),
             scenes.Draw("main", 3 * 150, 3 * 200, True, scenes.push_block(scenes.IDENTITY))]
    s = scenes.Scene("u16", 80, 112, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes, draws,
                     index_data=idx.tobytes(), index_type=INDEX_U16)
    _check(ctx, s, batched=batched)


@pytest.mark.parametrize("batched", [True, False])
def test_second_pass_loads_previous_contents(ctx, batched):
    """Clear + first half of the draws, then a LOAD pass with the second half: equals one pass
    with all draws."""
    s = _soup(ntris=400, w=72, h=72, seed=12)
    n = s.draws[0].count
    half = (n // 6) * 3
    a = scenes.Draw("main", half, 0, False, s.draws[0].push)
    b = scenes.Draw("main", n - half, half, False, s.draws[0].push)
    ref = refrast.render(s)
    _check(ctx, s, ref=ref, passes=[([a], True), ([b], False)], batched=batched)


def test_clear_values_outside_unit_range(ctx):
    """Clears store trunc(v * 255) without clamping (reference rasterizer.cpp:332-361)."""
    s = _soup(ntris=40, w=48, h=40)
    s.clear_color = (1.25, -0.25, 0.5, 2.0)
    s.clear_depth = 0.75
    _check(ctx, s)


SAMPLERS = [
    dict(mag_filter=FILTER_NEAREST, min_filter=FILTER_NEAREST, mipmap_mode=MIPMAP_NEAREST, address=ADDR_REPEAT),
    dict(mag_filter=FILTER_LINEAR, min_filter=FILTER_NEAREST, mipmap_mode=MIPMAP_LINEAR, address=ADDR_CLAMP),
    dict(mag_filter=FILTER_NEAREST, min_filter=FILTER_LINEAR, mipmap_mode=MIPMAP_NEAREST, address=ADDR_CLAMP,
         lod_bias=1.0),
    dict(mag_filter=FILTER_LINEAR, min_filter=FILTER_LINEAR, mipmap_mode=MIPMAP_LINEAR, address=ADDR_REPEAT,
         lod_bias=-1.5, min_lod=0.5),
]


@pytest.mark.parametrize("k", range(len(SAMPLERS)))
def test_sampler_modes(ctx, k):
    s = scenes.many_draws(ndraws=10, width=160, height=120, tex_size=64, max_tris=1024, seed=21)
    s.textures = [dataclasses.replace(t, **SAMPLERS[k]) for t in s.textures]
    _check(ctx, s)


BC_FORMATS = ["BC2_UNORM_BLOCK", "BC3_UNORM_BLOCK"]


def _bc_draws(fmt, k, seed):
    s = scenes.many_draws(ndraws=10, width=160, height=120, tex_size=64, max_tris=1024, seed=21)
    rng = scenes.Rng(seed)
    s.textures = [dataclasses.replace(t, format=FMT[fmt], data=scenes.bc_texture_data(rng, t.width, t.mips, t.layers),
                                      **SAMPLERS[k]) for t in s.textures]
    return s


@pytest.mark.parametrize("fmt", BC_FORMATS)
@pytest.mark.parametrize("k", [0, 3])
def test_bc_textures(ctx, fmt, k):
    """Random BC2/BC3 blocks (both colour modes, both alpha modes) with mips, sampled nearest or
    trilinear."""
    _check(ctx, _bc_draws(fmt, k, 40 + k))


@pytest.mark.parametrize("fmt", ["R8G8B8A8_UNORM"] + BC_FORMATS)
@pytest.mark.parametrize("batched", [True, False])
def test_cube_env_sphere(ctx, fmt, batched):
    """samplerCube lookups along the interpolated normal of a sphere: every face and the edges
    between faces."""
    _check(ctx, scenes.env_sphere(fmt=FMT[fmt]), batched=batched)


def test_bc_texture_rewritten_between_submits(ctx):
    """A BC texture rewritten by a buffer-to-image copy between two submits: the second render
    samples the new blocks."""
    s = _bc_draws("BC3_UNORM_BLOCK", 0, 50)
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=True, tile=8)
    r.render()
    first = r.read_color()
    assert np.array_equal(first, refrast.render(s)["color"])
    rng = scenes.Rng(51)
    for i, t in enumerate(s.textures):
        s.textures[i] = dataclasses.replace(t, data=scenes.bc_texture_data(rng, t.width, t.mips, t.layers))
        ctx.upload_texture(r.textures[i][0], s.textures[i])
    r.render()
    second = r.read_color()
    assert not np.array_equal(first, second)
    assert np.array_equal(second, refrast.render(s)["color"])


def test_resubmitted_recording_and_descriptor_update(ctx):
    """One recording submitted repeatedly renders the same bytes each time (its decoded form and
    pass tables are reused); a descriptor set pointed at another texture between two submits of
    the same recording is honoured; a re-recording with fewer draws renders only those."""
    import ctypes as C

    from visor_amd import vk
    s = scenes.many_draws(ndraws=8, width=128, height=96, tex_size=64, max_tris=512, seed=33)
    s.draws = [dataclasses.replace(d, texture=k % 2) for k, d in enumerate(s.draws)]  # textures 0 and 1
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=True, tile=8)
    want = refrast.render(s)["color"]
    for _ in range(3):
        r.render()
        assert np.array_equal(r.read_color(), want)
    # set 0 now holds texture 1 (same recording, no re-record)
    img, mem, view, samp = r.textures[1]
    dii = vk.DescriptorImageInfo(sampler=samp, imageView=view, imageLayout=5)
    w = vk.WriteDescriptorSet(dstSet=r.sets[0], dstBinding=0, descriptorCount=1,
                              descriptorType=vk.DESC_COMBINED_IMAGE_SAMPLER, pImageInfo=C.pointer(dii))
    ctx.icd.vkUpdateDescriptorSets(ctx.device, 1, C.byref(w), 0, None)
    s2 = dataclasses.replace(s, draws=[dataclasses.replace(d, texture=1 if d.texture == 0 else d.texture)
                                       for d in s.draws])
    r.render()
    assert np.array_equal(r.read_color(), refrast.render(s2)["color"])
    r.render(draws=s.draws[:3])
    assert np.array_equal(r.read_color(), refrast.render(s2, draws=s2.draws[:3])["color"])


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_slivers_and_far_vertices(ctx, tile):
    """Thin slivers across many tiles (tile masks drop the tiles they miss) and triangles with
    vertices far off screen, beyond the window-coordinate range where masks apply (every bbox tile
    kept): bit-exact either way, at every tile size."""
    rng = scenes.Rng(77)
    n = 300
    pos = np.zeros((3 * n, 4), np.float32)
    base = np.stack([rng.uniform(n, -1.0, 1.0), rng.uniform(n, -1.0, 1.0)], 1)
    ang = rng.uniform(n, 0.0, 6.283)
    length = rng.uniform(n, 0.2, 1.6)
    width = rng.uniform(n, 0.001, 0.02)
    d = np.stack([np.cos(ang), np.sin(ang)], 1) * length[:, None]
    nrm = np.stack([-np.sin(ang), np.cos(ang)], 1) * width[:, None]
    tri = np.stack([base, base + d, base + d * 0.5 + nrm], 1)  # (n, 3, 2)
    far = np.arange(n) % 5 == 0                                 # every 5th: one vertex far away
    tri[far, 2] = tri[far, 2] * 300.0                            # window coordinates ~ +-20000
    pos[:, :2] = tri.reshape(-1, 2)
    pos[:, 2] = np.repeat(rng.uniform(n, 0.0, 1.0), 3)
    pos[:, 3] = 1.0
    col = rng.bytes(3 * n * 4).reshape(3 * n, 4)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color")}
    s = scenes.Scene("slivers", 160, 120, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes,
                     [scenes.Draw("main", 3 * n, 0, False, scenes.push_block(scenes.IDENTITY))])
    ref = refrast.render(s)
    assert ref["counters"]["pixels_written"] > 0
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=True, tile=tile)
    r.render()
    assert np.array_equal(r.read_color(), ref["color"])
    assert np.array_equal(r.read_depth().view(np.uint32), ref["depth"].view(np.uint32))
    st = ctx.stats(0)
    assert st["pixels_written"] == ref["counters"]["pixels_written"]


def test_buffer_to_image_regions(ctx):
    """vkCmdCopyBufferToImage with image offsets, a source row longer than the region
    (bufferRowLength), a mip level and two layers at once writes exactly those texels of the
    linear image layout; a region past the image edge is rejected."""
    import ctypes as C

    from visor_amd import vk
    W, H, mips, layers = 20, 12, 2, 3
    img, mem, view, size = ctx.image(W, H, vk.FMT["R8G8B8A8_UNORM"], mips, layers, host_visible=True)
    lvl = [W * H * 4, (W // 2) * (H // 2) * 4]
    slice_bytes = sum(lvl)
    assert size == slice_bytes * layers
    C.memset(ctx.map(mem, size), 0, size)
    rng = np.random.default_rng(4)
    src = rng.integers(1, 256, 4096, dtype=np.uint8)
    stage, smem = ctx.buffer(src.size, host_visible=True)
    C.memmove(ctx.map(smem, src.size), src.tobytes(), src.size)
    want = np.zeros(size, np.uint8)

    def region(off, row, rows, mip, layer0, nlayers, x, y, w, h, expect=True):
        r = vk.BufferImageCopy(bufferOffset=off, bufferRowLength=row, bufferImageHeight=rows,
                               imageSubresource=vk.ImageSubresourceLayers(aspectMask=1, mipLevel=mip,
                                                                          baseArrayLayer=layer0, layerCount=nlayers),
                               imageOffset=vk.Offset3D(x=x, y=y, z=0), imageExtent=vk.Extent3D(width=w, height=h, depth=1))
        if not expect:
            return r
        lw = max(1, W >> mip)
        rr, rh = row or w, rows or h
        for layer in range(nlayers):
            base = (layer0 + layer) * slice_bytes + sum(lvl[:mip])
            for j in range(h):
                s0 = off + (layer * rr * rh + j * rr) * 4
                d0 = base + ((y + j) * lw + x) * 4
                want[d0:d0 + w * 4] = src[s0:s0 + w * 4]
        return r

    regions = [region(0, 9, 0, 0, 0, 1, 3, 2, 7, 5),      # offset region, longer source rows
               region(400, 0, 0, 1, 1, 2, 1, 1, 8, 4),    # level 1, layers 1-2
               region(1000, 11, 6, 0, 2, 1, 12, 6, 8, 6)]  # bufferImageHeight > height
    cb = ctx.begin()
    arr = (vk.BufferImageCopy * len(regions))(*regions)
    ctx.icd.vkCmdCopyBufferToImage(cb, stage, img, 7, len(regions), arr)
    ctx.end_submit(cb)
    got = np.frombuffer(C.string_at(ctx.map(mem, size), size), np.uint8)
    assert np.array_equal(got, want)
    # past the right edge of level 0: the submit fails and the image is unchanged
    bad = (vk.BufferImageCopy * 1)(region(0, 0, 0, 0, 0, 1, 15, 0, 8, 2, expect=False))
    cb = ctx.begin()
    ctx.icd.vkCmdCopyBufferToImage(cb, stage, img, 7, 1, bad)
    ctx.icd.vkEndCommandBuffer(cb)
    assert ctx.submit([cb]) != 0
    got2 = np.frombuffer(C.string_at(ctx.map(mem, size), size), np.uint8)
    diff = np.flatnonzero(got2 != want)
    assert diff.size == 0, (diff.size, diff[:8].tolist(), got2[diff[:8]].tolist(), want[diff[:8]].tolist())
