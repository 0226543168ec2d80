"""Pins the oracle (oracle/refrast.c) to what the reference's code fixes, without retyping it:
closed forms, brute force over tiny targets, truth tables and algebraic identities. Each check is
chosen so a plausible restatement mistake (floor for trunc, inclusive bbox max, a swapped edge,
a wrong blend factor, a wrong mip) fails it."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import refrast
from visor_amd import scenes
from visor_amd.vk import BF, CMP, CULL_BACK, CULL_FRONT, CULL_NONE, FRONT_CCW, FRONT_CW, TOPO_STRIP

from helpers import covered_mask, inside_closed_triangle, tri_scene

L = refrast.lib()


# ---- window transform (reference rasterizer.cpp:239-253) ----------------------------------------

def to_window(W, H, p):
    out = (C.c_int32 * 2)()
    refrast.lib().rr_to_window(W, H, (C.c_float * 4)(*p), out)
    return out[0], out[1]


def test_to_window_special_values_and_truncation():
    assert to_window(256, 256, (-1, 1, 0, 1)) == (0, 0)
    assert to_window(256, 256, (1, -1, 0, 1)) == (256, 256)
    assert to_window(256, 256, (0, 0, 0, 1)) == (128, 128)
    # y is flipped: clip +y is the top row
    assert to_window(256, 256, (0, 0.5, 0, 1))[1] == 64
    # truncation toward zero, not floor: (-1.001 + 1) * 128 = -0.128 -> 0
    assert to_window(256, 256, (-1.001, 0, 0, 1))[0] == 0
    assert to_window(256, 256, (-1.01, 0, 0, 1))[0] == -1  # -1.28 -> -1
    # perspective divide by w
    assert to_window(256, 256, (2, 0, 0, 2)) == (256, 128)


INT_MIN = -2 ** 31


def test_float_to_int_is_x86_truncation():
    """Float-to-int32 as the reference's x64 build converts (cvttss2si): truncation toward zero
    inside [-2^31, 2^31), and 0x80000000 for NaN, infinities and every value outside (C leaves
    that case undefined; a saturating conversion, which is what a GPU does natively, gives
    INT_MAX or 0 there). It decides far-off vertices (a vertex with w near 0, for which the
    reference does no clipping) and out-of-range clear values."""
    f2i = refrast.lib().rr_f2i
    for x, want in [(0.0, 0), (1.9, 1), (-1.9, -1), (2147483520.0, 2147483520), (-2147483648.0, INT_MIN),
                    (2147483648.0, INT_MIN), (1e12, INT_MIN), (-1e12, INT_MIN), (float("inf"), INT_MIN),
                    (float("-inf"), INT_MIN), (float("nan"), INT_MIN)]:
        assert f2i(x) == want, x
    # ToWindow of a vertex at w = 1e-7: x/w = 5e6 -> window 6.4e8 (in range); x = 100: out of range
    assert to_window(256, 256, (0.5, -0.5, 0, 1e-7)) == (640000128, 640000128)
    assert to_window(256, 256, (100.0, -100.0, 0, 1e-7)) == (INT_MIN, INT_MIN)
    assert to_window(256, 256, (float("nan"), 0.0, 0, 1.0))[0] == INT_MIN
    # the colour clear stores the low byte of the converted value (reference rasterizer.cpp:341-345)
    s = tri_scene([[(0, 0), (2, 0), (0, 2)]], W=16, H=16, clear_color=(1e10, 2.0, -0.5, -1e10))
    # bytes [2]=R: 1e10*255 -> INT_MIN -> 0; [1]=G: 510 -> 254; [0]=B: -127.5 -> -127 -> 129; [3]=A -> 0
    assert refrast.render(s)["color"][15, 15].tolist() == [129, 254, 0, 0]


def test_vertex_offset_is_not_applied():
    """vkCmdDrawIndexed's vertexOffset is recorded but GetIndex never adds it (reference
    rasterizer.cpp:100-119, cmd_exec.cpp:129-142): indices 0, 1, 2 with vertexOffset 3 draw
    vertices 0..2, a right triangle whose coverage has the closed form below, not vertices 3..5."""
    n = 10
    s = tri_scene([[(0, 0), (n, 0), (0, n)], [(16, 16), (30, 16), (16, 30)]], W=32, H=32)
    s.index_data = np.array([0, 1, 2], np.uint32).tobytes()
    s.index_type = 1
    s.draws = [scenes.Draw("p", 3, 0, True, s.draws[0].push, vertex_offset=3)]
    r = refrast.render(s)
    expect = np.zeros((32, 32), bool)
    for y in range(n):
        for x in range(n):
            expect[y, x] = x + y <= n
    assert np.array_equal(covered_mask(r["color"]), expect)


# ---- coverage: closed form, brute force, bbox rules ---------------------------------------------

@pytest.mark.parametrize("n", [1, 2, 5, 10, 17])
def test_right_triangle_closed_form(n):
    """(0,0),(n,0),(0,n): integer-corner samples, inclusive edges, exclusive bbox maximum give
    exactly n + n(n+1)/2 - 1 pixels: {x < n, y < n, x + y <= n}."""
    s = tri_scene([[(0, 0), (n, 0), (0, n)]], W=32, H=32)
    r = refrast.render(s)
    cov = covered_mask(r["color"])
    expect = np.zeros((32, 32), bool)
    for y in range(n):
        for x in range(n):
            if x + y <= n:
                expect[y, x] = True
    assert np.array_equal(cov, expect)
    assert r["counters"]["pixels_written"] == n + n * (n + 1) // 2 - 1
    assert r["counters"]["pixels_tested"] == n * n


def test_full_screen_pair_never_touches_last_column_or_row():
    W = H = 32
    s = tri_scene([[(0, 0), (W, 0), (0, H)], [(W, 0), (W, H), (0, H)]], W=W, H=H)
    cov = covered_mask(refrast.render(s)["color"])
    assert not cov[:, W - 1].any() and not cov[H - 1, :].any()
    assert cov[:H - 1, :W - 1].all()


def test_brute_force_coverage_small_targets():
    """Random triangles on a 16x16 target: the covered set equals the closed-triangle membership
    of integer corners restricted to [min, max) of the clamped bbox, by exact orientation tests."""
    rng = np.random.default_rng(7)
    W = H = 16
    for trial in range(60):
        pts = [tuple(int(v) for v in rng.integers(-6, 22, size=2)) for _ in range(3)]
        (ax, ay), (bx, by), (cx, cy) = pts
        area = (bx - ax) * (cy - ay) - (by - ay) * (cx - ax)
        s = tri_scene([pts], W=W, H=H)
        cov = covered_mask(refrast.render(s)["color"])
        expect = np.zeros((H, W), bool)
        if area != 0:
            minx, maxx = max(0, min(ax, bx, cx)), min(W - 1, max(ax, bx, cx))
            miny, maxy = max(0, min(ay, by, cy)), min(H - 1, max(ay, by, cy))
            for y in range(miny, maxy):
                for x in range(minx, maxx):
                    expect[y, x] = inside_closed_triangle(pts, x, y)
        assert np.array_equal(cov, expect), (trial, pts)


def test_shared_edge_is_drawn_twice():
    """No top-left rule: with additive blending, pixels on the shared diagonal get both triangles."""
    pipe = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", depth_op=CMP["ALWAYS"], blend=True, src=BF["ONE"], dst=BF["ONE"])
    c = (40, 40, 40, 255)
    s = tri_scene([[(0, 0), (8, 0), (0, 8)], [(8, 0), (8, 8), (0, 8)]], W=16, H=16, colors=[c] * 6, pipeline=pipe)
    col = refrast.render(s)["color"]
    # pixel (4,4) lies on x + y = 8 (both triangles), (1,1) only in the first, (6,6) only in the second
    assert col[4, 4, 0] == 80 and col[1, 1, 0] == 40 and col[6, 6, 0] == 40


# ---- facing and culling truth table (rasterizer.cpp:395-426) ------------------------------------

@pytest.mark.parametrize("front", [FRONT_CCW, FRONT_CW])
@pytest.mark.parametrize("cull", [CULL_NONE, CULL_FRONT, CULL_BACK, 3])
@pytest.mark.parametrize("positive_area", [True, False])
def test_cull_truth_table(front, cull, positive_area):
    tri = [(2, 2), (12, 2), (2, 12)]  # area2 = 100 > 0 in window coordinates
    if not positive_area:
        tri = [tri[0], tri[2], tri[1]]
    pipe = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", cull=cull, front=front)
    r = refrast.render(tri_scene([tri], W=16, H=16, pipeline=pipe))
    is_front = positive_area == (front == FRONT_CCW)
    culled = bool(cull & CULL_FRONT) if is_front else bool(cull & CULL_BACK)
    assert (r["counters"]["pixels_written"] == 0) == culled
    assert r["counters"]["tris_out"] == (0 if culled else 1)


def test_strip_keeps_winding():
    """A 6-vertex strip is 4 triangles; with back-face culling all survive only if every second
    triangle's vertex order is flipped as the reference does."""
    strip = [(0, 0), (0, 8), (8, 0), (8, 8), (16, 0), (16, 8)]
    pipe_b = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", topology=TOPO_STRIP, cull=CULL_BACK)
    pipe_f = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", topology=TOPO_STRIP, cull=CULL_FRONT)
    rb = refrast.render(tri_scene(None, W=32, H=32, strip=strip, pipeline=pipe_b))["counters"]
    rf = refrast.render(tri_scene(None, W=32, H=32, strip=strip, pipeline=pipe_f))["counters"]
    assert rb["tris_in"] == rf["tris_in"] == 4
    assert {rb["tris_out"], rf["tris_out"]} == {0, 4}


# ---- depth (rasterizer.cpp:551-576, 680-683) ------------------------------------------------------

def _two_overlapping(op, z_first, z_second, write=True, clear_depth=1.0):
    pipe = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", depth_op=CMP[op], depth_write=write)
    t = [(0, 0), (16, 0), (0, 16)]
    c1, c2 = (200, 0, 0, 255), (0, 0, 200, 255)
    s = tri_scene([t, t], W=16, H=16, colors=[c1] * 3 + [c2] * 3, depths=[z_first] * 3 + [z_second] * 3,
                  pipeline=pipe, clear_depth=clear_depth)
    return refrast.render(s)


def test_depth_ops():
    # colour bytes are B,G,R: red input lands in byte 2, blue in byte 0
    assert _two_overlapping("LESS", 0.3, 0.6)["color"][2, 2].tolist()[:3] == [0, 0, 200]
    assert _two_overlapping("LESS", 0.6, 0.3)["color"][2, 2].tolist()[:3] == [200, 0, 0]
    assert not covered_mask(_two_overlapping("GREATER", 0.6, 0.3)["color"]).any()  # nothing beats 1.0
    assert _two_overlapping("GREATER", 0.3, 0.6, clear_depth=0.0)["color"][2, 2].tolist()[:3] == [200, 0, 0]
    assert _two_overlapping("GEQUAL", 0.6, 0.3, clear_depth=0.0)["color"][2, 2].tolist()[:3] == [0, 0, 200]
    r = _two_overlapping("NEVER", 0.3, 0.6)
    assert not covered_mask(r["color"]).any() and r["counters"]["depth_passed"] == 0
    assert r["counters"]["pixels_written"] > 0  # counted before the depth test
    # ALWAYS: last writer wins and depth is not read
    assert _two_overlapping("ALWAYS", 0.3, 0.6)["color"][2, 2].tolist()[:3] == [200, 0, 0]
    # LESS without depth writes: both pass against the clear value, the second wins
    assert _two_overlapping("LESS", 0.3, 0.6, write=False)["color"][2, 2].tolist()[:3] == [200, 0, 0]
    # EQUAL: a flat triangle's interpolated depth equals what it wrote
    r = _two_overlapping("LEQUAL", 0.5, 0.5)
    assert r["color"][2, 2].tolist()[:3] == [200, 0, 0]


def test_flat_depth_is_close_to_vertex_depth():
    r = _two_overlapping("LESS", 0.25, 0.9)
    d = r["depth"][covered_mask(r["color"])]
    assert np.all(np.abs(d - 0.25) <= 2 * np.spacing(np.float32(0.25)))


# ---- blending and stores (rasterizer.cpp:593-676) ------------------------------------------------

def _blend(src, dst, alpha, under=(100, 150, 200, 255), over=(50, 60, 70, 255)):
    under_pipe = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", depth_op=CMP["ALWAYS"])
    over_pipe = scenes.Pipeline("vs_mvp_pos4_col", "fs_color", depth_op=CMP["ALWAYS"], blend=True, src=src, dst=dst)
    t = [(0, 0), (16, 0), (0, 16)]
    over = over[:3] + (alpha,)
    s = tri_scene([t, t], W=16, H=16, colors=[under] * 3 + [over] * 3, draws_per_tri=True,
                  pipelines={"u": under_pipe, "o": over_pipe}, draw_pipes=["u", "o"])
    return refrast.render(s)["color"][2, 2]


def test_blend_identities():
    under, over = (100, 150, 200, 255), (50, 60, 70, 255)
    # bytes are stored B, G, R
    opaque = [over[2], over[1], over[0]]
    assert _blend(BF["ONE"], BF["ZERO"], 255).tolist()[:3] == opaque
    assert _blend(BF["SRC_ALPHA"], BF["ONE_MINUS_SRC_ALPHA"], 255).tolist()[:3] == opaque
    assert _blend(BF["SRC_ALPHA"], BF["ONE_MINUS_SRC_ALPHA"], 0).tolist()[:3] == [under[2], under[1], under[0]]
    assert _blend(BF["ZERO"], BF["ONE"], 77).tolist()[:3] == [under[2], under[1], under[0]]
    add = _blend(BF["ONE"], BF["ONE"], 255).tolist()[:3]
    assert add == [min(255, under[2] + over[2]), min(255, under[1] + over[1]), min(255, under[0] + over[0])]


def test_alpha_byte_untouched_by_draws_and_clear_order():
    s = tri_scene([[(0, 0), (16, 0), (0, 16)]], W=16, H=16, colors=[(10, 20, 30, 40)] * 3,
                  clear_color=(0.5, 0.25, 1.0, 0.75))
    col = refrast.render(s)["color"]
    # clear: [2]=R, [1]=G, [0]=B, [3]=A, truncated: 127, 63, 255, 191
    assert col[15, 15].tolist() == [255, 63, 127, 191]
    # a draw writes B,G,R from the fragment colour and leaves alpha at the clear value
    assert col[2, 2].tolist() == [30, 20, 10, 191]


def test_unorm_store_truncates():
    # colour 254/255 interpolated and stored: trunc(clamp01(254/255) * 255) = 253 or 254, never 255
    s = tri_scene([[(0, 0), (16, 0), (0, 16)]], W=16, H=16, colors=[(254, 254, 254, 254)] * 3)
    v = refrast.render(s)["color"][3, 3, 0]
    assert v in (253, 254)


# ---- texture sampling (texture_sampling.cpp:139-184) and the level-selection extension -----------

def _tex_image(levels):
    """RGBA8 image from a list of (h, w, 4) uint8 levels, packed."""
    data = np.concatenate([l.reshape(-1) for l in levels]).astype(np.uint8)
    h, w = levels[0].shape[:2]
    return data, refrast.make_image(data, w, h, mips=len(levels))


def _sample(img, u, v, ddx=(0, 0), ddy=(0, 0), lod=False, mag=1, min_=1, mip=0, addr=0, max_lod=0.0, layer=0.0):
    s = refrast.Sampler(mag_filter=mag, min_filter=min_, mipmap_mode=mip, address_u=addr, address_v=addr,
                        mip_lod_bias=0.0, min_lod=0.0, max_lod=max_lod)
    out = (C.c_float * 4)()
    L.rr_sample(C.byref(img), C.byref(s), (C.c_float * 3)(u, v, layer), (C.c_float * 2)(*ddx), (C.c_float * 2)(*ddy),
                1 if lod else 0, out)
    return np.array(out[:])


def test_constant_texture_is_constant():
    lvl = np.full((8, 8, 4), 77, np.uint8)
    keep, img = _tex_image([lvl])
    for u, v in [(0.0, 0.0), (0.3, 0.9), (-2.7, 5.1), (0.999, 0.001)]:
        assert np.all(_sample(img, u, v) == np.float32(77) / np.float32(255))


def test_texel_exact_at_integer_coordinates_and_wrap():
    rng = np.random.default_rng(3)
    lvl = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    keep, img = _tex_image([lvl])
    for i in range(8):
        for j in range(8):
            got = _sample(img, i / 8.0, j / 8.0)
            assert np.array_equal(got, lvl[j, i].astype(np.float32) / np.float32(255)), (i, j)
            # period 1 in u and v (REPEAT), with exactly representable shifts
            assert np.array_equal(_sample(img, i / 8.0 + 3.0, j / 8.0 - 2.0), got)


def test_bilinear_matches_float64_bilinear():
    rng = np.random.default_rng(4)
    lvl = rng.integers(0, 256, (8, 8, 4), dtype=np.uint8)
    keep, img = _tex_image([lvl])
    t = lvl.astype(np.float64) / 255.0
    for _ in range(200):
        u, v = rng.random(), rng.random()
        U, V = u * 8, v * 8
        i0, j0 = int(U), int(V)
        fu, fv = U - i0, V - j0
        i1, j1 = (i0 + 1) % 8, (j0 + 1) % 8
        ref = (t[j0, i0] * (1 - fu) + t[j0, i1] * fu) * (1 - fv) + (t[j1, i0] * (1 - fu) + t[j1, i1] * fu) * fv
        assert np.allclose(_sample(img, np.float32(u), np.float32(v)), ref, atol=2e-6)


def test_clamp_and_nearest_extension():
    lvl = np.zeros((4, 4, 4), np.uint8)
    lvl[:, 3] = 200
    keep, img = _tex_image([lvl])
    # clamp: u beyond 1 sticks to the last column; repeat wraps to the first
    assert _sample(img, 1.5, 0.0, addr=2)[0] == np.float32(200) / np.float32(255)
    assert abs(_sample(img, 1.5, 0.1, addr=2)[0] - 200 / 255) < 1e-6
    assert _sample(img, 1.1, 0.1, addr=0, mag=0)[0] == 0.0
    # nearest picks one texel: u = 0.9 -> column 3
    assert _sample(img, 0.9, 0.1, mag=0)[0] == np.float32(200) / np.float32(255)


def test_level_selection_by_derivatives():
    """Levels filled with the constant 10*level: the LOD of a gradient of 2^k texels per pixel
    selects level k (NEAREST), and halfway gradients blend adjacent levels (LINEAR)."""
    size = 64
    levels = [np.full((size >> m, size >> m, 4), 10 * m, np.uint8) for m in range(7)]
    keep, img = _tex_image(levels)
    for k in range(7):
        d = (2.0 ** k) / size
        got = _sample(img, 0.3, 0.3, ddx=(d, 0), ddy=(0, d), lod=True, min_=1, mip=0, max_lod=6)
        assert abs(got[0] * 255 - 10 * k) < 1e-3, k
    # lambda = 2.5 with LINEAR: halfway between levels 2 and 3
    d = (2.0 ** 2.5) / size
    got = _sample(img, 0.3, 0.3, ddx=(d, 0), ddy=(0, d), lod=True, mip=1, max_lod=6)
    assert abs(got[0] * 255 - 25.0) < 0.01
    # max_lod clamps
    d = (2.0 ** 6) / size
    got = _sample(img, 0.3, 0.3, ddx=(d, 0), ddy=(0, d), lod=True, mip=0, max_lod=2)
    assert abs(got[0] * 255 - 20) < 1e-3


def test_array_layer_select():
    size = 4
    lay = [np.full((size, size, 4), 30 * (l + 1), np.uint8) for l in range(3)]
    data = np.concatenate([x.reshape(-1) for x in lay])
    img = refrast.make_image(data, size, size, mips=1, layers=3)
    for layer, expect in [(0.0, 30), (1.4, 60), (1.6, 90), (7.0, 90), (-3.0, 30)]:
        assert abs(_sample(img, 0.2, 0.2, layer=layer)[0] * 255 - expect) < 1e-3


def test_log2_accuracy_and_exact_powers():
    for e in range(-20, 21):
        assert L.rr_log2(2.0 ** e) == e
    xs = np.exp(np.random.default_rng(5).uniform(-30, 30, 2000)).astype(np.float32)
    err = max(abs(L.rr_log2(float(x)) - math.log2(float(x))) for x in xs)
    assert err < 2e-5
    assert L.rr_log2(0.0) == -128.0 and L.rr_log2(-1.0) == -128.0


def test_half_conversion_matches_numpy_rne():
    # every finite half round-trips
    for h in range(0, 0x7c00):
        for sign in (0, 0x8000):
            f = L.rr_f16_to_f32(h | sign)
            assert L.rr_f32_to_f16(f) == (h | sign)
    xs = np.concatenate([np.random.default_rng(6).standard_normal(20000).astype(np.float32) * 1e3,
                         np.float32([65504, 65519.99, 65520, 1e-8, 5.96e-8, 2.98e-8, -0.0, 1e30, np.inf, -np.inf])])
    for x in xs:
        assert L.rr_f32_to_f16(float(x)) == int(np.float32(x).astype(np.float16).view(np.uint16)), x


def test_subresource_offsets():
    img = refrast.make_image(np.zeros(1, np.uint8), 16, 8, mips=4, layers=3)
    # level sizes 16x8, 8x4, 4x2, 2x1 at 4 bytes
    sizes = [16 * 8 * 4, 8 * 4 * 4, 4 * 2 * 4, 2 * 1 * 4]
    for m in range(4):
        assert L.rr_subresource_offset(C.byref(img), m, 0) == sum(sizes[:m])
    assert L.rr_subresource_offset(C.byref(img), 2, 2) == 2 * sum(sizes) + sum(sizes[:2])


def test_threaded_mode_coverage_counters_match_serial():
    s = scenes.random_triangles(ntris=200, width=128, height=128, seed=9)
    a = refrast.render(s)["counters"]
    b = refrast.render(s, threads=4, time_only=True)["counters"]
    for k in ("tris_in", "tris_out", "blocks", "pixels_tested", "pixels_written"):
        assert a[k] == b[k], k


def test_random_triangles_expected_scale():
    """1000 random clip-space triangles on 256x256: about 5.0M covered, 0.3M shaded."""
    c = refrast.render(scenes.random_triangles())["counters"]
    assert 4.5e6 < c["pixels_written"] < 5.5e6
    assert 2.0e5 < c["depth_passed"] < 4.0e5
