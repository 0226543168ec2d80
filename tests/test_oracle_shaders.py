"""Pins the oracle's restated shader programs (RR_VS_*, RR_FS_*; reference spirv_compile.cpp
lowering :423-643, :1056-1888, :1912-2366) by closed forms and special cases rather than by
retyping them: interpolation against float64 perspective-correct barycentrics, the colour of the
vertex at an exactly sampled corner, Lambert at its extremes (normal along, against and across
the light), the push-constant alpha through blending, tap weights summing to one over a constant
array texture and matching the SPIR-V program's weight and layer tables over distinct layers, and
an MVP translation that moves the image by whole pixels.
Colour bytes are compared within 1 LSB where float rounding of an exact real value can go
either way before truncation."""
import dataclasses

import numpy as np

from oracle import refrast
from visor_amd import scenes, shaders
from visor_amd.vk import BF, CMP

W = H = 64
LIGHT = np.array([0.48, 0.6, 0.64])  # the lit programs' light direction (unit length)


def ndc(px, py):
    """NDC x, y whose window transform lands exactly on integer pixel corner (px, py)."""
    return 2.0 * px / W - 1.0, -(2.0 * py / H - 1.0)


def color_scene(corners, colours, w=(1.0, 1.0, 1.0), mvp=scenes.IDENTITY):
    pos = np.zeros((3, 4), np.float32)
    for k, (px, py) in enumerate(corners):
        x, y = ndc(px, py)
        pos[k] = (x * w[k], y * w[k], 0.5 * w[k], w[k])
    col = np.array(colours, np.uint8)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color", depth_op=CMP["ALWAYS"])}
    return scenes.Scene("tri", W, H, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes,
                        [scenes.Draw("main", 3, 0, False, scenes.push_block(mvp))])


def bgr(img, x, y):
    """Shader output (r, g, b) from a B8G8R8A8 target pixel (the reference stores [2]=R [1]=G [0]=B)."""
    p = img[y, x].astype(int)
    return np.array([p[2], p[1], p[0]])


COLOURS = [(200, 10, 30, 255), (5, 180, 60, 255), (90, 40, 250, 255)]


def test_vertex_colour_at_its_corner():
    """The pixel on the vertex at the bbox's (inclusive) minimum corner has barycentrics (1, 0, 0)
    up to rounding: its colour is that vertex's colour, whichever of the three slots it is in."""
    corner, a, b = (8, 8), (40, 12), (12, 44)
    for k in range(3):
        corners = [a, b]
        corners.insert(k, corner)
        img = refrast.render(color_scene(corners, COLOURS))["color"]
        got = bgr(img, *corner)
        assert np.abs(got - np.array(COLOURS[k][:3])).max() <= 1, (k, got)


def _bary64(corners, w, px, py):
    """Perspective-correct barycentrics at integer corner (px, py), in float64."""
    (x0, y0), (x1, y1), (x2, y2) = corners
    area = (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0)
    l1 = ((px - x0) * (y2 - y0) - (py - y0) * (x2 - x0)) / area
    l2 = ((x1 - x0) * (py - y0) - (y1 - y0) * (px - x0)) / area
    lam = np.array([1.0 - l1 - l2, l1, l2])
    persp = lam / np.array(w)
    return persp / persp.sum(), lam.min() >= 0


def test_interpolation_matches_float64_barycentrics():
    """Every covered pixel's colour equals the float64 perspective-correct mix of the vertex
    colours within 1 LSB, with w = 1 and with unequal w (perspective division)."""
    corners = [(6, 5), (58, 20), (15, 60)]
    for w in [(1.0, 1.0, 1.0), (1.0, 2.5, 0.4)]:
        img = refrast.render(color_scene(corners, COLOURS, w=w))["color"]
        cols = np.array(COLOURS, np.float64)[:, :3] / 255.0
        checked = 0
        for py in range(0, H, 3):
            for px in range(0, W, 3):
                b, inside = _bary64(corners, w, px, py)
                if not inside or min(b) < 0.02:  # skip edges, where coverage rules decide
                    continue
                if px >= max(c[0] for c in corners) or py >= max(c[1] for c in corners):
                    continue
                want = np.floor(np.clip(b @ cols, 0, 1) * 255.0)
                got = bgr(img, px, py)
                assert np.abs(got - want).max() <= 1, (w, px, py, got, want)
                checked += 1
        assert checked > 40


def depth_scene(corners, zndc, w):
    """One triangle at exact pixel corners with per-vertex NDC depth zndc and clip w (x, y and z
    scaled by w, so the perspective divide returns the NDC values)."""
    pos = np.zeros((3, 4), np.float32)
    for k, (px, py) in enumerate(corners):
        x, y = ndc(px, py)
        pos[k] = (x * w[k], y * w[k], zndc[k] * w[k], w[k])
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color", depth_op=CMP["ALWAYS"], depth_write=True)}
    return scenes.Scene("tri", W, H, scenes.interleave_pos_col(pos, np.array(COLOURS, np.uint8)), 20,
                        scenes.VATTRS_POS_COL, pipes, [scenes.Draw("main", 3, 0, False, scenes.push_block(scenes.IDENTITY))])


def test_depth_matches_screen_affine_interpolation():
    """Stored depth is the screen-space (pre-perspective) barycentric mix of the vertices' z/w
    (reference rasterizer.cpp:551-558): with sloped z and strongly unequal w it equals the
    float64 affine interpolation within a few float32 ulps, which a perspective-corrected depth
    (weights divided by w) misses by orders of magnitude."""
    corners = [(6, 5), (58, 20), (15, 60)]
    zndc = np.array([0.1, 0.55, 0.9], np.float32)
    w = (1.0, 2.5, 0.375)  # exact in binary32, so x * w / w returns x and the corners stay exact
    r = refrast.render(depth_scene(corners, zndc, w))
    dep = r["depth"]
    checked, spread = 0, []
    for py in range(H):
        for px in range(W):
            _, inside = _bary64(corners, (1.0, 1.0, 1.0), px, py)
            if not inside or px >= 58 or py >= 60:
                continue
            lam, _ = _bary64(corners, (1.0, 1.0, 1.0), px, py)  # w = 1: the screen-affine weights
            want = float(lam @ zndc.astype(np.float64))
            got = float(dep[py, px])
            assert abs(got - want) <= 4 * np.spacing(np.float32(1.0)) / 2, (px, py, got, want)
            checked += 1
            spread.append(got)
    assert checked > 1000 and max(spread) - min(spread) > 0.5


def test_mvp_translation_moves_the_image_by_whole_pixels():
    """A translation of 8 pixels (0.25 in NDC at 64 px) in the push-constant MVP, column-major as
    the shaders read it, moves colour and coverage by exactly 8 pixels."""
    corners = [(6, 5), (40, 14), (12, 44)]
    base = refrast.render(color_scene(corners, COLOURS))["color"]
    mvp = np.eye(4)
    mvp[0, 3] = 0.25   # +8 px in x
    mvp[1, 3] = -0.25  # window y grows downwards: +8 px in y
    moved = refrast.render(color_scene(corners, COLOURS, mvp=mvp))["color"]
    assert np.array_equal(moved[8:, 8:], base[:-8, :-8])
    assert not moved[:8].any() and not moved[:, :8].any()


def ubo_scene(corners, colours, mvp, tint, bias):
    """color_scene drawn with the uniform-buffer programs: MVP, tint and bias from set 0 binding 1."""
    s = color_scene(corners, colours)
    s.pipelines = {"main": scenes.Pipeline("vs_ubo_pos4_col", "fs_ubo_tint", depth_op=CMP["ALWAYS"])}
    s.uniforms = [shaders.ubo_block_bytes(mvp, tint, bias)]
    s.draws = [dataclasses.replace(s.draws[0], push=b"\0" * 128, uniform=0)]
    return s


def test_uniform_buffer_programs():
    """The UBO-fed programs read their block through the descriptor (reference
    spirv_compile.cpp:552-558): with tint 1 and bias 0 they render the push-constant programs'
    image bit for bit (also under an 8-pixel MVP translation read from the buffer), and a flat
    colour c comes out as trunc(clamp01(c/255 * tint + bias) * 255) within 1 LSB."""
    corners = [(6, 5), (40, 14), (12, 44)]
    mvp = np.eye(4)
    mvp[0, 3], mvp[1, 3] = 0.25, -0.25
    for m in (np.eye(4), mvp):
        want = refrast.render(color_scene(corners, COLOURS, mvp=m))["color"]
        got = refrast.render(ubo_scene(corners, COLOURS, m, (1, 1, 1, 1), (0, 0, 0, 0)))["color"]
        assert np.array_equal(got, want)
    flat = [(200, 100, 40, 255)] * 3
    tint, bias = (0.5, 0.25, 1.5, 1.0), (0.25, 0.5, -0.125, 0.0)
    img = refrast.render(ubo_scene(corners, flat, np.eye(4), tint, bias))["color"]
    want = np.floor(np.clip(np.array(flat[0][:3]) / 255.0 * tint[:3] + bias[:3], 0, 1) * 255)
    _check_uniform(img, want)


def lit_scene(fs, normals, tex_value=(100, 150, 200, 255), nmat=np.eye(3), alpha=1.0, blend=False, layers=1,
              tex_size=8):
    corners = [(4, 4), (60, 8), (10, 58)]
    pos = np.zeros((3, 3), np.float32)
    for k, (px, py) in enumerate(corners):
        pos[k, :2] = ndc(px, py)
        pos[k, 2] = 0.5
    nrm = np.array(normals, np.float32)
    uv = np.array([(0.1, 0.2), (0.7, 0.3), (0.4, 0.9)], np.float32)
    texels = np.tile(np.array(tex_value, np.uint8), (layers * tex_size * tex_size, 1))
    tex = scenes.Texture(tex_size, tex_size, 1, layers, texels.tobytes())
    pipe = scenes.Pipeline("vs_mvp_normal_uv", fs, depth_op=CMP["ALWAYS"], blend=blend,
                           src=BF["SRC_ALPHA"] if blend else BF["ONE"],
                           dst=BF["ONE_MINUS_SRC_ALPHA"] if blend else BF["ZERO"])
    return scenes.Scene("lit", W, H, scenes.interleave_pnu(pos, nrm, uv), 32, scenes.VATTRS_PNU, {"p": pipe},
                        [scenes.Draw("p", 3, 0, False, scenes.push_block(scenes.IDENTITY, nmat, alpha), texture=0)],
                        textures=[tex])


def _covered(img):
    return img[..., :3].any(-1)


def _check_uniform(img, want_rgb):
    cov = _covered(img)
    assert cov.sum() > 500
    got = img[cov][:, [2, 1, 0]].astype(int)
    assert np.abs(got - np.array(want_rgb)).max() <= 1, (got.min(0), got.max(0), want_rgb)


def test_lambert_extremes():
    """light = max(dot(normalize(n), L), 0) * 0.75 + 0.25 times a constant texture: 1 along L
    (also for a non-unit normal and through a scaling normal matrix), 0.25 against L and across
    it; alpha comes out 1 (no blending here, so only colour is checked)."""
    t = np.array([100, 150, 200])
    for normals, nmat, scale in [([LIGHT] * 3, np.eye(3), 1.0), ([3.0 * LIGHT] * 3, np.eye(3), 1.0),
                                 ([LIGHT] * 3, 2.0 * np.eye(3), 1.0), ([-LIGHT] * 3, np.eye(3), 0.25),
                                 ([(0.8, -0.64, 0.0)] * 3, np.eye(3), 0.25)]:
        img = refrast.render(lit_scene("fs_lambert_tex", normals, nmat=nmat))["color"]
        _check_uniform(img, np.floor(t * scale))


def test_lambert_through_a_rotation_normal_matrix():
    """The normal matrix is applied as M * n with M's columns 16 bytes apart in the push block
    (reference spirv_compile.cpp:433-441): through a general rotation (not symmetric, so M^T n
    differs) each lit result equals the float64 Lambert of normalize(M n) within 1 LSB."""
    t = np.array([100, 150, 200])
    rot = scenes.rotation(0.4, -0.9, 1.3)
    for n in [(0.6, 0.0, 0.8), (0.0, -0.28, 0.96), (-0.36, 0.48, 0.8)]:
        m = rot @ np.array(n)
        m /= np.linalg.norm(m)
        light = max(float(m @ LIGHT), 0.0) * 0.75 + 0.25
        img = refrast.render(lit_scene("fs_lambert_tex", [n] * 3, nmat=rot))["color"]
        _check_uniform(img, np.floor(t * light))


def test_push_constant_alpha_blends():
    """fs_lambert_tex_a takes alpha from push-constant offset 112: over a black clear with
    SRC_ALPHA / ONE_MINUS_SRC_ALPHA the result is alpha x colour."""
    t = np.array([100, 150, 200])
    for alpha in (0.5, 0.25):
        img = refrast.render(lit_scene("fs_lambert_tex_a", [LIGHT] * 3, alpha=alpha, blend=True))["color"]
        _check_uniform(img, np.floor(t * alpha))


def test_texture_only_and_multitap_weights_sum_to_one():
    """fs_tex_only over a constant texture is the texel; fs_multitap8's eight weighted taps over a
    constant 4-layer array texture sum to the texel (weights 1/4, 5 x 1/8, 2 x 1/16)."""
    t = np.array([100, 150, 200])
    img = refrast.render(lit_scene("fs_tex_only", [LIGHT] * 3))["color"]
    _check_uniform(img, t)
    img = refrast.render(lit_scene("fs_multitap8", [LIGHT] * 3, layers=4))["color"]
    _check_uniform(img, t)


def test_multitap_layer_weights():
    """Layers of distinct constant colour: fs_multitap8's result is the sum of each tap's weight
    times its layer's colour, with the weights and layers of the SPIR-V program's tables."""
    from visor_amd import shaders
    layer_rgb = np.array([(40, 200, 10), (80, 20, 120), (160, 90, 240), (240, 140, 60)], np.float64)
    tex_size = 8
    texels = np.zeros((4, tex_size * tex_size, 4), np.uint8)
    for layer in range(4):
        texels[layer, :, :3] = layer_rgb[layer]
        texels[layer, :, 3] = 255
    s = lit_scene("fs_multitap8", [LIGHT] * 3, layers=4)
    s.textures = [dataclasses.replace(s.textures[0], data=texels.tobytes())]
    want = np.zeros(3)
    for w, layer in zip(shaders.TAP_WEIGHT, shaders.TAP_LAYER):
        want += w * layer_rgb[int(layer)]
    img = refrast.render(s)["color"]
    _check_uniform(img, np.floor(want))
