"""The SPIR-V translator's semantics on the GPU: every test of test_translator_semantics.py runs
again with its fragment program translated, built with hiprtc for gfx950 and executed by the
raster kernel through the ICD (vkCreateGraphicsPipelines, vkCmdDraw, vkQueueSubmit), and its
outputs are held to the same expectations (numpy float32 / int32 semantics, the oracle's samplers).

How the outputs come back exactly: each program's final OpStore is replaced by one that bitcasts the
output component chosen by an extra flat input to uint32 and writes its four bytes as the floats
0..255 into an RGBA16F target (exact in half precision). Each (input row, component) is one pixel,
drawn as the triangle (X, 0), (X + 1, 0), (X, 1) in window coordinates: its only pixel is X's, whose
corner is vertex 0, so the interpolation weights are exactly (1, 0, 0) as in the host harness and
the neighbour weights that feed derivatives are (0, 1, 0) and (0, 0, 1) (flat inputs: derivatives 0).
"""
import functools
import itertools

import numpy as np
import pytest

import test_translator_semantics as T
from helpers import clip_from_window
from test_translator_semantics import lib  # noqa: F401  (fixture used by the wrapped tests)
from visor_amd import render, scenes, shaders, spirv
from visor_amd.vk import CMP, FMT

pytestmark = pytest.mark.gpu

_names = itertools.count()


def packed_finish(self, comps):
    """FS.finish for the GPU: output = the 4 bytes of comps[j], j from the input after the
    program's own (flat, exact)."""
    m = self.m
    sel = m.input(self.v4, location=len(self.ins))
    j = m.op("CompositeExtract", self.f, m.load(self.v4, sel), 0)
    chosen = comps[3]
    for k in (2, 1, 0):
        chosen = m.op("Select", self.f, m.op("FOrdEqual", self.b, j, self.c(float(k))), comps[k], chosen)
    bits = m.op("Bitcast", self.u, chosen)
    outs = []
    for k in range(4):
        byte = m.op("BitwiseAnd", self.u, m.op("ShiftRightLogical", self.u, bits, m.const_u(8 * k)), m.const_u(255))
        outs.append(m.op("ConvertUToF", self.f, byte))
    m.store(self.out, m.op("CompositeConstruct", self.v4, *outs))
    m.end_main("Fragment", "main", self.ins + [sel, self.out], frag=True)
    return m.to_bytes()


def passthrough_vs(n):
    """gl_Position = input 0; outputs 0..n-1 = inputs 1..n (vec4)."""
    m = spirv.Module()
    v4 = m.vec(4)
    pos = m.input(v4, location=0)
    ins = [m.input(v4, location=1 + k) for k in range(n)]
    outs = [m.output(v4, location=k) for k in range(n)]
    pv = shaders._per_vertex_position(m)
    m.begin_main()
    m.store(m.access("Output", v4, pv, m.const_i(0)), m.load(v4, pos))
    for a, o in zip(ins, outs):
        m.store(o, m.load(v4, a))
    m.end_main("Vertex", "main", [pos] + ins + outs + [pv])
    return m.to_bytes()


def gpu_run_fs_rows(ctx, L, tmp_path, spv, rows, push=b"", tex=None):
    """run_fs_rows on the GPU (same arguments and result: outputs of shape (len(rows), 4))."""
    nin = len(rows[0])
    R = len(rows)
    W = 1 << (4 * R).bit_length()  # a power of two above 4R: the last pixel is never the last column
    H = 2
    k = next(_names)
    vs_name, fs_name = "gputest_vs_%d" % k, "gputest_fs_%d" % k
    vs = passthrough_vs(nin + 1)
    shaders.PROGRAMS[vs_name] = lambda: vs
    shaders.PROGRAMS[fs_name] = lambda: spv
    verts = []
    for r, row in enumerate(rows):
        vals = [float(np.float32(x)) for v in row for x in (list(v) + [0.0] * 4)[:4]]
        for j in range(4):
            X = 4 * r + j
            for vx, vy in ((X, 0), (X + 1, 0), (X, 1)):
                x, y = clip_from_window(vx, vy, W, H)
                verts.append([x, y, 0.5, 1.0] + vals + [float(j), 0.0, 0.0, 0.0])
    vdata = np.array(verts, np.float32)
    vattrs = [(loc, FMT["R32G32B32A32_SFLOAT"], 16 * loc) for loc in range(nin + 2)]
    textures = []
    if tex is not None:
        fmt = {0: FMT["R8G8B8A8_UNORM"], 1: FMT["BC2_UNORM_BLOCK"], 2: FMT["BC3_UNORM_BLOCK"]}[tex["fmt"]]
        filt = tex.get("filter", 1)
        textures.append(scenes.Texture(tex["width"], tex["height"], 1, tex["layers"], tex["data"], mag_filter=filt,
                                       min_filter=filt, address=scenes.ADDR_REPEAT, format=fmt,
                                       cube=tex["layers"] == 6 and tex["width"] == tex["height"]))
    pipe = scenes.Pipeline(vs_name, fs_name, depth_op=CMP["ALWAYS"], depth_write=False)
    draw = scenes.Draw("p", len(verts), 0, False, bytes(push) + b"\0" * (128 - len(push)),
                       texture=0 if textures else None)
    scene = scenes.Scene("translator", W, H, vdata.tobytes(), vdata.shape[1] * 4, vattrs, {"p": pipe}, [draw],
                         textures=textures, color_format=FMT["R16G16B16A16_SFLOAT"], clear_depth=None)
    rr = render.SceneRenderer(ctx, scene)
    ctx.set_mode(batched=True, tile=8)
    rr.render()
    px = rr.read_color().view(np.float16)[0, :4 * R].astype(np.float32)  # (4R, 4) bytes as floats
    b = px.astype(np.uint32).reshape(R, 4, 4)
    bits = b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16) | (b[..., 3] << 24)
    return bits.view(np.float32)


@pytest.fixture
def on_gpu(ctx, monkeypatch):
    monkeypatch.setattr(T, "run_fs_rows", functools.partial(gpu_run_fs_rows, ctx))
    monkeypatch.setattr(T.FS, "finish", packed_finish)


def _on_gpu(fn):
    @functools.wraps(fn)
    def run(*args, **kwargs):
        return fn(*args, **kwargs)
    return pytest.mark.usefixtures("on_gpu")(run)


for _name, _fn in list(vars(T).items()):
    if _name.startswith("test_") and callable(_fn):
        globals()[_name] = _on_gpu(_fn)
