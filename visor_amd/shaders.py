"""The shader programs the draw path is exercised and benchmarked with, as SPIR-V.

Each function returns SPIR-V bytes for one entry point named "main". The same programs are
restated in C in the oracle (oracle/refrast.c, RR_VS_* / RR_FS_*), lowered by the reference's
rules, so a parity failure points at the translator or the raster kernels.

Push-constant layout shared by the vertex and fragment stages (std430 offsets):
    0   mat4  mvp
    64  mat3  nmat   (MatrixStride 16, 48 bytes)
    112 float alpha
"""

from .spirv import Module

LIGHT_DIR = (0.48, 0.6, 0.64)  # unit length exactly: 0.2304 + 0.36 + 0.4096 = 1

# 8-tap array-texture program constants (mirrors kTap* in the oracle)
TAP_SCALE = (1.0, 2.0, 0.5, 4.0, 1.0, 3.0, 0.25, 8.0)
TAP_OFFX = (0.0, 0.125, 0.25, 0.5, 0.75, 0.375, 0.0625, 0.5)
TAP_OFFY = (0.0, 0.25, 0.125, 0.75, 0.5, 0.0625, 0.375, 0.25)
TAP_LAYER = (0.0, 1.0, 2.0, 3.0, 0.0, 1.0, 2.0, 3.0)
TAP_WEIGHT = (0.25, 0.125, 0.125, 0.125, 0.125, 0.125, 0.0625, 0.0625)


def _push_block(m, with_nmat):
    f = m.float()
    mat4 = m.mat(4, 4)
    if with_nmat:
        mat3 = m.mat(3, 3)
        st, var = m.push_block([mat4, mat3, f], [0, 64, 112], key="PC")
    else:
        st, var = m.push_block([mat4], [0], key="PC4")
    return st, var


def _per_vertex_position(m):
    """gl_PerVertex-style output block with Position as member 0 (what GLSL front ends emit)."""
    v4 = m.vec(4)
    st = m.struct([v4], key="gl_PerVertex")
    m.member_decorate(st, 0, "BuiltIn", 0)
    m.decorate(st, "Block")
    var = m.variable("Output", st)
    return var


def vs_mvp_pos4_col():
    """gl_Position = pc.mvp * pos; out0 = col.  (oracle RR_VS_MVP_POS4_COL)"""
    m = Module()
    v4 = m.vec(4)
    mat4 = m.mat(4, 4)
    _, pc = _push_block(m, with_nmat=False)
    pos = m.input(v4, location=0)
    col = m.input(v4, location=1)
    out_col = m.output(v4, location=0)
    pv = _per_vertex_position(m)
    m.begin_main()
    i0 = m.const_i(0)
    mvp = m.load(mat4, m.access("PushConstant", mat4, pc, i0))
    p = m.load(v4, pos)
    gp = m.op("MatrixTimesVector", v4, mvp, p)
    m.store(m.access("Output", v4, pv, i0), gp)
    m.store(out_col, m.load(v4, col))
    m.end_main("Vertex", "main", [pos, col, out_col, pv])
    return m.to_bytes()


def vs_mvp_normal_uv():
    """gl_Position = mvp * vec4(pos, 1); out0 = nmat * nrm; out1 = uv.  (RR_VS_MVP_NORMAL_UV)"""
    m = Module()
    f = m.float()
    v2, v3, v4 = m.vec(2), m.vec(3), m.vec(4)
    mat4, mat3 = m.mat(4, 4), m.mat(3, 3)
    _, pc = _push_block(m, with_nmat=True)
    pos = m.input(v3, location=0)
    nrm = m.input(v3, location=1)
    uv = m.input(v2, location=2)
    o_n = m.output(v3, location=0)
    o_uv = m.output(v2, location=1)
    pv = _per_vertex_position(m)
    m.begin_main()
    i0, i1 = m.const_i(0), m.const_i(1)
    mvp = m.load(mat4, m.access("PushConstant", mat4, pc, i0))
    p = m.load(v3, pos)
    px = m.op("CompositeExtract", f, p, 0)
    py = m.op("CompositeExtract", f, p, 1)
    pz = m.op("CompositeExtract", f, p, 2)
    p4 = m.op("CompositeConstruct", v4, px, py, pz, m.const_f(1.0))
    m.store(m.access("Output", v4, pv, i0), m.op("MatrixTimesVector", v4, mvp, p4))
    nmat = m.load(mat3, m.access("PushConstant", mat3, pc, i1))
    m.store(o_n, m.op("MatrixTimesVector", v3, nmat, m.load(v3, nrm)))
    m.store(o_uv, m.load(v2, uv))
    m.end_main("Vertex", "main", [pos, nrm, uv, o_n, o_uv, pv])
    return m.to_bytes()


def fs_color():
    """out = in0.  (RR_FS_COLOR)"""
    m = Module()
    v4 = m.vec(4)
    c = m.input(v4, location=0)
    o = m.output(v4, location=0)
    m.begin_main()
    m.store(o, m.load(v4, c))
    m.end_main("Fragment", "main", [c, o], frag=True)
    return m.to_bytes()


def _lambert(m, nrm_var):
    f, v3 = m.float(), m.vec(3)
    n = m.ext(v3, "Normalize", m.load(v3, nrm_var))
    ndl = m.op("Dot", f, n, m.const_vec(LIGHT_DIR))
    ndl = m.ext(f, "FMax", ndl, m.const_f(0.0))
    return m.op("FAdd", f, m.op("FMul", f, ndl, m.const_f(0.75)), m.const_f(0.25))


def fs_lambert_tex(alpha_from_push=False):
    """Lambert x texture; alpha 1 or pc.alpha.  (RR_FS_LAMBERT_TEX / RR_FS_LAMBERT_TEX_A)"""
    m = Module()
    f, v2, v3, v4 = m.float(), m.vec(2), m.vec(3), m.vec(4)
    pc = None
    if alpha_from_push:
        _, pc = _push_block(m, with_nmat=True)
    nrm = m.input(v3, location=0)
    uv = m.input(v2, location=1)
    tex = m.sampler(0, 0)
    o = m.output(v4, location=0)
    m.begin_main()
    light = _lambert(m, nrm)
    t = m.op("ImageSampleImplicitLod", v4, m.load(m.sampled_image(), tex), m.load(v2, uv))
    rgb = m.op("VectorShuffle", v3, t, t, 0, 1, 2)
    rgb = m.op("VectorTimesScalar", v3, rgb, light)
    if alpha_from_push:
        a = m.load(f, m.access("PushConstant", f, pc, m.const_i(2)))
    else:
        a = m.const_f(1.0)
    r = m.op("CompositeExtract", f, rgb, 0)
    g = m.op("CompositeExtract", f, rgb, 1)
    b = m.op("CompositeExtract", f, rgb, 2)
    m.store(o, m.op("CompositeConstruct", v4, r, g, b, a))
    iface = [nrm, uv, o]
    m.end_main("Fragment", "main", iface, frag=True)
    return m.to_bytes()


def fs_tex_only():
    """out = texture(tex, in1).  (RR_FS_TEX_ONLY)"""
    m = Module()
    v2, v4 = m.vec(2), m.vec(4)
    uv = m.input(v2, location=1)
    tex = m.sampler(0, 0)
    o = m.output(v4, location=0)
    m.begin_main()
    m.store(o, m.op("ImageSampleImplicitLod", v4, m.load(m.sampled_image(), tex), m.load(v2, uv)))
    m.end_main("Fragment", "main", [uv, o], frag=True)
    return m.to_bytes()


def fs_cube_env():
    """out = texture(samplerCube, in0.xyz): an environment lookup by the interpolated normal.
    (RR_FS_CUBE_ENV)"""
    m = Module()
    v3, v4 = m.vec(3), m.vec(4)
    nrm = m.input(v3, location=0)
    tex = m.sampler(0, 0, cube=True)
    o = m.output(v4, location=0)
    m.begin_main()
    m.store(o, m.op("ImageSampleImplicitLod", v4, m.load(m.sampled_image(cube=True), tex), m.load(v3, nrm)))
    m.end_main("Fragment", "main", [nrm, o], frag=True)
    return m.to_bytes()


def fs_multitap8():
    """Sum of 8 weighted array-texture taps in FP32.  (RR_FS_MULTITAP8)"""
    m = Module()
    f, v2, v3, v4 = m.float(), m.vec(2), m.vec(3), m.vec(4)
    uv = m.input(v2, location=1)
    tex = m.sampler(0, 0, arrayed=True)
    o = m.output(v4, location=0)
    m.begin_main()
    img = m.load(m.sampled_image(arrayed=True), tex)
    uvv = m.load(v2, uv)
    acc = m.const_vec((0.0, 0.0, 0.0, 0.0))
    for k in range(8):
        st = m.op("VectorTimesScalar", v2, uvv, m.const_f(TAP_SCALE[k]))
        st = m.op("FAdd", v2, st, m.const_vec((TAP_OFFX[k], TAP_OFFY[k])))
        cx = m.op("CompositeExtract", f, st, 0)
        cy = m.op("CompositeExtract", f, st, 1)
        c = m.op("CompositeConstruct", v3, cx, cy, m.const_f(TAP_LAYER[k]))
        t = m.op("ImageSampleImplicitLod", v4, img, c)
        acc = m.op("FAdd", v4, acc, m.op("VectorTimesScalar", v4, t, m.const_f(TAP_WEIGHT[k])))
    m.store(o, acc)
    m.end_main("Fragment", "main", [uv, o], frag=True)
    return m.to_bytes()


# Uniform block of the UBO-fed programs: set 0, binding 1 (std140 offsets)
#     0   mat4 mvp
#     64  vec4 tint
#     80  vec4 bias
UBO_BINDING = 1


def _ubo_block(m):
    return m.uniform_block([m.mat(4, 4), m.vec(4), m.vec(4)], [0, 64, 80], key="UBO", set_=0, binding=UBO_BINDING)


def vs_ubo_pos4_col():
    """gl_Position = ubo.mvp * pos; out0 = col.  (oracle RR_VS_UBO_POS4_COL)"""
    m = Module()
    v4, mat4 = m.vec(4), m.mat(4, 4)
    _, ubo = _ubo_block(m)
    pos = m.input(v4, location=0)
    col = m.input(v4, location=1)
    out_col = m.output(v4, location=0)
    pv = _per_vertex_position(m)
    m.begin_main()
    i0 = m.const_i(0)
    mvp = m.load(mat4, m.access("Uniform", mat4, ubo, i0))
    m.store(m.access("Output", v4, pv, i0), m.op("MatrixTimesVector", v4, mvp, m.load(v4, pos)))
    m.store(out_col, m.load(v4, col))
    m.end_main("Vertex", "main", [pos, col, out_col, pv])
    return m.to_bytes()


def fs_ubo_tint():
    """out = in0 * ubo.tint + ubo.bias (component-wise).  (oracle RR_FS_UBO_TINT)"""
    m = Module()
    v4 = m.vec(4)
    _, ubo = _ubo_block(m)
    c = m.input(v4, location=0)
    o = m.output(v4, location=0)
    m.begin_main()
    tint = m.load(v4, m.access("Uniform", v4, ubo, m.const_i(1)))
    bias = m.load(v4, m.access("Uniform", v4, ubo, m.const_i(2)))
    m.store(o, m.op("FAdd", v4, m.op("FMul", v4, m.load(v4, c), tint), bias))
    m.end_main("Fragment", "main", [c, o], frag=True)
    return m.to_bytes()


def ubo_block_bytes(mvp, tint, bias):
    """The uniform block's bytes: column-major mvp (given row-major), tint, bias."""
    import numpy as np
    b = bytearray(96)
    b[0:64] = np.asarray(mvp, dtype=np.float32).T.tobytes()
    b[64:80] = np.asarray(tint, dtype=np.float32).tobytes()
    b[80:96] = np.asarray(bias, dtype=np.float32).tobytes()
    return bytes(b)


# oracle program ids (oracle/refrast.h)
ORACLE_ID = {
    "vs_mvp_pos4_col": 1, "vs_mvp_normal_uv": 2, "fs_color": 101, "fs_lambert_tex": 102,
    "fs_lambert_tex_a": 103, "fs_multitap8": 104, "fs_tex_only": 105, "fs_cube_env": 106,
    "vs_ubo_pos4_col": 3, "fs_ubo_tint": 107,
}

PROGRAMS = {
    "vs_mvp_pos4_col": vs_mvp_pos4_col,
    "vs_mvp_normal_uv": vs_mvp_normal_uv,
    "fs_color": fs_color,
    "fs_lambert_tex": lambda: fs_lambert_tex(False),
    "fs_lambert_tex_a": lambda: fs_lambert_tex(True),
    "fs_multitap8": fs_multitap8,
    "fs_tex_only": fs_tex_only,
    "fs_cube_env": fs_cube_env,
    "vs_ubo_pos4_col": vs_ubo_pos4_col,
    "fs_ubo_tint": fs_ubo_tint,
}
