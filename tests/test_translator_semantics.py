"""SPIR-V translator semantics on the host (reference spirv_compile.cpp:1056-1888 is the op set).

Fragment programs are written with the SPIR-V builder, translated by the library
(vh_translate_spirv), compiled with g++ together with the device prelude (pipeline_prelude.hip)
and run on chosen inputs; the outputs are compared with numpy float32 / int32 computations of the
SPIR-V specification's semantics. Interpolation weights are (1, 0, 0), so each input equals its
vertex-0 value exactly. Ops the reference lowers incorrectly (ShiftLeftLogical, Cross,
MatrixInverse) are checked against their specified meaning, which is what the translator emits.
"""
import array
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from visor_amd import spirv, vk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "visor_amd", "csrc")

SHIM = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#define __device__
#define __forceinline__ inline
static inline float __uint_as_float(unsigned u) { float f; memcpy(&f, &u, 4); return f; }
static inline unsigned __float_as_uint(float f) { unsigned u; memcpy(&u, &f, 4); return u; }
"""


@pytest.fixture(scope="module")
def lib():
    L = C.CDLL(vk.LIB_PATH)
    L.vh_translate_spirv.restype = C.c_int
    L.vh_translate_spirv.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.c_int, C.c_char_p, C.c_size_t]
    return L


def _translate(L, spv):
    w = array.array("I", spv)
    out = C.create_string_buffer(1 << 20)
    n = L.vh_translate_spirv((C.c_uint32 * len(w))(*w), len(w), 4, out, len(out))
    assert n > 0, out.value.decode()
    return out.value.decode()


def run_fs_rows(L, tmp_path, spv, rows, push=b"", tex=None):
    """Runs the fragment program once per row of vec4 inputs (locations 0..n-1); returns the
    outputs, shape (len(rows), 4). `tex` = dict(data, width, height, layers, fmt, bpp, filter)
    binds a host-memory texture (one level, REPEAT) at binding 0, read through its RGBA8 transcode
    as the executor does (fmt 0 plain, 1 BC2, 2 BC3)."""
    code = _translate(L, spv)
    vary = []
    for inputs in rows:
        for v in inputs:
            for x in v:
                vary += [float(np.float32(x)), 0.0, 0.0]
    nv = max(1, len(vary) // max(1, len(rows)))
    vary = vary or [0.0]
    push = bytes(push) + b"\0" * (128 - len(push))
    tex_init = ""
    if tex is not None:
        # the executor's transcode (texel_decode.h) fills the RGBA8 copy the samplers read
        t = tex
        tex_init = (r"""
  static const unsigned char texdata[] = {""" + ",".join(str(b) for b in t["data"]) + r"""};
  static uint32_t shadow[%d * %d * %d];
  const uint32_t W = %d, H = %d, L = %d, BPP = %d, FMT = %d;
  for(uint32_t l = 0; l < L; l++)
    for(uint32_t y = 0; y < H; y++)
      for(uint32_t x = 0; x < W; x++)
        shadow[(l * H + y) * W + x] = va_texel_bytes(texdata, sizeof(texdata), (uint64_t)l * W * H * BPP, W, BPP, FMT, x, y);
  VaTex &t = d.tex[0];
  t.base = (unsigned long long)(uintptr_t)shadow;
  t.width = W; t.height = H; t.mips = 1; t.layers = L; t.bpp = 4;
  t.slice_bytes = (unsigned long long)W * H * 4;
  t.mag_filter = t.min_filter = %d; t.mipmap_mode = 0; t.addr_u = t.addr_v = 0;
  t.lod_bias = 0.0f; t.min_lod = 0.0f; t.max_lod = 0.0f;
""" % (t["width"], t["height"], t["layers"], t["width"], t["height"], t["layers"], t["bpp"], t["fmt"],
       t.get("filter", 1)))
    src = (SHIM + open(os.path.join(CSRC, "device_abi.h")).read() + "\n" +
           open(os.path.join(CSRC, "texel_decode.h")).read() + "\n" +
           open(os.path.join(CSRC, "pipeline_prelude.hip")).read() + "\n" + code + r"""
int main(void) {
  static VaDraw d;
  memset(&d, 0, sizeof(d));
  static const unsigned char push[128] = {""" + ",".join(str(b) for b in push) + r"""};
  memcpy(d.push, push, 128);""" + tex_init + r"""
  static const float vary[] = {""" + ",".join(repr(x) + "f" for x in vary) + r"""};
  const float b[3][4] = {{1.0f, 0.0f, 0.0f, 0.0f}, {1.0f, 0.0f, 0.0f, 0.0f}, {1.0f, 0.0f, 0.0f, 0.0f}};
  for(int r = 0; r < %d; r++) {
    float out[4];
    va_shader(d, vary + r * %d, b, out);
    for(int i = 0; i < 4; i++) printf("%%08x\n", __float_as_uint(out[i]));
  }
  return 0;
}
""" % (max(1, len(rows)), nv))
    path = tmp_path / "fs.cpp"
    path.write_text(src)
    exe = tmp_path / "fs"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-w", str(path), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return np.array([int(x, 16) for x in lines], dtype=np.uint32).view(np.float32).reshape(-1, 4)


def run_fs(L, tmp_path, spv, inputs, push=b""):
    """Runs the fragment program on vec4 inputs (locations 0..n-1); returns its 4 outputs."""
    return run_fs_rows(L, tmp_path, spv, [inputs], push)[0]


class FS:
    """A fragment program: vec4 inputs at locations 0..n-1, one vec4 output."""

    def __init__(self, ninputs=2):
        self.m = m = spirv.Module()
        self.f, self.i, self.u, self.b = m.float(), m.int(), m.int(signed=False), m.bool()
        self.v4 = m.vec(4)
        self.ins = [m.input(self.v4, location=k) for k in range(ninputs)]
        self.out = m.output(self.v4, location=0)
        m.begin_main()
        self.vals = [m.load(self.v4, p) for p in self.ins]

    def x(self, k, c):
        return self.m.op("CompositeExtract", self.f, self.vals[k], c)

    def c(self, v):
        return self.m.const_f(v)

    def finish(self, comps):
        m = self.m
        m.store(self.out, m.op("CompositeConstruct", self.v4, *comps))
        m.end_main("Fragment", "main", self.ins + [self.out], frag=True)
        return m.to_bytes()


f32 = np.float32
A = [1.75, -2.5, 0.3, 7.0]
B = [0.5, 3.25, -1.2, 2.0]


def test_arithmetic(lib, tmp_path):
    p = FS()
    m = p.m
    out = [m.op("FAdd", p.f, p.x(0, 0), p.x(1, 0)), m.op("FMul", p.f, p.x(0, 1), p.x(1, 1)),
           m.op("FSub", p.f, p.x(0, 2), p.x(1, 2)), m.op("FDiv", p.f, p.x(0, 3), p.x(1, 3))]
    got = run_fs(lib, tmp_path, p.finish(out), [A, B])
    a, b = np.float32(A), np.float32(B)
    assert np.array_equal(got, [a[0] + b[0], a[1] * b[1], a[2] - b[2], a[3] / b[3]])


def test_vector_ops_and_negate(lib, tmp_path):
    p = FS()
    m = p.m
    s = m.op("FAdd", p.v4, p.vals[0], p.vals[1])
    s = m.op("VectorTimesScalar", p.v4, s, p.c(3.0))
    s = m.op("FNegate", p.v4, s)
    d = m.op("Dot", p.f, p.vals[0], p.vals[1])
    out = [m.op("CompositeExtract", p.f, s, k) for k in range(3)] + [d]
    got = run_fs(lib, tmp_path, p.finish(out), [A, B])
    a, b = np.float32(A), np.float32(B)
    s = -((a + b) * f32(3.0))
    # the reference's dot: (((a0 b0) + a1 b1) + a2 b2) + a3 b3, rounded after every step
    d = f32(f32(f32(a[0] * b[0]) + f32(a[1] * b[1])) + f32(a[2] * b[2])) + f32(a[3] * b[3])
    assert np.array_equal(got, [s[0], s[1], s[2], d])


@pytest.mark.parametrize("names", [("FMin", "FMax", "Floor", "Ceil"), ("Fract", "Trunc", "FAbs", "FSign"),
                                   ("Round", "RoundEven", "Sqrt", "InverseSqrt")])
def test_glsl_unary_binary_exact(lib, tmp_path, names):
    p = FS()
    m = p.m
    binary = {"FMin", "FMax"}
    xs = [-2.5, 3.5, 0.75, 6.25]
    ys = [1.5, -0.5, 2.0, 4.0]
    out = []
    for k, n in enumerate(names):
        if n in binary:
            out.append(m.ext(p.f, n, p.x(0, k), p.x(1, k)))
        else:
            out.append(m.ext(p.f, n, p.x(0, k)))
    got = run_fs(lib, tmp_path, p.finish(out), [xs, ys])
    x, y = np.float32(xs), np.float32(ys)
    ref = {"FMin": lambda k: np.minimum(x[k], y[k]), "FMax": lambda k: np.maximum(x[k], y[k]),
           "Floor": lambda k: np.floor(x[k]), "Ceil": lambda k: np.ceil(x[k]), "Fract": lambda k: x[k] - np.floor(x[k]),
           "Trunc": lambda k: np.trunc(x[k]), "FAbs": lambda k: np.abs(x[k]), "FSign": lambda k: np.sign(x[k]),
           "Round": lambda k: np.rint(x[k]), "RoundEven": lambda k: np.rint(x[k]),
           "Sqrt": lambda k: np.sqrt(x[k]),
           "InverseSqrt": lambda k: f32(1.0) / np.sqrt(x[k])}
    with np.errstate(invalid="ignore"):
        expect = np.array([ref[n](k) for k, n in enumerate(names)], dtype=np.float32)
    assert np.array_equal(got, expect, equal_nan=True), (got, expect)


def test_round_even_halfway(lib, tmp_path):
    p = FS()
    m = p.m
    out = [m.ext(p.f, "RoundEven", p.x(0, k)) for k in range(4)]
    got = run_fs(lib, tmp_path, p.finish(out), [[0.5, 1.5, 2.5, -3.5], B])
    assert np.array_equal(got, np.float32([0.0, 2.0, 2.0, -4.0]))


def test_glsl_transcendental(lib, tmp_path):
    p = FS()
    m = p.m
    out = [m.ext(p.f, "Sin", p.x(0, 0)), m.ext(p.f, "Cos", p.x(0, 1)), m.ext(p.f, "Pow", p.x(0, 2), p.x(1, 2)),
           m.ext(p.f, "Exp2", p.x(0, 3))]
    xs, ys = [0.7, -1.3, 2.5, 3.3], [0.0, 0.0, 1.7, 0.0]
    got = run_fs(lib, tmp_path, p.finish(out), [xs, ys])
    expect = np.float32([np.sin(0.7), np.cos(-1.3), 2.5 ** 1.7, 2.0 ** 3.3])
    assert np.allclose(got, expect, rtol=2e-6)


def test_glsl_mix_clamp_step_smoothstep(lib, tmp_path):
    p = FS(3)
    m = p.m
    out = [m.ext(p.f, "FClamp", p.x(0, 0), p.x(1, 0), p.x(2, 0)), m.ext(p.f, "FMix", p.x(0, 1), p.x(1, 1), p.x(2, 1)),
           m.ext(p.f, "Step", p.x(1, 2), p.x(0, 2)), m.ext(p.f, "SmoothStep", p.x(1, 3), p.x(2, 3), p.x(0, 3))]
    a, b, c = [3.0, 1.0, 0.25, 0.4], [-1.0, 5.0, 0.5, 0.0], [2.0, 0.25, 0.0, 1.0]
    got = run_fs(lib, tmp_path, p.finish(out), [a, b, c])
    a, b, c = np.float32(a), np.float32(b), np.float32(c)
    t = np.clip((a[3] - b[3]) / (c[3] - b[3]), 0, 1)
    expect = np.float32([min(max(a[0], b[0]), c[0]), a[1] * (1 - c[1]) + b[1] * c[1], 0.0 if a[2] < b[2] else 1.0,
                         t * t * (3 - 2 * t)])
    assert np.allclose(got, expect, rtol=1e-6, atol=1e-7)


def test_geometric(lib, tmp_path):
    p = FS()
    m = p.m
    v3 = m.vec(3)
    a3 = m.op("VectorShuffle", v3, p.vals[0], p.vals[0], 0, 1, 2)
    b3 = m.op("VectorShuffle", v3, p.vals[1], p.vals[1], 0, 1, 2)
    cr = m.ext(v3, "Cross", a3, b3)
    out = [m.op("CompositeExtract", p.f, cr, 0), m.op("CompositeExtract", p.f, cr, 1),
           m.ext(p.f, "Length", p.vals[0]), m.ext(p.f, "Distance", p.vals[0], p.vals[1])]
    got = run_fs(lib, tmp_path, p.finish(out), [A, B])
    a, b = np.float64(A), np.float64(B)
    cr = np.cross(a[:3], b[:3])
    expect = [cr[0], cr[1], np.linalg.norm(a), np.linalg.norm(a - b)]
    assert np.allclose(got, expect, rtol=2e-6)


def test_normalize_reflect(lib, tmp_path):
    p = FS()
    m = p.m
    n = m.ext(p.v4, "Normalize", p.vals[0])
    r = m.ext(p.v4, "Reflect", p.vals[1], n)
    out = [m.op("CompositeExtract", p.f, n, 0), m.op("CompositeExtract", p.f, n, 3),
           m.op("CompositeExtract", p.f, r, 0), m.op("CompositeExtract", p.f, r, 2)]
    got = run_fs(lib, tmp_path, p.finish(out), [A, B])
    a, b = np.float64(A), np.float64(B)
    nn = a / np.linalg.norm(a)
    rr = b - 2 * np.dot(nn, b) * nn
    assert np.allclose(got, [nn[0], nn[3], rr[0], rr[2]], rtol=2e-6, atol=1e-6)


def test_integer_arithmetic(lib, tmp_path):
    p = FS()
    m = p.m
    i = m.op("ConvertFToS", p.i, p.x(0, 0))
    j = m.op("ConvertFToS", p.i, p.x(1, 0))
    r = [m.op("IAdd", p.i, i, j), m.op("IMul", p.i, i, j), m.op("SDiv", p.i, i, j), m.op("SRem", p.i, i, j)]
    out = [m.op("ConvertSToF", p.f, v) for v in r]
    got = run_fs(lib, tmp_path, p.finish(out), [[-17.9, 0, 0, 0], [5.2, 0, 0, 0]])
    assert np.array_equal(got, np.float32([-17 + 5, -17 * 5, -3, -2]))  # C truncation: -17 / 5 = -3, -17 % 5 = -2


def test_integer_mod_bits_and_shifts(lib, tmp_path):
    p = FS()
    m = p.m
    i = m.op("ConvertFToS", p.i, p.x(0, 0))
    j = m.op("ConvertFToS", p.i, p.x(1, 0))
    three = m.const_i(3)
    r = [m.op("SMod", p.i, i, j),
         m.op("ShiftLeftLogical", p.i, i, three),
         m.op("ShiftRightArithmetic", p.i, i, m.const_i(1)),
         m.op("BitwiseXor", p.i, m.op("BitwiseAnd", p.i, i, m.const_i(0xFF)), m.op("BitwiseOr", p.i, j, m.const_i(0x30)))]
    out = [m.op("ConvertSToF", p.f, v) for v in r]
    got = run_fs(lib, tmp_path, p.finish(out), [[-17.0, 0, 0, 0], [5.0, 0, 0, 0]])
    i, j = -17, 5
    # SMod takes the sign of the divisor (Python %); shifts and bit operations on two's complement
    expect = [i % j, i * 8, i >> 1, (i & 0xFF) ^ (j | 0x30)]
    assert np.array_equal(got, np.float32(expect))


def test_integer_min_max_abs_not_and_unsigned(lib, tmp_path):
    p = FS()
    m = p.m
    i = m.op("ConvertFToS", p.i, p.x(0, 0))
    j = m.op("ConvertFToS", p.i, p.x(1, 0))
    u = m.op("ConvertFToU", p.u, p.x(0, 1))
    r = [m.op("ConvertSToF", p.f, m.ext(p.i, "SClamp", i, m.op("SNegate", p.i, j), j)),
         m.op("ConvertSToF", p.f, m.ext(p.i, "SAbs", i)),
         m.op("ConvertSToF", p.f, m.op("Not", p.i, i)),
         m.op("ConvertUToF", p.f, m.op("ShiftRightLogical", p.u, u, m.const_u(2)))]
    got = run_fs(lib, tmp_path, p.finish(r), [[-17.0, 1000.0, 0, 0], [5.0, 0, 0, 0]])
    assert np.array_equal(got, np.float32([-5, 17, ~(-17), 1000 >> 2]))


def test_float_to_int_conversions_follow_x64(lib, tmp_path):
    """ConvertFToS / ConvertFToU of out-of-range values give what the reference's x64 JIT code
    gives (LLVM fptosi -> cvttss2si: 0x80000000 outside int32 range; fptoui to i32 -> a 64-bit
    signed conversion whose low 32 bits are kept), not the saturated values of a GPU conversion."""
    p = FS()
    m = p.m
    r = [m.op("ConvertSToF", p.f, m.op("ConvertFToS", p.i, p.x(0, 0))),
         m.op("ConvertSToF", p.f, m.op("ConvertFToS", p.i, p.x(0, 1))),
         m.op("ConvertUToF", p.f, m.op("ConvertFToU", p.u, p.x(0, 2))),
         m.op("ConvertUToF", p.f, m.op("ConvertFToU", p.u, p.x(0, 3)))]
    got = run_fs_rows(lib, tmp_path, p.finish(r), [[[3e9, -2.5, -1.0, 5e9], [0, 0, 0, 0]],
                                                    [[-1e12, 2147483520.0, 1e20, 7.9], [0, 0, 0, 0]]])
    want = np.float32([[-2147483648.0, -2.0, 4294967295.0, 5e9 % 2 ** 32],
                       [-2147483648.0, 2147483520.0, 0.0, 7.0]])
    assert np.array_equal(got, want), got


def test_compare_select_logic(lib, tmp_path):
    p = FS()
    m = p.m
    lt = m.op("FOrdLessThan", p.b, p.x(0, 0), p.x(1, 0))
    ge = m.op("FOrdGreaterThanEqual", p.b, p.x(0, 1), p.x(1, 1))
    both = m.op("LogicalAnd", p.b, lt, m.op("LogicalNot", p.b, ge))
    either = m.op("LogicalOr", p.b, lt, ge)
    out = [m.op("Select", p.f, lt, p.c(1.0), p.c(2.0)), m.op("Select", p.f, ge, p.c(3.0), p.c(4.0)),
           m.op("Select", p.f, both, p.c(5.0), p.c(6.0)), m.op("Select", p.f, either, p.c(7.0), p.c(8.0))]
    got = run_fs(lib, tmp_path, p.finish(out), [[1.0, 1.0, 0, 0], [2.0, 3.0, 0, 0]])
    assert np.array_equal(got, np.float32([1.0, 4.0, 5.0, 7.0]))


def test_branch_with_phi(lib, tmp_path):
    """if (a.x > b.x) r = a.x * 2 else r = b.x + 1, merged with OpPhi."""
    for a0, expect in [(5.0, 10.0), (-1.0, 3.0)]:
        p = FS()
        m = p.m
        ax, bx = p.x(0, 0), p.x(1, 0)
        cond = m.op("FOrdGreaterThan", p.b, ax, bx)
        l_then, l_else, l_merge = m._id(), m._id(), m._id()
        m.op_void("SelectionMerge", l_merge, 0)
        m.op_void("BranchConditional", cond, l_then, l_else)
        m.body += m._inst("Label", [l_then])
        t = m.op("FMul", p.f, ax, p.c(2.0))
        m.op_void("Branch", l_merge)
        m.body += m._inst("Label", [l_else])
        e = m.op("FAdd", p.f, bx, p.c(1.0))
        m.op_void("Branch", l_merge)
        m.body += m._inst("Label", [l_merge])
        r = m.op("Phi", p.f, t, l_then, e, l_else)
        got = run_fs(lib, tmp_path, p.finish([r, p.c(0.0), p.c(0.0), p.c(0.0)]), [[a0, 0, 0, 0], [2.0, 0, 0, 0]])
        assert got[0] == expect


def test_loop_with_phi(lib, tmp_path):
    """sum = 0; for (i = 0; i < n; i++) sum += a.x * i."""
    p = FS()
    m = p.m
    ax = p.x(0, 0)
    n = m.op("ConvertFToS", p.i, p.x(1, 0))
    zero_i, one_i, zero_f = m.const_i(0), m.const_i(1), p.c(0.0)
    l_entry = m._id()
    m.op_void("Branch", l_entry)
    m.body += m._inst("Label", [l_entry])
    l_head, l_body, l_cont, l_merge = m._id(), m._id(), m._id(), m._id()
    m.op_void("Branch", l_head)
    m.body += m._inst("Label", [l_head])
    i_phi, s_phi = m._id(), m._id()
    i_next, s_next = m._id(), m._id()
    m.body += m._inst("Phi", [p.i, i_phi, zero_i, l_entry, i_next, l_cont])
    m.body += m._inst("Phi", [p.f, s_phi, zero_f, l_entry, s_next, l_cont])
    m.op_void("LoopMerge", l_merge, l_cont, 0)
    cond = m.op("SLessThan", p.b, i_phi, n)
    m.op_void("BranchConditional", cond, l_body, l_merge)
    m.body += m._inst("Label", [l_body])
    term = m.op("FMul", p.f, ax, m.op("ConvertSToF", p.f, i_phi))
    m.body += m._inst("FAdd", [p.f, s_next, s_phi, term])
    m.op_void("Branch", l_cont)
    m.body += m._inst("Label", [l_cont])
    m.body += m._inst("IAdd", [p.i, i_next, i_phi, one_i])
    m.op_void("Branch", l_head)
    m.body += m._inst("Label", [l_merge])
    got = run_fs(lib, tmp_path, p.finish([s_phi, m.op("ConvertSToF", p.f, i_phi), p.c(0.0), p.c(0.0)]),
                 [[1.5, 0, 0, 0], [6.0, 0, 0, 0]])
    s = f32(0)
    for i in range(6):
        s = f32(s + f32(f32(1.5) * f32(i)))
    assert np.array_equal(got[:2], np.float32([s, 6.0]))


def test_function_call(lib, tmp_path):
    """A helper function with a parameter, called twice (reference OpFunctionCall inlining)."""
    p = FS()
    m = p.m
    # float helper(float x) { return x * 3 + 1; } emitted before main's body
    fty = m.func_type(p.f, [p.f])
    fn, par, lab = m._id(), m._id(), m._id()
    t1, t2 = m._id(), m._id()
    three, one = p.c(3.0), p.c(1.0)
    m.funcs += (m._inst("Function", [p.f, fn, 0, fty]) + m._inst("FunctionParameter", [p.f, par]) +
                m._inst("Label", [lab]) + m._inst("FMul", [p.f, t1, par, three]) + m._inst("FAdd", [p.f, t2, t1, one]) +
                m._inst("ReturnValue", [t2]) + m._inst("FunctionEnd", []))
    r1 = m.op("FunctionCall", p.f, fn, p.x(0, 0))
    r2 = m.op("FunctionCall", p.f, fn, p.x(1, 0))
    got = run_fs(lib, tmp_path, p.finish([r1, r2, p.c(0.0), p.c(0.0)]), [[2.0, 0, 0, 0], [-1.5, 0, 0, 0]])
    assert np.array_equal(got[:2], np.float32([7.0, -3.5]))


def test_matrix_ops_from_push_constants(lib, tmp_path):
    """mat4 from push constants (column-major, MatrixStride 16): M v, v M, (M M) v, transpose."""
    M = np.arange(16, dtype=np.float32).reshape(4, 4) * np.float32(0.25) - np.float32(1.0)  # M[row][col]
    push = M.T.astype(np.float32).tobytes()  # columns
    p = FS()
    m = p.m
    mat4 = m.mat(4, 4)
    _, pc = m.push_block([mat4], [0], key="PCM")
    mm = m.load(mat4, m.access("PushConstant", mat4, pc, m.const_i(0)))
    mv = m.op("MatrixTimesVector", p.v4, mm, p.vals[0])
    vm = m.op("VectorTimesMatrix", p.v4, p.vals[0], mm)
    m2 = m.op("MatrixTimesMatrix", mat4, mm, mm)
    m2v = m.op("MatrixTimesVector", p.v4, m2, p.vals[1])
    tv = m.op("MatrixTimesVector", p.v4, m.op("Transpose", mat4, mm), p.vals[1])
    out = [m.op("CompositeExtract", p.f, mv, 1), m.op("CompositeExtract", p.f, vm, 2),
           m.op("CompositeExtract", p.f, m2v, 3), m.op("CompositeExtract", p.f, tv, 0)]
    got = run_fs(lib, tmp_path, p.finish(out), [A, B], push=push)
    a, b = np.float64(A), np.float64(B)
    Md = M.astype(np.float64)
    expect = [(Md @ a)[1], (a @ Md)[2], (Md @ Md @ b)[3], (Md.T @ b)[0]]
    assert np.allclose(got, expect, rtol=1e-6, atol=1e-5)


def test_matrix_inverse(lib, tmp_path):
    """inverse(M) v for a well-conditioned mat4 from push constants, against numpy's inverse."""
    rng = np.random.default_rng(3)
    M = (np.eye(4) * 2.0 + rng.uniform(-0.5, 0.5, (4, 4))).astype(np.float32)
    p = FS()
    m = p.m
    mat4 = m.mat(4, 4)
    _, pc = m.push_block([mat4], [0], key="PCI")
    mm = m.load(mat4, m.access("PushConstant", mat4, pc, m.const_i(0)))
    iv = m.op("MatrixTimesVector", p.v4, m.ext(mat4, "MatrixInverse", mm), p.vals[0])
    out = [m.op("CompositeExtract", p.f, iv, k) for k in range(4)]
    got = run_fs(lib, tmp_path, p.finish(out), [A, B], push=M.T.tobytes())
    expect = np.linalg.inv(M.astype(np.float64)) @ np.float64(A)
    assert np.allclose(got, expect, rtol=1e-5, atol=1e-5)


# --- texture sampling of the device prelude, run on the host, against the oracle --------------------

def _cube_fs(cube=True):
    p = FS(1)
    m = p.m
    v3 = m.vec(3)
    tex = m.sampler(0, 0, cube=cube)
    d = m.op("VectorShuffle", v3, p.vals[0], p.vals[0], 0, 1, 2)
    if not cube:
        d = m.op("VectorShuffle", m.vec(2), p.vals[0], p.vals[0], 0, 1)
    t = m.op("ImageSampleImplicitLod", p.v4, m.load(m.sampled_image(cube=cube), tex), d)
    return p.finish([m.op("CompositeExtract", p.f, t, k) for k in range(4)])


@pytest.mark.parametrize("fmt", ["R8G8B8A8_UNORM", "BC2_UNORM_BLOCK", "BC3_UNORM_BLOCK"])
def test_prelude_cube_sampling_matches_oracle(lib, tmp_path, fmt):
    """Translated samplerCube lookup, the device prelude's face selection and bilinear filter, and
    the executor's texel transcode (BC decode), compiled for the host: bit-exact with the oracle's
    sample_cube_wrapped (ties, axes and random directions)."""
    from oracle import refrast
    from visor_amd.vk import FMT
    W = 16
    code = FMT[fmt]
    bpp = 1 if "BC" in fmt else 4
    rng = np.random.default_rng(17)
    data = rng.integers(0, 256, 6 * W * W * bpp, dtype=np.uint8)
    dirs = rng.normal(size=(150, 3)).astype(np.float32).tolist()
    dirs += [[1, 1, 1], [1, 1, 0], [-1, -1, 0], [1, 0, -1], [0, 0, 1], [0, -1, 0], [-1, 0, 0], [0.5, -0.5, 0.5]]
    got = run_fs_rows(lib, tmp_path, _cube_fs(), [[d + [0.0]] for d in dirs],
                      tex=dict(data=data.tobytes(), width=W, height=W, layers=6, bpp=bpp,
                               fmt={37: 0, 135: 1, 137: 2}[code]))
    img = refrast.make_image(data, W, W, 1, 6, bpp=bpp, fmt=code)
    out = (C.c_float * 4)()
    L = refrast.lib()
    for k, d in enumerate(dirs):
        L.rr_sample_cube_wrapped(float(np.float32(d[0])), float(np.float32(d[1])), float(np.float32(d[2])),
                                 C.byref(img), out)
        want = np.frombuffer(bytes(out), np.float32)
        assert np.array_equal(got[k], want), (fmt, d, got[k], want)


@pytest.mark.parametrize("fmt", ["BC2_UNORM_BLOCK", "BC3_UNORM_BLOCK"])
@pytest.mark.parametrize("filt", [0, 1])
def test_prelude_bc_2d_sampling_matches_oracle(lib, tmp_path, fmt, filt):
    """2D BC2/BC3 sampling (texel transcode, then the device prelude's nearest or bilinear filter,
    repeat) on the host against the oracle's sampler, including coordinates outside [0, 1)."""
    from oracle import refrast
    from visor_amd.vk import FMT
    W, H = 20, 12
    code = FMT[fmt]
    rng = np.random.default_rng(23)
    data = rng.integers(0, 256, W * H, dtype=np.uint8)
    uvs = rng.uniform(-1.5, 2.5, size=(200, 2)).astype(np.float32).tolist()
    got = run_fs_rows(lib, tmp_path, _cube_fs(cube=False), [[uv + [0.0, 0.0]] for uv in uvs],
                      tex=dict(data=data.tobytes(), width=W, height=H, layers=1, bpp=1,
                               fmt={135: 1, 137: 2}[code], filter=filt))
    img = refrast.make_image(data, W, H, 1, 1, bpp=1, fmt=code)
    smp = refrast.Sampler(mag_filter=filt, min_filter=filt, mipmap_mode=0, address_u=0, address_v=0,
                          mip_lod_bias=0.0, min_lod=0.0, max_lod=0.0)
    out = (C.c_float * 4)()
    L = refrast.lib()
    for k, uv in enumerate(uvs):
        L.rr_sample(C.byref(img), C.byref(smp), (C.c_float * 3)(uv[0], uv[1], 0.0), (C.c_float * 2)(),
                    (C.c_float * 2)(), 0, out)
        want = np.frombuffer(bytes(out), np.float32)
        assert np.array_equal(got[k], want), (fmt, uv, got[k], want)
