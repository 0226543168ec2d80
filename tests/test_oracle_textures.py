"""Pins the oracle's BC2/BC3 block decoder and cube-map sampling (reference
texture_sampling.cpp:34-250 and its vendored decoder 3rdparty/decompress.c).

* The block decoder is compared with the reference's own decoder, compiled from its C source into
  oracle/_ref (skipped where that build is absent), on random and hand-made blocks; endpoint
  widening and BC2 alpha are pinned by their closed forms.
* Cube sampling is pinned by the Vulkan specification's major-axis table (face choice and the
  (sc, tc) orientation of every face), by the reference's tie order, and by equality of the
  full-sampler path with the reference-structured sample_cube_wrapped at the default sampler.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from oracle import refrast
from visor_amd import scenes
from visor_amd.vk import FMT

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_BC = os.path.join(ROOT, "oracle", "_ref", "libbcdecode.so")
BC2, BC3 = FMT["BC2_UNORM_BLOCK"], FMT["BC3_UNORM_BLOCK"]
f32 = C.c_float


@pytest.fixture(scope="module")
def refbc():
    if not os.path.exists(REF_BC):
        pytest.skip("reference BC decoder not built (needs the reference source tree)")
    L = C.CDLL(REF_BC)
    for name in ("DecompressBlockBC2", "DecompressBlockBC3"):
        getattr(L, name).argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_char_p, C.c_void_p]
        getattr(L, name).restype = None
    return L


def ref_decode(L, fmt, block):
    out = (C.c_uint8 * 64)()
    fn = L.DecompressBlockBC2 if fmt == BC2 else L.DecompressBlockBC3
    fn(0, 0, 16, bytes(block), out)
    return np.frombuffer(bytes(out), dtype=np.uint8).reshape(16, 4)


def _special_blocks():
    """Blocks at the mode boundaries: c0 == c1, c0 < c1, c0 > c1; a0 == a1, a0 < a1 with
    selectors 6 and 7, a0 > a1; all-zero and all-one bytes."""
    out = []
    for c0, c1 in [(0x1234, 0x1234), (0x0000, 0xFFFF), (0xFFFF, 0x0000), (0xF800, 0x07E0), (0x001F, 0x001F)]:
        for a0, a1 in [(0, 255), (255, 0), (128, 128), (7, 200), (200, 7)]:
            for sel in [0, 0xFFFFFFFF, 0x1B1B1B1B, 0xE4E4E4E4]:
                b = bytearray(16)
                b[0], b[1] = a0, a1
                b[2:8] = (0xFAC688 * (1 + (sel & 7))).to_bytes(6, "little")  # every 3-bit selector
                b[8:10] = c0.to_bytes(2, "little")
                b[10:12] = c1.to_bytes(2, "little")
                b[12:16] = sel.to_bytes(4, "little")
                out.append(bytes(b))
    out.append(bytes(16))
    out.append(b"\xff" * 16)
    return out


@pytest.mark.parametrize("fmt", [BC2, BC3])
def test_bc_decode_matches_reference_decoder(refbc, fmt):
    rng = np.random.default_rng(11 + fmt)
    blocks = [bytes(b) for b in rng.integers(0, 256, (4000, 16), dtype=np.uint8)] + _special_blocks()
    for blk in blocks:
        got = refrast.decode_bc_block(fmt, blk)
        want = ref_decode(refbc, fmt, blk)
        assert np.array_equal(got, want), (fmt, blk.hex(), got.tolist(), want.tolist())


# The vendored decoder widens with (t / 2^n + t) / 2^n, t = v * 255 + 2^(n-1): exact rounding of
# v * 255 / (2^n - 1) (no halfway cases exist for n = 5, 6) except at these endpoint values, where
# it is one less. The reference-decoder comparison below covers every value.
WIDEN_BELOW_ROUND = {5: {20, 25, 29}, 6: {53}}


def _solid(c):
    return bytes(8) + c.to_bytes(2, "little") * 2 + bytes(4)


@pytest.mark.parametrize("nbits,shift", [(5, 11), (6, 5), (5, 0)])
def test_bc_endpoint_widening_is_rounded_scaling(nbits, shift):
    """A solid block (c0 == c1, all selectors 0) decodes each channel to round(v * 255 / (2^n - 1)),
    one less at the decoder's known exceptions."""
    chan = {11: 0, 5: 1, 0: 2}[shift]
    for v in range(1 << nbits):
        want = (2 * v * 255 + (1 << nbits) - 1) // (2 * ((1 << nbits) - 1))
        want -= v in WIDEN_BELOW_ROUND[nbits]
        for fmt in (BC2, BC3):
            got = refrast.decode_bc_block(fmt, _solid(v << shift))
            assert (got[:, chan] == want).all(), (fmt, v, got[0], want)


def test_bc_every_endpoint_value_matches_reference_decoder(refbc):
    for fmt in (BC2, BC3):
        for c in list(range(0, 1 << 16, 257)) + [v << 11 for v in range(32)] + [v << 5 for v in range(64)] + list(range(32)):
            blk = _solid(c)
            assert np.array_equal(refrast.decode_bc_block(fmt, blk), ref_decode(refbc, fmt, blk)), (fmt, hex(c))


def test_bc2_alpha_is_nibble_times_17():
    nib = np.arange(16, dtype=np.uint64)
    word = int(sum(int(n) << (4 * i) for i, n in enumerate(nib)))
    blk = word.to_bytes(8, "little") + bytes(8)
    got = refrast.decode_bc_block(BC2, blk)
    assert got[:, 3].tolist() == [int(n) * 255 // 15 for n in nib]


def test_bc_palette_modes():
    """c0 > c1: selectors 2 and 3 lie strictly between the endpoints (four colours). BC2 with
    c0 <= c1: selector 3 is black (BC1's three-colour rule); BC3 keeps four colours."""
    hi, lo = 0xFFFF, 0x0000
    sel = sum(k << (2 * k) for k in range(4))  # texels 0..3 use selectors 0..3
    for fmt, c0, c1 in [(BC2, hi, lo), (BC3, hi, lo), (BC2, lo, hi), (BC3, lo, hi)]:
        blk = bytes(8) + c0.to_bytes(2, "little") + c1.to_bytes(2, "little") + sel.to_bytes(4, "little")
        px = refrast.decode_bc_block(fmt, blk)[:4, 0].astype(int)
        e0, e1 = (255, 0) if c0 == hi else (0, 255)
        assert px[0] == e0 and px[1] == e1
        if fmt == BC2 and c0 <= c1:
            assert px[2] == (e0 + e1) // 2 and px[3] == 0
        else:
            assert sorted([e0, px[2], px[3], e1]) == ([e1, px[3], px[2], e0] if e0 > e1 else [e0, px[2], px[3], e1])
            assert abs(px[2] - (2 * e0 + e1) / 3) < 1 and abs(px[3] - (e0 + 2 * e1) / 3) < 1


def _image(arr, w, h, mips, layers, fmt):
    return refrast.make_image(arr, w, h, mips, layers, bpp=1 if fmt in (BC2, BC3) else 4, fmt=fmt)


def _sampler(filt=1, addr=0):
    return refrast.Sampler(mag_filter=filt, min_filter=filt, mipmap_mode=0, address_u=addr, address_v=addr,
                           mip_lod_bias=0.0, min_lod=0.0, max_lod=0.0)


@pytest.mark.parametrize("fmt", [BC2, BC3])
def test_bc_texture_addressing(refbc, fmt):
    """NEAREST sampling of every texel of a 16 x 12 BC texture returns the block-decoded texel:
    pins the block row pitch (width / 4 blocks) and the texel position inside a block."""
    W, H = 16, 12
    rng = np.random.default_rng(5)
    data = rng.integers(0, 256, W * H, dtype=np.uint8)
    full = np.zeros((H, W, 4), np.uint8)
    for by in range(H // 4):
        for bx in range(W // 4):
            k = by * (W // 4) + bx
            full[by * 4:by * 4 + 4, bx * 4:bx * 4 + 4] = ref_decode(refbc, fmt, data[k * 16:k * 16 + 16]).reshape(4, 4, 4)
    img = _image(data, W, H, 1, 1, fmt)
    smp = _sampler(filt=0)
    L = refrast.lib()
    out = (f32 * 4)()
    for y in range(H):
        for x in range(W):
            coord = (f32 * 3)((x + 0.25) / W, (y + 0.25) / H, 0.0)
            L.rr_sample(C.byref(img), C.byref(smp), coord, (f32 * 2)(), (f32 * 2)(), 0, out)
            want = full[y, x].astype(np.float32) / np.float32(255.0)
            assert np.array_equal(np.frombuffer(bytes(out), np.float32), want), (x, y)


# --- cube maps -----------------------------------------------------------------------------------

def _vulkan_face(d):
    """Vulkan 1.0 spec table 'Cube map face selection' for a direction with a unique major axis:
    (face, sc, tc, |ma|)."""
    x, y, z = d
    ax = np.argmax(np.abs(d))
    if ax == 0:
        return (0, -z, -y, abs(x)) if x > 0 else (1, z, -y, abs(x))
    if ax == 1:
        return (2, x, z, abs(y)) if y > 0 else (3, x, -z, abs(y))
    return (4, x, -y, abs(z)) if z > 0 else (5, -x, -y, abs(z))


def _cube_probe(W=16):
    """6 faces; texel (i, j) of face f holds R = 16 i, G = 16 j, B = 40 f."""
    face = np.zeros((6, W, W, 4), np.uint8)
    i = np.arange(W)
    face[..., 0] = (16 * i)[None, None, :]
    face[..., 1] = (16 * i)[None, :, None]
    face[..., 2] = (40 * np.arange(6))[:, None, None]
    face[..., 3] = 255
    return face.reshape(-1)


def test_cube_face_and_orientation_follow_the_spec():
    W = 16
    data = _cube_probe(W)
    img = _image(data, W, W, 1, 6, FMT["R8G8B8A8_UNORM"])
    smp = _sampler()
    L = refrast.lib()
    rng = np.random.default_rng(9)
    checked = 0
    out = (f32 * 4)()
    for d in rng.normal(size=(3000, 3)).astype(np.float32):
        a = np.sort(np.abs(d))
        if a[2] - a[1] < 1e-3:
            continue  # near-ties are covered by the tie-order test
        f, sc, tc, ma = _vulkan_face(d.astype(np.float64))
        U = 0.5 * (sc / ma + 1.0) * W
        V = 0.5 * (tc / ma + 1.0) * W
        for wrapped in (False, True):
            if wrapped:
                L.rr_sample_cube_wrapped(float(d[0]), float(d[1]), float(d[2]), C.byref(img), out)
            else:
                L.rr_sample_cube(C.byref(img), C.byref(smp), (f32 * 3)(*d), out)
            o = np.frombuffer(bytes(out), np.float32) * 255.0
            assert round(o[2] / 40) == f, (d, f, o)
            if U < W - 1 and V < W - 1:  # away from the wrap column / row, bilinear of i is U
                assert abs(o[0] / 16 - U) < 2e-3 and abs(o[1] / 16 - V) < 2e-3, (d, U, V, o)
                checked += 1
    assert checked > 1000


@pytest.mark.parametrize("d,face", [((1, 1, 1), 4), ((1, 1, 0), 2), ((-1, -1, 0), 3), ((1, 0, -1), 5),
                                    ((-2, 2, 2), 4), ((0, 0, 0), 5), ((-0.5, 0.5, 0.25), 2)])
def test_cube_tie_order(d, face):
    """The reference tests +X, -X, +Y, -Y, +Z, -Z in order, each overriding the previous ones:
    ties go to the Z faces, then Y; the zero vector ends on -Z."""
    W = 16
    img = _image(_cube_probe(W), W, W, 1, 6, FMT["R8G8B8A8_UNORM"])
    out = (f32 * 4)()
    refrast.lib().rr_sample_cube(C.byref(img), C.byref(_sampler(filt=0)), (f32 * 3)(*d), out)
    got = np.frombuffer(bytes(out), np.float32)[2] * 255.0
    if d == (0, 0, 0):
        return  # 0 / 0 coordinates: only the face is defined; NaN coordinates sample texel 0
    assert round(got / 40) == face


def test_cube_full_sampler_equals_sample_cube_wrapped():
    """At the reference's only sampler state (REPEAT, LINEAR, one level) the full-sampler path is
    sample_cube_wrapped bit for bit, for RGBA8 and BC3 cubes."""
    rng = np.random.default_rng(3)
    L = refrast.lib()
    out_a, out_b = (f32 * 4)(), (f32 * 4)()
    for fmt in (FMT["R8G8B8A8_UNORM"], BC3):
        W = 32
        n = 6 * W * W * (1 if fmt == BC3 else 4)
        data = rng.integers(0, 256, n, dtype=np.uint8)
        img = _image(data, W, W, 1, 6, fmt)
        smp = _sampler()
        for d in rng.normal(size=(2000, 3)).astype(np.float32):
            L.rr_sample_cube_wrapped(float(d[0]), float(d[1]), float(d[2]), C.byref(img), out_a)
            L.rr_sample_cube(C.byref(img), C.byref(smp), (f32 * 3)(*d), out_b)
            assert bytes(out_a) == bytes(out_b), (fmt, d)


def test_env_sphere_samples_every_face():
    s = scenes.env_sphere(stacks=30, slices=45, width=128, height=96, tex_size=16)
    s.textures[0].data = _cube_probe(16).tobytes()
    col = refrast.render(s)["color"]  # B8G8R8A8 target: byte 0 holds the shader's blue
    faces = set(np.unique(np.round(col[..., 0][col[..., 0] > 0] / 40.0)).astype(int).tolist())
    assert len(faces & {0, 1, 2, 3, 4, 5}) >= 4  # a sphere seen from one side shows 4-5 faces


# --- the executor's texel transcode (visor_amd/csrc/texel_decode.h), compiled for the host -------

def _transcode_host(tmp_path, data, w, h, mips, layers, bpp, fmt_code):
    """Every texel word of every level and layer, as the transcode kernel computes it."""
    src = r"""
#include <cstdio>
#include <cstdint>
#include <vector>
#define VA_TD_FN static inline
""" + open(os.path.join(ROOT, "visor_amd", "csrc", "texel_decode.h")).read() + r"""
int main() {
  std::vector<unsigned char> img;
  int c;
  while((c = getchar()) != EOF) img.push_back((unsigned char)c);
  const uint32_t W = %d, H = %d, M = %d, L = %d, BPP = %d, FMT = %d;
  uint64_t slice = 0;
  for(uint32_t m = 0, w = W, h = H; m < M; m++, w = w > 1 ? w >> 1 : 1, h = h > 1 ? h >> 1 : 1) slice += (uint64_t)w * h * BPP;
  for(uint32_t l = 0; l < L; l++) {
    uint64_t off = 0;
    for(uint32_t m = 0, w = W, h = H; m < M; m++) {
      for(uint32_t y = 0; y < h; y++)
        for(uint32_t x = 0; x < w; x++)
          printf("%%08x\n", va_texel_bytes(img.data(), img.size(), l * slice + off, w, BPP, FMT, x, y));
      off += (uint64_t)w * h * BPP;
      w = w > 1 ? w >> 1 : 1; h = h > 1 ? h >> 1 : 1;
    }
  }
  return 0;
}
""" % (w, h, mips, layers, bpp, fmt_code)
    path = tmp_path / "td.cpp"
    path.write_text(src)
    exe = tmp_path / "td"
    subprocess.run(["g++", "-std=c++17", "-O1", "-w", str(path), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], input=bytes(data), check=True, capture_output=True).stdout.split()
    return np.array([int(x, 16) for x in out], dtype=np.uint32)


@pytest.mark.parametrize("fmt_name,bpp,code", [("BC2_UNORM_BLOCK", 1, 1), ("BC3_UNORM_BLOCK", 1, 2),
                                               ("R8_UNORM", 1, 0), ("R16G16B16A16_SFLOAT", 8, 0)])
def test_transcode_matches_oracle_fetch(tmp_path, fmt_name, bpp, code):
    """Each shadow texel equals the oracle's NEAREST fetch of that texel at every level and layer,
    down to the 2x2 and 1x1 levels whose BC block runs past the level (and the image) end."""
    fmt = FMT[fmt_name]
    W, H, M, Lyr = 16, 8, 5, 2
    slice_ = sum(max(1, W >> m) * max(1, H >> m) * bpp for m in range(M))
    rng = np.random.default_rng(31 + code + bpp)
    data = rng.integers(0, 256, slice_ * Lyr, dtype=np.uint8)
    words = _transcode_host(tmp_path, data, W, H, M, Lyr, bpp, code)
    img = refrast.make_image(data, W, H, M, Lyr, bpp=bpp, fmt=fmt)
    L = refrast.lib()
    out = (f32 * 4)()
    k = 0
    for layer in range(Lyr):
        for m in range(M):
            w, h = max(1, W >> m), max(1, H >> m)
            smp = refrast.Sampler(mag_filter=0, min_filter=0, mipmap_mode=0, address_u=0, address_v=0,
                                  mip_lod_bias=0.0, min_lod=float(m), max_lod=float(m))
            for y in range(h):
                for x in range(w):
                    coord = (f32 * 3)((x + 0.25) / w, (y + 0.25) / h, float(layer))
                    L.rr_sample(C.byref(img), C.byref(smp), coord, (f32 * 2)(), (f32 * 2)(), 0, out)
                    want = np.frombuffer(bytes(out), np.float32)
                    got = np.frombuffer(int(words[k]).to_bytes(4, "little"), np.uint8).astype(np.float32) / np.float32(255.0)
                    assert np.array_equal(got, want), (fmt_name, layer, m, x, y, got, want)
                    k += 1
    assert k == len(words)
