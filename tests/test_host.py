"""Host-side logic without a GPU: workload generation, the SPIR-V translator, the offline
pipeline build, and the exactness of the device's division-free byte-to-float conversion."""
import array
import ctypes as C
import hashlib
import os
import re
import subprocess

import numpy as np
import pytest

from visor_amd import scenes, shaders, vk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    L = C.CDLL(vk.LIB_PATH)
    L.vh_translate_spirv.restype = C.c_int
    L.vh_translate_spirv.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.c_int, C.c_char_p, C.c_size_t]
    L.vh_build_pipeline_offline.restype = C.c_int
    L.vh_build_pipeline_offline.argtypes = [C.POINTER(C.c_uint32), C.c_size_t, C.POINTER(C.c_uint32), C.c_size_t,
                                            C.c_char_p, C.c_char_p, C.c_size_t]
    return L


def _words(name):
    return list(array.array("I", shaders.PROGRAMS[name]()))


def _translate(L, words, stage):
    w = (C.c_uint32 * len(words))(*words)
    out = C.create_string_buffer(1 << 20)
    n = L.vh_translate_spirv(w, len(words), stage, out, len(out))
    return n, out.value.decode()


# ---- workload generation ------------------------------------------------------------------------

def test_splitmix64_known_answers():
    # published SplitMix64 outputs for state 0
    got = [int(x) for x in scenes.splitmix64(0, 4)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F, 0xF88BB8A8724C81EC]
    # offsets continue the same stream
    assert [int(x) for x in scenes.splitmix64(0, 2, offset=2)] == got[2:]


def test_rng_uniform_range_and_stream():
    r = scenes.Rng(9)
    u = r.uniform(100000, -2.0, 3.0)
    assert u.dtype == np.float32 and u.min() >= -2.0 and u.max() < 3.0
    assert abs(float(u.mean()) - 0.5) < 0.02
    r2 = scenes.Rng(9)
    assert np.array_equal(r2.uniform(100000, -2.0, 3.0), u)


def _digest(scene):
    h = hashlib.sha256(scene.vertex_data)
    if scene.index_data is not None:
        h.update(scene.index_data)
    for t in scene.textures:
        h.update(t.data)
    for d in scene.draws:
        h.update(repr((d.pipeline, d.count, d.first, d.indexed)).encode() + bytes(d.push))
    return h.hexdigest()


def test_workloads_are_deterministic():
    assert _digest(scenes.random_triangles(200, 64, 64, seed=4)) == _digest(scenes.random_triangles(200, 64, 64, seed=4))
    assert _digest(scenes.random_triangles(200, 64, 64, seed=4)) != _digest(scenes.random_triangles(200, 64, 64, seed=5))


def test_256_draw_workload_shape():
    s = scenes.many_draws()
    assert (s.width, s.height, len(s.draws)) == (1920, 1080, 256)
    counts = [d.count // 3 for d in s.draws]
    assert min(counts) >= 16 and max(counts) <= 65536
    assert s.triangles == sum(counts)
    assert len(s.textures) == 4 and all(t.width == 512 and t.mips == 10 for t in s.textures)
    # even draws opaque, odd draws blended without depth writes
    for i, d in enumerate(s.draws):
        p = s.pipelines[d.pipeline]
        assert bool(p.blend) == (i % 2 == 1)
        assert p.depth_write == (i % 2 == 0)


def test_mip_chain_is_box_filtered():
    base = np.arange(4 * 4 * 4, dtype=np.uint8).reshape(4, 4, 4)
    packed = np.frombuffer(scenes.mip_chain(base, 3), np.uint8)
    assert packed.size == (16 + 4 + 1) * 4
    assert np.array_equal(packed[:64], base.reshape(-1))
    b = base.astype(np.uint32)
    l1 = (b[0::2, 0::2] + b[1::2, 0::2] + b[0::2, 1::2] + b[1::2, 1::2] + 2) // 4
    assert np.array_equal(packed[64:80], l1.reshape(-1))
    l2 = (l1[0, 0] + l1[1, 0] + l1[0, 1] + l1[1, 1] + 2) // 4
    assert np.array_equal(packed[80:84], l2.reshape(-1))


# ---- SPIR-V translator (reference spirv_compile.cpp) ------------------------------------------------

@pytest.mark.parametrize("name", sorted(shaders.PROGRAMS))
def test_translator_accepts_every_program(lib, name):
    words = _words(name)
    stage = 0 if name.startswith("vs_") else 4
    n, text = _translate(lib, words, stage)
    assert n > 0, text
    assert n == len(text)
    # no host-side math library calls: the translated code uses the device prelude only
    assert "std::" not in text


def test_translator_rejects_bad_modules(lib):
    words = _words("fs_color")
    n, text = _translate(lib, [0xDEADBEEF] + words[1:], 4)
    assert n == -1 and text
    n, text = _translate(lib, words[:3], 4)
    assert n == -1 and text
    # a function left open (OpReturn and OpFunctionEnd cut off)
    n, text = _translate(lib, words[: len(words) - 2], 4)
    assert n == -1 and "OpFunctionEnd" in text
    # an instruction whose word count runs past the end of the module
    idx = 5
    while words[idx] >> 16 < 3:
        idx += words[idx] >> 16
    n, text = _translate(lib, words[: idx + 2], 4)
    assert n == -1 and "truncated" in text
    # wrong stage for the entry point
    n, text = _translate(lib, words, 0)
    assert n == -1 and text


def test_translator_rejects_unknown_opcode(lib):
    words = _words("fs_color")
    # find OpReturn (opcode 253, one word) and replace it by an unassigned opcode of the same length
    idx = 5
    while idx < len(words):
        wc, op = words[idx] >> 16, words[idx] & 0xFFFF
        if op == 253:
            break
        idx += wc
    assert idx < len(words)
    bad = words[:idx] + [(1 << 16) | 4095] + words[idx + 1:]
    n, text = _translate(lib, bad, 4)
    assert n == -1 and "4095" in text


def test_translator_output_is_deterministic(lib):
    words = _words("fs_multitap8")
    assert _translate(lib, words, 4) == _translate(lib, words, 4)


def test_offline_pipeline_build(lib, tmp_path, monkeypatch):
    """hiprtc compiles a pair for gfx950 without a GPU; the second call hits the on-disk cache."""
    monkeypatch.setenv("VISOR_CACHE_DIR", str(tmp_path))
    vw = array.array("I", shaders.PROGRAMS["vs_mvp_pos4_col"]())
    fw = array.array("I", shaders.PROGRAMS["fs_color"]())
    err = C.create_string_buffer(4096)
    args = ((C.c_uint32 * len(vw))(*vw), len(vw), (C.c_uint32 * len(fw))(*fw), len(fw), b"gfx950", err, 4096)
    n = lib.vh_build_pipeline_offline(*args)
    assert n > 0, err.value
    files = os.listdir(tmp_path)
    assert any(f.endswith(".hsaco") for f in files)
    assert lib.vh_build_pipeline_offline(*args) == n


# ---- the device's byte -> float conversion --------------------------------------------------------

def test_unorm8_division_free_is_exact(tmp_path):
    """va_unorm8 (pipeline_prelude.hip) replaces float(b) / 255.0f by a product and an FMA with a
    split constant. Its source is compiled for the host and compared with the IEEE quotient for
    every byte value."""
    src = open(os.path.join(ROOT, "visor_amd", "csrc", "pipeline_prelude.hip")).read()
    m = re.search(r"VA_INLINE float va_unorm8\(u32 b\)\s*\{.*?\n\}", src, flags=re.S)
    assert m
    prog = ("#include <stdio.h>\n#include <string.h>\ntypedef unsigned u32;\n#define VA_INLINE static\n" + m.group(0)
            + "\nint main(void){int bad=0;for(u32 b=0;b<256;b++){volatile float d=255.0f;float q=(float)b/d;"
              "float g=va_unorm8(b);if(memcmp(&q,&g,4))bad++;}printf(\"%d\\n\",bad);return 0;}\n")
    (tmp_path / "u.c").write_text(prog)
    exe = tmp_path / "u"
    subprocess.run(["gcc", "-O0", "-ffp-contract=off", "-std=c11", str(tmp_path / "u.c"), "-o", str(exe), "-lm"],
                   check=True)
    assert subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.strip() == "0"
