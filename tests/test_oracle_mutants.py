"""Planted mistakes in the oracle, each of which a pin must catch.

Every case copies oracle/refrast.c, plants one plausible restatement error (the kind a pin exists
to rule out), builds the copy with the oracle's own Makefile and flags, and runs the named pins
against it: each pin must fail on the mutant (and the rest of the suite shows it passes on the real
oracle). A mutant no pin catches would mean that part of the oracle is not pinned.

Mistakes planted (reference lines they would break):
  depth_after_perspective  depth interpolated with the perspective-corrected weights
                           (rasterizer.cpp:551-558 uses the pre-perspective ones)
  depth_after_invw         depth taken after the weights are multiplied by 1/w
  mat3_transposed          normal matrix applied as M^T n (spirv_compile.cpp:433-441)
  ubo_offsets_swapped      uniform block members read at each other's offsets (spirv_compile.cpp:552-558)
  mat4_transposed          MVP applied as M^T v (spirv_compile.cpp:423-431)
  vertex_offset_applied    GetIndex adds vkCmdDrawIndexed's vertexOffset (rasterizer.cpp:100-119)
  saturating_f2i           float->int saturates as a GPU converts natively, instead of x64's
                           0x80000000 (rasterizer.cpp:248-249, 341-345)
  floor_window             ToWindow floors instead of truncating (rasterizer.cpp:248-249)
  inclusive_bbox_max       the block loop includes the bbox maximum row (rasterizer.cpp:538-540)
  src_alpha_factor         SRC_ALPHA blend factor returns 1 - alpha (rasterizer.cpp:600-640)
"""
import functools
import os
import shutil
import subprocess

import pytest

import test_oracle
import test_oracle_shaders
from oracle import refrast

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

MUTANTS = [
    ("depth_after_perspective",
     """  if(pixdepth)
    *pixdepth = n0 * w->depth[0] + n1 * w->depth[1] + n2 * w->depth[2];
  n0 *= w->invw[0];
  n1 *= w->invw[1];
  n2 *= w->invw[2];
  float invlen = 1.0f / (n0 + n1 + n2);""",
     """  n0 *= w->invw[0];
  n1 *= w->invw[1];
  n2 *= w->invw[2];
  float invlen = 1.0f / (n0 + n1 + n2);
  if(pixdepth)
    *pixdepth = (n0 * invlen) * w->depth[0] + (n1 * invlen) * w->depth[1] + (n2 * invlen) * w->depth[2];""",
     [test_oracle_shaders.test_depth_matches_screen_affine_interpolation]),
    ("depth_after_invw",
     """  if(pixdepth)
    *pixdepth = n0 * w->depth[0] + n1 * w->depth[1] + n2 * w->depth[2];
  n0 *= w->invw[0];
  n1 *= w->invw[1];
  n2 *= w->invw[2];""",
     """  n0 *= w->invw[0];
  n1 *= w->invw[1];
  n2 *= w->invw[2];
  if(pixdepth)
    *pixdepth = n0 * w->depth[0] + n1 * w->depth[1] + n2 * w->depth[2];""",
     [test_oracle_shaders.test_depth_matches_screen_affine_interpolation]),
    ("mat3_transposed",
     """    for(int c = 0; c < 3; c++)
      acc += m[c * 4 + r] * v[c];""",
     """    for(int c = 0; c < 3; c++)
      acc += m[r * 4 + c] * v[c];""",
     [test_oracle_shaders.test_lambert_through_a_rotation_normal_matrix]),
    ("ubo_offsets_swapped",
     """      ubo_float(st, RR_UBO_BINDING, 64, tint, 4);
      ubo_float(st, RR_UBO_BINDING, 80, bias, 4);""",
     """      ubo_float(st, RR_UBO_BINDING, 80, tint, 4);
      ubo_float(st, RR_UBO_BINDING, 64, bias, 4);""",
     [test_oracle_shaders.test_uniform_buffer_programs]),
    ("mat4_transposed",
     """    for(int c = 0; c < 4; c++)
      acc += m[c * 4 + r] * v[c];""",
     """    for(int c = 0; c < 4; c++)
      acc += m[r * 4 + c] * v[c];""",
     [test_oracle_shaders.test_mvp_translation_moves_the_image_by_whole_pixels]),
    ("vertex_offset_applied",
     """  uint32_t v;
  memcpy(&v, st->ib + 4ull * i, 4);
  return v;
}""",
     """  uint32_t v;
  memcpy(&v, st->ib + 4ull * i, 4);
  return v + (uint32_t)st->vertex_offset;
}""",
     [test_oracle.test_vertex_offset_is_not_applied]),
    ("saturating_f2i",
     """    return (int32_t)f;
  return INT32_MIN;""",
     """    return (int32_t)f;
  return f != f ? 0 : (f < 0.0f ? INT32_MIN : INT32_MAX);""",
     [test_oracle.test_float_to_int_is_x86_truncation]),
    ("floor_window",
     """  out[0] = rr_f2i((p[0] / p[3] + 1.0f) * 0.5f * (float)w);""",
     """  out[0] = rr_f2i(floorf((p[0] / p[3] + 1.0f) * 0.5f * (float)w));""",
     [test_oracle.test_to_window_special_values_and_truncation]),
    ("inclusive_bbox_max",
     """  for(int y = w->miny; y < w->maxy; y++)""",
     """  for(int y = w->miny; y <= w->maxy; y++)""",
     [functools.partial(test_oracle.test_right_triangle_closed_form, 5)]),
    ("src_alpha_factor",
     """    case RR_BF_SRC_ALPHA: return src_alpha;""",
     """    case RR_BF_SRC_ALPHA: return 1.0f - src_alpha;""",
     [test_oracle.test_blend_identities]),
]


@pytest.fixture(scope="module")
def source():
    with open(os.path.join(ORACLE, "refrast.c")) as f:
        return f.read()


@pytest.mark.parametrize("name,old,new,pins", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_fails_its_pin(name, old, new, pins, source, tmp_path):
    assert source.count(old) == 1, "mutation site of %s not found exactly once" % name
    for fn in ("Makefile", "refrast.h"):
        shutil.copy(os.path.join(ORACLE, fn), tmp_path / fn)
    (tmp_path / "refrast.c").write_text(source.replace(old, new))
    r = subprocess.run(["make", "-s", "-C", str(tmp_path)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    with refrast.use_library(str(tmp_path / "_build" / "librefrast.so")):
        for pin in pins:
            with pytest.raises(AssertionError):
                pin()
