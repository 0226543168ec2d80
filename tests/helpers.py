"""Small explicit scenes for the tests: triangles given in window-exact clip coordinates."""
import struct

import numpy as np

from visor_amd import scenes
from visor_amd.vk import BF, CMP, CULL_NONE, FRONT_CCW, TOPO_LIST


def clip_from_window(X, Y, W, H):
    """Clip-space (x, y) with w = 1 that the window transform maps exactly to integer (X, Y)
    (W, H powers of two, so every step is exact in binary32)."""
    return X * 2.0 / W - 1.0, 1.0 - Y * 2.0 / H


def tri_scene(tris, W=32, H=32, colors=None, depths=None, pipeline=None, topology=TOPO_LIST, clear_depth=1.0,
              clear_color=(0.0, 0.0, 0.0, 0.0), draws_per_tri=False, strip=None, pipelines=None, draw_pipes=None):
    """Scene of window-space triangles [(X0,Y0),(X1,Y1),(X2,Y2)], colour passthrough."""
    verts = []
    cols = []
    if strip is not None:
        pts = strip
    else:
        pts = [p for t in tris for p in t]
    for i, (X, Y) in enumerate(pts):
        x, y = clip_from_window(X, Y, W, H)
        z = depths[i] if depths is not None else 0.5
        verts.append((x, y, z, 1.0))
        c = colors[i] if colors is not None else (255, 255, 255, 255)
        cols.append(c)
    pos = np.array(verts, dtype=np.float32)
    col = np.array(cols, dtype=np.uint8)
    pipes = pipelines or {"p": pipeline or scenes.Pipeline("vs_mvp_pos4_col", "fs_color", topology=topology)}
    push = scenes.push_block(np.eye(4))
    if draws_per_tri:
        draws = [scenes.Draw(draw_pipes[i] if draw_pipes else "p", 3, 3 * i, False, push) for i in range(len(pts) // 3)]
    else:
        draws = [scenes.Draw("p", len(pts), 0, False, push)]
    return scenes.Scene("test", W, H, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes, draws,
                        clear_color=clear_color, clear_depth=clear_depth)


def covered_mask(color):
    """Pixels whose colour bytes changed from the zero clear."""
    return (color[..., :3] != 0).any(-1)


def inside_closed_triangle(tri, x, y):
    """Exact point-in-closed-triangle test by orientation signs (Python integers, no overflow)."""
    (ax, ay), (bx, by), (cx, cy) = tri
    d1 = (bx - ax) * (y - ay) - (by - ay) * (x - ax)
    d2 = (cx - bx) * (y - by) - (cy - by) * (x - bx)
    d3 = (ax - cx) * (y - cy) - (ay - cy) * (x - cx)
    neg = d1 < 0 or d2 < 0 or d3 < 0
    pos = d1 > 0 or d2 > 0 or d3 > 0
    return not (neg and pos)
