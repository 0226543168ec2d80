"""Seeded workloads for the draw path (one description drives the GPU path and the oracle).

A Scene is an application-level description: one interleaved vertex buffer, an optional index
buffer, textures with their sampler state, pipelines (shader programs + fixed-function state) and
an ordered list of draws inside one render pass that clears colour and depth.

Random numbers come from SplitMix64 with 24-bit uniform floats, so every workload is fixed by its
seed and reproducible on any host.
"""

import math
import struct
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import shaders
from .vk import BF, CMP, CULL_BACK, CULL_NONE, FMT, FRONT_CCW, TOPO_LIST, ADDR_CLAMP, ADDR_REPEAT, \
    FILTER_LINEAR, MIPMAP_LINEAR, MIPMAP_NEAREST

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n, offset=0):
    """n consecutive SplitMix64 outputs starting after `offset` draws of stream `seed`."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GOLDEN
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


class Rng:
    def __init__(self, seed):
        self.seed = seed
        self.pos = 0

    def raw(self, n):
        r = splitmix64(self.seed, n, self.pos)
        self.pos += n
        return r

    def uniform(self, n, lo=0.0, hi=1.0):
        u = (self.raw(n) >> np.uint64(40)).astype(np.float64) / float(1 << 24)
        return (lo + (hi - lo) * u).astype(np.float32)

    def bytes(self, n):
        return (self.raw(n) >> np.uint64(56)).astype(np.uint8)


@dataclass
class Texture:
    width: int
    height: int
    mips: int
    layers: int
    data: bytes  # texels in `format`, layer-major, mips packed inside each layer
    mag_filter: int = FILTER_LINEAR
    min_filter: int = FILTER_LINEAR
    mipmap_mode: int = MIPMAP_NEAREST
    address: int = ADDR_REPEAT
    lod_bias: float = 0.0
    min_lod: float = 0.0
    max_lod: float = 0.0
    format: int = FMT["R8G8B8A8_UNORM"]  # R8G8B8A8_UNORM, BC2_UNORM_BLOCK or BC3_UNORM_BLOCK
    cube: bool = False  # 6 layers viewed as a cube (+X, -X, +Y, -Y, +Z, -Z)


@dataclass
class Pipeline:
    vs: str
    fs: str
    topology: int = TOPO_LIST
    cull: int = CULL_NONE
    front: int = FRONT_CCW
    depth_op: int = CMP["LESS"]
    depth_write: bool = True
    blend: bool = False
    src: int = BF["ONE"]
    dst: int = BF["ZERO"]


@dataclass
class Draw:
    pipeline: str
    count: int
    first: int = 0
    indexed: bool = False
    push: bytes = b"\0" * 128
    texture: Optional[int] = None
    vertex_offset: int = 0   # vkCmdDrawIndexed's vertexOffset (recorded; the reference never applies it)
    uniform: Optional[int] = None  # index into Scene.uniforms, bound at set 0 binding 1


@dataclass
class Scene:
    name: str
    width: int
    height: int
    vertex_data: bytes
    vertex_stride: int
    vattrs: list  # (location, format, offset)
    pipelines: dict
    draws: List[Draw]
    index_data: Optional[bytes] = None
    index_type: int = 1
    textures: List[Texture] = field(default_factory=list)
    color_format: int = FMT["B8G8R8A8_UNORM"]
    clear_color: tuple = (0.0, 0.0, 0.0, 0.0)
    clear_depth: Optional[float] = 1.0
    uniforms: List[bytes] = field(default_factory=list)  # uniform-buffer contents the draws bind

    @property
    def triangles(self):
        n = 0
        for d in self.draws:
            topo = self.pipelines[d.pipeline].topology
            n += d.count // 3 if topo == TOPO_LIST else max(d.count - 2, 0)
        return n

    @property
    def color_bpp(self):
        return 8 if self.color_format == FMT["R16G16B16A16_SFLOAT"] else 4


# ---- helpers ----------------------------------------------------------------------------------

def mat4_bytes(m):
    """Column-major float32 bytes of a 4x4 matrix given row-major (m[row][col])."""
    return np.asarray(m, dtype=np.float32).T.tobytes()


def mat3_std140(m):
    """mat3 with 16-byte column stride (std430/std140 MatrixStride 16)."""
    a = np.zeros((3, 4), dtype=np.float32)
    a[:, :3] = np.asarray(m, dtype=np.float32).T
    return a.tobytes()


def push_block(mvp, nmat=None, alpha=1.0):
    b = bytearray(128)
    b[0:64] = mat4_bytes(mvp)
    if nmat is not None:
        b[64:112] = mat3_std140(nmat)
    b[112:116] = struct.pack("<f", alpha)
    return bytes(b)


def perspective(fovy, aspect, near, far):
    f = 1.0 / math.tan(fovy / 2)
    # depth in [0, 1], clip y kept as is (the window transform flips it)
    return np.array([[f / aspect, 0, 0, 0],
                     [0, f, 0, 0],
                     [0, 0, far / (near - far), near * far / (near - far)],
                     [0, 0, -1, 0]], dtype=np.float64)


def rotation(ax, ay, az):
    cx, sx, cy, sy, cz, sz = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay), math.cos(az), math.sin(az)
    rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    rz = np.array([[cz, -sz, 0], [sz, cz, 0], [0, 0, 1]])
    return rz @ ry @ rx


def model_matrix(rot, scale, trans):
    m = np.eye(4)
    m[:3, :3] = rot * scale
    m[:3, 3] = trans
    return m


def mip_chain(base, levels):
    """Box-filtered mip chain of an (H, W, 4) uint8 image, packed level after level."""
    out = [base]
    cur = base.astype(np.uint32)
    for _ in range(1, levels):
        h, w = cur.shape[:2]
        nh, nw = max(1, h // 2), max(1, w // 2)
        if h > 1 and w > 1:
            cur = (cur[0:2 * nh:2, 0:2 * nw:2] + cur[1:2 * nh:2, 0:2 * nw:2] + cur[0:2 * nh:2, 1:2 * nw:2] +
                   cur[1:2 * nh:2, 1:2 * nw:2] + 2) // 4
        elif h > 1:
            cur = (cur[0:2 * nh:2] + cur[1:2 * nh:2] + 1) // 2
        else:
            cur = (cur[:, 0:2 * nw:2] + cur[:, 1:2 * nw:2] + 1) // 2
        out.append(cur.astype(np.uint8))
    return b"".join(o.tobytes() for o in out)


def noise_texture(rng, size, levels, layers=1, smooth=True):
    data = b""
    for _ in range(layers):
        base = rng.bytes(size * size * 4).reshape(size, size, 4)
        if smooth:
            # a few box blurs make the noise band-limited so that mips matter but stay varied
            f = base.astype(np.float32)
            for _ in range(2):
                f = (f + np.roll(f, 1, 0) + np.roll(f, 1, 1) + np.roll(np.roll(f, 1, 0), 1, 1)) / 4
            base = np.clip(f, 0, 255).astype(np.uint8)
        base[..., 3] = 255
        data += mip_chain(base, levels)
    return data


def full_mips(size):
    return int(math.log2(size)) + 1


VATTRS_POS_COL = [(0, FMT["R32G32B32A32_SFLOAT"], 0), (1, FMT["R8G8B8A8_UNORM"], 16)]
VATTRS_PNU = [(0, FMT["R32G32B32_SFLOAT"], 0), (1, FMT["R32G32B32_SFLOAT"], 12), (2, FMT["R32G32_SFLOAT"], 24)]
IDENTITY = np.eye(4)


def interleave_pos_col(pos, col):
    n = pos.shape[0]
    v = np.zeros(n, dtype=[("p", "<f4", 4), ("c", "u1", 4)])
    v["p"] = pos
    v["c"] = col
    return v.tobytes()


def interleave_pnu(pos, nrm, uv):
    n = pos.shape[0]
    v = np.zeros(n, dtype=[("p", "<f4", 3), ("n", "<f4", 3), ("t", "<f4", 2)])
    v["p"], v["n"], v["t"] = pos, nrm, uv
    return v.tobytes()


# ---- configurations ------------------------------------------------------------------------------

def random_triangles(ntris=1000, width=256, height=256, seed=1):
    """256x256 RGBA8 + D32, 1000 non-indexed random clip-space triangles, colour passthrough."""
    rng = Rng(seed)
    nv = 3 * ntris
    pos = np.zeros((nv, 4), dtype=np.float32)
    pos[:, 0] = rng.uniform(nv, -1.0, 1.0)
    pos[:, 1] = rng.uniform(nv, -1.0, 1.0)
    pos[:, 2] = rng.uniform(nv, 0.0, 1.0)
    pos[:, 3] = 1.0
    col = rng.bytes(nv * 4).reshape(nv, 4)
    pipes = {"main": Pipeline("vs_mvp_pos4_col", "fs_color")}
    draws = [Draw("main", nv, 0, False, push_block(IDENTITY))]
    return Scene("random_triangles", width, height, interleave_pos_col(pos, col), 20, VATTRS_POS_COL, pipes, draws)


def uv_sphere(stacks, slices, radius, rng=None, jitter=0.0, center=(0, 0, 0)):
    """Positions, normals, uvs and CCW (outward) triangle indices of a UV sphere grid."""
    i = np.arange(stacks + 1, dtype=np.float64)
    j = np.arange(slices + 1, dtype=np.float64)
    th = i[:, None] / stacks * math.pi            # polar
    ph = j[None, :] / slices * 2 * math.pi        # azimuth
    x = np.sin(th) * np.cos(ph)
    y = np.cos(th) * np.ones_like(ph)
    z = np.sin(th) * np.sin(ph)
    nrm = np.stack([x, y, z], -1).reshape(-1, 3)
    r = np.full(nrm.shape[0], radius)
    if rng is not None and jitter:
        r = r * (1.0 + jitter * rng.uniform(nrm.shape[0], -1.0, 1.0).astype(np.float64))
    pos = nrm * r[:, None] + np.asarray(center)
    uv = np.stack([(j[None, :] / slices) * np.ones_like(th), (i[:, None] / stacks) * np.ones_like(ph)], -1).reshape(-1, 2)
    a = (np.arange(stacks)[:, None] * (slices + 1) + np.arange(slices)[None, :]).reshape(-1)
    b = a + slices + 1
    tri = np.stack([np.stack([a, a + 1, b], -1), np.stack([a + 1, b + 1, b], -1)], 1).reshape(-1, 3)
    return pos.astype(np.float32), nrm.astype(np.float32), uv.astype(np.float32), tri.astype(np.uint32)


def _orient_front_ccw(tri, pos, mvp, width, height):
    """Order triangle vertices so that faces pointing at the camera have positive window area
    (the front-face rule of the reference's setup with frontFace = COUNTER_CLOCKWISE)."""
    p = np.c_[pos.astype(np.float64), np.ones(len(pos))] @ np.asarray(mvp).T
    wx = (p[:, 0] / p[:, 3] + 1) * 0.5 * width
    wy = (-p[:, 1] / p[:, 3] + 1) * 0.5 * height
    ax, ay = wx[tri[:, 0]], wy[tri[:, 0]]
    area = (wx[tri[:, 1]] - ax) * (wy[tri[:, 2]] - ay) - (wy[tri[:, 1]] - ay) * (wx[tri[:, 2]] - ax)
    # the nearest quarter of the faces of a convex mesh all face the camera
    w = (p[tri[:, 0], 3] + p[tri[:, 1], 3] + p[tri[:, 2], 3]) / 3
    near = w <= np.quantile(w, 0.25)
    sel = area[near & (area != 0)]
    if sel.size and np.median(sel) < 0:
        tri = tri[:, [0, 2, 1]]
    return tri


def lit_sphere(stacks=250, slices=400, width=1920, height=1080, tex_size=1024, seed=2):
    """One indexed draw: 200k-triangle perturbed UV sphere, MVP + normal matrix, Lambert x
    mipmapped noise texture (bilinear, clamp), LESS, back-face cull."""
    rng = Rng(seed)
    pos, nrm, uv, tri = uv_sphere(stacks, slices, 1.0, rng, jitter=0.05)
    proj = perspective(math.radians(60), width / height, 0.5, 10.0)
    model = model_matrix(rotation(0.3, 0.5, 0.0), 1.0, (0.0, 0.0, -3.2))
    mvp = proj @ model
    nmat = np.linalg.inv(model[:3, :3]).T
    tri = _orient_front_ccw(tri, pos, mvp, width, height)
    levels = full_mips(tex_size)
    tex = Texture(tex_size, tex_size, levels, 1, noise_texture(rng, tex_size, levels), address=ADDR_CLAMP,
                  mipmap_mode=MIPMAP_NEAREST, max_lod=float(levels - 1))
    pipes = {"lit": Pipeline("vs_mvp_normal_uv", "fs_lambert_tex", cull=CULL_BACK)}
    draws = [Draw("lit", tri.size, 0, True, push_block(mvp, nmat), texture=0)]
    return Scene("lit_sphere", width, height, interleave_pnu(pos, nrm, uv), 32, VATTRS_PNU, pipes, draws,
                 index_data=tri.astype(np.uint32).tobytes(), textures=[tex])


def bc_texture_data(rng, size, levels, layers=1):
    """Random BC2/BC3 blocks for a size x size texture (size a multiple of 4): every 16-byte
    pattern is a valid block, and random endpoints cover both colour-palette and both
    alpha-palette modes. Level m holds (size >> m)^2 bytes, the linear layout's 1 byte per texel."""
    out = []
    for _ in range(layers):
        for m in range(levels):
            n = max(1, size >> m)
            out.append(rng.bytes(n * n).tobytes())
    return b"".join(out)


def env_sphere(stacks=60, slices=90, width=256, height=192, tex_size=64, fmt=FMT["R8G8B8A8_UNORM"], seed=4):
    """The perturbed sphere of lit_sphere, shaded by a cube-map lookup along the interpolated
    normal (fs_cube_env): every face and the edges between faces are sampled. The cube texture is
    RGBA8 noise or random BC2/BC3 blocks, one level, bilinear, repeat."""
    s = lit_sphere(stacks, slices, width, height, tex_size, seed)
    rng = Rng(seed + 100)
    if fmt == FMT["R8G8B8A8_UNORM"]:
        data = noise_texture(rng, tex_size, 1, layers=6, smooth=False)
    else:
        data = bc_texture_data(rng, tex_size, 1, layers=6)
    s.textures = [Texture(tex_size, tex_size, 1, 6, data, format=fmt, cube=True)]
    s.pipelines = {"lit": Pipeline("vs_mvp_normal_uv", "fs_cube_env", cull=CULL_BACK)}
    s.name = "env_sphere"
    return s


def many_draws(ndraws=256, width=1920, height=1080, tex_size=512, seed=3, min_tris=16, max_tris=65536):
    """256 draws of log-uniform 16..65536 triangles (about 2M), per-draw MVP, 4 trilinear repeat
    textures; even draws opaque LESS, odd draws alpha-blended (depth write off)."""
    rng = Rng(seed)
    counts = np.exp(rng.uniform(ndraws, math.log(min_tris), math.log(max_tris)).astype(np.float64))
    counts = np.maximum(np.round(counts).astype(np.int64), min_tris)
    proj = perspective(math.radians(60), width / height, 0.5, 50.0)
    levels = full_mips(tex_size)
    textures = [Texture(tex_size, tex_size, levels, 1, noise_texture(rng, tex_size, levels), address=ADDR_REPEAT,
                        mipmap_mode=MIPMAP_LINEAR, max_lod=float(levels - 1)) for _ in range(4)]
    pos_all, nrm_all, uv_all, idx_all, draws = [], [], [], [], []
    vbase = 0
    ibase = 0
    pipes = {"opaque": Pipeline("vs_mvp_normal_uv", "fs_lambert_tex_a"),
             "blend": Pipeline("vs_mvp_normal_uv", "fs_lambert_tex_a", depth_write=False, blend=True,
                               src=BF["SRC_ALPHA"], dst=BF["ONE_MINUS_SRC_ALPHA"])}
    params = rng.uniform(ndraws * 9)
    for d in range(ndraws):
        n = int(counts[d])
        stacks = max(2, int(round(math.sqrt(n / 4.0))))
        slices = max(3, int(round(n / (2.0 * stacks))))
        pos, nrm, uv, tri = uv_sphere(stacks, slices, 1.0)
        p = params[d * 9:(d + 1) * 9].astype(np.float64)
        rot = rotation(p[0] * 6.283, p[1] * 6.283, p[2] * 6.283)
        scale = 0.3 + 1.2 * p[3]
        z = -(4.0 + 20.0 * p[4])
        half_h = math.tan(math.radians(30)) * -z
        trans = ((p[5] * 2 - 1) * half_h * width / height * 0.9, (p[6] * 2 - 1) * half_h * 0.9, z)
        model = model_matrix(rot, scale, trans)
        mvp = proj @ model
        nmat = np.linalg.inv(model[:3, :3]).T
        tri = _orient_front_ccw(tri, pos, mvp, width, height)
        alpha = 1.0 if d % 2 == 0 else 0.2 + 0.7 * float(p[7])
        tex = int(p[8] * 4) % 4
        pos_all.append(pos)
        nrm_all.append(nrm)
        uv_all.append(uv * np.float32(1.0 + 3.0 * p[3]))
        idx_all.append((tri + vbase).astype(np.uint32).reshape(-1))
        draws.append(Draw("opaque" if d % 2 == 0 else "blend", tri.size, ibase, True, push_block(mvp, nmat, alpha),
                          texture=tex))
        vbase += len(pos)
        ibase += tri.size
    vdata = interleave_pnu(np.concatenate(pos_all), np.concatenate(nrm_all), np.concatenate(uv_all))
    return Scene("many_draws", width, height, vdata, 32, VATTRS_PNU, pipes, draws,
                 index_data=np.concatenate(idx_all).tobytes(), textures=textures)


def micro_triangles(ntris=4_000_000, ndraws=16, width=3840, height=2160, tex_size=2048, layers=4, seed=5):
    """Micro-triangles: 4M indexed triangles of about 4 px in 50 depth layers on a 4K RGBA16F
    target; 8 trilinear taps of a 4-layer mipmapped array texture; 25% of draws additive."""
    rng = Rng(seed)
    per = ntris // ndraws
    levels = full_mips(tex_size)
    tex = Texture(tex_size, tex_size, levels, layers, noise_texture(rng, tex_size, levels, layers, smooth=False),
                  address=ADDR_REPEAT, mipmap_mode=MIPMAP_LINEAR, max_lod=float(levels - 1))
    n = per * ndraws
    cx = rng.uniform(n, -1.0, 1.0)
    cy = rng.uniform(n, -1.0, 1.0)
    layer = np.floor(rng.uniform(n, 0.0, 50.0)).astype(np.float32)
    ang = rng.uniform(n, 0.0, 6.2831853)
    # equilateral triangles of circumradius 1.75 px: area 1.3 * 1.75^2 ~ 4 px
    sx, sy = np.float32(2.0 * 1.75 / width), np.float32(2.0 * 1.75 / height)
    pos = np.zeros((n, 3, 3), dtype=np.float32)
    for k, off in enumerate((0.0, 2.094, 4.189)):
        pos[:, k, 0] = cx + sx * np.cos(ang + off)
        pos[:, k, 1] = cy + sy * np.sin(ang + off)
        pos[:, k, 2] = (layer + 0.5) / 50.0
    nrm = np.zeros((n * 3, 3), dtype=np.float32)
    nrm[:, 2] = 1.0
    uv = np.stack([(pos[..., 0] + 1) * 0.5, (pos[..., 1] + 1) * 0.5], -1).reshape(-1, 2).astype(np.float32)
    idx = np.arange(n * 3, dtype=np.uint32)
    # orient every triangle so it is not culled (cull none anyway); window area sign is irrelevant
    pipes = {"opaque": Pipeline("vs_mvp_normal_uv", "fs_multitap8"),
             "add": Pipeline("vs_mvp_normal_uv", "fs_multitap8", depth_write=False, blend=True, src=BF["ONE"], dst=BF["ONE"])}
    draws = []
    for d in range(ndraws):
        draws.append(Draw("add" if d % 4 == 3 else "opaque", per * 3, d * per * 3, True,
                          push_block(IDENTITY, np.eye(3)), texture=0))
    return Scene("micro_triangles", width, height, interleave_pnu(pos.reshape(-1, 3), nrm, uv), 32, VATTRS_PNU, pipes, draws,
                 index_data=idx.tobytes(), textures=[tex], color_format=FMT["R16G16B16A16_SFLOAT"])


WORKLOADS = {"random_triangles": random_triangles, "lit_sphere": lit_sphere, "many_draws": many_draws,
             "micro_triangles": micro_triangles}
