"""Parity of the GPU draw path with the oracle (bit-exact coverage, depth and colour)."""
import numpy as np
import pytest

from oracle import refrast
from visor_amd import render, scenes

pytestmark = pytest.mark.gpu


def _compare(scene, out, ref, color_tol=0):
    col, dep = out["color"], out["depth"]
    assert col.shape == ref["color"].shape
    if color_tol == 0:
        diff = np.argwhere((col != ref["color"]).any(-1))
        assert diff.size == 0, "colour differs at %d pixels, first %s gpu %s ref %s" % (
            len(diff), diff[:3].tolist(), col[tuple(diff[0])].tolist(), ref["color"][tuple(diff[0])].tolist())
    else:
        d = np.abs(col.astype(np.int32) - ref["color"].astype(np.int32)).max()
        assert d <= color_tol
    if dep is not None:
        bad = np.argwhere(dep.view(np.uint32) != ref["depth"].view(np.uint32))
        assert bad.size == 0, "depth differs at %d pixels, first %s" % (len(bad), bad[:3].tolist())


def _render_gpu(ctx, scene, batched=True, tile=16):
    r = render.SceneRenderer(ctx, scene)
    ctx.set_mode(batched=batched, tile=tile)
    r.render()
    out = {"color": r.read_color(), "depth": r.read_depth() if r.has_depth else None, "stats": ctx.stats(0)}
    return out


@pytest.mark.parametrize("batched", [True, False])
def test_random_triangles_bit_exact(ctx, batched):
    s = scenes.random_triangles()
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, batched=batched)
    _compare(s, out, ref)
    st = out["stats"]
    c = ref["counters"]
    assert st["pixels_written"] == c["pixels_written"]
    assert st["depth_passed"] == c["depth_passed"]
    assert st["tris_in"] == c["tris_in"] and st["tris_out"] == c["tris_out"]
    assert st["pixels_tested"] == c["pixels_tested"]


def test_lit_sphere_small_bit_exact(ctx):
    """Indexed sphere, MVP + normal matrix, Lambert x mipmapped texture, clamp, back-face cull."""
    s = scenes.lit_sphere(stacks=40, slices=64, width=320, height=180, tex_size=64)
    ref = refrast.render(s)
    out = _render_gpu(ctx, s)
    _compare(s, out, ref)
    assert out["stats"]["pixels_written"] == ref["counters"]["pixels_written"] > 0
    assert out["stats"]["tris_out"] == ref["counters"]["tris_out"] < ref["counters"]["tris_in"]


@pytest.mark.parametrize("tile", [8, 16, 32])
def test_many_draws_small_bit_exact(ctx, tile):
    """Many draws, per-draw MVP, trilinear repeat, alternating opaque / alpha-blended."""
    s = scenes.many_draws(ndraws=24, width=320, height=240, tex_size=64, max_tris=2048)
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, tile=tile)
    _compare(s, out, ref)
    assert out["stats"]["depth_passed"] == ref["counters"]["depth_passed"]


@pytest.mark.parametrize("seed", range(20, 28))
def test_many_draws_seeds_odd_sizes_bit_exact(ctx, seed):
    """Seeded variants on a 333x197 target, which no tile size divides: partial edge tiles, and
    tile counts with a remainder past the last whole round of per-XCD tile runs (1050 tiles at
    8x8, 273 at 16x16, 77 at 32x32); the tile size cycles with the seed."""
    tile = (8, 16, 32)[seed % 3]
    s = scenes.many_draws(ndraws=16, width=333, height=197, tex_size=32, seed=seed, max_tris=1500)
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, tile=tile)
    _compare(s, out, ref)
    st, c = out["stats"], ref["counters"]
    assert (st["pixels_written"], st["depth_passed"], st["tris_out"]) == (
        c["pixels_written"], c["depth_passed"], c["tris_out"])


@pytest.mark.parametrize("seed", range(2, 8))
def test_random_triangles_seeds_bit_exact(ctx, seed):
    """Seeded random clip-space triangles (large, overlapping, both windings) on odd widths."""
    s = scenes.random_triangles(ntris=800, width=257 + seed, height=131, seed=seed)
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, tile=(8, 16, 32)[seed % 3])
    _compare(s, out, ref)
    assert out["stats"]["pixels_written"] == ref["counters"]["pixels_written"]


def _stacked_triangles(ntris, width, height, seed):
    """ntris triangles that each cover the whole target, depths on 8 levels (LESS: the first
    triangle of a level wins, so the result depends on the order inside every bin), random
    colours, alternating winding."""
    rng = np.random.default_rng(seed)
    a = np.array([[-1.0, -1.0], [3.0, -1.0], [-1.0, 3.0]], np.float32)
    b = a[[0, 2, 1]]
    xy = np.where((np.arange(ntris) % 2 == 0)[:, None, None], a[None], b[None]).reshape(-1, 2)
    pos = np.zeros((3 * ntris, 4), np.float32)
    pos[:, :2] = xy
    pos[:, 2] = np.repeat((rng.integers(0, 8, ntris) / 8.0).astype(np.float32), 3)
    pos[:, 3] = 1.0
    col = rng.integers(0, 256, (3 * ntris, 4), dtype=np.uint8)
    pipes = {"main": scenes.Pipeline("vs_mvp_pos4_col", "fs_color")}
    draws = [scenes.Draw("main", 3 * ntris, 0, False, scenes.push_block(scenes.IDENTITY))]
    return scenes.Scene("stacked", width, height, scenes.interleave_pos_col(pos, col), 20, scenes.VATTRS_POS_COL, pipes,
                        draws)


@pytest.mark.parametrize("ntris", [50, 500, 3000, 17000, 70000])
def test_stacked_bins_keep_api_order(ctx, ntris):
    """Every tile's bin holds all ntris triangles (bins of 50 to 70000 pairs, 70000 primitives
    needing three 8-bit digits of the primitive index); the colours equal the oracle's only if
    every bin is walked in API order."""
    s = _stacked_triangles(ntris, 24, 16, seed=ntris)
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, tile=8)
    _compare(s, out, ref)
    assert out["stats"]["pixels_written"] == ref["counters"]["pixels_written"]
    assert out["stats"]["depth_passed"] == ref["counters"]["depth_passed"]


@pytest.mark.parametrize("chunk", [32, 64])
def test_pixel_parallel_coverage_path(monkeypatch, chunk):
    """A queue whose small-triangle threshold is 16 pixels (VISOR_SMALL; the default tests every
    triangle triangle-parallel) sends larger bbox overlaps to the pixel-parallel test with one
    ballot per triangle: random large triangles and the 256-draw mix at reduced size equal the
    oracle through that path too, for chunks of 32 (two lanes per triangle) and 64."""
    monkeypatch.setenv("VISOR_SMALL", "16")
    monkeypatch.setenv("VISOR_CHUNK", str(chunk))
    q = render.Context()  # the thresholds are read when the queue is created
    for s in (scenes.random_triangles(ntris=500, width=200, height=136, seed=21),
              scenes.many_draws(ndraws=24, width=320, height=240, tex_size=64, max_tris=2048)):
        ref = refrast.render(s)
        r = render.SceneRenderer(q, s)
        q.set_mode(batched=True, tile=8)
        r.render()
        _compare(s, {"color": r.read_color(), "depth": r.read_depth()}, ref)
        assert q.stats(0)["chunk"] == chunk


def test_bin_overflow_reruns_the_pass(ctx):
    """2000 large random triangles on 640x480 with 8x8 tiles make about 2.5M (tile, triangle)
    pairs, over twice the first bin capacity of a new queue (max(2^20, 2 x triangles)): the first
    try's raster writes nothing, the pass runs again with larger bins, and the target equals the
    oracle's (so the failed try left it untouched). A second submit runs at the grown capacity."""
    fresh = render.Context()  # a queue whose bin capacity has not grown yet
    s = scenes.random_triangles(ntris=2000, width=640, height=480, seed=9)
    ref = refrast.render(s)
    r = render.SceneRenderer(fresh, s)
    fresh.set_mode(batched=True, tile=8)
    r.render()
    first = {"color": r.read_color(), "depth": r.read_depth(), "stats": fresh.stats(0)}
    assert first["stats"]["bin_pairs"] > 1 << 20
    _compare(s, first, ref)
    assert first["stats"]["pixels_written"] == ref["counters"]["pixels_written"]
    r.render()
    assert np.array_equal(r.read_color(), first["color"])
    assert np.array_equal(r.read_depth().view(np.uint32), first["depth"].view(np.uint32))


def test_micro_triangles_small_bit_exact(ctx):
    """RGBA16F target, micro-triangles, 8 trilinear taps of an array texture, additive draws."""
    s = scenes.micro_triangles(ntris=20000, ndraws=4, width=256, height=256, tex_size=64, layers=4)
    ref = refrast.render(s)
    out = _render_gpu(ctx, s)
    _compare(s, out, ref)


def test_many_draws_full_batched_equals_plain(ctx):
    """At full size: batched and plain per-draw execution give identical bytes."""
    s = scenes.many_draws()
    a = _render_gpu(ctx, s, batched=True)
    b = _render_gpu(ctx, s, batched=False)
    assert np.array_equal(a["color"], b["color"])
    assert np.array_equal(a["depth"].view(np.uint32), b["depth"].view(np.uint32))
    assert a["stats"]["pixels_written"] == b["stats"]["pixels_written"]


def test_graph_replays_match_eager(ctx):
    """A batched pass captured into a HIP graph and replayed, with other (eager, per-draw) work
    and host synchronisation between replays, gives the eager result on every replay."""
    s = scenes.many_draws(ndraws=24, width=320, height=240, tex_size=64, max_tris=2048)
    r = render.SceneRenderer(ctx, s)
    o = render.SceneRenderer(ctx, scenes.random_triangles(300, 96, 96, seed=7))
    ctx.set_mode(batched=True, tile=8, graphs=False)
    r.render()
    ref_c, ref_d = r.read_color().copy(), r.read_depth().copy()
    ctx.set_mode(batched=True, tile=8, graphs=True)
    ctx.icd.vh_reset_stats(ctx.queue)
    for i in range(4):
        r.render()
        assert np.array_equal(r.read_color(), ref_c), i
        assert np.array_equal(r.read_depth().view(np.uint32), ref_d.view(np.uint32)), i
        ctx.set_mode(batched=False)
        o.render()
        ctx.set_mode(batched=True)
    assert ctx.stats(1)["graph_launches"] >= 4


def test_autotuner_shapes_render_identical_bytes(ctx):
    """With the autotuner on, the first repetitions of a pass run under each candidate launch
    shape (tile size, waves per workgroup, LDS chunk, persistent or one workgroup per tile); every
    repetition must give the same bytes, and the tuner settles on one candidate."""
    s = scenes.many_draws(ndraws=24, width=320, height=240, tex_size=64, max_tris=2048)
    ref = refrast.render(s)
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=True, tile=0, graphs=True)
    shapes = []
    for i in range(16):
        r.render()
        _compare(s, {"color": r.read_color(), "depth": r.read_depth()}, ref)
        st = ctx.stats(0)
        shapes.append((st["tile_size"], st["chunk"], st["persistent"]))
    assert len(set(shapes[:14])) == 7
    assert shapes[14] == shapes[15] and shapes[14] in set(shapes[:14])
    ctx.set_mode(batched=True, tile=8)


def test_lit_sphere_full_bit_exact(ctx):
    """Full size: 200k-triangle indexed sphere at 1920x1080, against the oracle."""
    s = scenes.lit_sphere()
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, tile=8)
    _compare(s, out, ref)
    assert out["stats"]["depth_passed"] == ref["counters"]["depth_passed"]


def test_many_draws_full_bit_exact(ctx):
    """The 256-draw workload at full size (1.85M triangles, 1920x1080) against the oracle, with
    the autotuner choosing the launch shape."""
    s = scenes.many_draws()
    ref = refrast.render(s)
    r = render.SceneRenderer(ctx, s)
    ctx.set_mode(batched=True, tile=0, graphs=True)
    r.render()
    _compare(s, {"color": r.read_color(), "depth": r.read_depth()}, ref)
    st = ctx.stats(0)
    assert st["pixels_written"] == ref["counters"]["pixels_written"]
    assert st["depth_passed"] == ref["counters"]["depth_passed"]
    ctx.set_mode(batched=True, tile=8)


def test_micro_triangles_4k_bit_exact(ctx):
    """The micro-triangle workload at its full size: 3840x2160 RGBA16F target, 4M micro-triangles
    in 50 depth layers, 8 trilinear taps of a 2048^2 x 4 array texture, a quarter of the draws
    additive, against the oracle (about a minute of oracle time), counters included."""
    s = scenes.micro_triangles()
    assert s.triangles == 4_000_000
    ref = refrast.render(s)
    out = _render_gpu(ctx, s, tile=8)
    _compare(s, out, ref)
    for k in ("pixels_written", "depth_passed", "tris_in", "tris_out", "blocks", "pixels_tested"):
        assert out["stats"][k] == ref["counters"][k], k
