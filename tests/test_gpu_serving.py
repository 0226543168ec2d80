"""Serving: continuous batching of many clients' command buffers (vh_server_*)."""
import numpy as np
import pytest

from oracle import refrast
from visor_amd import serving

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def world(ctx):
    return serving.ServingWorld(ctx, nclients=6, seed=11, target_sizes=((96, 64), (128, 128), (200, 120)),
                                texture_sizes=(32, 64), mesh_grids=((6, 10), (10, 20), (20, 40)),
                                tris_range=(200, 6000), draws_range=(1, 6))


def test_served_results_equal_alone_and_queue(ctx, world):
    srv = serving.Server(ctx, slots=2, max_batch=8)
    try:
        alone = []
        for c in world.clients:
            srv.wait(srv.submit(c.cb))
            alone.append(world.read_color(c))
        srv.reset_stats()
        for rep in range(3):
            tickets = [srv.submit(c.cb) for c in world.clients]
            for t in tickets:
                srv.wait(t)
            for c, a in zip(world.clients, alone):
                assert np.array_equal(world.read_color(c), a), (rep, c.index)
        st = srv.stats()
        assert st["completed"] == 3 * len(world.clients)
        assert st["mean_batch"] > 1.0  # requests shared launches
        assert st["p99_ms"] >= st["p50_ms"] > 0.0
    finally:
        srv.close()
    # the ordinary queue path gives the same bytes
    ctx.set_mode(batched=True)
    for c, a in zip(world.clients, alone):
        ctx.icd.check(ctx.submit([c.cb]), "vkQueueSubmit")
        ctx.icd.check(ctx.icd.vkQueueWaitIdle(ctx.queue), "vkQueueWaitIdle")
        assert np.array_equal(world.read_color(c), a), c.index


def test_served_result_matches_oracle(ctx, world):
    srv = serving.Server(ctx, slots=2, max_batch=8)
    try:
        tickets = [srv.submit(c.cb) for c in world.clients]
        for t in tickets:
            srv.wait(t)
        for c in world.clients[:3]:
            ref = refrast.render(world.scene(c))
            got = world.read_color(c)
            diff = np.argwhere((got != ref["color"]).any(-1))
            assert diff.size == 0, (c.index, len(diff))
    finally:
        srv.close()


def test_admission_by_memory_budget(ctx, world):
    """With a budget smaller than any request, requests run one at a time (a request that
    fits nothing is admitted only onto an idle GPU), and every one completes."""
    srv = serving.Server(ctx, hbm_budget_bytes=1, slots=2, max_batch=8)
    try:
        alone = []
        tickets = [srv.submit(c.cb) for c in world.clients]
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        assert st["completed"] == len(world.clients)
        assert st["mean_batch"] == 1.0
        assert st["batches"] == len(world.clients)
    finally:
        srv.close()


@pytest.fixture(scope="module")
def world64(ctx):
    """The 64-client serving workload at its own shape: targets 256^2, 512^2, 1024^2 and
    1920x1080, textures 256^2..2048^2, 1-32 draws of 1k..500k triangles in total per request
    (depth made host-visible so that it can be compared too)."""
    return serving.ServingWorld(ctx, nclients=64, seed=4, host_visible_depth=True)


def test_64_clients_mix_equals_alone_and_oracle(ctx, world64):
    w = world64
    assert {(c.width, c.height) for c in w.clients} == set(serving.TARGET_SIZES)
    tris = [c.triangles for c in w.clients]
    assert 1000 <= min(tris) and max(tris) <= 500_000 and max(tris) > 100_000
    srv = serving.Server(ctx, slots=2, max_batch=64)
    try:
        # the mix: every client's request at once, so admitted requests share launches
        tickets = [srv.submit(c.cb) for c in w.clients]
        for t in tickets:
            srv.wait(t)
        st = srv.stats()
        assert st["completed"] == len(w.clients) and st["mean_batch"] > 1.0
        mix = [(w.read_color(c), w.read_depth(c)) for c in w.clients]
        # alone: one request on the GPU at a time
        for c, (col, dep) in zip(w.clients, mix):
            srv.wait(srv.submit(c.cb))
            assert np.array_equal(w.read_color(c), col), c.index
            assert np.array_equal(w.read_depth(c).view(np.uint32), dep.view(np.uint32)), c.index
    finally:
        srv.close()
    # the oracle, for the client with the most triangles of each target size
    for size in serving.TARGET_SIZES:
        c = max((c for c in w.clients if (c.width, c.height) == size), key=lambda c: c.triangles)
        ref = refrast.render(w.scene(c))
        col, dep = mix[c.index]
        assert np.array_equal(col, ref["color"]), (size, c.index)
        assert np.array_equal(dep.view(np.uint32), ref["depth"].view(np.uint32)), (size, c.index)
