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
