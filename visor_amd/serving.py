"""Serving workload: many clients submitting command buffers to one GPU through vh_server_*.

Resources every client shares stay resident in HBM across requests: one device-local pool of
UV-sphere meshes of several sizes (one vertex and one index buffer), four mipmapped noise
textures of 256^2 .. 2048^2 sampled bilinear with clamp, and one lit pipeline (MVP + normal
matrix, Lambert x texture, LESS, back-face cull). Each client owns its targets (colour in
host-visible memory, so a completed request has its pixels on the host; depth device-local)
and one recorded command buffer: a render pass that clears both and issues 1..32 indexed draws
whose triangle total is log-uniform in [1k, 500k], each with its own MVP and texture.
"""

import ctypes as C
import math
import random
import threading
import time

import numpy as np

from . import scenes, vk
from .render import SceneRenderer
from .vk import Handle

TARGET_SIZES = ((256, 256), (512, 512), (1024, 1024), (1920, 1080))
TEXTURE_SIZES = (256, 512, 1024, 2048)
MESH_GRIDS = ((10, 25), (20, 50), (40, 100), (80, 200), (160, 400), (250, 1000))  # 2*s*t triangles


class Client:
    def __init__(self, index, width, height, draws):
        self.index, self.width, self.height, self.draws = index, width, height, draws
        self.cb = None

    @property
    def triangles(self):
        return sum(d.count // 3 for d in self.draws)


class ServingWorld:
    def __init__(self, ctx, nclients=64, seed=4, target_sizes=TARGET_SIZES, texture_sizes=TEXTURE_SIZES,
                 mesh_grids=MESH_GRIDS, tris_range=(1000, 500000), draws_range=(1, 32), host_visible_depth=False):
        self.ctx = ctx
        self.host_visible_depth = host_visible_depth
        rng = random.Random(seed)
        nrng = scenes.Rng(seed)
        # mesh pool: one vertex and one index buffer, each mesh a contiguous index range
        pos_l, nrm_l, uv_l, tri_l, self.meshes = [], [], [], [], []
        vbase = ibase = 0
        for (st, sl) in mesh_grids:
            pos, nrm, uv, tri = scenes.uv_sphere(st, sl, 1.0, nrng, jitter=0.03)
            pos_l.append(pos), nrm_l.append(nrm), uv_l.append(uv)
            tri_l.append(tri + vbase)
            self.meshes.append((ibase, tri.size))  # (firstIndex, indexCount)
            vbase += pos.shape[0]
            ibase += tri.size
        pos, nrm, uv = np.concatenate(pos_l), np.concatenate(nrm_l), np.concatenate(uv_l)
        tri = np.concatenate(tri_l)
        # one winding for every draw: all MVPs below are rotations, positive scales, translations
        # and the same projection, so the outward faces keep one orientation on screen
        proj = scenes.perspective(math.radians(60), 1.0, 0.5, 20.0)
        tri = scenes._orient_front_ccw(tri, pos, proj @ scenes.model_matrix(np.eye(3), 1.0, (0, 0, -4)), 256, 256)
        textures = []
        for sz in texture_sizes:
            lv = scenes.full_mips(sz)
            textures.append(scenes.Texture(sz, sz, lv, 1, scenes.noise_texture(nrng, sz, lv), address=vk.ADDR_CLAMP,
                                           mipmap_mode=vk.MIPMAP_NEAREST, max_lod=float(lv - 1)))
        # clients: target size, draws and their MVPs
        self.clients = []
        lo, hi = math.log(tris_range[0]), math.log(tris_range[1])
        for c in range(nclients):
            w, h = target_sizes[rng.randrange(len(target_sizes))]
            nd = rng.randint(draws_range[0], draws_range[1])
            total = math.exp(rng.uniform(lo, hi))
            # each draw takes `per` triangles (a leading index range) of the smallest mesh that has
            # that many, so the request's total stays within tris_range
            per = min(max(round(total / nd), -(-tris_range[0] // nd)), tris_range[1] // nd)
            first, mesh_count = min((m for m in self.meshes if m[1] >= 3 * per), key=lambda m: m[1],
                                    default=self.meshes[-1])
            count = min(3 * per, mesh_count)
            proj = scenes.perspective(math.radians(60), w / h, 0.5, 20.0)
            draws = []
            for _ in range(nd):
                rot = scenes.rotation(rng.uniform(0, 6.283), rng.uniform(0, 6.283), rng.uniform(0, 6.283))
                z = rng.uniform(-7.0, -3.0)
                model = scenes.model_matrix(rot, rng.uniform(0.4, 1.0),
                                            (rng.uniform(-0.45, 0.45) * -z * w / h, rng.uniform(-0.45, 0.45) * -z, z))
                nmat = np.linalg.inv(model[:3, :3]).T
                draws.append(scenes.Draw("lit", count, first, True, scenes.push_block(proj @ model, nmat),
                                         texture=rng.randrange(len(textures))))
            self.clients.append(Client(c, w, h, draws))
        self.scene_data = dict(vertex_data=scenes.interleave_pnu(pos, nrm, uv), index_data=tri.astype(np.uint32).tobytes(),
                               textures=textures)
        self._build()

    def scene(self, client):
        """The client's request as a scenes.Scene (for the oracle and the ordinary queue path)."""
        d = self.scene_data
        return scenes.Scene("serve%d" % client.index, client.width, client.height, d["vertex_data"], 32,
                            scenes.VATTRS_PNU, {"lit": scenes.Pipeline("vs_mvp_normal_uv", "fs_lambert_tex",
                                                                       cull=vk.CULL_BACK)},
                            client.draws, index_data=d["index_data"], textures=d["textures"])

    def _build(self):
        """Vulkan objects: the shared ones once (through a SceneRenderer of client 0, whose
        buffers, textures, layout, render pass and pipeline every client reuses), then each
        client's targets, framebuffer and command buffer."""
        ctx, i, dev = self.ctx, self.ctx.icd, self.ctx.device
        c0 = self.clients[0]
        self.base = SceneRenderer(ctx, self.scene(c0), host_visible_targets=True)
        b = self.base
        for c in self.clients:
            c.color, c.cmem, c.cview, _ = ctx.image(c.width, c.height, vk.FMT["B8G8R8A8_UNORM"], host_visible=True)
            c.depth, c.dmem, c.dview, _ = ctx.image(c.width, c.height, vk.FMT["D32_SFLOAT"],
                                                    host_visible=self.host_visible_depth)
            views = (Handle * 2)(c.cview, c.dview)
            fci = vk.FramebufferCreateInfo(renderPass=b.rp, attachmentCount=2, pAttachments=views, width=c.width,
                                           height=c.height, layers=1)
            c.fb = Handle()
            i.check(i.vkCreateFramebuffer(dev, C.byref(fci), None, C.byref(c.fb)), "vkCreateFramebuffer")
            c.cb = self._record(c)

    def _record(self, c):
        ctx, i, b = self.ctx, self.ctx.icd, self.base
        cb = ctx.begin()
        clears = (vk.ClearValue * 2)()
        clears[1].depthStencil[0] = 1.0
        rbi = vk.RenderPassBeginInfo(renderPass=b.rp, framebuffer=c.fb,
                                     renderArea=vk.Rect2D(extent=vk.Extent2D(width=c.width, height=c.height)),
                                     clearValueCount=2, pClearValues=clears)
        i.vkCmdBeginRenderPass(cb, C.byref(rbi), 0)
        i.vkCmdBindVertexBuffers(cb, 0, 1, (Handle * 1)(b.vb), (vk.u64 * 1)(0))
        i.vkCmdBindIndexBuffer(cb, b.ib, 0, vk.INDEX_U32)
        i.vkCmdBindPipeline(cb, 0, b.pipelines["lit"])
        for d in c.draws:
            i.vkCmdBindDescriptorSets(cb, 0, b.layout, 0, 1, (Handle * 1)(b.sets[d.texture]), 0, None)
            pc = C.create_string_buffer(d.push, 128)
            i.vkCmdPushConstants(cb, b.layout, vk.STAGE_VERTEX | vk.STAGE_FRAGMENT, 0, 128, pc)
            i.vkCmdDrawIndexed(cb, d.count, 1, d.first, 0, 0)
        i.vkCmdEndRenderPass(cb)
        i.check(i.vkEndCommandBuffer(cb), "vkEndCommandBuffer")
        return cb

    def read_color(self, c):
        n = c.width * c.height * 4
        p = self.ctx.map(c.cmem, n)
        return np.frombuffer((C.c_uint8 * n).from_address(p), dtype=np.uint8).copy().reshape(c.height, c.width, 4)


    def read_depth(self, c):
        """The client's depth target (a world built with host_visible_depth=True)."""
        n = c.width * c.height * 4
        p = self.ctx.map(c.dmem, n)
        return np.frombuffer((C.c_uint8 * n).from_address(p), dtype=np.float32).copy().reshape(c.height, c.width)


class Server:
    """vh_server_* wrapper."""

    def __init__(self, ctx, hbm_budget_bytes=16 << 30, slots=2, max_batch=64):
        self.icd = ctx.icd
        cfg = vk.VhServerConfig(hbm_budget_bytes=hbm_budget_bytes, slots=slots, max_batch=max_batch)
        self.h = vk.vp()
        self.icd.check(self.icd.vh_server_create(ctx.device, C.byref(cfg), C.byref(self.h)), "vh_server_create")

    def submit(self, cb):
        t = vk.u64()
        self.icd.check(self.icd.vh_server_submit(self.h, cb, C.byref(t)), "vh_server_submit")
        return t.value

    def wait(self, ticket, timeout_s=None):
        lat = C.c_double()
        ns = (1 << 64) - 1 if timeout_s is None else int(timeout_s * 1e9)
        self.icd.check(self.icd.vh_server_wait(self.h, ticket, ns, C.byref(lat)), "vh_server_wait")
        return lat.value

    def stats(self):
        s = vk.VhServerStats()
        self.icd.vh_server_stats(self.h, C.byref(s))
        return {k: getattr(s, k) for k, _ in vk.VhServerStats._fields_}

    def reset_stats(self):
        self.icd.vh_server_reset_stats(self.h)

    def close(self):
        if self.h:
            self.icd.vh_server_destroy(self.h)
            self.h = None


def run_clients(world, server, rate_rps, seconds, seed=0):
    """Open-loop Poisson arrivals at an aggregate rate over all clients, each client with at most
    one request outstanding (a client whose previous frame is not done waits for it, then
    submits at once). Returns the server statistics of the run and its wall time."""
    n = len(world.clients)
    per_client = rate_rps / n
    t_end = time.perf_counter() + seconds
    server.reset_stats()
    errors = []

    def loop(c):
        r = random.Random(seed * 1000003 + c.index)
        nxt = time.perf_counter() + r.expovariate(per_client)
        pending = None
        try:
            # arrivals that fell behind (the client was waiting for its previous frame) are
            # submitted at once; no submission starts after the end of the window
            while nxt < t_end and time.perf_counter() < t_end:
                delay = nxt - time.perf_counter()
                if delay > 0:
                    time.sleep(delay)
                if pending is not None:
                    server.wait(pending)
                pending = server.submit(c.cb)
                nxt += r.expovariate(per_client)
            if pending is not None:
                server.wait(pending)
        except Exception as e:  # reported by the caller
            errors.append(e)

    t0 = time.perf_counter()
    threads = [threading.Thread(target=loop, args=(c,), daemon=True) for c in world.clients]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    return server.stats(), wall


def run_sequential(world, ctx, seconds):
    """The same clients' requests through the ordinary queue, one at a time: a synchronous
    vkQueueSubmit per request (the reference's blocking submit, cmd_exec.cpp:187-201), round-robin
    over the clients for `seconds`. Returns completed requests, covered fragments, wall time and
    the per-request latencies (submit -> the target's bytes in host memory) in ms."""
    lat, frags, done = [], 0, 0
    t0 = time.perf_counter()
    t_end = t0 + seconds
    while time.perf_counter() < t_end:
        for c in world.clients:
            t = time.perf_counter()
            ctx.icd.check(ctx.submit([c.cb]), "vkQueueSubmit")
            ctx.icd.check(ctx.icd.vkQueueWaitIdle(ctx.queue), "vkQueueWaitIdle")
            lat.append((time.perf_counter() - t) * 1e3)
            frags += ctx.stats(0)["pixels_written"]
            done += 1
            if time.perf_counter() >= t_end:
                break
    return done, frags, time.perf_counter() - t0, lat
