"""GPU tests of the batched launch (device work queue), hipGraph replay and the
continuous-batching serving loop.  Bar: bit-exact -- an item traced inside a
mixed batch, a replayed graph and a request served in a mix give exactly the
results of the oracle / of the same work alone."""
import numpy as np
import pytest

import grt
import pyoracle as orc
from grt import workloads as W
from test_gpu_parity import _random_rays, assert_same

pytestmark = pytest.mark.gpu


def _scene(ctx, v, n, lam=8.0):
    sc = grt.Scene(ctx)
    sc.upload(v, n)
    sc.build_grid(lam)
    return sc


def _oracle_grid(v, lam=8.0):
    fr = orc.grid_frame(v, lam)
    _, offs, refs = orc.grid_build(v, fr)
    return fr, offs, refs


@pytest.fixture(scope="module")
def three_scenes(gpu_ctx):
    import torch  # noqa: F401
    a = grt.gen_spheres(11, 16, 0.05, 0.15)
    b = grt.gen_spheres(W.SEED_SPHERES, 64, 0.02, 0.08)
    c = grt.gen_stadium(W.SEED_STADIUM, 8, 1024, 8 * 1024 - 2)[:2]
    out = []
    for v, n in (a, b, c):
        out.append((_scene(gpu_ctx, v, n), v, n, _oracle_grid(v)))
    return out


def _primary(width, height, eye=(0.5, 0.5, -1.5), center=(0.5, 0.5, 0.0)):
    cam = grt.camera_lookat(eye, center, (0, 1, 0), 60.0, width, height)
    return orc.primary_rays(grt.camera_dict(cam))


def test_batch_items_match_oracle(gpu_ctx, three_scenes):
    import torch
    specs = [  # (scene, rays, any, image_width)
        (0, _primary(96, 64), False, 96),
        (1, _random_rays(1, 5), False, 0),
        (1, _random_rays(65, 6), True, 0),
        (2, _primary(100, 37, (0.5, 0.6, -1.5), (0.5, 0.3, 0.5)), False, 100),
        (2, _random_rays(5000, 7), True, 0),
        (0, _random_rays(777, 8), False, 0),
        (1, _primary(64, 64), True, 64),
    ]
    want, items, outs = [], [], []
    for si, rays, anyhit, width in specs:
        sc, v, _, (fr, offs, refs) = three_scenes[si]
        rb = grt.RayBatch.from_numpy(rays)
        n = rb.n
        if anyhit:
            occ = torch.full((n,), 7, dtype=torch.uint8, device="cuda:0")
            items.append((sc, rb, None, occ, width))
            outs.append(occ)
            want.append(orc.trace_any(fr, offs, refs, v, rays)["occluded"])
        else:
            hb = grt.HitBatch.empty(n)
            items.append((sc, rb, hb, None, width))
            outs.append(hb)
            want.append(orc.trace_closest(fr, offs, refs, v, rays))
    # every batched kernel (one ray per lane, persistent, several lanes per
    # ray), cooperative pass on and off, two item orders
    for mode, param, coop, rev in (("simple", 16, 16, False), ("simple", 16, 0, True),
                                   ("pooled", 16, 16, True), ("persistent", 16, 16, True),
                                   ("persistent", 1, 1, False), ("split", 2, 16, False),
                                   ("split", 8, 0, True), ("split", 16, 16, False)):
        gpu_ctx.set_trace_mode(mode, param)
        gpu_ctx.set_coop(coop)
        order = list(range(len(items)))
        if rev:
            order.reverse()
        for out in outs:   # stale results must not pass for new ones
            if isinstance(out, grt.HitBatch):
                out.f["id"].fill_(-3)
            else:
                out.fill_(7)
        batch = grt.Batch(gpu_ctx, 16)
        batch.set([items[i] for i in order])
        assert batch.slots() % 64 == 0
        batch.trace()
        torch.cuda.synchronize()
        tag = f"[{mode} param={param} coop={coop}]"
        for k, (spec, out, w) in enumerate(zip(specs, outs, want)):
            if spec[2]:
                assert_same(out.cpu().numpy(), w, f"item {k} occluded {tag}")
            else:
                g = out.numpy()
                for f in ("id", "t", "u", "v"):
                    assert_same(g[f], w[f], f"item {k} hit.{f} {tag}")
        batch.close()
    gpu_ctx.set_trace_mode(False, 16)
    gpu_ctx.set_coop(16)


def test_batch_rejects_bad_items(gpu_ctx, three_scenes):
    batch = grt.Batch(gpu_ctx, 2)
    sc = three_scenes[0][0]
    rb = grt.RayBatch.empty(10)
    with pytest.raises(grt.GrtError):
        batch.set([(sc, rb, grt.HitBatch.empty(10), None, 0)] * 3)   # over capacity
    empty = grt.Scene(gpu_ctx)
    with pytest.raises(grt.GrtError):
        batch.set([(empty, rb, grt.HitBatch.empty(10), None, 0)])   # no grid
    batch.close()


def test_graph_replay_matches_eager(gpu_ctx, three_scenes):
    import torch
    sc, v, nrm, _ = three_scenes[1]
    cam = grt.camera_lookat((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0, 1, 0), 60.0, 160, 120)
    n = cam.width * cam.height
    stream = grt.stream_create()
    torch.cuda.synchronize()
    gpu_ctx.set_stream(stream)
    try:
        rb, hb, sb = grt.RayBatch.empty(n), grt.HitBatch.empty(n), grt.RayBatch.empty(n)
        occ = torch.empty(n, dtype=torch.uint8, device="cuda:0")
        torch.cuda.synchronize()

        def work():
            gpu_ctx.gen_primary(cam, rb)
            sc.trace_closest(rb, hb, image_width=cam.width)
            sc.gen_shadow(rb, hb, W.LIGHT, 1e-3, sb)
            sc.trace_any(sb, occ, image_width=cam.width)

        work()
        gpu_ctx.synchronize()
        eager_h, eager_o = hb.numpy(), occ.cpu().numpy()
        g = grt.Graph(gpu_ctx)
        g.begin()
        work()
        g.end()
        for rep in range(3):
            for t in hb.f.values():
                t.fill_(-3)
            occ.fill_(9)
            torch.cuda.synchronize()
            g.launch()
            gpu_ctx.synchronize()
            got = hb.numpy()
            for f in ("id", "t", "u", "v"):
                assert_same(got[f], eager_h[f], f"replay {rep} hit.{f}")
            assert_same(occ.cpu().numpy(), eager_o, f"replay {rep} occluded")
            sc.trace_closest(rb, grt.HitBatch.empty(n))   # eager work between replays
            gpu_ctx.synchronize()
        g.close()
    finally:
        gpu_ctx.synchronize()
        gpu_ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        grt.stream_destroy(stream)


def test_serve_result_same_alone_and_in_mix(gpu_ctx, three_scenes):
    import torch
    reqs = [(0, 64, 64, 1), (1, 128, 96, 3), (2, 200, 100, 2), (1, 96, 96, 2), (0, 80, 48, 3),
            (2, 64, 64, 1)]

    def make(batched_list, batched, arrivals):
        outs, rq = [], []
        for k, (si, w, h, d) in enumerate(batched_list):
            cam = W.camera(w, h) if si != 2 else grt.camera_lookat(
                (0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, w, h)
            hb = grt.HitBatch.empty(w * h)
            occ = torch.full((w * h,), 5, dtype=torch.uint8, device="cuda:0")
            outs.append((hb, occ, cam))
            rq.append((three_scenes[si][0], cam, d, arrivals[k], (hb, occ)))
        torch.cuda.synchronize()
        res = grt.serve_run(gpu_ctx, rq, batched=batched, memory_budget=256 << 20, max_active=8)
        torch.cuda.synchronize()
        return outs, res

    mix, res_mix = make(reqs, True, [0.0] * len(reqs))
    assert np.all(res_mix[2] >= res_mix[1]) and np.all(res_mix[3] > 0)
    # the persistent, pooled and split batched kernels serve the same mix
    # identically
    for mode, param in (("persistent", 16), ("pooled", 16), ("split", 4)):
        gpu_ctx.set_trace_mode(mode, param)
        try:
            mixp, _ = make(reqs, True, [0.0] * len(reqs))
        finally:
            gpu_ctx.set_trace_mode("simple", 16)
        for k in range(len(reqs)):
            a, b = mix[k][0].numpy(), mixp[k][0].numpy()
            for f in ("id", "t", "u", "v"):
                assert_same(a[f], b[f], f"request {k} primary hit.{f} ({mode} batch)")
            assert_same(mix[k][1].cpu().numpy(), mixp[k][1].cpu().numpy(),
                        f"request {k} shadow ({mode} batch)")
    for k, r in enumerate(reqs):
        alone, res_alone = make([r], False, [0.0])
        assert res_alone[3][0] == res_mix[3][k], "same active-ray count alone and in the mix"
        a, b = mix[k][0].numpy(), alone[0][0].numpy()
        for f in ("id", "t", "u", "v"):
            assert_same(a[f], b[f], f"request {k} primary hit.{f}")
        assert_same(mix[k][1].cpu().numpy(), alone[0][1].cpu().numpy(), f"request {k} shadow")
    # the plain path under the split schedule (several lanes per ray)
    gpu_ctx.set_trace_mode("split", 4)
    try:
        alone, _ = make([reqs[1]], False, [0.0])
    finally:
        gpu_ctx.set_trace_mode("simple", 16)
    for f in ("id", "t", "u", "v"):
        assert_same(mix[1][0].numpy()[f], alone[0][0].numpy()[f], f"request 1 hit.{f} (split)")
    assert_same(mix[1][1].cpu().numpy(), alone[0][1].cpu().numpy(), "request 1 shadow (split)")
    # and request 0 against the oracle
    sc, v, nrm, (fr, offs, refs) = three_scenes[0]
    rays = orc.primary_rays(grt.camera_dict(mix[0][2]))
    o = orc.trace_closest(fr, offs, refs, v, rays)
    g = mix[0][0].numpy()
    for f in ("id", "t", "u", "v"):
        assert_same(g[f], o[f], f"oracle hit.{f}")
    sh = orc.gen_shadow(rays, o, nrm, W.LIGHT, 1e-4)
    assert_same(mix[0][1].cpu().numpy(), orc.trace_any(fr, offs, refs, v, sh)["occluded"],
                "oracle shadow")


def test_serve_budget_admission(gpu_ctx, three_scenes):
    """A budget that fits one 256x256 request at a time still completes every
    request (admission waits for memory), FIFO."""
    reqs = [(three_scenes[0][0], W.camera(256, 256), 2, 0.0) for _ in range(4)]
    arrival, start, finish, rays, wall = grt.serve_run(gpu_ctx, reqs, batched=True,
                                                       memory_budget=12 << 20, max_active=8)
    assert np.all(finish > 0) and np.all(np.diff(start) >= 0)
    # each admission waits for the previous request to leave the budget
    assert np.all(start[1:] >= finish[:-1] - 1e-9)
    with pytest.raises(grt.GrtError):   # a request larger than the whole budget
        grt.serve_run(gpu_ctx, [(three_scenes[0][0], W.camera(1024, 1024), 1, 0.0)],
                      batched=True, memory_budget=1 << 20)


def test_server_concurrent_submit_same_as_alone(gpu_ctx, three_scenes):
    """grt_server: four host threads submit 16 requests of mixed size and depth
    at once to one server (6 request slots, so admission waits for slots) and
    each waits on its own tickets.  Every request's primary hits, shadow flags
    and active-ray count equal the same request served alone on the plain
    path; a polled ticket reports "running" (None) until done; an unknown
    ticket is an error; after close() the context serves normally again."""
    import threading
    import torch
    specs = [(k % 3, 64 + 16 * (k % 5), 48 + 16 * (k % 4), 1 + k % 3) for k in range(16)]

    def cam_of(si, w, h):
        if si != 2:
            return W.camera(w, h)
        return grt.camera_lookat((0.5, 0.6, -1.5), (0.5, 0.3, 0.5), (0, 1, 0), 60.0, w, h)

    bufs = []
    for si, w, h, d in specs:
        bufs.append((grt.HitBatch.empty(w * h),
                     torch.full((w * h,), 5, dtype=torch.uint8, device="cuda:0"),
                     cam_of(si, w, h)))
    torch.cuda.synchronize()
    results, errors = [None] * len(specs), []
    srv = grt.Server(gpu_ctx, memory_budget=64 << 20, max_active=6)
    try:
        def worker(ks):
            try:
                mine = [(k, srv.submit(three_scenes[specs[k][0]][0], bufs[k][2], specs[k][3],
                                       bufs[k][0], bufs[k][1])) for k in ks]
                for k, t in mine:
                    results[k] = srv.wait(t)
            except Exception as e:   # reported below
                errors.append(e)

        threads = [threading.Thread(target=worker, args=(range(i, len(specs), 4),))
                   for i in range(4)]
        for th in threads:
            th.start()
        for th in threads:
            th.join(timeout=60)
        assert not any(th.is_alive() for th in threads), "a submitting thread hung"
        assert not errors, errors
        # polling: one more request, queried until complete
        hb, occ = grt.HitBatch.empty(64 * 64), torch.empty(64 * 64, dtype=torch.uint8,
                                                           device="cuda:0")
        t = srv.submit(three_scenes[0][0], W.camera(64, 64), 2, hb, occ)
        polled = None
        for _ in range(200000):
            polled = srv.query(t)
            if polled is not None:
                break
        assert polled is not None and polled["rays"] >= 64 * 64
        with pytest.raises(grt.GrtError):
            srv.query(t)           # taken once
        with pytest.raises(grt.GrtError):
            srv.query(10 ** 12)    # never issued
    finally:
        srv.close()
    for k, (si, w, h, d) in enumerate(specs):
        r = results[k]
        assert r is not None and r["finish_s"] >= r["start_s"] >= 0.0, (k, r)
        hb, occ = grt.HitBatch.empty(w * h), torch.full((w * h,), 5, dtype=torch.uint8,
                                                        device="cuda:0")
        torch.cuda.synchronize()
        res = grt.serve_run(gpu_ctx, [(three_scenes[si][0], bufs[k][2], d, 0.0, (hb, occ))],
                            batched=False, memory_budget=64 << 20)
        torch.cuda.synchronize()
        assert int(res[3][0]) == r["rays"], (k, int(res[3][0]), r["rays"])
        a, b = bufs[k][0].numpy(), hb.numpy()
        for f in ("id", "t", "u", "v"):
            assert_same(a[f], b[f], f"request {k} primary hit.{f} (server vs alone)")
        assert_same(bufs[k][1].cpu().numpy(), occ.cpu().numpy(), f"request {k} shadow")
