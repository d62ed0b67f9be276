"""grt_compact_rays_noalloc (ray-list compaction with its tile sums in a fixed
per-GPU device array) against grt_compact_rays, through the C ABI.

Bar: bit-exact -- the same count, list, occlusion and miss-hit outputs on two
calls in a row, the second with a smaller input (the fixed array keeps the
first call's contents), and the list equal to the active ray indices in slot
order.  Sizes beyond the fixed array are refused before any launch.
"""
import numpy as np
import pytest
import torch

import grt

pytestmark = pytest.mark.gpu


def _rays(n, seed, frac_active):
    rng = np.random.default_rng(seed)
    d = {k: rng.random(n, dtype=np.float32) for k in grt.RAY_FIELDS}
    dead = rng.random(n) > frac_active
    d["tmin"] = np.where(dead, np.float32(1.0), np.float32(0.0)).astype(np.float32)
    d["tmax"] = np.where(dead, np.float32(-1.0), np.float32(np.inf)).astype(np.float32)
    return grt.RayBatch.from_numpy(d), ~dead


def _outputs(n):
    lst = torch.full((n,), -9, dtype=torch.int32, device="cuda:0")
    count = torch.full((1,), -9, dtype=torch.int32, device="cuda:0")
    occ = torch.full((n,), 7, dtype=torch.uint8, device="cuda:0")
    hits = grt.HitBatch.empty(n)
    hits.f["id"].fill_(-5)
    for k in ("t", "u", "v"):
        hits.f[k].fill_(3.5)
    return lst, count, occ, hits


def _slot_order(n, width):
    """Ray index of every slot: ray order, or 8x8 tiles of a row-major image."""
    if width == 0:
        return np.arange(n)
    h = (n + width - 1) // width
    tx, ty = (width + 7) // 8, (h + 7) // 8
    t = np.arange(tx * ty * 64)
    tile, lane = t // 64, t % 64
    x = (tile % tx) * 8 + lane % 8
    y = (tile // tx) * 8 + lane // 8
    i = y * width + x
    return i[(x < width) & (i < n)]


@pytest.mark.parametrize("order", ["large_then_small", "image_then_rays"])
def test_noalloc_compaction_matches_allocating(gpu_ctx, order):
    calls = ([(1920 * 1080, 1920, 0.6, 1), (5000 + 13, 0, 0.3, 2)]
             if order == "large_then_small" else
             [(1000 * 700 + 3, 1000, 0.5, 3), (70 * 45, 70, 0.8, 4)])
    for n, width, frac, seed in calls:
        rb, active = _rays(n, seed, frac)
        a = _outputs(n)
        b = _outputs(n)
        gpu_ctx.compact_rays(rb, width, a[0], a[1], occluded=a[2], miss_hits=a[3])
        gpu_ctx.compact_rays_noalloc(rb, width, b[0], b[1], occluded=b[2], miss_hits=b[3])
        torch.cuda.synchronize()
        ca, cb = int(a[1].item()), int(b[1].item())
        assert ca == cb == int(active.sum()), (n, width, ca, cb)
        la, lb = a[0][:ca].cpu().numpy(), b[0][:cb].cpu().numpy()
        assert np.array_equal(la, lb), f"list differs (n={n}, width={width})"
        slots = _slot_order(n, width)
        assert np.array_equal(lb, slots[active[slots]]), "list is not the active rays in slot order"
        assert np.array_equal(a[2].cpu().numpy(), b[2].cpu().numpy()), "occlusion output differs"
        occ = b[2].cpu().numpy()
        assert np.all(occ[~active] == 0) and np.all(occ[active] == 7)
        ha, hb = a[3].numpy(), b[3].numpy()
        for k in ("id", "t", "u", "v"):
            assert np.array_equal(ha[k].view(np.uint32), hb[k].view(np.uint32)), f"miss hit.{k}"
        assert np.all(hb["id"][~active] == -1) and np.all(hb["id"][active] == -5)


def test_noalloc_compaction_refuses_oversized(gpu_ctx):
    """More slots than the fixed array covers: refused with GRT_CAPACITY
    before anything is launched (the buffers are never touched)."""
    rb, _ = _rays(64, 5, 0.5)
    lst, count, occ, hits = _outputs(64)
    with pytest.raises(grt.GrtError) as e:
        gpu_ctx.compact_rays_noalloc(rb, 0, lst, count, occluded=occ, n=70_000_000)
    assert e.value.code == grt.GRT_CAPACITY
    torch.cuda.synchronize()
    assert int(count.item()) == -9
    assert np.all(occ.cpu().numpy() == 7)
