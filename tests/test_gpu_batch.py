"""Batched multi-scene tracing through the C ABI: every scene's records must equal, bit for bit,
the CPU oracle and the single-scene path (set_tris + nearest_hits of that scene alone), whatever
the per-bucket parameters; graph replays must equal the eager run."""
import numpy as np
import pytest

import oracle
import portablert_amd as prt
from portablert_amd import scenes
from helpers import assert_same

pytestmark = pytest.mark.gpu

FULL = prt.mask_of(prt.ALL_TAGS)


def _torch():
    import torch
    assert torch.cuda.is_available()
    return torch


def soup(rng, n, centre=0.0, spread=1.0):
    c = rng.uniform(-spread, spread, (n, 1, 3)) + centre
    return (c + rng.uniform(-0.05, 0.05, (n, 3, 3))).reshape(n, 9).astype(np.float32)


def ray_set(rng, n, centre=0.0):
    o = rng.uniform(-1.2, 1.2, (n, 3)) + centre
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.concatenate([o, d], 1).astype(np.float32)
    if n >= 8:  # axis-parallel, zero and NaN directions, a ray from outside
        r[0, 3:] = (0, 0, 1)
        r[1, 3:] = (1, 0, 0)
        r[2, 3:] = (0, 0, 0)
        r[3, 3:] = (np.nan, 0, 1)
        r[4, :3] = (centre - 5.0, centre, centre)
        r[4, 3:] = (1, 0, 0)
    return r


class Device:
    """Concatenated scenes on the device and a record buffer."""

    def __init__(self, tris_list, rays_list, mask=FULL):
        torch = _torch()
        self.tris_list, self.rays_list, self.mask = tris_list, rays_list, mask
        self.size = prt.layout(mask)[0]
        self.tc = [len(t) for t in tris_list]
        self.rc = [len(r) for r in rays_list]
        tris = np.concatenate([t.reshape(-1, 9) for t in tris_list]).astype(np.float32)
        rays = np.concatenate([r.reshape(-1, 6) for r in rays_list]).astype(np.float32)
        self.d_tris = torch.from_numpy(tris).cuda()
        self.d_rays = torch.from_numpy(rays).cuda()
        self.n_rays = len(rays)
        self.ray_off = np.concatenate([[0], np.cumsum(self.rc)]).astype(np.int64)

    def out(self):
        torch = _torch()
        o = torch.zeros(max(1, self.n_rays) * self.size, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()  # the library works on its own stream
        return o

    def records(self, d_out, s):
        a, b = self.ray_off[s], self.ray_off[s + 1]
        raw = d_out[a * self.size:b * self.size].cpu().numpy()
        return raw.view(prt.hit_dtype(self.mask)).reshape(-1)


def run_eager(ctx, dev, batch):
    d_out = dev.out()
    batch.run(dev.d_tris.data_ptr(), dev.d_rays.data_ptr(), d_out.data_ptr())
    batch.check()
    return d_out


def test_mixed_scenes_vs_oracle(gpu_ctx):
    """Empty scene, one-triangle scene, a scene without rays, scenes of every size bucket, with
    default and changed per-bucket parameters."""
    rng = np.random.default_rng(7)
    sizes = [300, 0, 1, 3000, 5000, 20000, 70000, 40]
    tris_list = [soup(rng, n, centre=3.0 * k) for k, n in enumerate(sizes)]
    tris_list[2] = np.array([[-1, -1, 6.0, 1, -1, 6.0, 0, 1, 6.0]], np.float32)
    nr = [700, 300, 200, 900, 1000, 1500, 2000, 0]
    rays_list = [ray_set(rng, n, centre=3.0 * k) for k, n in enumerate(nr)]
    rays_list[2][:, :3] = rng.uniform(-0.2, 0.2, (nr[2], 3)) + np.array([0, 0, 5.0])
    rays_list[2][:, 3:] = (0, 0, 1)
    dev = Device(tris_list, rays_list)
    print("mixed batch:", sizes, nr, flush=True)
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, FULL)
    want = [oracle.trace(t, r, oracle.LOWEST_ID) for t, r in zip(tris_list, rays_list)]
    for params in (None, [(4, 1, 17), (6, 4, 1), (2, 3, 100), (8, 1, 64)]):
        if params:
            for k, v in enumerate(params):
                batch.set_params(k, v)
        d_out = run_eager(gpu_ctx, dev, batch)
        for s in range(len(sizes)):
            assert_same(dev.records(d_out, s), want[s], what=f"scene {s} ({sizes[s]} tris) params {params}")
    assert (dev.records(d_out, 2)["valid"] == 1).all()
    assert (dev.records(d_out, 1)["valid"] == 0).all()


def test_extreme_scales_in_one_batch(gpu_ctx):
    """Scenes of very different scales built and traced together: one scaled by 1e12, one by
    1e-12, and an ordinary one with a single far triangle at 2^61 (its box is beyond 2^60, so its
    nodes take the exact two-step test instead of the one-fma form)."""
    rng = np.random.default_rng(29)
    base = [soup(rng, n) for n in (500, 2000, 500)]
    rays = [ray_set(rng, n) for n in (600, 600, 600)]
    far = np.float32(2.0 ** 61)
    tris_list, rays_list = [], []
    for t, r, k in zip(base, rays, (1e12, 1.0, 1e-12)):
        rr = r.copy()
        rr[:, :3] *= np.float32(k)
        tt = (t * np.float32(k)).astype(np.float32)
        if k == 1.0:
            tt = np.concatenate([tt, np.array([[far, far, far, 1.5 * far, far, far, far, 1.5 * far, far]],
                                              np.float32)])
        tris_list.append(tt)
        rays_list.append(rr)
    dev = Device(tris_list, rays_list)
    print("extreme batch:", [len(t) for t in tris_list], flush=True)
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, FULL)
    d_out = run_eager(gpu_ctx, dev, batch)
    for s, (t, r) in enumerate(zip(tris_list, rays_list)):
        want = oracle.trace(t, r, oracle.LOWEST_ID)
        got = dev.records(d_out, s)
        assert_same(got, want, what=f"scene {s}")
        assert (got["valid"] == 1).sum() > 10


def test_non_finite_vertex_reported(gpu_ctx):
    rng = np.random.default_rng(3)
    t0, t1 = soup(rng, 50), soup(rng, 60)
    t1[7, 4] = np.inf
    dev = Device([t0, t1], [ray_set(rng, 64), ray_set(rng, 64)])
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, FULL)
    d_out = dev.out()
    batch.run(dev.d_tris.data_ptr(), dev.d_rays.data_ptr(), d_out.data_ptr())
    with pytest.raises(prt.PrtError) as e:
        batch.check()
    assert e.value.code == 7


@pytest.fixture(scope="module")
def batch_workload():
    n = scenes.n_scenes(3)
    tris_list = [scenes.scene(3, s) for s in range(n)]
    rays_list = [scenes.rays(3, s) for s in range(n)]
    return tris_list, rays_list


def test_batch_workload_vs_single_scene_and_oracle(gpu_ctx, batch_workload):
    """64 scenes, 2.5M triangles, 24M rays: every record equals the single-scene path; samples of
    every fourth scene and of the smallest and largest equal the oracle."""
    torch = _torch()
    tris_list, rays_list = batch_workload
    dev = Device(tris_list, rays_list)
    print(f"batch workload: {sum(dev.tc)} tris, {dev.n_rays} rays", flush=True)
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, FULL)
    d_out = run_eager(gpu_ctx, dev, batch)
    single = dev.out()
    tri_off = np.concatenate([[0], np.cumsum(dev.tc)]).astype(np.int64)
    for s in range(len(tris_list)):
        gpu_ctx.set_tris_device(dev.d_tris[tri_off[s]:tri_off[s + 1]].data_ptr(), dev.tc[s])
        a = int(dev.ray_off[s])
        gpu_ctx.nearest_hits_device(dev.d_rays[a:].data_ptr(), dev.rc[s], FULL, single[a * dev.size:].data_ptr())
    gpu_ctx.synchronize()
    torch.cuda.synchronize()
    if not torch.equal(d_out, single):
        rec = d_out.view(-1, dev.size).ne(single.view(-1, dev.size)).any(1)
        first = int(rec.nonzero()[0])
        s = int(np.searchsorted(dev.ray_off, first, side="right") - 1)
        raise AssertionError(f"{int(rec.sum())} records differ from the single-scene path; first ray {first} scene {s}")
    rng = np.random.default_rng(11)
    for s in list(range(0, len(tris_list), 4)) + [int(np.argmin(dev.tc)), int(np.argmax(dev.tc))]:
        idx = np.unique(np.concatenate([np.arange(64), rng.integers(0, dev.rc[s], 192)]))
        got = dev.records(d_out, s)[idx]
        want = oracle.trace(tris_list[s], rays_list[s][idx], oracle.LOWEST_ID)
        assert_same(got, want, what=f"workload scene {s} sample")


def test_graph_replays_equal_eager(gpu_ctx, batch_workload):
    """Capture once, replay several times with eager work and synchronisation in between; every
    replay equals the eager result (the graph holds no memset nodes)."""
    torch = _torch()
    tris_list, rays_list = batch_workload
    sub = list(range(0, 64, 4))  # 16 scenes keep the test short
    dev = Device([tris_list[s] for s in sub], [rays_list[s][:200000] for s in sub],
                 mask=prt.T | prt.PRIMITIVE_ID)
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, dev.mask)
    eager = run_eager(gpu_ctx, dev, batch)
    g_out = dev.out()
    batch.capture(dev.d_tris.data_ptr(), dev.d_rays.data_ptr(), g_out.data_ptr())
    for k in range(3):
        g_out.zero_()
        other = run_eager(gpu_ctx, dev, batch)  # eager work between replays
        assert torch.equal(other, eager)
        torch.cuda.synchronize()
        batch.replay()
        batch.check()
        assert torch.equal(g_out, eager), f"replay {k} differs from the eager run"


def test_graph_replay_follows_moved_geometry(gpu_ctx, batch_workload):
    """The captured graph holds the builds too: triangles moved in place between replays (same
    counts, as an animation would) are rebuilt and traced by the next replay, equal to an eager
    run on the moved geometry and to the oracle."""
    torch = _torch()
    tris_list, rays_list = batch_workload
    sub = list(range(2, 64, 8))
    tl = [tris_list[s] for s in sub]
    rl = [rays_list[s][:50000] for s in sub]
    dev = Device(tl, rl, mask=prt.T | prt.PRIMITIVE_ID)
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, dev.mask)
    run_eager(gpu_ctx, dev, batch)
    g_out = dev.out()
    batch.capture(dev.d_tris.data_ptr(), dev.d_rays.data_ptr(), g_out.data_ptr())
    rng = np.random.default_rng(17)
    for frame in range(2):
        moved = [(t.reshape(-1, 3, 3) + rng.uniform(-0.02, 0.02, (len(t), 1, 3))).reshape(-1, 9).astype(np.float32)
                 for t in tl]
        dev.d_tris.copy_(torch.from_numpy(np.concatenate(moved)).cuda())
        torch.cuda.synchronize()
        print(f"moved frame {frame}", flush=True)
        batch.replay()
        batch.check()
        eager = run_eager(gpu_ctx, dev, batch)
        assert torch.equal(g_out, eager), f"frame {frame}: replay differs from the eager run"
        for s in (0, len(sub) - 1):
            got = dev.records(g_out, s)
            idx = np.arange(0, len(rl[s]), 13)
            want = oracle.trace(moved[s], rl[s][idx], oracle.LOWEST_ID)
            assert_same(got[idx], want, fields=("t", "primitive_id"), what=f"frame {frame} scene {s}")


def test_tuner_keeps_results(gpu_ctx, batch_workload):
    torch = _torch()
    tris_list, rays_list = batch_workload
    sub = list(range(1, 64, 8))
    dev = Device([tris_list[s] for s in sub], [rays_list[s][:100000] for s in sub], mask=prt.T | prt.PRIMITIVE_ID)
    batch = prt.Batch(gpu_ctx, dev.tc, dev.rc, dev.mask)
    eager = run_eager(gpu_ctx, dev, batch)
    d_out = dev.out()
    before, after = batch.tune(dev.d_tris.data_ptr(), dev.d_rays.data_ptr(), d_out.data_ptr(), reps=2)
    print(f"tuner: {before:.3f} ms -> {after:.3f} ms;", [batch.params(k) for k in range(-1, 4)], flush=True)
    assert after <= before * 1.15
    assert torch.equal(d_out, eager)
    again = run_eager(gpu_ctx, dev, batch)
    assert torch.equal(again, eager)
    for k in range(4):  # back to defaults for later tests
        batch.set_params(k, (8, 2, 64))
    batch.set_params(-1, (128, 32, 8))
