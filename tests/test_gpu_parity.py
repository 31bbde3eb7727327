"""Parity of the gfx950 nearest-hit path with the CPU oracle, through the C ABI.

The bar is bit-exact: t, u, v, p bit patterns, valid and primitive_id (lowest id at equal t)
equal to the oracle's lowest-id mode; against the reference's own golden records (depth-first
tie order) every field but primitive_id is bit-exact and primitive_id is equal wherever the
reference result has no equal-t tie. u, v and primitive_id are compared only on hits (the
reference leaves them indeterminate on misses).
"""
import os

import numpy as np
import pytest

import oracle
import portablert_amd as prt
from portablert_amd import scenes
from conftest import GOLDEN, load_golden
from helpers import as_full, assert_same

pytestmark = pytest.mark.gpu

FULL = prt.mask_of(prt.ALL_TAGS)
VARIANTS = [(fmt, cull) for fmt in (1, 0) for cull in (1, 0)]  # (node format, culling)
EDGE = sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("edge_"))


def configure(ctx, fmt=1, cull=1):
    ctx.set_option("node_format", fmt)
    ctx.set_option("cull", cull)


def trace(ctx, tris, rays, mask=FULL, variants=VARIANTS, leaf_size=None):
    """Runs every variant; asserts they agree bit for bit; returns the first result."""
    if leaf_size is not None:
        ctx.set_option("leaf_size", leaf_size)
    ctx.set_tris(tris)
    first = None
    for fmt, cull in variants:
        configure(ctx, fmt, cull)
        got = ctx.nearest_hits(rays, mask)
        if first is None:
            first = got
        else:
            assert first.tobytes() == got.tobytes(), f"variant fmt={fmt} cull={cull} differs"
    configure(ctx)
    return first


def test_known_answer(gpu_ctx):
    """examples/validation/main.cpp:66-111 with its own tolerance (1e-4) and exactly."""
    tri = np.array([[-1, -1, 0, 1, -1, 0, 0, 1, 0]], np.float32)
    hit_ray = np.array([[0.1, 0, -1, 0, 0, 1]], np.float32)
    miss_ray = np.array([[-2, 0, -1, 0, 0, 1]], np.float32)
    gpu_ctx.set_tris(tri)
    h = gpu_ctx.nearest_hits(hit_ray)[0]
    assert h["valid"] == 1 and h["primitive_id"] == 0
    for f, want in (("t", 1.0), ("u", 0.3), ("v", 0.5), ("px", 0.1), ("py", 0.0), ("pz", 0.0)):
        assert abs(float(h[f]) - want) <= 1e-4, (f, h[f])
    m = gpu_ctx.nearest_hits(miss_ray)[0]
    assert m["valid"] == 0 and np.isinf(m["t"])
    # filtered record HitReg<valid>
    assert gpu_ctx.nearest_hits(hit_ray, prt.VALID)["valid"][0] == 1
    assert gpu_ctx.nearest_hits(miss_ray, prt.VALID)["valid"][0] == 0
    # exact agreement with the oracle on both
    both = np.concatenate([hit_ray, miss_ray])
    assert_same(gpu_ctx.nearest_hits(both), oracle.trace(tri, both), what="known")


@pytest.mark.parametrize("name", EDGE)
def test_edge_fixtures(gpu_ctx, name):
    g = load_golden("edge_" + name)
    tris, rays, ref = g["tris"], g["rays"], as_full(g["hits"])
    print(f"edge {name}: {len(tris)} tris {len(rays)} rays", flush=True)
    for leaf in (1, 2, 4):
        got = trace(gpu_ctx, tris, rays, leaf_size=leaf)
        want = oracle.trace(tris, rays, oracle.LOWEST_ID)
        assert_same(got, want, what=f"{name} leaf={leaf} vs oracle")
        assert_same(got, ref, ids=False, what=f"{name} vs reference")
    gpu_ctx.set_option("leaf_size", 3)  # back to the default


@pytest.mark.parametrize("name", ["workload0", "workload1_primary", "workload1_bounce", "workload2"])
def test_workload_fixtures(gpu_ctx, name):
    g = load_golden(name)
    w = int(name[8])
    tris = scenes.scene(w)
    assert scenes.sha256(tris) == str(g["scene_sha256"]), "scene generator is not deterministic"
    rays, ref = g["rays"], as_full(g["hits"])
    print(f"{name}: {len(tris)} tris {len(rays)} rays", flush=True)
    got = trace(gpu_ctx, tris, rays)
    want = oracle.trace(tris, rays, oracle.LOWEST_ID)
    assert_same(got, want, what=f"{name} vs oracle lowest-id")
    assert_same(got, ref, ids=False, what=f"{name} vs reference records")
    tie_free = oracle.trace(tris, rays, oracle.DFS)["primitive_id"] == want["primitive_id"]
    valid = ref["t"] < np.inf
    assert (got["primitive_id"][valid & tie_free] == ref["primitive_id"][valid & tie_free]).all()


def test_workload0_full(gpu_ctx):
    """All 1,048,576 primary rays of the 16,384-triangle soup, t + primitive_id."""
    tris, rays = scenes.scene(0), scenes.rays(0)
    mask = prt.T | prt.PRIMITIVE_ID
    got = trace(gpu_ctx, tris, rays, mask)
    want = oracle.trace(tris, rays, oracle.LOWEST_ID)
    assert_same(got, want, fields=("t", "primitive_id"), what="workload0 full")
    print(f"hits {int((got['t'] < np.inf).sum())}", flush=True)


def test_workload2_full(gpu_ctx):
    """All 8,388,608 incoherent rays in the 262,144-triangle room, all four variants (fp32 /
    quantized nodes, culling / exact), equal to the oracle on every ray."""
    tris, rays = scenes.scene(2), scenes.rays(2)
    mask = prt.T | prt.PRIMITIVE_ID
    got = trace(gpu_ctx, tris, rays, mask)
    want = oracle.trace(tris, rays, oracle.LOWEST_ID, min(16, os.cpu_count() or 1))
    assert_same(got, want, fields=("t", "primitive_id"), what="workload2 full")
    neg = float(((got["t"] < 0)).mean())
    print(f"hit rate {float((got['t'] < np.inf).mean()):.4f} negative-t rate {neg:.4f}", flush=True)


@pytest.mark.parametrize("opts", [
    dict(block=64, lds_stack=8, waves_per_cu=8),
    dict(block=256, lds_stack=16, waves_per_cu=32),
    dict(chunk=64, leaf_phase=1, refill=1),
    dict(chunk=4096, leaf_phase=64, refill=64),
    dict(refill=33, leaf_phase=40, lds_stack=8),
])
def test_scheduling_options(gpu_ctx, opts):
    """Launch shape, stack size, ray chunks, triangle-phase and refill thresholds change only
    the schedule: records equal the oracle's for each setting (8-entry stacks spill)."""
    tris = scenes.scene(2)
    rays = scenes.rays(2, first=1000000, count=300001)
    mask = prt.T | prt.PRIMITIVE_ID | prt.UV
    gpu_ctx.set_tris(tris)
    saved = {k: gpu_ctx.get_option(k) for k in opts}
    try:
        for k, v in opts.items():
            gpu_ctx.set_option(k, v)
        got = gpu_ctx.nearest_hits(rays, mask)
    finally:
        for k, v in saved.items():
            gpu_ctx.set_option(k, v)
    want = oracle.trace(tris, rays, oracle.LOWEST_ID, min(16, os.cpu_count() or 1))
    assert_same(got, want, fields=("t", "primitive_id", "u", "v"), what=str(opts))


def test_default_wave_count(gpu_ctx):
    """The default persistent-wave count (option 0) is chosen from the ray count: records equal
    those of an explicit 32 waves per CU and the oracle's at sizes that pick 16, 22 and 32."""
    tris = scenes.scene(2)
    gpu_ctx.set_tris(tris)
    assert gpu_ctx.get_option("waves_per_cu") == 0
    mask = prt.T | prt.PRIMITIVE_ID
    for count in (65536, 2073600, 4194304):
        rays = scenes.rays(2, first=7, count=count)
        print(f"default wave count, {count} rays", flush=True)
        auto = gpu_ctx.nearest_hits(rays, mask)
        try:
            gpu_ctx.set_option("waves_per_cu", 32)
            fixed = gpu_ctx.nearest_hits(rays, mask)
        finally:
            gpu_ctx.set_option("waves_per_cu", 0)
        assert auto.tobytes() == fixed.tobytes()
        idx = np.random.default_rng(count).choice(count, 20000, replace=False)
        want = oracle.trace(tris, rays[idx], oracle.LOWEST_ID)
        assert_same(auto[idx], want, fields=("t", "primitive_id"), what=f"{count} rays")


def test_ray_order_does_not_matter(gpu_ctx):
    """All 8,388,608 room rays in a random order give the same record for every ray (a full-size
    property: which wave takes a ray and what runs beside it change nothing)."""
    import torch
    tris, rays = scenes.scene(2), scenes.rays(2)
    mask = prt.T | prt.PRIMITIVE_ID | prt.UV
    size, _ = prt.layout(mask)
    gpu_ctx.set_tris(tris)
    perm = np.random.default_rng(5).permutation(len(rays))
    d_rays = torch.from_numpy(rays).cuda()
    d_perm = torch.from_numpy(np.ascontiguousarray(rays[perm])).cuda()
    out_a = torch.empty(len(rays) * size, dtype=torch.uint8, device="cuda")
    out_b = torch.empty_like(out_a)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream().cuda_stream
    gpu_ctx.nearest_hits_device(d_rays.data_ptr(), len(rays), mask, out_a.data_ptr(), st)
    gpu_ctx.nearest_hits_device(d_perm.data_ptr(), len(rays), mask, out_b.data_ptr(), st)
    torch.cuda.synchronize()
    a = out_a.cpu().numpy().reshape(len(rays), size)
    b = out_b.cpu().numpy().reshape(len(rays), size)
    assert np.array_equal(a[perm], b)


def test_all_field_masks(gpu_ctx):
    """Every HitReg<Tags...> layout holds the same values as the full record."""
    tris = scenes.scene(0)
    rays = scenes.rays(0, first=400000, count=3000)
    gpu_ctx.set_tris(tris)
    full = gpu_ctx.nearest_hits(rays, FULL)
    for mask in range(32):
        got = gpu_ctx.nearest_hits(rays, mask)
        size, off = prt.layout(mask)
        assert got.dtype.itemsize == size
        raw = got.view(np.uint8).reshape(len(rays), size)
        covered = np.zeros(size, bool)
        for name in got.dtype.names:
            a, b = got[name], full[name]
            if name in ("u", "v", "primitive_id"):
                hit = full["t"] < np.inf
                a, b = a[hit], b[hit]
            assert a.tobytes() == b.tobytes(), (mask, name)
            o = got.dtype.fields[name][1]
            covered[o:o + got.dtype.fields[name][0].itemsize] = True
        assert (raw[:, ~covered] == 0).all(), f"padding bytes of mask {mask} not zero"


def test_device_entry_point_matches_host(gpu_ctx):
    import torch
    tris = scenes.scene(0)
    rays = scenes.rays(0, first=300000, count=70001)
    mask = prt.T | prt.PRIMITIVE_ID | prt.UV
    gpu_ctx.set_tris(tris)
    host = gpu_ctx.nearest_hits(rays, mask)
    size, _ = prt.layout(mask)
    d_rays = torch.from_numpy(rays).cuda()
    d_out = torch.zeros(len(rays) * size, dtype=torch.uint8, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    gpu_ctx.nearest_hits_device(d_rays.data_ptr(), len(rays), mask, d_out.data_ptr(), stream)
    torch.cuda.synchronize()
    dev = d_out.cpu().numpy().view(prt.hit_dtype(mask))
    assert dev.tobytes() == host.tobytes()


def test_calls_on_two_streams(gpu_ctx):
    """Device calls enqueued back to back on two streams with no host synchronisation between
    them: the context orders them (they share its work counter and stack scratch), and a host
    call afterwards waits for both."""
    import torch
    tris = scenes.scene(2)
    gpu_ctx.set_tris(tris)
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    sets = [scenes.rays(2, first=k * 600000, count=600000 - 77 * k) for k in range(3)]
    d_rays = [torch.from_numpy(r).cuda() for r in sets]
    d_out = [torch.zeros(len(r) * size, dtype=torch.uint8, device="cuda") for r in sets]
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]  # default priority
    torch.cuda.synchronize()
    print("two streams: 3 calls", [len(r) for r in sets], flush=True)
    for k in range(3):
        gpu_ctx.nearest_hits_device(d_rays[k].data_ptr(), len(sets[k]), mask, d_out[k].data_ptr(),
                                    streams[k % 2].cuda_stream)
    host = gpu_ctx.nearest_hits(sets[0][:1000], mask)
    torch.cuda.synchronize()
    for k in range(3):
        got = d_out[k].cpu().numpy().view(prt.hit_dtype(mask))
        idx = np.arange(0, len(sets[k]), 97)
        want = oracle.trace(tris, sets[k][idx], oracle.LOWEST_ID)
        assert_same(got[idx], want, what=f"call {k} (stream {k % 2})")
    assert host.tobytes() == d_out[0].cpu().numpy().view(prt.hit_dtype(mask))[:1000].tobytes()


def test_cpp_api_on_device():
    """The header-only C++ layer on the device (tests/cpp/api_test.cpp, built in-tree by
    build()): registration and selection, set_tris, the free and member nearest_hits with several
    tag lists and orders, the validation example's known answer, slice(), refit, the move
    overload of set_tris, a call above one host-pipeline chunk, and the error path."""
    import subprocess
    exe = os.path.join(os.path.dirname(prt.LIB), "api_test")
    assert os.path.exists(exe), "api_test is not built (python portablert_amd/build.py)"
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "backend MI355X on" in r.stdout and "all API checks passed" in r.stdout


def test_host_entry_point_in_chunks(gpu_ctx):
    """Host calls above one chunk (1,048,576 rays) are pipelined through pinned buffers: a ragged
    count with the full record layout equals the device entry point byte for byte."""
    import torch
    tris = scenes.scene(2)
    rays = scenes.rays(2, first=0, count=2_500_001)
    gpu_ctx.set_tris(tris)
    host = gpu_ctx.nearest_hits(rays, FULL)
    size, _ = prt.layout(FULL)
    d_rays = torch.from_numpy(rays).cuda()
    d_out = torch.zeros(len(rays) * size, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    gpu_ctx.nearest_hits_device(d_rays.data_ptr(), len(rays), FULL, d_out.data_ptr(),
                                torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert d_out.cpu().numpy().tobytes() == host.tobytes()
    sample = np.arange(0, len(rays), 97)
    assert_same(host[sample], oracle.trace(tris, rays[sample]), what="chunked host path sample")


def test_sizes_and_empty(gpu_ctx):
    tris = scenes.scene(0)[:1000]
    gpu_ctx.set_tris(tris)
    for n in (1, 2, 63, 64, 65, 1000, 4097):
        rays = scenes.rays(0, first=500000, count=n)
        assert_same(gpu_ctx.nearest_hits(rays), oracle.trace(tris, rays), what=f"n={n}")
    assert len(gpu_ctx.nearest_hits(np.zeros((0, 6), np.float32))) == 0
    # empty scene: every ray misses
    gpu_ctx.set_tris(np.zeros((0, 9), np.float32))
    out = gpu_ctx.nearest_hits(scenes.rays(0, first=500000, count=100))
    assert (out["valid"] == 0).all() and np.isinf(out["t"]).all()
    assert (out["primitive_id"] == 0xFFFFFFFF).all()


def test_error_paths(gpu_ctx):
    fresh = prt.Context(0)
    rays = scenes.rays(0, first=0, count=10)
    with pytest.raises(prt.PrtError) as e:
        fresh.nearest_hits(rays)
    assert e.value.code == prt.ERR_STATE
    bad = scenes.scene(0)[:10].copy()
    bad[3, 4] = np.nan
    with pytest.raises(prt.PrtError) as e:
        fresh.set_tris(bad)
    assert e.value.code == prt.ERR_INPUT
    fresh.set_option("alloc_limit", 1000)
    with pytest.raises(prt.PrtError) as e:
        fresh.set_tris(scenes.scene(0))
    assert e.value.code == prt.ERR_OOM
    fresh.set_option("alloc_limit", 0)
    fresh.set_tris(scenes.scene(0)[:100])
    # layout mismatch through the raw entry point
    import ctypes
    out = np.zeros(10 * 32, np.uint8)
    offs = np.zeros(8, np.int32)
    rc = prt.lib().prt_mi355x_nearest_hits(fresh.h, rays.ctypes.data_as(ctypes.c_void_p), 10, 31,
                                            out.ctypes.data_as(ctypes.c_void_p), 32,
                                            offs.ctypes.data_as(ctypes.c_void_p))
    assert rc == prt.ERR_LAYOUT
    # still usable after errors
    assert_same(fresh.nearest_hits(rays), oracle.trace(scenes.scene(0)[:100], rays), what="after errors")
    fresh.close()


def check_bvh(ctx, tris):
    """Exported BVH: every triangle once with its own vertices; quantized child boxes contain the
    fp32 ones; internal child boxes equal the union of the child's slots; leaf boxes contain
    their triangles."""
    nq = ctx.export("nodes_q")
    nf = ctx.export("nodes_f")
    tr = ctx.export("tris").view(np.uint32).reshape(-1, 12)
    ids = tr[:, 9]
    assert np.array_equal(np.sort(ids), np.arange(len(tris), dtype=np.uint32))
    assert np.array_equal(tr[:, :9].view(np.float32), tris[ids])
    n_nodes = len(nf) // 208
    assert len(nq) // 80 == n_nodes
    f = nf.reshape(n_nodes, 208)
    meta = f[:, 8:16]
    lo = f[:, 16:112].copy().view(np.float32).reshape(n_nodes, 3, 8)
    hi = f[:, 112:208].copy().view(np.float32).reshape(n_nodes, 3, 8)
    q = nq.reshape(n_nodes, 80)
    p = q[:, 0:12].copy().view(np.float32)
    e = q[:, 12:15].view(np.int8).astype(np.int32)
    scale = (2.0 ** e).astype(np.float32)
    qlo = q[:, 32:56].reshape(n_nodes, 3, 8).astype(np.float32)
    qhi = q[:, 56:80].reshape(n_nodes, 3, 8).astype(np.float32)
    used = meta != 0xFF
    # fp32 model of fmaf(q, 2^e, p): the product is exact, so one rounding of the sum
    dlo = (qlo * scale[:, :, None]).astype(np.float64) + p[:, :, None].astype(np.float64)
    dhi = (qhi * scale[:, :, None]).astype(np.float64) + p[:, :, None].astype(np.float64)
    dlo, dhi = dlo.astype(np.float32), dhi.astype(np.float32)
    u3 = np.repeat(used[:, None, :], 3, 1)
    assert (dlo[u3] <= lo[u3]).all() and (dhi[u3] >= hi[u3]).all()
    cbase = f[:, 0:4].copy().view(np.uint32)[:, 0]
    inner = used & ((meta & 0x80) != 0)
    node, slot = np.nonzero(inner)
    child = cbase[node] + (meta[node, slot] & 0x7F)
    assert (child < n_nodes).all() and (child > node).all()
    cu = used[child]
    clo = np.where(cu[:, None, :], lo[child], np.inf).min(axis=2)
    chi = np.where(cu[:, None, :], hi[child], -np.inf).max(axis=2)
    assert np.array_equal(clo, lo[node, :, slot]) and np.array_equal(chi, hi[node, :, slot])
    tbase = f[:, 4:8].copy().view(np.uint32)[:, 0]
    v = tr[:, :9].view(np.float32).reshape(-1, 3, 3)
    for k in range(8):
        m = meta[:, k]
        leaf = (m & 0x80) == 0
        cnt, off = (m & 3).astype(np.int64) + 1, (m >> 2).astype(np.int64)
        nodes = np.flatnonzero(leaf)
        first = tbase[nodes].astype(np.int64) + off[nodes]
        llo = np.full((len(nodes), 3), np.inf, np.float32)
        lhi = np.full((len(nodes), 3), -np.inf, np.float32)
        for j in range(4):  # leaves hold 1..4 triangles
            sel = cnt[nodes] > j
            tv = v[first[sel] + j]
            llo[sel] = np.minimum(llo[sel], tv.min(axis=1))
            lhi[sel] = np.maximum(lhi[sel], tv.max(axis=1))
        # leaf boxes are exactly the union of their triangles' boxes
        assert np.array_equal(llo, lo[nodes, :, k]) and np.array_equal(lhi, hi[nodes, :, k])
    return n_nodes


def moved(tris, seed, scale):
    """Every vertex displaced by up to `scale` (a deformation keeping triangle ids)."""
    rng = np.random.default_rng(seed)
    return (tris + rng.uniform(-scale, scale, tris.shape)).astype(np.float32)


@pytest.mark.parametrize("builder,collapse", [(1, 1), (0, 0)])
def test_refit_matches_rebuild(gpu_ctx, builder, collapse):
    """Refit to moved geometry: the tree keeps its nodes, every box is recomputed, and the
    records equal the oracle's (and a fresh build's) on the new positions, for a small motion,
    a large one, and geometry swapped for another scene of the same size; for the default tree
    (clustering, least-cost collapse) and the batch's (radix tree, greedy collapse)."""
    tris = scenes.scene(0)
    rays = scenes.rays(0, first=200000, count=150000)
    mask = prt.mask_of(prt.ALL_TAGS)
    gpu_ctx.set_option("builder", builder)
    gpu_ctx.set_option("collapse", collapse)
    gpu_ctx.set_tris(tris)
    gpu_ctx.set_option("builder", 1)
    gpu_ctx.set_option("collapse", 1)
    n0 = check_bvh(gpu_ctx, tris)
    fresh = prt.Context(0)
    for step, new in enumerate((moved(tris, 1, 0.002), moved(tris, 2, 0.2),
                                scenes.scene(2)[:len(tris)] * np.float32(0.1))):
        gpu_ctx.refit(new)
        assert check_bvh(gpu_ctx, new) == n0  # same tree
        got = gpu_ctx.nearest_hits(rays, mask)
        assert_same(got, oracle.trace(new, rays), what=f"refit step {step}")
        fresh.set_tris(new)
        assert got.tobytes() == fresh.nearest_hits(rays, mask).tobytes()
        for fmt, cull in VARIANTS[1:]:
            configure(gpu_ctx, fmt, cull)
            assert got.tobytes() == gpu_ctx.nearest_hits(rays, mask).tobytes(), (step, fmt, cull)
        configure(gpu_ctx)
        print(f"refit step {step}: {gpu_ctx.stat('refit_ms'):.3f} ms", flush=True)
    fresh.close()


def test_refit_device_matches_host_refit(gpu_ctx):
    """Refit from device memory (set_tris_device + refit_device) gives the records of the
    host-buffer refit on a second context, and a wrong triangle count is refused."""
    import torch
    tris = scenes.scene(2)
    new = moved(tris, 3, 0.05)
    rays = scenes.rays(2, first=3000000, count=200000)
    mask = prt.mask_of(prt.ALL_TAGS)
    d_tris = torch.from_numpy(tris).cuda()
    d_new = torch.from_numpy(new).cuda()
    torch.cuda.synchronize()
    gpu_ctx.set_tris_device(d_tris.data_ptr(), len(tris))
    gpu_ctx.refit_device(d_new.data_ptr(), len(new))
    got = gpu_ctx.nearest_hits(rays, mask)
    host = prt.Context(0)
    host.set_tris(tris)
    host.refit(new)
    assert got.tobytes() == host.nearest_hits(rays, mask).tobytes()
    host.close()
    idx = np.arange(0, len(rays), 11)
    assert_same(got[idx], oracle.trace(new, rays[idx]), what="refit_device")
    with pytest.raises(prt.PrtError) as e:
        gpu_ctx.refit_device(d_new.data_ptr(), len(new) - 1)
    assert e.value.code == prt.ERR_ARG


def test_refit_errors(gpu_ctx):
    fresh = prt.Context(0)
    tris = scenes.scene(0)[:1000]
    with pytest.raises(prt.PrtError) as e:
        fresh.refit(tris)
    assert e.value.code == prt.ERR_STATE
    fresh.set_tris(tris)
    with pytest.raises(prt.PrtError) as e:
        fresh.refit(tris[:999])
    assert e.value.code == prt.ERR_ARG
    bad = tris.copy()
    bad[10, 2] = np.inf
    with pytest.raises(prt.PrtError) as e:
        fresh.refit(bad)
    assert e.value.code == prt.ERR_INPUT
    rays = scenes.rays(0, first=0, count=10)
    with pytest.raises(prt.PrtError) as e:  # the scene was cleared
        fresh.nearest_hits(rays)
    assert e.value.code == prt.ERR_STATE
    fresh.set_tris(tris)
    fresh.refit(tris)
    assert_same(fresh.nearest_hits(rays), oracle.trace(tris, rays), what="after refit errors")
    fresh.close()


def test_bvh_invariants(gpu_ctx):
    """Exported BVH of the room: every triangle once; quantized child boxes contain the fp32
    ones; internal child boxes are the union of the child's slots; leaf boxes are the union of
    their triangles."""
    tris = scenes.scene(2)
    gpu_ctx.set_tris(tris)
    check_bvh(gpu_ctx, tris)



@pytest.mark.parametrize("builder,collapse", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_builders_and_collapse(gpu_ctx, builder, collapse):
    """Radix-tree and clustering builders, greedy and least-cost collapse: a valid BVH of the room
    and the same records as the oracle and the reference on its golden sample."""
    gpu_ctx.set_option("builder", builder)
    gpu_ctx.set_option("collapse", collapse)
    assert gpu_ctx.get_option("collapse") == collapse
    with pytest.raises(prt.PrtError) as e:
        gpu_ctx.set_option("collapse", 2)
    assert e.value.code == prt.ERR_ARG
    try:
        tris = scenes.scene(2)
        g = load_golden("workload2")
        print(f"builder {builder} collapse {collapse}", flush=True)
        gpu_ctx.set_tris(tris)
        check_bvh(gpu_ctx, tris)
        got = gpu_ctx.nearest_hits(g["rays"])
        assert_same(got, oracle.trace(tris, g["rays"], oracle.LOWEST_ID), what="vs oracle")
        assert_same(got, as_full(g["hits"]), ids=False, what="vs reference")
    finally:
        gpu_ctx.set_option("builder", 1)
        gpu_ctx.set_option("collapse", 1)


# ------------------------------------------------------------- watertight leaf test (optional)

@pytest.mark.parametrize("name", EDGE + ["workload0", "workload1_primary", "workload1_bounce", "workload2"])
def test_watertight_parity(gpu_ctx, name):
    """PRT_OPT_LEAF_TEST = 1 (Woop-Benthin-Wald): records equal the oracle's watertight mode
    (oracle.WATERTIGHT, lowest id at equal t) bit for bit, with quantized and fp32 nodes."""
    g = load_golden(name if name.startswith("workload") else "edge_" + name)
    tris = scenes.scene(int(name[8])) if name.startswith("workload") else g["tris"]
    rays = g["rays"]
    want = oracle.trace(tris, rays, oracle.LOWEST_ID | oracle.WATERTIGHT)
    gpu_ctx.set_tris(tris)
    gpu_ctx.set_option("leaf_test", 1)
    try:
        for fmt in (1, 0):
            gpu_ctx.set_option("node_format", fmt)
            got = gpu_ctx.nearest_hits(rays, FULL)
            assert_same(got, want, what=f"{name} watertight fmt={fmt}")
    finally:
        gpu_ctx.set_option("leaf_test", 0)
        gpu_ctx.set_option("node_format", 1)


def test_watertight_closed_mesh(gpu_ctx):
    """Rays from the centre of the closed icosphere through 16,384 of its vertices (shared by
    five or six triangles): every one hits with the watertight test, equal to the oracle's brute
    force; the default Moller-Trumbore test lets some through (the reference's behaviour)."""
    tris = scenes.scene(1)
    verts = np.unique(tris.reshape(-1, 3), axis=0)[:16384]
    rays = np.concatenate([np.zeros_like(verts), verts], 1).astype(np.float32)
    gpu_ctx.set_tris(tris)
    gpu_ctx.set_option("leaf_test", 1)
    try:
        got = gpu_ctx.nearest_hits(rays, FULL)
    finally:
        gpu_ctx.set_option("leaf_test", 0)
    assert (got["t"] < np.inf).all(), f"{int((got['t'] == np.inf).sum())} rays leaked"
    assert_same(got, oracle.brute(tris, rays, mode=oracle.LOWEST_ID | oracle.WATERTIGHT), what="closed mesh")
    mt = gpu_ctx.nearest_hits(rays, FULL)
    assert_same(mt, oracle.trace(tris, rays, oracle.LOWEST_ID), what="closed mesh, Moller-Trumbore")
    print(f"Moller-Trumbore: {int((mt['t'] == np.inf).sum())} of {len(rays)} vertex rays pass through", flush=True)


def test_watertight_refused_where_unsupported(gpu_ctx):
    """The learned traversal keeps the reference's leaf test: it refuses the watertight option."""
    gpu_ctx.set_tris(scenes.scene(0))
    gpu_ctx.set_option("leaf_test", 1)
    try:
        with pytest.raises(prt.PrtError):
            rays = scenes.rays(0, count=64)
            import torch
            d = torch.from_numpy(rays).cuda()
            out = torch.empty(64 * 16, dtype=torch.uint8, device="cuda")
            prt.learned_setup(gpu_ctx, 1)
            prt.nearest_hits_learned_device(gpu_ctx, d.data_ptr(), 64, prt.T | prt.PRIMITIVE_ID, out.data_ptr())
    finally:
        gpu_ctx.set_option("leaf_test", 0)


def test_bunny_validation_camera(gpu_ctx):
    """The reference's scene-level validation (examples/validation/main.cpp:113-234) on the GPU:
    its 1024 x 1024 camera (directions divided by the sensor position's length, so not unit;
    zero components on the centre row and column) on a closed 5,120-triangle stand-in for the
    bunny mesh (the mesh itself has no licence file). Every ray, every field, all four variants,
    equal to the oracle; set_tris and nearest_hits through the host entry points within the
    example's 1000 ms windows (validation/main.cpp:230-233)."""
    import time
    tris, rays = scenes.scene(5), scenes.rays(5)
    assert (rays[:, 3:] == 0).any(axis=1).sum() >= 2047  # the zero-component row and column
    t0 = time.perf_counter()
    gpu_ctx.set_tris(tris)
    build_ms = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    hits = gpu_ctx.nearest_hits(rays, FULL)
    trace_ms = (time.perf_counter() - t0) * 1e3
    print(f"bunny camera: build {build_ms:.2f} ms, traversal {trace_ms:.2f} ms (host entry points), "
          f"{int((hits['t'] < np.inf).sum())} hits", flush=True)
    assert 0.0 <= build_ms <= 1000.0 and 0.0 <= trace_ms <= 1000.0
    got = trace(gpu_ctx, tris, rays)
    assert got.tobytes() == hits.tobytes()
    assert_same(got, oracle.trace(tris, rays, oracle.LOWEST_ID), what="bunny camera vs oracle")


def test_workload1_full(gpu_ctx):
    """The closed-mesh configuration at its stated size: all 2,073,600 primary rays of the
    1920 x 1080 camera and the 4 cosine-weighted bounce rays of every hit (from the oracle's
    primary hits), every record field, all four variants, equal to the oracle on every ray."""
    tris, prim = scenes.scene(1), scenes.rays(1)
    assert len(prim) == 2073600
    got = trace(gpu_ctx, tris, prim)
    want = oracle.trace(tris, prim, oracle.LOWEST_ID)
    assert_same(got, want, what="workload1 primaries")
    bounce = scenes.bounce_rays(tris, prim, want["t"], want["primitive_id"],
                                np.stack([want["px"], want["py"], want["pz"]], 1))
    print(f"{len(prim)} primaries, {int((want['t'] < np.inf).sum())} hits, {len(bounce)} bounce rays", flush=True)
    assert len(bounce) == 4 * int((want["t"] < np.inf).sum())
    got_b = trace(gpu_ctx, tris, bounce)
    assert_same(got_b, oracle.trace(tris, bounce, oracle.LOWEST_ID), what="workload1 bounce rays")
