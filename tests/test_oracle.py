"""Pins the CPU oracle (oracle/oracle.c) before it is trusted as the checker.

Each oracle function is checked against something that does not restate it:
  * orc_build + orc_trace (DFS mode): the reference's own CPU backend -- its golden records
    committed under tests/golden (bit-exact, including the depth-first tie order), and, when
    the reference sources are present, the compiled reference on fresh random scenes.
  * orc_mt: the closed-form known answer of examples/validation/main.cpp:86-99, the textbook
    method in double precision (orc_fp64, itself pinned by exact rational cases), and the
    rejection rules (det == 0, inclusive u/v bounds, negative t accepted).
  * orc_slab: the textbook slab method on generic boxes and hand-derived NaN-order outcomes of
    std::min/std::max (bvh.hpp:210-217).
  * orc_trace LOWEST_ID and orc_brute: equal to each other on generic inputs (the result does
    not depend on the BVH), and equal to DFS except at equal-t ties.
  * a planted mistake in the oracle's tie rule is caught by the golden tie fixture.
"""
import os

import numpy as np
import pytest

import oracle
from portablert_amd import scenes
from conftest import GOLDEN, load_golden
from helpers import as_full, assert_same, diff_records

EDGE = sorted(f[5:-4] for f in os.listdir(GOLDEN) if f.startswith("edge_"))


@pytest.mark.parametrize("name", EDGE)
def test_dfs_equals_reference_golden_edge(name):
    g = load_golden("edge_" + name)
    got = oracle.trace(g["tris"], g["rays"], oracle.DFS)
    assert_same(got, as_full(g["hits"]), what=name)


@pytest.mark.parametrize("name", ["workload0", "workload1_primary", "workload1_bounce", "workload2"])
def test_dfs_equals_reference_golden_workloads(name):
    g = load_golden(name)
    tris = scenes.scene(int(name[8]))
    assert scenes.sha256(tris) == str(g["scene_sha256"])
    got = oracle.trace(tris, g["rays"], oracle.DFS)
    assert_same(got, as_full(g["hits"]), what=name)


def test_golden_rays_regenerate():
    """The committed ray samples are what the generators produce (same on every machine)."""
    for name, w in (("workload0", 0), ("workload2", 2), ("workload1_primary", 1)):
        g = load_golden(name)
        idx = g["index"][:300]
        assert np.array_equal(scenes.ray_sample(w, idx), g["rays"][:300]), name


@pytest.mark.skipif(not oracle.have_ref(), reason="reference sources not present")
@pytest.mark.parametrize("seed", range(6))
def test_dfs_equals_compiled_reference_random(seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 3000))
    c = rng.uniform(-1, 1, (n, 1, 3))
    tris = (c + rng.uniform(-0.2, 0.2, (n, 3, 3))).reshape(n, 9).astype(np.float32)
    if seed % 2:  # snap to a coarse grid: exact ties, shared edges, axis-parallel triangles
        tris = np.round(tris * 8) / 8
    o = rng.uniform(-1.3, 1.3, (3000, 3))
    d = rng.normal(size=(3000, 3))
    if seed % 3 == 0:
        d[::7, 0] = 0.0  # zero direction components
        d[::11, 1] = -0.0
    rays = np.concatenate([o, d], 1).astype(np.float32)
    assert_same(oracle.trace(tris, rays, oracle.DFS), oracle.ref_trace(tris, rays), what=f"seed {seed}")


def test_known_answer_closed_form():
    tri = np.array([-1, -1, 0, 1, -1, 0, 0, 1, 0], np.float32)
    hit, t, u, v = oracle.mt(tri, [0.1, 0, -1, 0, 0, 1])
    assert hit and (t, u, v) == (np.float32(1.0), np.float32(0.3), np.float32(0.5))
    h = oracle.trace(tri[None], np.array([[0.1, 0, -1, 0, 0, 1]], np.float32))[0]
    assert (h["px"], h["py"], h["pz"], h["valid"], h["primitive_id"]) == (np.float32(0.1), 0, 0, 1, 0)
    assert not oracle.mt(tri, [-2, 0, -1, 0, 0, 1])[0]


def test_mt_rules():
    tri = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0], np.float32)
    # ray in the triangle's plane: det == 0 exactly, rejected
    assert not oracle.mt(tri, [0.2, 0.2, 0, 1, 0, 0])[0]
    # vertices and edges are inside (inclusive bounds)
    for x, y in ((0, 0), (1, 0), (0, 1), (0.5, 0.5), (0.5, 0)):
        assert oracle.mt(tri, [x, y, -1, 0, 0, 1])[0], (x, y)
    # behind the origin: accepted with negative t (no range check, core.hpp:63)
    hit, t, _, _ = oracle.mt(tri, [0.2, 0.2, 1, 0, 0, 1])
    assert hit and t == -1.0


def test_fp64_exact_rationals():
    tri = np.array([[0, 0, 0, 2, 0, 0, 0, 4, 0]], np.float32)
    rays = np.array([[0.5, 1.0, -3, 0, 0, 1]], np.float32)
    t, u, v = oracle.fp64(tri, rays, [0])
    assert (t[0], u[0], v[0]) == (3.0, 0.25, 0.25)


def test_mt_close_to_fp64():
    rng = np.random.default_rng(7)
    n = 20000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    rays = np.concatenate([rng.uniform(-2, 2, (n, 3)), rng.normal(size=(n, 3))], 1).astype(np.float32)
    t64, u64, v64 = oracle.fp64(tris, rays, np.arange(n))
    hits = 0
    for i in range(n):
        hit, t, u, v = oracle.mt(tris[i], rays[i])
        in64 = u64[i] >= 0 and v64[i] >= 0 and u64[i] + v64[i] <= 1
        if hit:
            hits += 1
            scale = max(1.0, abs(t64[i]))
            assert abs(t - t64[i]) <= 1e-3 * scale, (i, t, t64[i])
        elif abs(u64[i]) > 1e-4 and abs(v64[i]) > 1e-4 and abs(u64[i] + v64[i] - 1) > 1e-4:
            assert not in64, i  # a clear interior hit in fp64 is a hit in fp32
    assert hits > 300


def test_slab_textbook_and_nan_order():
    lo, hi = [0, 0, 0], [1, 1, 1]
    assert oracle.slab([0.5, 0.5, -1, 0.1, 0.1, 1], lo, hi)
    assert not oracle.slab([0.5, 0.5, 2, 0, 0.1, 1], lo, hi)         # box behind: tmax < 0
    assert oracle.slab([0.5, 0.5, 0.5, 0.3, -0.2, 0.4], lo, hi)      # origin inside
    assert not oracle.slab([2, 2, -1, 0.1, 0.1, 1], lo, hi)          # passes beside
    # zero x direction with origin on the box's min-x face: t1 = 0 * inf = NaN poisons both
    # chains (first argument of std::max / std::min), so the test passes even though y misses
    assert oracle.slab([0, 5, 0.5, 0, 0.1, 1], lo, hi)
    # on the max-x face instead: t2 is NaN, dropped, x interval [-inf, -inf] -> rejected
    assert not oracle.slab([1, 0.5, -1, 0, 0.1, 1], lo, hi)
    # zero y direction on the min-y face: the y pair is dropped (second argument), x/z decide
    assert oracle.slab([0.5, 0, -1, 0.1, 0, 1], lo, hi)
    assert not oracle.slab([5, 0, -1, 0.1, 0, 1], lo, hi)


def test_lowest_id_equals_brute_force():
    """F7: on generic inputs the result does not depend on the BVH."""
    rng = np.random.default_rng(3)
    for n in (1, 2, 5, 64, 700):
        c = rng.uniform(-1, 1, (n, 1, 3))
        tris = (c + rng.uniform(-0.3, 0.3, (n, 3, 3))).reshape(n, 9).astype(np.float32)
        rays = np.concatenate([rng.uniform(-1.5, 1.5, (2000, 3)), rng.normal(size=(2000, 3))], 1).astype(np.float32)
        assert_same(oracle.trace(tris, rays, oracle.LOWEST_ID), oracle.brute(tris, rays), what=f"n={n}")


def test_dfs_differs_from_lowest_id_only_on_ties():
    g = load_golden("edge_tie_swapped")
    dfs = oracle.trace(g["tris"], g["rays"], oracle.DFS)
    low = oracle.trace(g["tris"], g["rays"], oracle.LOWEST_ID)
    assert (dfs["t"] == low["t"]).all()
    assert (low["primitive_id"] == 0).all() and (dfs["primitive_id"] == 1).all()


def test_planted_tie_mistake_is_caught():
    """If the oracle's tie rule were inverted (highest id), the golden tie fixture catches it."""
    g = load_golden("edge_tie")
    low = oracle.trace(g["tris"], g["rays"], oracle.LOWEST_ID)
    planted = low.copy()
    planted["primitive_id"] = 1 - planted["primitive_id"]  # what a highest-id rule returns here
    assert diff_records(planted, oracle.brute(g["tris"], g["rays"]))
    assert not diff_records(low, oracle.brute(g["tris"], g["rays"]))


def test_negative_t_rate_is_reported_by_reference():
    """F1 is the reference's behaviour: the golden room sample contains negative-t hits."""
    g = load_golden("workload2")
    ref = as_full(g["hits"])
    assert ((ref["t"] < 0) & (ref["valid"] == 1)).sum() > 0


BUNNY_OBJ = os.environ.get("PRT_BUNNY_OBJ", "/root/reference/examples/common/bunny.obj")


@pytest.mark.skipif(not os.path.exists(BUNNY_OBJ), reason="the bunny mesh ships only with the reference sources")
def test_bunny_validation_scene():
    """The reference's scene-level validation (examples/validation/main.cpp:113-234): the bunny
    under the 1024 x 1024 camera. The oracle's depth-first mode reproduces the hit mask the
    compiled reference backend produced (its golden bunny.png is not in the reference) and the
    sampled records bit for bit; the lowest-id mode and brute force give the same mask."""
    g = load_golden("bunny")
    tris = scenes.load_obj(BUNNY_OBJ)
    assert len(tris) == int(g["n_tris"]) and scenes.sha256(tris) == str(g["mesh_sha256"])
    rays = scenes.rays(5)
    got = oracle.trace(tris, rays, oracle.DFS, min(8, os.cpu_count() or 1))
    assert np.array_equal(np.packbits(got["valid"] != 0), g["mask"])
    idx = g["index"].astype(np.int64)
    assert_same(got[idx], as_full(g["hits"]), what="bunny sample")
    low = oracle.trace(tris, rays, oracle.LOWEST_ID, min(8, os.cpu_count() or 1))
    assert np.array_equal(low["valid"], got["valid"]) and np.array_equal(low["t"].view(np.uint32), got["t"].view(np.uint32))
    sub = idx[::8]
    assert_same(oracle.brute(tris, rays[sub]), low[sub], what="bunny brute force")


def _entry_f32(ray, tri):
    """Entry distance of the reference slab test (bvh.hpp:195-222) on a triangle's own box."""
    lo = np.minimum(np.minimum(tri[0:3], tri[3:6]), tri[6:9])
    hi = np.maximum(np.maximum(tri[0:3], tri[3:6]), tri[6:9])
    with np.errstate(all="ignore"):
        inv = np.float32(1) / ray[3:]
        t1, t2 = (lo - ray[:3]) * inv, (hi - ray[:3]) * inv
    return float(np.max(np.minimum(t1, t2)))


@pytest.mark.parametrize("name", ["sliver_gap", "sliver_cloud", "sliver_tilted"])
def test_sliver_fixtures_are_adversarial(name):
    """The distance-culling fixtures hold what they claim, on the oracle alone: for every ray the
    nearest candidate is a triangle whose Moller-Trumbore t lies more than 2^-9 (relative) below
    the entry of its own box, and another triangle (the competitor) is a candidate between that
    t and that box entry, so a traversal that culls by box distance after finding the competitor
    would return the wrong triangle. The winner agrees with brute force."""
    g = load_golden("edge_" + name)
    tris, rays = g["tris"], g["rays"]
    want = oracle.trace(tris, rays, oracle.LOWEST_ID)
    assert_same(want, oracle.brute(tris, rays), what=name)
    for i, ray in enumerate(rays):
        win = int(want["primitive_id"][i])
        t_win = float(want["t"][i])
        entry = _entry_f32(ray, tris[win])
        assert (entry - t_win) / entry > 2.0 ** -9, (i, t_win, entry)
        between = []
        for j, tri in enumerate(tris):
            hit, t, _, _ = oracle.mt(tri, ray)
            if j != win and hit and t_win < t < entry:
                between.append(j)
        assert between, f"ray {i}: no competitor between {t_win} and {entry}"
