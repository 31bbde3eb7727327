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


# ---------------------------------------------------------------- watertight leaf test (orc_wt)

def test_wt_known_answer():
    """The validation example's known answer (examples/validation/main.cpp:86-99: t = 1,
    u = 0.3, v = 0.5) and its miss ray, with the watertight test; u, v are the weights of
    vertices 1 and 2 as in Moller-Trumbore."""
    tri = [-1, -1, 0, 1, -1, 0, 0, 1, 0]
    hit, t, u, v = oracle.wt(tri, [0.1, 0, -1, 0, 0, 1])
    assert hit and t == 1.0 and abs(u - 0.3) < 1e-6 and abs(v - 0.5) < 1e-6
    assert not oracle.wt(tri, [-2, 0, -1, 0, 0, 1])[0]
    # both windings, t without a range check (behind the origin), like the reference
    hit, t, _, _ = oracle.wt(tri[3:6] + tri[0:3] + tri[6:9], [0.1, 0, 1, 0, 0, 1])
    assert hit and t == -1.0
    # every axis as the dominant one, both signs: rays aimed at an interior point give the FP64
    # re-intersection's t, u, v (the weights of vertices 1 and 2)
    rng = np.random.default_rng(3)
    for _ in range(200):
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        w = rng.dirichlet([1, 1, 1])
        p = w[0] * tri[0:3] + w[1] * tri[3:6] + w[2] * tri[6:9]
        d = rng.normal(size=3)
        ray = np.concatenate([p - 3.0 * d, d]).astype(np.float32)
        t64, u64, v64 = (x[0] for x in oracle.fp64(tri[None], ray[None], np.zeros(1, np.uint32)))
        e1, e2 = tri[3:6].astype(np.float64) - tri[0:3], tri[6:9].astype(np.float64) - tri[0:3]
        nrm = np.cross(e1, e2)
        cos = abs(nrm @ d) / (np.linalg.norm(nrm) * np.linalg.norm(d))
        if min(u64, v64, 1 - u64 - v64) > 1e-2 and cos > 0.3 and np.linalg.norm(nrm) > 0.05:
            hit, t, u, v = oracle.wt(tri, ray)
            assert hit and abs(t - t64) < 1e-5 * abs(t64) and abs(u - u64) < 1e-4 and abs(v - v64) < 1e-4


def _sheared(tri, ray):
    """The watertight test's sheared 2-D vertex coordinates, float32 (for the exact-sign check)."""
    f = np.float32
    o, d = ray[:3].astype(f), ray[3:].astype(f)
    kz = int(np.argmax(np.abs(d))) if np.abs(d).max() > 0 else 0
    kz = 0 if abs(d[1]) <= abs(d[0]) and abs(d[2]) <= abs(d[0]) else (1 if abs(d[2]) <= abs(d[1]) else 2)
    kx, ky = (kz + 1) % 3, (kz + 2) % 3
    if d[kz] < 0:
        kx, ky = ky, kx
    sx, sy = d[kx] / d[kz], d[ky] / d[kz]
    out = []
    for i in range(3):
        a = tri[3 * i:3 * i + 3].astype(f) - o
        out.append((a[kx] - sx * a[kz], a[ky] - sy * a[kz]))
    return out


def test_wt_edge_signs_are_exact():
    """Accept/reject of rays aimed at a shared edge equals the sign test evaluated in exact
    rational arithmetic on the same float32 sheared coordinates (a plain float32 evaluation of
    the edge functions, without the fma-exact fallback, gives zeros where the exact value is
    not zero and fails this); and a ray through a shared edge is accepted by at least one of the
    two triangles."""
    from fractions import Fraction as Fr
    rng = np.random.default_rng(11)
    checked = zero_rounded = 0
    for _ in range(3000):
        a, b = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        d = rng.normal(size=3)
        # c1 and c2 on opposite sides of the edge as seen along the ray: together the two
        # triangles cover a neighbourhood of every interior point of the edge
        c1 = rng.uniform(-1, 1, 3)
        side = np.cross(b - a, c1 - a) @ d
        while True:
            c2 = rng.uniform(-1, 1, 3)
            if (np.cross(b - a, c2 - a) @ d) * side < 0:
                break
        t1 = np.concatenate([a, b, c1]).astype(np.float32)
        t2 = np.concatenate([b, a, c2]).astype(np.float32)  # shares edge (a, b), opposite order
        s = rng.uniform(0.05, 0.95)
        p = t1[0:3].astype(np.float64) * (1 - s) + t1[3:6].astype(np.float64) * s
        ray = np.concatenate([p - 2.0 * d, d]).astype(np.float32)
        hits = []
        for tri in (t1, t2):
            (ax, ay), (bx, by), (cx, cy) = _sheared(tri, ray)
            U = Fr(float(cx)) * Fr(float(by)) - Fr(float(cy)) * Fr(float(bx))
            V = Fr(float(ax)) * Fr(float(cy)) - Fr(float(ay)) * Fr(float(cx))
            W = Fr(float(bx)) * Fr(float(ay)) - Fr(float(by)) * Fr(float(ax))
            if np.float32(cx) * np.float32(by) - np.float32(cy) * np.float32(bx) == 0 and U != 0:
                zero_rounded += 1
            want = not ((U < 0 or V < 0 or W < 0) and (U > 0 or V > 0 or W > 0)) and (U + V + W) != 0
            got = oracle.wt(tri, ray)[0]
            # det == 0 is tested on the rounded sum; only same-sign (accepted) cases matter here
            if U + V + W != 0:
                assert got == want, (tri, ray, U, V, W)
                checked += 1
            hits.append(got)
        assert hits[0] or hits[1], "a ray through the shared edge went through both triangles"
    assert checked > 5000


def test_wt_closed_mesh_has_no_leaks():
    """Rays from the centre of the closed icosphere (81,920 triangles) through every vertex
    shared by five or six triangles: each hits the mesh with the watertight test (brute force
    over all triangles); the reference's Moller-Trumbore lets some of them through."""
    tris = scenes.scene(1)
    verts = np.unique(tris.reshape(-1, 3), axis=0)[:4096]
    rays = np.concatenate([np.zeros_like(verts), verts], 1).astype(np.float32)
    wt = oracle.brute(tris, rays, mode=oracle.LOWEST_ID | oracle.WATERTIGHT)
    assert (wt["t"] < np.inf).all(), f"{int((wt['t'] == np.inf).sum())} rays leaked"
    assert np.allclose(wt["t"], 1.0, atol=1e-5)  # they reach the mesh at the vertex
    mt = oracle.brute(tris, rays)
    print(f"Moller-Trumbore misses {int((mt['t'] == np.inf).sum())} of {len(rays)} vertex rays")


def test_wt_agrees_with_mt_and_fp64_away_from_edges():
    """On random (ray, triangle) pairs whose hit lies clearly inside the triangle, the watertight
    test accepts exactly when Moller-Trumbore does and both agree with the FP64 re-intersection
    (t to 2e-5 relative, u and v to 5e-5), for rays at least 11.5 degrees off the plane."""
    rng = np.random.default_rng(5)
    n = 4000
    tris = rng.uniform(-1, 1, (n, 9)).astype(np.float32)
    # aim at barycentric points in [-0.2, 1.2]: inside and outside the triangle
    w = rng.uniform(-0.2, 1.2, (n, 2))
    tv = tris.astype(np.float64)
    aim = tv[:, 0:3] + w[:, :1] * (tv[:, 3:6] - tv[:, 0:3]) + w[:, 1:] * (tv[:, 6:9] - tv[:, 0:3])
    d = rng.normal(size=(n, 3))
    rays = np.concatenate([aim - rng.uniform(0.5, 4, (n, 1)) * d, d], 1).astype(np.float32)
    t64, u64, v64 = oracle.fp64(tris, rays, np.arange(n, dtype=np.uint32))
    e1, e2 = tv[:, 3:6] - tv[:, 0:3], tv[:, 6:9] - tv[:, 0:3]
    nrm = np.cross(e1, e2)
    cos = np.abs((nrm * d).sum(1)) / (np.linalg.norm(nrm, axis=1) * np.linalg.norm(d, axis=1))
    both = 0
    for i in range(n):
        if cos[i] < 0.2:  # grazing: both tests' rounding is large, acceptance near edges differs
            continue
        inside = min(u64[i], v64[i], 1 - u64[i] - v64[i]) > 1e-4
        outside = min(u64[i], v64[i], 1 - u64[i] - v64[i]) < -1e-4
        hw, tw, uw, vw = oracle.wt(tris[i], rays[i])
        hm, tm, um, vm = oracle.mt(tris[i], rays[i])
        if inside:
            assert hw and hm
            assert abs(tw - t64[i]) <= 2e-5 * max(1.0, abs(t64[i])) and abs(uw - u64[i]) < 5e-5 and abs(vw - v64[i]) < 5e-5
            both += 1
        elif outside:
            assert not hw and not hm
    assert both > 300


def test_wt_traversal_equals_brute_force():
    """The BVH2 traversal with the watertight leaf test equals brute force (lowest id) on random
    soups, rays through vertices included."""
    rng = np.random.default_rng(9)
    for n in (7, 60, 500):
        tris = (rng.uniform(-1, 1, (n, 1, 3)) + rng.uniform(-0.3, 0.3, (n, 3, 3))).reshape(n, 9).astype(np.float32)
        o = rng.uniform(-1.5, 1.5, (600, 3))
        d = np.concatenate([rng.normal(size=(400, 3)), tris[np.arange(200) % n, 0:3] - o[400:]], 0)
        rays = np.concatenate([o, d], 1).astype(np.float32)
        mode = oracle.LOWEST_ID | oracle.WATERTIGHT
        assert_same(oracle.trace(tris, rays, mode), oracle.brute(tris, rays, mode=mode), what=f"wt n={n}")


@pytest.mark.skipif(not oracle.have_ref(), reason="compiled reference not built (needs the reference sources)")
def test_bunny_stand_in_equals_reference():
    """The oracle on the validation camera (workload 5, non-unit directions, zero components)
    and the closed stand-in mesh equals the compiled reference backend on every ray (every
    field; primitive_id where the reference has no equal-t tie)."""
    tris, rays = scenes.scene(5), scenes.rays(5)
    ref = oracle.ref_trace(tris, rays)
    assert_same(oracle.trace(tris, rays, oracle.DFS), ref, what="bunny stand-in, depth-first")
    assert_same(oracle.trace(tris, rays, oracle.LOWEST_ID), ref, ids=False, what="bunny stand-in, lowest id")
    assert 0.05 < float((ref["t"] < np.inf).mean()) < 0.2
