"""Writes the golden vectors under tests/golden/ from the reference CPU backend.

Runs the reference's own CPU backend (compiled from its sources into oracle/_ref by
`make -C oracle`) on:
  * the known-answer and edge-case scenes below (hand-made, tiny),
  * samples of the deterministic workloads 0, 1 (primaries and bounces) and 2 at full scene
    size (the scenes are regenerated from their seeds; only their SHA-256 is stored),
  * the bunny validation scene (mesh read from the reference's examples; only its hash, the
    regenerated hit mask and a record sample are stored).
Each fixture stores the rays (float32 x 6), the reference FullHitReg records (32 bytes each,
t/u/v/p bit patterns, valid, and the id the reference's depth-first order picks) and the
scene hash. Miss records keep the reference's indeterminate u/v/id bytes; tests ignore them.

Usage: python tests/golden/make_golden.py   (needs /root/reference)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from portablert_amd import scenes  # noqa: E402


def edge_scenes():
    """Small scenes with the cases the reference's tests leave open."""
    out = {}
    kt = np.array([[-1, -1, 0, 1, -1, 0, 0, 1, 0]], np.float32)
    # known answer (examples/validation/main.cpp:68-101)
    out["known"] = (kt, np.array([[0.1, 0, -1, 0, 0, 1], [-2, 0, -1, 0, 0, 1]], np.float32))
    # quad split on its diagonal, rays on the diagonal: equal-t ties (F2)
    quad = np.array([[0, 0, 0, 1, 0, 0, 1, 1, 0], [0, 0, 0, 1, 1, 0, 0, 1, 0]], np.float32)
    diag = np.array([[x, x, -1, 0, 0, 1] for x in (0.125, 0.25, 0.5, 0.75)] +
                    [[x, x, 1, 0, 0, -1] for x in (0.25, 0.5)], np.float32)
    out["tie"] = (quad, diag)
    out["tie_swapped"] = (quad[::-1].copy(), diag)
    # triangle behind the origin, origin inside its box: negative t (F1)
    tilt = np.array([[-1, -1, -0.5, 1, -1, 0.5, 0, 1, 0.0]], np.float32)
    neg = np.array([[0.0, -0.2, 0.1, 0, 0, 1], [0.2, 0.0, 0.3, 0, 0, 1], [0.0, 0.0, 2.0, 0, 0, 1],
                    [0.1, -0.1, 0.05, 0.3, 0.2, 0.9]], np.float32)
    out["negative_t"] = (tilt, neg)
    # axis-parallel rays, zero direction components, rays in the triangle plane (det == 0)
    ax = np.array([[0.1, 0, -1, 0, 0, 1], [0.1, 0, 1, 0, 0, -1], [-1, -1, -1, 0, 0, 1],
                   [0, 0, -1, 0, 0, 1], [-5, 0, 0, 1, 0, 0], [0, -5, 0, 0, 1, 0],
                   [1, -1, -1, 0, 0, 1], [0.5, 0.0, 0.0, 0, 0, 1], [0, 0, 0, 0, 0, 0],
                   [0.3, 0.2, -2, -0.0, 0.0, 1]], np.float32)
    out["axis_parallel"] = (kt, ax)
    # two coplanar copies and a stack of parallel triangles (ties across ids, nearest of many)
    stack = np.concatenate([kt + np.array([0, 0, z, 0, 0, z, 0, 0, z], np.float32) for z in
                            (0.5, 0.0, 0.25, 0.0, -0.25, 0.5)])
    sr = np.array([[0.1, 0.0, -1, 0, 0, 1], [0.1, 0.0, 1, 0, 0, -1], [0.0, 0.0, 0.1, 0, 0, 1],
                   [0.0, 0.0, 0.1, 0, 0, -1], [0.05, -0.5, -2, 0.01, 0.02, 1]], np.float32)
    out["stack"] = (stack, sr)
    # single triangle (bvh.hpp:165-181) and rays missing it
    out["single"] = (kt, np.array([[0, 0, -1, 0.1, 0.1, 1], [5, 5, -1, 0, 0, 1], [0, 0, 1, 0, 0, 1]], np.float32))
    # random small soups with random rays, including origins inside boxes
    rng = np.random.default_rng(12345)
    for k, n in enumerate((2, 3, 7, 33, 200)):
        c = rng.uniform(-1, 1, (n, 1, 3))
        tris = (c + rng.uniform(-0.3, 0.3, (n, 3, 3))).reshape(n, 9).astype(np.float32)
        o = rng.uniform(-1.2, 1.2, (256, 3))
        d = rng.normal(size=(256, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        out[f"soup{n}"] = (tris, np.concatenate([o, d], 1).astype(np.float32))
    # degenerate triangles among ordinary ones: zero area (collinear), a point, exact duplicates
    # (equal t on different ids) and triangles sharing an edge
    base = (rng.uniform(-1, 1, (40, 1, 3)) + rng.uniform(-0.3, 0.3, (40, 3, 3))).reshape(40, 9)
    col = np.array([[-1, -1, 0, 0, 0, 0, 1, 1, 0], [0.2, 0.2, 0.2] * 3, [-1, 0, 0.3, 1, 0, 0.3, 0, 0, 0.3]])
    dup = np.concatenate([base[:5], base[:5], col, base[5:]]).astype(np.float32)
    o = rng.uniform(-1.2, 1.2, (256, 3))
    d = rng.normal(size=(256, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    out["degenerate"] = (dup, np.concatenate([o, d], 1).astype(np.float32))
    # non-finite and degenerate rays: NaN or inf in origin or direction, a zero direction
    nf = np.array([[np.nan, 0, -1, 0, 0, 1], [0, np.nan, -1, 0, 0, 1], [0.1, 0, -1, np.nan, 0, 1],
                   [0.1, 0, -1, 0, 0, np.nan], [np.inf, 0, -1, 0, 0, 1], [0.1, 0, -np.inf, 0, 0, 1],
                   [0.1, 0, -1, np.inf, 0, 1], [0.1, 0, -1, 0, 0, np.inf], [0.1, 0, -1, 0, 0, -np.inf],
                   [0.1, 0, -1, 0, 0, 0], [0, 0, 0, 0, 0, 0], [0.1, 0, -1, 0, 0, 1e-40]], np.float32)
    out["nonfinite_rays"] = (np.concatenate([kt, dup]), nf)
    # the soups scaled far up and far down: coordinates beyond 2^60 (no fast node bounds) and
    # near the bottom of the normal range. At 1e19 Moller-Trumbore overflows for most rays and at
    # 1e-19 1/det does for all of them, so the same soups scaled by 1e12 and 1e-12 (which still
    # hit) and the soup with one far triangle at 2^61 (a scene box beyond 2^60) join them
    s200, r200 = out["soup200"]
    for name, k in (("huge", 1e19), ("tiny", 1e-19), ("scaled_up", 1e12), ("scaled_down", 1e-12)):
        r = r200.copy()
        r[:, :3] *= np.float32(k)
        out[name] = ((s200 * np.float32(k)).astype(np.float32), r)
    far = np.float32(2.0 ** 61)
    ft = np.array([[far, far, far, 1.5 * far, far, far, far, 1.5 * far, far]], np.float32)
    out["far_triangle"] = (np.concatenate([s200, ft]), r200.copy())
    return out


def _mt_f32(tri, o, d):
    """The reference leaf test (core.hpp:27-65) vectorised in float32, every operation rounded on
    its own (numpy does not fuse): (accepted, t). Used only to search candidates; every case
    kept is confirmed with the oracle's own oracle.mt."""
    f32 = np.float32
    e1, e2 = tri[:, 3:6] - tri[:, 0:3], tri[:, 6:9] - tri[:, 0:3]
    px = d[:, 1] * e2[:, 2] - d[:, 2] * e2[:, 1]
    py = d[:, 2] * e2[:, 0] - d[:, 0] * e2[:, 2]
    pz = d[:, 0] * e2[:, 1] - d[:, 1] * e2[:, 0]
    det = e1[:, 0] * px + e1[:, 1] * py + e1[:, 2] * pz
    with np.errstate(all="ignore"):
        inv = f32(1) / det
        s = o - tri[:, 0:3]
        u = (s[:, 0] * px + s[:, 1] * py + s[:, 2] * pz) * inv
        qx = s[:, 1] * e1[:, 2] - s[:, 2] * e1[:, 1]
        qy = s[:, 2] * e1[:, 0] - s[:, 0] * e1[:, 2]
        qz = s[:, 0] * e1[:, 1] - s[:, 1] * e1[:, 0]
        v = (d[:, 0] * qx + d[:, 1] * qy + d[:, 2] * qz) * inv
        t = (e2[:, 0] * qx + e2[:, 1] * qy + e2[:, 2] * qz) * inv
        ok = (det != 0) & ~((u < 0) | (u > 1)) & ~((v < 0) | (u + v > 1))
    return ok, t


def _box_entry_f32(o, d, tri):
    """Entry distance of the reference slab test (bvh.hpp:195-222) on each triangle's own box,
    NaN where the test rejects."""
    lo = np.minimum(np.minimum(tri[:, 0:3], tri[:, 3:6]), tri[:, 6:9])
    hi = np.maximum(np.maximum(tri[:, 0:3], tri[:, 3:6]), tri[:, 6:9])
    with np.errstate(all="ignore"):
        inv = np.float32(1) / d
        t1, t2 = (lo - o) * inv, (hi - o) * inv
    mn, mx = np.where(t2 < t1, t2, t1), np.where(t1 < t2, t2, t1)
    tmin = np.maximum(np.maximum(mn[:, 0], mn[:, 1]), mn[:, 2])
    tmax = np.minimum(np.minimum(mx[:, 0], mx[:, 1]), mx[:, 2])
    return np.where(~(tmax < 0) & ~(tmin > tmax), tmin, np.nan)


def _sliver_case(rng, centre, tilted):
    """A sliver triangle near `centre` and a ray for which the reference leaf test accepts it at
    a t more than 2^-9 (relative) below the entry distance of the sliver's own box: searched
    with _mt_f32, confirmed with oracle.mt and oracle.slab. Axis-aligned slivers lie in a plane
    y = const; tilted ones in a random plane."""
    n = 20000
    while True:
        v0 = centre + rng.uniform(-20, 20, (n, 3))
        if tilted:
            a = rng.normal(size=(n, 3))
            a /= np.linalg.norm(a, axis=1, keepdims=True)
            b = np.cross(a, rng.normal(size=(n, 3)))
            b /= np.linalg.norm(b, axis=1, keepdims=True)
        else:
            ang = rng.uniform(0, 2 * np.pi, n)
            a = np.stack([np.cos(ang), np.zeros(n), np.sin(ang)], 1)
            b = np.stack([-np.sin(ang), np.zeros(n), np.cos(ang)], 1)
        length = 10 ** rng.uniform(-0.5, 1, n)
        h = length * 10 ** rng.uniform(-6.5, -4, n)
        frac = rng.uniform(0.2, 0.8, n)
        tri = np.concatenate([v0, v0 + length[:, None] * a, v0 + (length * frac)[:, None] * a + h[:, None] * b],
                             1).astype(np.float32)
        tv = tri.astype(np.float64)
        b1 = rng.uniform(0.05, 0.9, n)
        b2 = rng.uniform(0.05, 0.95 - b1)
        aim = tv[:, 0:3] + b1[:, None] * (tv[:, 3:6] - tv[:, 0:3]) + b2[:, None] * (tv[:, 6:9] - tv[:, 0:3])
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        o = (aim - rng.uniform(8, 30, n)[:, None] * d).astype(np.float32)
        d = d.astype(np.float32)
        ok, t = _mt_f32(tri, o, d)
        te = _box_entry_f32(o, d, tri)
        with np.errstate(all="ignore"):
            under = (te - t) / te
        good = np.flatnonzero(ok & (t > 5) & (under > 2.0 ** -9) & (under < 0.05))
        for i in good:
            ray = np.concatenate([o[i], d[i]])
            hit, tt, _, _ = oracle.mt(tri[i], ray)
            lo = np.minimum(np.minimum(tri[i, 0:3], tri[i, 3:6]), tri[i, 6:9])
            hi = np.maximum(np.maximum(tri[i, 0:3], tri[i, 3:6]), tri[i, 6:9])
            if hit and np.float32(tt) == t[i] and oracle.slab(ray, lo, hi):
                return tri[i], ray, float(t[i]), float(te[i])


def _facing_triangle(rng, ray, t_at, radius, hit):
    """A small triangle in the plane perpendicular to the ray at distance t_at: hit=True puts
    the ray inside it (an ordinary, well-conditioned triangle), hit=False outside it but inside
    its box (the ray crosses the box, the leaf test rejects it)."""
    o, d = ray[:3].astype(np.float64), ray[3:].astype(np.float64)
    x = o + t_at * d
    u1 = np.cross(d, rng.normal(size=3))
    u1 /= np.linalg.norm(u1)
    u2 = np.cross(d, u1)
    for _ in range(10000):
        if hit:
            ang = rng.uniform(0, 2 * np.pi) + np.array([0, 2 * np.pi / 3, 4 * np.pi / 3])
        else:  # three directions within a half-plane: the centre is outside the triangle
            ang = rng.uniform(0, 2 * np.pi) + np.sort(rng.uniform(0, np.pi * 0.9, 3))
        vs = [x + radius * (np.cos(g) * u1 + np.sin(g) * u2) for g in ang]
        tri = np.concatenate(vs).astype(np.float32)
        h, tt, _, _ = oracle.mt(tri, ray)
        lo = np.minimum(np.minimum(tri[0:3], tri[3:6]), tri[6:9])
        hi = np.maximum(np.maximum(tri[0:3], tri[3:6]), tri[6:9])
        if hit == bool(h) and oracle.slab(ray, lo, hi):
            return tri
    raise RuntimeError("no facing triangle found")


def sliver_scenes():
    """Adversarial scenes for distance culling (no reference test covers them): per cluster, a
    sliver whose Moller-Trumbore t undershoots its own box entry by more than 2^-9, an ordinary
    competitor the ray hits between the two distances (the reference returns the sliver, which
    is nearer), and padding so that the competitor is tested before the sliver's box is popped:
    'gap' one triangle whose box the ray crosses between them, 'cloud' 24 triangles around the
    competitor, off the ray. Clusters sit 400 units apart so rays meet only their own."""
    rng = np.random.default_rng(4242)
    out = {}
    for name, tilted, cloud, gap in (("sliver_gap", False, 0, True), ("sliver_cloud", False, 24, False),
                                     ("sliver_tilted", True, 12, True)):
        tris, rays = [], []
        for c in range(6):
            centre = np.array([400.0 * c, 0.0, 0.0])
            sl, ray, t_c, t_e = _sliver_case(rng, centre, tilted)
            t_comp = t_c + 0.5 * (t_e - t_c)
            group = [sl, _facing_triangle(rng, ray, t_comp, 0.02 * t_comp, True)]
            if gap:
                group.append(_facing_triangle(rng, ray, t_c + 0.8 * (t_e - t_c), 0.02 * t_comp, False))
            o, d = ray[:3].astype(np.float64), ray[3:].astype(np.float64)
            for _ in range(cloud):
                x = o + t_comp * d + rng.normal(size=3) * 0.1 * t_comp / 20
                group.append((np.tile(x, 3) + rng.uniform(-0.01, 0.01, 9) * t_comp / 20).astype(np.float32))
            tris.extend(group)
            rays.append(ray)
        tris = np.stack(tris).astype(np.float32)
        rays = np.stack(rays).astype(np.float32)
        # shuffle the triangles (ids): the sliver is not always the lowest id of its cluster
        tris = tris[rng.permutation(len(tris))]
        out[name] = (tris, rays)
    return out


def workload_samples():
    rng = np.random.default_rng(777)
    res = {}
    for w, count in ((0, 4096), (2, 4096)):
        tris = scenes.scene(w)
        total = scenes.n_rays(w)
        idx = np.sort(rng.choice(total, count, replace=False))
        res[f"workload{w}"] = (tris, scenes.ray_sample(w, idx), idx)
    # workload 1: primaries and the bounce rays they spawn (traced with the reference)
    tris = scenes.scene(1)
    total = scenes.n_rays(1)
    idx = np.sort(rng.choice(total, 4096, replace=False))
    prim = scenes.ray_sample(1, idx)
    res["workload1_primary"] = (tris, prim, idx)
    hits = oracle.ref_trace(tris, prim)
    p = np.stack([hits["px"], hits["py"], hits["pz"]], 1)
    bounce = scenes.bounce_rays(tris, prim, hits["t"], hits["primitive_id"], p)
    res["workload1_bounce"] = (tris, bounce[:4096], None)
    return res


BUNNY_OBJ = "/root/reference/examples/common/bunny.obj"


def bunny_fixture():
    """The reference's scene-level validation (examples/validation/main.cpp:113-234): the bunny
    mesh seen by the 1024 x 1024 camera (workload 5). Its golden image (bunny.png) is not in the
    reference, so the hit mask is regenerated with the compiled reference backend. Only the
    mesh hash, the mask (1 bit per ray) and a sample of full records are stored."""
    tris = scenes.load_obj(BUNNY_OBJ)
    rays = scenes.rays(5)
    hits = oracle.ref_trace(tris, rays)
    idx = np.arange(0, len(rays), 256, dtype=np.uint64)
    np.savez_compressed(os.path.join(HERE, "bunny.npz"), mesh_sha256=np.array(scenes.sha256(tris)),
                        n_tris=np.array(len(tris)), mask=np.packbits(hits["valid"] != 0), index=idx,
                        hits=hits[idx.astype(np.int64)].view(np.uint8).reshape(-1, 32))


def main():
    if not oracle.have_ref():
        raise SystemExit("build the reference shim first: make -C oracle")
    if os.path.exists(BUNNY_OBJ):
        bunny_fixture()
    for name, (tris, rays) in list(edge_scenes().items()) + list(sliver_scenes().items()):
        hits = oracle.ref_trace(tris, rays)
        np.savez_compressed(os.path.join(HERE, f"edge_{name}.npz"), tris=tris, rays=rays,
                            hits=hits.view(np.uint8).reshape(-1, 32))
    for name, (tris, rays, idx) in workload_samples().items():
        hits = oracle.ref_trace(tris, rays)
        extra = {} if idx is None else {"index": idx.astype(np.uint64)}
        np.savez_compressed(os.path.join(HERE, f"{name}.npz"), scene_sha256=np.array(scenes.sha256(tris)),
                            rays=rays, hits=hits.view(np.uint8).reshape(-1, 32), **extra)
    # HitReg layouts of the reference headers for all 32 field masks
    lay = np.stack([oracle.ref_layout(m) if m else np.array([8] + [-1] * 8, np.int32) for m in range(32)])
    np.save(os.path.join(HERE, "hitreg_layouts.npy"), lay)
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith((".npz", ".npy"))))


if __name__ == "__main__":
    main()
