"""Nearest-hit throughput on one MI355X: 262,144-triangle room, 8,388,608 incoherent rays.

A step is one nearest_hits pass over the 8,388,608 rays (t + primitive_id records, 40 bytes
of algorithmic traffic per ray), with rays and records resident in HBM (the C ABI's device
entry point). set_tris (the GPU BVH build) runs once before the timed region and is reported
as build_ms. Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s nearest-hit at 262k tris, 8M incoherent rays: plain vs learned-seeded"
HBM_PEAK_GBS = 8000.0
BYTES_PER_RAY = 24 + 16  # sizeof(Ray) + sizeof(HitReg<t, primitive_id>)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_info():
    model, smt = "unknown", None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        with open("/sys/devices/system/cpu/smt/active") as f:
            smt = f.read().strip() == "1"
    except OSError:
        pass
    return {"cpu": model, "nproc": os.cpu_count(), "smt": smt,
            "affinity": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None}


def host_threads():
    """Every CPU this process may run on (the reference backend uses hardware_concurrency(),
    intersect_cpu.hpp:26; on a shared host the process's affinity set is what it can use)."""
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def cpu_baseline(tris, rays, target_s=10.0):
    """The oracle's restatement of the reference CPU backend (binned-SAH BVH2, depth-first
    traversal, static thread split as intersect_cpu.hpp:20-43), timed on this host: a prefix of
    the ray set sized from a short probe so that each of 3 runs takes about target_s / 3 seconds
    (repeated whole passes when one pass is shorter); the median run is reported."""
    import oracle
    threads = host_threads()
    bvh = oracle.BVH2(tris)
    n = min(len(rays), 1 << 20)  # large enough that thread start-up does not skew the probe
    t0 = time.perf_counter()
    bvh.trace(rays[:n], oracle.DFS, threads)
    rate = n / max(time.perf_counter() - t0, 1e-6)
    want = int(rate * target_s / 3)
    reps = max(1, want // len(rays))
    count = min(len(rays), want) if reps == 1 else len(rays)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        for _ in range(reps):
            bvh.trace(rays[:count], oracle.DFS, threads)
        runs.append(time.perf_counter() - t0)
    dt = float(np.median(runs))
    port = {"value": round(reps * count / dt / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"median of 3 runs, each {reps} pass(es) over the first {count} of the 8,388,608 rays, same "
                      f"scene, oracle depth-first mode (the reference CPU backend restated), {threads} threads, "
                      f"{dt:.2f} s per run",
            "host": host_info()}
    if not oracle.have_ref():
        return port
    # the reference's own CPU backend, compiled from its sources (its BVH2 build, traversal and
    # hardware_concurrency() thread split), on the same scene and every ray
    sec = oracle.ref_time(tris, rays, 3)
    return {"value": round(len(rays) / sec / 1e6, 4), "unit": "Mrays/s", "cores": threads, "kind": "reference",
            "sample": f"best of 3 nearest_hits<t, primitive_id> calls over all 8,388,608 rays after set_tris, "
                      f"same scene, the reference CPU backend compiled from its sources (threads: "
                      f"hardware_concurrency() = {os.cpu_count()}, on the {threads} CPUs this process may use), "
                      f"{sec:.2f} s per call",
            "port_value": port["value"], "port_sample": port["sample"], "host": port["host"]}


def fp64_report(tris, rays, rec):
    """Error of every fp32 hit record against a double-precision re-intersection of the same
    (ray, triangle) pair: for t the max relative error and the distance in fp32 ulps, for the
    barycentrics u and v (in [0, 1]) the absolute error, for the hit point p = o + t d (taken
    from the record, or computed from t as the reference does, bvh.hpp:261-263, when the record
    has no p) the error relative to |p|_inf and in ulps of its largest component."""
    import oracle
    hit = np.isfinite(rec["t"]) & (rec["primitive_id"] != 0xFFFFFFFF)
    if not hit.any():
        return None
    r = rays[hit]
    t64, u64, v64 = oracle.fp64(tris, r, rec["primitive_id"][hit])
    out = {"hits": int(hit.sum())}
    fields = [("t", t64)] + ([("u", u64), ("v", v64)] if "u" in rec.dtype.names else [])
    for name, ref in fields:
        got = rec[name][hit].astype(np.float64)
        err = np.abs(got - ref)
        if name == "t":  # relative to t, and in fp32 ulps of the result
            ulp = err / np.spacing(np.abs(rec[name][hit])).astype(np.float64)
            out[name] = {"max_rel": float((err / np.maximum(np.abs(ref), 1e-30)).max()),
                         "max_ulp": float(ulp.max()), "p99_ulp": float(np.percentile(ulp, 99))}
        else:  # barycentrics lie in [0, 1]: absolute error
            out[name] = {"max_abs": float(err.max()), "p99_abs": float(np.percentile(err, 99))}
    if "px" in rec.dtype.names:
        p32 = np.stack([rec["px"][hit], rec["py"][hit], rec["pz"][hit]], 1).astype(np.float64)
    else:
        t32 = rec["t"][hit]
        p32 = (r[:, :3] + t32[:, None] * r[:, 3:]).astype(np.float64)  # float32, two roundings
    p64 = r[:, :3].astype(np.float64) + t64[:, None] * r[:, 3:].astype(np.float64)
    scale = np.abs(p64).max(1)
    err = np.abs(p32 - p64).max(1)
    ulp = err / np.spacing(scale.astype(np.float32)).astype(np.float64)
    out["p"] = {"max_rel": float((err / np.maximum(scale, 1e-30)).max()), "max_ulp": float(ulp.max()),
                "p99_ulp": float(np.percentile(ulp, 99)), "from_record": "px" in rec.dtype.names}
    return out


def learned_leg(ctx, tris, steps, warmup, epochs, train_rays=None):
    """Learned t_max seeding on the room scene: train on 16,777,216 rays traced by the plain
    method (or their first train_rays), then time plain and learned-seeded traversal on
    8,388,608 fresh rays (same records, same rays). Training time is reported, not part of the
    traversal rate."""
    import torch

    import portablert_amd as prt
    from portablert_amd import scenes
    train = scenes.rays(4, stream=1, count=train_rays)
    test = scenes.rays(4, stream=2)
    n = len(test)
    d_train = torch.from_numpy(train).cuda()
    d_test = torch.from_numpy(test).cuda()
    torch.cuda.synchronize()  # the library trains on its own stream
    del train
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    out_p = torch.empty(n * size, dtype=torch.uint8, device="cuda")
    out_l = torch.empty_like(out_p)
    stream = torch.cuda.current_stream()
    sp = stream.cuda_stream
    prt.learned_setup(ctx, 2024)
    untrained = prt.learned_eval(ctx, d_test.data_ptr(), n)
    t0 = time.perf_counter()
    prt.learned_train(ctx, d_train.data_ptr(), len(d_train), epochs=epochs, batch=8192, lr=1e-3)
    train_s = time.perf_counter() - t0
    top1, top4 = prt.learned_eval(ctx, d_test.data_ptr(), n)
    log(f"learned: trained {epochs} epoch(s) on {len(d_train)} rays in {train_s:.1f} s, loss "
        f"{ctx.stat('train_loss'):.3f}; top-1 {top1:.3f} top-4 {top4:.3f} (untrained {untrained[1]:.4f})")

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    plain_s = timed(lambda: ctx.nearest_hits_device(d_test.data_ptr(), n, mask, out_p.data_ptr(), sp))
    # the number of predicted clusters searched first: the fastest seeded traversal of 1..4
    seeded_by_n = {}
    ctx.set_option("time_kernel", 1)
    for k in (1, 2, 4):
        ctx.set_option("seeds", k)
        ms = []
        for _ in range(3):
            prt.nearest_hits_learned_device(ctx, d_test.data_ptr(), n, mask, out_l.data_ptr(), sp)
            ms.append(ctx.stat("seeded_ms"))
        seeded_by_n[k] = float(np.median(ms))
    ctx.set_option("time_kernel", 0)
    best_k = min(seeded_by_n, key=seeded_by_n.get)
    ctx.set_option("seeds", best_k)
    learn_s = timed(lambda: prt.nearest_hits_learned_device(ctx, d_test.data_ptr(), n, mask, out_l.data_ptr(), sp))
    identical = bool(torch.equal(out_p, out_l))
    ctx.set_option("time_kernel", 1)
    prt.nearest_hits_learned_device(ctx, d_test.data_ptr(), n, mask, out_l.data_ptr(), sp)
    mlp_ms, seeded_ms = ctx.stat("mlp_ms"), ctx.stat("seeded_ms")
    ctx.nearest_hits_device(d_test.data_ptr(), n, mask, out_p.data_ptr(), sp)
    plain_kernel_ms = ctx.stat("trace_ms")
    ctx.set_option("time_kernel", 0)
    ctx.set_option("count_stats", 1)
    ctx.nearest_hits_device(d_test.data_ptr(), n, mask, out_p.data_ptr(), sp)
    ctx.synchronize()
    nodes_plain = ctx.stat("node_visits") / n
    prt.nearest_hits_learned_device(ctx, d_test.data_ptr(), n, mask, out_l.data_ptr(), sp)
    ctx.synchronize()
    nodes_seeded = ctx.stat("node_visits") / n
    ctx.set_option("count_stats", 0)
    flops = 2 * (192 * 256 + 2 * 256 * 256 + 256 * 4097) * n
    peak_ms = flops / 2.5e15 * 1e3  # the network at the dense BF16 MFMA peak
    res = {
        "rays": f"room scene; {len(d_train):,} training rays, 8,388,608 fresh rays (uniform origins, uniform directions)",
        "epochs": epochs,
        "train_s": round(train_s, 2),
        "train_loss": round(ctx.stat("train_loss"), 4),
        "plain_mrays_s": round(n / plain_s / 1e6, 2),
        "learned_mrays_s": round(n / learn_s / 1e6, 2),
        "speedup": round(plain_s / learn_s, 4),
        "identical_records": identical,
        "mlp_ms": round(mlp_ms, 3),
        "seeded_traversal_ms": round(seeded_ms, 3),
        "mlp_tflops": round(flops / (mlp_ms * 1e-3) / 1e12, 1) if mlp_ms > 0 else None,
        "mlp_frac_of_bf16_peak": round(flops / (mlp_ms * 1e-3) / 2.5e15, 4) if mlp_ms > 0 else None,
        "seeds_searched_first": best_k,
        "seeded_traversal_ms_by_seeds": {str(k): round(v, 3) for k, v in seeded_by_n.items()},
        "plain_traversal_kernel_ms": round(plain_kernel_ms, 3),
        # learned wins only when the plain traversal takes longer than the learned call: with the
        # network as measured, and with the network at the dense BF16 peak (2.5 PFLOP/s)
        "break_even_plain_ms": round(mlp_ms + seeded_ms, 3),
        "break_even_plain_ms_at_mfma_peak": round(peak_ms + seeded_ms, 3),
        "seed_top1": round(top1, 4),
        "seed_top4": round(top4, 4),
        "seed_top4_untrained": round(untrained[1], 5),
        "node_visits_per_ray_plain": round(nodes_plain, 2),
        "node_visits_per_ray_seeded": round(nodes_seeded, 2),
    }
    del d_train
    torch.cuda.empty_cache()
    return res


def precision_leg(ctx, d_rays, n, mask, d_ref, stream, steps):
    """fp32 BVH8 nodes (208 B) against 8-bit quantized nodes (80 B) on the same BVH topology:
    traversal time, records compared byte for byte, and counted node + triangle bytes per ray."""
    import torch
    sp = stream.cuda_stream
    res = {}
    fmt0 = ctx.get_option("node_format")
    for fmt, name in ((1, "quantized8"), (0, "fp32")):
        ctx.set_option("node_format", fmt)
        out = torch.zeros_like(d_ref)
        torch.cuda.synchronize()
        ms = _timed(lambda: ctx.nearest_hits_device(d_rays.data_ptr(), n, mask, out.data_ptr(), sp), steps, 1, stream)
        ctx.set_option("count_stats", 1)
        ctx.nearest_hits_device(d_rays.data_ptr(), n, mask, out.data_ptr(), sp)
        ctx.synchronize()
        ctx.set_option("count_stats", 0)
        nodes, tris_t = ctx.stat("node_visits") / n, ctx.stat("tri_tests") / n
        nb = ctx.stat("node_bytes")
        res[name] = {"ms": round(ms, 4), "mrays_s": round(n / ms / 1e3, 2), "node_bytes": int(nb),
                     "node_visits_per_ray": round(nodes, 3), "tri_tests_per_ray": round(tris_t, 3),
                     "counted_bytes_per_ray": round(nodes * nb + tris_t * 48, 1),
                     "identical_to_timed_run": bool(torch.equal(out, d_ref))}
    ctx.set_option("node_format", fmt0)
    q, f = res["quantized8"], res["fp32"]
    res["node_byte_saving"] = round(1 - (q["node_visits_per_ray"] * q["node_bytes"]) /
                                    (f["node_visits_per_ray"] * f["node_bytes"]), 4)
    res["speedup_quantized"] = round(f["ms"] / q["ms"], 3)
    log(f"precision: quantized {q['ms']:.3f} ms, fp32 {f['ms']:.3f} ms, node bytes saved {res['node_byte_saving']:.1%}")
    return res


def refit_leg(ctx, tris, d_rays, n, mask, stream, steps):
    """Dynamic geometry on the room: the triangles moved (every vertex displaced uniformly by up
    to 1 cm, then up to 50 cm), the tree refit instead of rebuilt. Reports refit and build device
    times and the median traversal time on each tree; records of the refit tree must equal the
    rebuilt tree's."""
    import numpy as np
    import torch
    sp = stream.cuda_stream
    rng = np.random.default_rng(11)
    d_tris = torch.from_numpy(tris).cuda()
    torch.cuda.synchronize()  # the library reads it on its own stream
    ctx.set_tris_device(d_tris.data_ptr(), len(tris))
    out_r = torch.empty(n * ctx_layout_size(mask), dtype=torch.uint8, device="cuda")
    out_b = torch.empty_like(out_r)
    res = {"workload": "room, 262,144 triangles, 8,388,608 rays; vertices displaced uniformly by up to d"}
    for name, d in (("moved_1cm", 0.01), ("moved_50cm", 0.5)):
        moved = (tris + rng.uniform(-d, d, tris.shape)).astype(np.float32)
        d_moved = torch.from_numpy(moved).cuda()
        torch.cuda.synchronize()
        ctx.set_tris_device(d_tris.data_ptr(), len(tris))  # refit always starts from the original tree
        refit = []
        for _ in range(3):
            ctx.refit_device(d_moved.data_ptr(), len(moved))
            refit.append(ctx.stat("refit_ms"))
        ms_r = _timed_median(lambda: ctx.nearest_hits_device(d_rays.data_ptr(), n, mask, out_r.data_ptr(), sp), steps, 1, stream)
        ctx.set_tris_device(d_moved.data_ptr(), len(moved))
        build = ctx.stat("build_ms")
        ms_b = _timed_median(lambda: ctx.nearest_hits_device(d_rays.data_ptr(), n, mask, out_b.data_ptr(), sp), steps, 1, stream)
        res[name] = {"refit_ms": round(float(np.median(refit)), 3), "build_ms": round(build, 3),
                     "trace_ms_refit_tree": round(ms_r, 4), "trace_ms_rebuilt_tree": round(ms_b, 4),
                     "identical": bool(torch.equal(out_r, out_b))}
        log(f"refit {name}: refit {res[name]['refit_ms']:.3f} ms vs build {build:.3f} ms; traversal "
            f"{ms_r:.3f} ms (refit tree) vs {ms_b:.3f} ms (rebuilt); identical {res[name]['identical']}")
    ctx.set_tris_device(d_tris.data_ptr(), len(tris))
    return res


def ctx_layout_size(mask):
    import portablert_amd as prt
    return prt.layout(mask)[0]


def _timed_median(fn, steps, warmup, stream):
    """Median ms of single calls of fn, each bracketed by its own events on `stream`."""
    import numpy as np
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, b in ev:
        a.record(stream)
        fn()
        b.record(stream)
    torch.cuda.synchronize()
    return float(np.median([a.elapsed_time(b) for a, b in ev]))


def _timed(fn, steps, warmup, stream):
    """Mean ms per call of fn over `steps` calls, bracketed by events on `stream`."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(stream)
    for _ in range(steps):
        fn()
    b.record(stream)
    torch.cuda.synchronize()
    return a.elapsed_time(b) / steps


def batch_leg(ctx, steps, warmup, stream):
    """64 scenes of 1,011..191,723 triangles with 10,948..1,979,349 rays each (2,516,661
    triangles, 24,428,377 rays): every BVH build and the one persistent traversal launch replayed
    as one hipGraph, after the autotuner. Compared with the same work scene by scene through
    set_tris + nearest_hits with the same builder and collapse (radix tree, greedy collapse), a host
    synchronisation per build."""
    import torch

    import portablert_amd as prt
    from portablert_amd import scenes
    n_sc = scenes.n_scenes(3)
    tl = [scenes.scene(3, s) for s in range(n_sc)]
    rl = [scenes.rays(3, s) for s in range(n_sc)]
    tc, rc = [len(t) for t in tl], [len(r) for r in rl]
    d_tris = torch.from_numpy(np.concatenate(tl)).cuda()
    d_rays = torch.from_numpy(np.concatenate(rl)).cuda()
    del tl, rl
    n = int(sum(rc))
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    d_out = torch.zeros(n * size, dtype=torch.uint8, device="cuda")
    d_seq = torch.zeros_like(d_out)
    torch.cuda.synchronize()
    sp = stream.cuda_stream
    batch = prt.Batch(ctx, tc, rc, mask)
    default_params = [batch.params(k) for k in range(-1, prt.BATCH_BUCKETS)]
    batch.run(d_tris.data_ptr(), d_rays.data_ptr(), d_out.data_ptr(), sp)
    batch.check()
    t0 = time.perf_counter()
    before, after = batch.tune(d_tris.data_ptr(), d_rays.data_ptr(), d_out.data_ptr(), reps=2)
    tune_s = time.perf_counter() - t0
    params = [batch.params(k) for k in range(-1, prt.BATCH_BUCKETS)]
    eager_ms = _timed(lambda: batch.run(d_tris.data_ptr(), d_rays.data_ptr(), d_out.data_ptr(), sp), steps, warmup,
                      stream)
    batch.check()
    batch.capture(d_tris.data_ptr(), d_rays.data_ptr(), d_out.data_ptr())
    graph_ms = _timed(lambda: batch.replay(sp), steps, warmup, stream)
    batch.check()
    # the same work one scene at a time through the single-scene entry points
    tri_off = np.concatenate([[0], np.cumsum(tc)])
    ray_off = np.concatenate([[0], np.cumsum(rc)])

    builder0, collapse0 = ctx.get_option("builder"), ctx.get_option("collapse")
    ctx.set_option("builder", 0)  # the batch builds radix trees with the greedy collapse:
    ctx.set_option("collapse", 0)  # compare like with like

    def sequential():
        for s in range(n_sc):
            ctx.set_tris_device(d_tris[tri_off[s]:].data_ptr(), tc[s])
            ctx.nearest_hits_device(d_rays[ray_off[s]:].data_ptr(), rc[s], mask, d_seq[ray_off[s] * size:].data_ptr(), sp)
        torch.cuda.synchronize()

    sequential()
    t0 = time.perf_counter()
    reps = 2
    for _ in range(reps):
        sequential()
    seq_ms = (time.perf_counter() - t0) * 1e3 / reps
    ctx.set_option("builder", builder0)
    ctx.set_option("collapse", collapse0)
    identical = bool(torch.equal(d_out, d_seq))
    log(f"batch: graph {graph_ms:.2f} ms, eager {eager_ms:.2f} ms, scene by scene {seq_ms:.2f} ms; "
        f"tuner {before:.2f} -> {after:.2f} ms in {tune_s:.1f} s; identical {identical}")
    res = {
        "workload": f"{n_sc} scenes, {int(sum(tc))} triangles, {n} rays (uniform-sphere rays from interior points), "
                    "HitReg<t, primitive_id>; builds + traversal per step",
        "graph_ms": round(graph_ms, 3),
        "graph_mrays_s": round(n / graph_ms / 1e3, 2),
        "eager_ms": round(eager_ms, 3),
        "scene_by_scene_ms": round(seq_ms, 3),
        "speedup_vs_scene_by_scene": round(seq_ms / graph_ms, 2),
        "identical_to_scene_by_scene": identical,
        "tuner_before_ms": round(before, 3),
        "tuner_after_ms": round(after, 3),
        "tuner_s": round(tune_s, 2),
        "params_default": {"launch": default_params[0], "buckets": default_params[1:]},
        "params_tuned": {"launch [block, waves/CU, LDS stack]": params[0],
                         "buckets [width, leaf, tile]": params[1:]},
    }
    del d_tris, d_rays, d_out, d_seq, batch
    torch.cuda.empty_cache()
    return res


def icosphere_leg(ctx, steps, warmup, stream):
    """Closed displaced icosphere (81,920 triangles): 1920x1080 primary rays, then 4
    cosine-weighted bounce rays per hit, every HitReg field (32-byte records)."""
    import torch

    import portablert_amd as prt
    from portablert_amd import scenes
    tris = scenes.scene(1)
    prim = scenes.rays(1)
    mask = prt.mask_of(prt.ALL_TAGS)
    ctx.set_tris(tris)
    hits = ctx.nearest_hits(prim, mask)
    bounce = scenes.bounce_rays(tris, prim, hits["t"], hits["primitive_id"],
                                np.stack([hits["px"], hits["py"], hits["pz"]], 1))
    size, _ = prt.layout(mask)
    sp = stream.cuda_stream
    res = {"workload": f"icosphere {len(tris)} triangles, {len(prim)} primary + {len(bounce)} bounce rays, "
                       "all HitReg fields"}
    for name, r in (("primary", prim), ("bounce", bounce)):
        d_r = torch.from_numpy(r).cuda()
        d_o = torch.zeros(len(r) * size, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        ms = _timed(lambda: ctx.nearest_hits_device(d_r.data_ptr(), len(r), mask, d_o.data_ptr(), sp), steps, warmup,
                    stream)
        res[f"{name}_rays"] = len(r)
        res[f"{name}_ms"] = round(ms, 3)
        res[f"{name}_mrays_s"] = round(len(r) / ms / 1e3, 2)
        rec = d_o.cpu().numpy().view(prt.hit_dtype(mask))
        res[f"{name}_negative_t_rate"] = float(np.mean(rec["t"] < 0))
        try:
            res[f"{name}_fp64_error"] = fp64_report(tris, r, rec)
        except Exception as e:  # checker unavailable: report, do not fail the measurement
            log("fp64 check skipped:", e)
    total = len(prim) + len(bounce)
    res["total_mrays_s"] = round(total / (res["primary_ms"] + res["bounce_ms"]) / 1e3, 2)
    log(f"icosphere: primary {res['primary_mrays_s']} Mrays/s, bounce {res['bounce_mrays_s']} Mrays/s")
    return res


def watertight_leg(ctx, steps, warmup, stream):
    """The optional watertight leaf test (PRT_OPT_LEAF_TEST = 1, no distance culling) beside the
    reference's Moller-Trumbore on the soup (config 0), the icosphere primaries (config 1) and
    the room (config 2): traversal time of each, and the fraction of rays whose nearest hit
    differs (a hit against a miss, or another triangle) or whose t differs."""
    import torch

    import portablert_amd as prt
    from portablert_amd import scenes
    sp = stream.cuda_stream
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    res = {}
    for w, name in ((0, "soup"), (1, "icosphere_primary"), (2, "room")):
        tris, rays = scenes.scene(w), scenes.rays(w)
        n = len(rays)
        ctx.set_tris(tris)
        d_r = torch.from_numpy(rays).cuda()
        outs, ms = {}, {}
        for lt in (0, 1):
            ctx.set_option("leaf_test", lt)
            out = torch.zeros(n * size, dtype=torch.uint8, device="cuda")
            torch.cuda.synchronize()
            ms[lt] = _timed(lambda: ctx.nearest_hits_device(d_r.data_ptr(), n, mask, out.data_ptr(), sp), steps, warmup,
                            stream)
            outs[lt] = out.cpu().numpy().view(prt.hit_dtype(mask))
        ctx.set_option("leaf_test", 0)
        a, b = outs[0], outs[1]
        va, vb = a["t"] < np.inf, b["t"] < np.inf
        differ = (va != vb) | (va & vb & (a["primitive_id"] != b["primitive_id"]))
        res[name] = {"rays": n, "mt_ms": round(ms[0], 4), "watertight_ms": round(ms[1], 4),
                     "watertight_mrays_s": round(n / ms[1] / 1e3, 2),
                     "hit_differs": float(differ.mean()), "mt_hit_wt_miss": float((va & ~vb).mean()),
                     "wt_hit_mt_miss": float((vb & ~va).mean()),
                     "t_bits_differ": float((a["t"].view(np.uint32) != b["t"].view(np.uint32)).mean())}
        log(f"watertight {name}: {ms[1]:.3f} ms vs {ms[0]:.3f} ms; hit differs on {differ.mean():.2e} of rays")
        del d_r
    torch.cuda.empty_cache()
    return res


def bunny_leg(ctx, steps, warmup, stream):
    """The reference's bunny benchmark (examples/bunny/main.cpp:84-140: 4096 x 4096 unit-direction
    primary rays, FullHitReg records) and validation timing windows (validation/main.cpp:230-233:
    set_tris and nearest_hits each within 1000 ms) on the closed 5,120-triangle stand-in for the
    bunny mesh. Host entry points timed by wall clock as the examples do (PCIe copies of 16.8 M
    rays and records included), and the device-resident traversal."""
    import torch

    import portablert_amd as prt
    from portablert_amd import scenes
    tris, rays = scenes.scene(6), scenes.rays(6)
    n = len(rays)
    mask = prt.mask_of(prt.ALL_TAGS)
    build = []
    for _ in range(3):
        t0 = time.perf_counter()
        ctx.set_tris(tris)
        build.append((time.perf_counter() - t0) * 1e3)
    api = []
    for _ in range(3):
        t0 = time.perf_counter()
        hits = ctx.nearest_hits(rays, mask)
        api.append((time.perf_counter() - t0) * 1e3)
    size, _ = prt.layout(mask)
    d_r = torch.from_numpy(rays).cuda()
    d_o = torch.empty(n * size, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    ms = _timed(lambda: ctx.nearest_hits_device(d_r.data_ptr(), n, mask, d_o.data_ptr(), stream.cuda_stream), steps,
                warmup, stream)
    res = {"workload": "bunny benchmark camera, 4096 x 4096 rays (eye (0, -0.02, -0.3)), closed 5,120-triangle "
                       "stand-in mesh, FullHitReg",
           "rays": n, "hits": int((hits["t"] < np.inf).sum()),
           "set_tris_ms": round(float(np.median(build)), 3),
           "nearest_hits_host_ms": round(float(np.median(api)), 3),
           "nearest_hits_host_mrays_s": round(n / float(np.median(api)) / 1e3, 2),
           "device_traversal_ms": round(ms, 4), "device_mrays_s": round(n / ms / 1e3, 2),
           "within_reference_windows_1000ms": bool(max(build) <= 1000.0 and max(api) <= 1000.0)}
    log(f"bunny 4096^2: set_tris {res['set_tris_ms']} ms, nearest_hits {res['nearest_hits_host_ms']} ms (host), "
        f"{ms:.3f} ms (device)")
    del d_r, d_o
    torch.cuda.empty_cache()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20, help="untimed steps first (the clock settles within about 20)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-watertight", action="store_true")
    ap.add_argument("--no-bunny", action="store_true")
    ap.add_argument("--no-precision", action="store_true", help="skip the fp32-node leg")
    ap.add_argument("--no-api", action="store_true", help="skip the host-buffer (PCIe) entry point leg")
    ap.add_argument("--format", type=int, default=1, help="0 fp32 nodes, 1 quantized nodes")
    ap.add_argument("--cull", type=int, default=1)
    ap.add_argument("--calib", action="store_true",
                    help="also run a 512 MiB elementwise copy kernel (HBM counter calibration)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-seeding leg")
    ap.add_argument("--epochs", type=int, default=2, help="training epochs of the learned leg")
    ap.add_argument("--train-rays", type=int, default=None,
                    help="train the learned leg on the first N of the 16,777,216 training rays")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batch leg")
    ap.add_argument("--no-icosphere", action="store_true", help="skip the icosphere primary + bounce leg")
    ap.add_argument("--no-refit", action="store_true", help="skip the dynamic-geometry (refit) leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (sharding rays or scenes "
              "across GPUs is designed but not built)", file=sys.stderr)
        sys.exit(2)

    import torch

    import portablert_amd as prt
    from portablert_amd import scenes

    n_dev, dev_name = prt.probe()
    if n_dev == 0:
        raise SystemExit("no gfx950 device")
    tris = scenes.scene(2)
    rays = scenes.rays(2)
    n = len(rays)
    log(f"scene {len(tris)} tris, {n} rays on {dev_name}")

    ctx = prt.Context(0)
    ctx.set_option("node_format", args.format)
    ctx.set_option("cull", args.cull)
    ctx.set_tris(tris)
    build_ms = ctx.stat("build_ms")
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    d_rays = torch.from_numpy(rays).cuda()
    d_out = torch.empty(n * size, dtype=torch.uint8, device="cuda")
    # a stream of our own (default priority): a non-null handle, so the library launches on it
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sp = stream.cuda_stream

    def step():
        ctx.nearest_hits_device(d_rays.data_ptr(), n, mask, d_out.data_ptr(), sp)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # one event pair around the timed steps (an event between steps is a queue marker with cache
    # maintenance, which costs each step tens of microseconds): the mean step on the stream
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        step()
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = ev0.elapsed_time(ev1) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = n / (wall / args.steps) / 1e6
    # the traversal kernel alone, timed by the library's HIP events (outside the timed region)
    ctx.set_option("time_kernel", 1)
    lib_ms = []
    for _ in range(5):  # median of 5, after the warm-up and timed steps
        step()
        lib_ms.append(ctx.stat("trace_ms"))
    ctx.set_option("time_kernel", 0)
    lib_kernel_ms = float(np.median(lib_ms))
    log(f"{ms_per_step:.3f} ms/step wall, {kernel_ms:.3f} ms/step (stream events), "
        f"{lib_kernel_ms:.3f} ms traversal kernel, build {build_ms:.2f} ms")
    if args.calib:
        a = torch.empty(128 << 20, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        for _ in range(3):
            torch.add(a, 0.0, out=b)  # reads and writes 512 MiB
        torch.cuda.synchronize()
        log("calibration: 3 x elementwise add over 512 MiB in, 512 MiB out")

    # agreement of the timed output with the oracle on every ray, and the error of every hit
    # against a double-precision re-intersection (reported, not timed)
    hits = d_out.cpu().numpy().view(prt.hit_dtype(mask))
    agree, fp64 = None, None
    try:
        import oracle
        want = oracle.trace(tris, rays, oracle.LOWEST_ID, host_threads())
        agree = float(np.mean((hits["t"].view(np.uint32) == want["t"].view(np.uint32)) &
                              ((want["t"] == np.inf) | (hits["primitive_id"] == want["primitive_id"]))))
        del want
        fp64 = fp64_report(tris, rays, hits)
        log(f"oracle agreement {agree} on {n} rays; fp64 error over {fp64['hits'] if fp64 else 0} hits")
    except Exception as e:  # checker unavailable: report, do not fail the measurement
        log("oracle check skipped:", e)
    neg_t = float(np.mean(hits["t"] < 0))

    # the host-buffer entry point: PCIe copies of rays in and records out included
    api_ms = None
    if not args.no_api:
        api = []
        for _ in range(5):  # median of 5; the call above was the warm-up
            ctx.nearest_hits(rays, mask)
            api.append(ctx.stat("api_ms"))
        api_ms = float(np.median(api))

    # precision: fp32 BVH8 nodes against 8-bit quantized nodes, same records
    precision = None if args.no_precision else precision_leg(ctx, d_rays, n, mask, d_out, stream,
                                                             max(5, args.steps // 2))
    refit = None if args.no_refit else refit_leg(ctx, tris, d_rays, n, mask, stream, max(3, args.steps // 4))

    achieved = BYTES_PER_RAY * n / (kernel_ms * 1e-3) / 1e9
    traffic, valu = None, None
    tf = os.path.join(ROOT, "profiles", "traffic_room.json")
    if os.path.exists(tf) and args.format == 1 and args.cull:
        with open(tf) as f:
            traffic = json.load(f).get("bytes_per_launch")  # counters of the committed profile
    vf = os.path.join(ROOT, "profiles", "valu_room.json")
    if os.path.exists(vf) and args.format == 1 and args.cull:
        with open(vf) as f:
            vj = json.load(f)
        # the bound that holds: VALU issue. Instructions per launch from the committed counters of
        # this workload; the ceiling weights the measured int32 and fp32 issue rates of this GPU by
        # the kernel's instruction mix (gfx950 issues 32-bit integer adds about 1.8x faster than
        # fp32 work).
        # The counters leave 42% of the instructions unclassified: at the int32 rate they give
        # the upper bound used as `peak` (frac is then a lower bound), at the fp32 rate the lower
        # bound `peak_lower`
        achieved_valu = vj["sq_insts_valu_per_dispatch"] / (kernel_ms * 1e-3) / 1e9
        peak_v = vj["peak_for_mix_g_per_s_upper"]
        valu = {"bound": "valu_issue", "achieved": round(achieved_valu, 1), "peak": peak_v,
                "unit": "G wave-instructions/s", "frac": round(achieved_valu / peak_v, 4),
                "peak_lower": vj["peak_for_mix_g_per_s"],
                "int32_share": vj["int32_share"], "unclassified_share": vj["unclassified_share"],
                "lane_activity": vj["lane_activity"],
                "source": "instructions and mix: profiles/valu_room.json (rocprofv3 counters of this "
                          "workload); issue rates: profiles/valu_issue_rate.md; achieved uses kernel_ms"}
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "room: 262,144 triangles (boxes, 32-segment cylinders, slabs in a 30x12x18 room), "
                               "8,388,608 rays, origins uniform in the room, directions uniform on the sphere, "
                               "HitReg<t, primitive_id>",
                   "node_format": "quantized8" if args.format else "fp32", "cull": bool(args.cull),
                   "traversal": "plain"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "formula": "achieved = 40 B/ray x 8,388,608 rays / kernel_ms (stream events over the timed "
                                "steps, the per-step counter reset included); frac = achieved / peak",
                     "traffic_source": "FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 per launch from "
                                       "profiles/traffic_room.json (committed counters, not this run)"},
        "valu_roofline": valu,
        "build_ms": round(build_ms, 3),
        "kernel_ms": round(kernel_ms, 4),
        "traversal_kernel_ms": round(lib_kernel_ms, 4),
        "plain_mrays_s": round(value, 2),
        "learned_mrays_s": None,
        "oracle_agreement": agree,
        "fp64_error": fp64,
        "negative_t_rate": neg_t,
        "api_boundary_ms": round(api_ms, 3) if api_ms else None,
        "api_boundary_mrays_s": round(n / api_ms / 1e3, 2) if api_ms else None,
        "precision": precision,
    }
    if refit is not None:
        out["refit"] = refit
    if not args.no_learned:
        lr = learned_leg(ctx, tris, max(3, args.steps // 4), 1, args.epochs, args.train_rays)
        out["learned_mrays_s"] = lr["learned_mrays_s"]
        out["learned"] = lr
    if not args.no_batch:
        out["batch"] = batch_leg(ctx, max(3, args.steps // 4), 1, stream)
    if not args.no_icosphere:
        out["icosphere"] = icosphere_leg(ctx, max(3, args.steps // 4), 1, stream)
    if not args.no_bunny:
        out["bunny"] = bunny_leg(ctx, max(3, args.steps // 4), 1, stream)
    if not args.no_watertight:
        out["watertight"] = watertight_leg(ctx, max(3, args.steps // 4), 1, stream)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(tris, rays)
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
