"""The learned t_max seeding path on the GPU: the MFMA forward/backward kernels against the CPU
oracle (oracle/mlp_ref.py), the fused top-4 against the full logits, the cluster cut, and the
seeded traversal returning exactly the plain traversal's records."""
import numpy as np
import pytest

import oracle
import portablert_amd as prt
from oracle import mlp_ref
from portablert_amd import scenes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def room(gpu_ctx):
    tris = scenes.scene(2)
    gpu_ctx.set_tris(tris)
    prt.learned_setup(gpu_ctx, 7)
    return tris


def test_clusters_partition_the_triangles(gpu_ctx, room):
    count, tc = prt.learned_clusters(gpu_ctx, len(room))
    assert 3000 < count <= prt.CLUSTERS
    assert (tc < count).all()
    sizes = np.bincount(tc, minlength=count)
    assert (sizes > 0).all() and sizes.sum() == len(room)
    print(f"{count} clusters, {sizes.min()}..{sizes.max()} triangles", flush=True)


def test_gradients_match_oracle(gpu_ctx, room):
    rng = np.random.default_rng(3)
    rays = scenes.rays(4, stream=1, first=12345, count=256)
    labels = rng.integers(0, prt.MLP_OUT, len(rays)).astype(np.uint32)
    params = [prt.mlp_params(gpu_ctx, layer) for layer in range(4)]
    loss, x0_bits, grads = prt.mlp_grad(gpu_ctx, rays, labels)
    x0 = mlp_ref.bf16_bits_to_f64(x0_bits)
    # the encoding itself: sin/cos of the scaled inputs within bf16 resolution
    assert np.abs(x0).max() <= 1.0
    want_loss, want = mlp_ref.loss_and_grads(x0, params, labels.astype(np.int64), bf16=True)
    assert abs(loss - want_loss) <= 1e-3 * abs(want_loss), (loss, want_loss)
    for layer in range(4):
        for k, name in ((0, "W"), (1, "b")):
            g, w = grads[layer][k].astype(np.float64), want[layer][k]
            scale = np.abs(w).max()
            err = np.abs(g - w).max()
            assert err <= 2e-2 * scale + 1e-7, (layer, name, err, scale)
            cos = float((g * w).sum() / (np.linalg.norm(g) * np.linalg.norm(w) + 1e-30))
            assert cos > 0.9999, (layer, name, cos)


def check_top4(ctx, rays):
    """The fused kernel's top-4 equals the stable argsort of the full logits on every row whose
    4th and 5th logits are clearly apart."""
    top, logits = prt.mlp_top4(ctx, rays, with_logits=True)
    order = np.argsort(-logits, axis=1, kind="stable")[:, :4]
    gap = np.sort(logits, axis=1)[:, ::-1]
    clear = (gap[:, 3] - gap[:, 4]) > 1e-3 * np.abs(gap[:, 3]).clip(1e-3)
    assert clear.mean() > 0.9
    assert np.array_equal(top[clear], order[clear].astype(np.uint32))


def test_top4_matches_full_logits(gpu_ctx, room):
    # 517 rows: eight full 64-ray workgroups and a ragged one with 5 rays
    check_top4(gpu_ctx, scenes.rays(4, stream=2, first=777, count=517))


def test_learned_traversal_identical_to_plain(gpu_ctx, room):
    import torch
    train = scenes.rays(4, stream=1, count=262144)
    test = scenes.rays(4, stream=2, count=1 << 20)
    d_train = torch.from_numpy(train).cuda()
    d_test = torch.from_numpy(test).cuda()
    mask = prt.T | prt.PRIMITIVE_ID | prt.UV
    size, _ = prt.layout(mask)
    before = prt.learned_eval(gpu_ctx, d_test.data_ptr(), len(test))
    loss0 = None
    prt.learned_train(gpu_ctx, d_train.data_ptr(), len(train), epochs=1, batch=8192, lr=2e-3)
    loss0 = gpu_ctx.stat("train_loss")
    prt.learned_train(gpu_ctx, d_train.data_ptr(), len(train), epochs=2, batch=8192, lr=1e-3)
    loss1 = gpu_ctx.stat("train_loss")
    after = prt.learned_eval(gpu_ctx, d_test.data_ptr(), len(test))
    print(f"loss {loss0:.3f} -> {loss1:.3f}; top1/top4 before {before} after {after}", flush=True)
    assert loss1 < np.log(prt.MLP_OUT)
    assert after[1] > before[1] + 0.05
    check_top4(gpu_ctx, test[:517])  # the trained network's peaked logits
    out_plain = torch.zeros(len(test) * size, dtype=torch.uint8, device="cuda")
    out_learn = torch.zeros_like(out_plain)
    torch.cuda.synchronize()
    for fmt in (1, 0):
        gpu_ctx.set_option("node_format", fmt)
        gpu_ctx.nearest_hits_device(d_test.data_ptr(), len(test), mask, out_plain.data_ptr())
        prt.nearest_hits_learned_device(gpu_ctx, d_test.data_ptr(), len(test), mask, out_learn.data_ptr())
        gpu_ctx.synchronize()
        assert torch.equal(out_plain, out_learn), f"learned differs from plain (node format {fmt})"
    gpu_ctx.set_option("node_format", 1)
    # every record field, with and without distance culling (without it the seeds only order the
    # search, and the clusters searched first must still be skipped exactly once)
    full = prt.mask_of(prt.ALL_TAGS)
    fsize, _ = prt.layout(full)
    nf = 262144
    a = torch.zeros(nf * fsize, dtype=torch.uint8, device="cuda")
    b = torch.zeros_like(a)
    torch.cuda.synchronize()  # the library works on its own stream
    try:
        for cull in (1, 0):
            gpu_ctx.set_option("cull", cull)
            gpu_ctx.nearest_hits_device(d_test.data_ptr(), nf, full, a.data_ptr())
            for seeds in (1, 4):  # one seed, and the several-seed kernel
                gpu_ctx.set_option("seeds", seeds)
                b.zero_()
                prt.nearest_hits_learned_device(gpu_ctx, d_test.data_ptr(), nf, full, b.data_ptr())
                gpu_ctx.synchronize()
                assert torch.equal(a, b), f"learned differs from plain (all fields, cull {cull}, {seeds} seeds)"
    finally:
        gpu_ctx.set_option("cull", 1)
        gpu_ctx.set_option("seeds", 1)
    got = out_learn.cpu().numpy().view(prt.hit_dtype(mask))
    idx = np.random.default_rng(0).choice(len(test), 8192, replace=False)
    want = oracle.trace(room, test[idx], oracle.LOWEST_ID)
    assert (got["t"][idx].view(np.uint32) == want["t"].view(np.uint32)).all()
    hit = want["t"] < np.inf
    assert (got["primitive_id"][idx][hit] == want["primitive_id"][hit]).all()


def test_seed_count_keeps_records(gpu_ctx, room):
    """Searching the best 1, 2, 3 or 4 predicted clusters first changes only the order of the
    search: records equal the plain traversal's byte for byte."""
    import torch
    test = scenes.rays(4, stream=2, first=4096, count=262144)
    d_test = torch.from_numpy(test).cuda()
    mask = prt.mask_of(prt.ALL_TAGS)
    size, _ = prt.layout(mask)
    plain = torch.zeros(len(test) * size, dtype=torch.uint8, device="cuda")
    learn = torch.zeros_like(plain)
    torch.cuda.synchronize()
    gpu_ctx.nearest_hits_device(d_test.data_ptr(), len(test), mask, plain.data_ptr())
    try:
        for k in (1, 2, 3, 4):
            gpu_ctx.set_option("seeds", k)
            prt.nearest_hits_learned_device(gpu_ctx, d_test.data_ptr(), len(test), mask, learn.data_ptr())
            gpu_ctx.synchronize()
            assert torch.equal(plain, learn), f"learned with {k} seeds differs from plain"
    finally:
        gpu_ctx.set_option("seeds", 1)


def test_untrained_seeds_keep_records(gpu_ctx, room):
    """With an untrained network the predicted clusters are mostly wrong (many rays miss the
    best cluster's box, which the one-seed traversal enters without testing it): records still
    equal the plain traversal's byte for byte, for 1 and 4 seeds, with and without culling."""
    import torch
    prt.learned_setup(gpu_ctx, 11)
    try:
        test = scenes.rays(4, stream=2, first=65536, count=131072)
        d_test = torch.from_numpy(test).cuda()
        mask = prt.mask_of(prt.ALL_TAGS)
        size, _ = prt.layout(mask)
        plain = torch.zeros(len(test) * size, dtype=torch.uint8, device="cuda")
        learn = torch.zeros_like(plain)
        torch.cuda.synchronize()
        for cull in (1, 0):
            gpu_ctx.set_option("cull", cull)
            gpu_ctx.nearest_hits_device(d_test.data_ptr(), len(test), mask, plain.data_ptr())
            for k in (1, 4):
                gpu_ctx.set_option("seeds", k)
                learn.zero_()
                prt.nearest_hits_learned_device(gpu_ctx, d_test.data_ptr(), len(test), mask, learn.data_ptr())
                gpu_ctx.synchronize()
                assert torch.equal(plain, learn), f"untrained, {k} seeds, cull {cull}: differs from plain"
    finally:
        gpu_ctx.set_option("cull", 1)
        gpu_ctx.set_option("seeds", 1)
        prt.learned_setup(gpu_ctx, 7)


def test_refit_requires_learned_setup_again(gpu_ctx, room):
    """Refit changes every box, so the cluster boxes of the learned state are stale: learned
    calls report ERR_STATE until learned_setup runs again."""
    import torch
    rays = scenes.rays(4, stream=2, count=4096)
    d_rays = torch.from_numpy(rays).cuda()
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    d_out = torch.zeros(len(rays) * size, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    gpu_ctx.refit(room)
    with pytest.raises(prt.PrtError) as e:
        prt.nearest_hits_learned_device(gpu_ctx, d_rays.data_ptr(), len(rays), mask, d_out.data_ptr())
    assert e.value.code == prt.ERR_STATE
    prt.learned_setup(gpu_ctx, 7)
    prt.nearest_hits_learned_device(gpu_ctx, d_rays.data_ptr(), len(rays), mask, d_out.data_ptr(),
                                    torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    got = d_out.cpu().numpy().view(prt.hit_dtype(mask))
    want = oracle.trace(room, rays, oracle.LOWEST_ID)
    assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
    hit = want["t"] < np.inf
    assert np.array_equal(got["primitive_id"][hit], want["primitive_id"][hit])


def test_learned_config_size(gpu_ctx, room):
    """The learned configuration at its stated size: trained on all 16,777,216 training rays
    (one epoch), then the learned-seeded traversal equals the plain traversal byte for byte on
    all 8,388,608 fresh rays, and both equal the oracle on every ray."""
    import torch
    prt.learned_setup(gpu_ctx, 7)
    train = scenes.rays(4, stream=1)
    assert len(train) == 16777216
    d_train = torch.from_numpy(train).cuda()
    torch.cuda.synchronize()
    del train
    print("training on 16,777,216 rays", flush=True)
    prt.learned_train(gpu_ctx, d_train.data_ptr(), len(d_train), epochs=1, batch=8192, lr=1e-3)
    del d_train
    torch.cuda.empty_cache()
    test = scenes.rays(4, stream=2)
    assert len(test) == 8388608
    d_test = torch.from_numpy(test).cuda()
    top1, top4 = prt.learned_eval(gpu_ctx, d_test.data_ptr(), len(test))
    print(f"trained: top-1 {top1:.3f} top-4 {top4:.3f}", flush=True)
    assert top4 > 0.5
    mask = prt.T | prt.PRIMITIVE_ID
    size, _ = prt.layout(mask)
    out_plain = torch.zeros(len(test) * size, dtype=torch.uint8, device="cuda")
    out_learn = torch.zeros_like(out_plain)
    torch.cuda.synchronize()
    gpu_ctx.nearest_hits_device(d_test.data_ptr(), len(test), mask, out_plain.data_ptr())
    prt.nearest_hits_learned_device(gpu_ctx, d_test.data_ptr(), len(test), mask, out_learn.data_ptr())
    gpu_ctx.synchronize()
    assert torch.equal(out_plain, out_learn), "learned differs from plain on the 8,388,608 fresh rays"
    got = out_learn.cpu().numpy().view(prt.hit_dtype(mask))
    want = oracle.trace(room, test, oracle.LOWEST_ID)
    assert np.array_equal(got["t"].view(np.uint32), want["t"].view(np.uint32))
    hit = want["t"] < np.inf
    assert np.array_equal(got["primitive_id"][hit], want["primitive_id"][hit])
