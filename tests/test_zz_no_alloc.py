"""The no-allocation traversal entry point (prt_mi355x_nearest_hits_device_noalloc): records equal
prt_mi355x_nearest_hits_device's byte for byte, on two calls in a row, the second with fewer rays;
and the oracle on a sample."""
import ctypes

import numpy as np
import pytest

import oracle
import portablert_amd as prt
from portablert_amd import scenes

pytestmark = pytest.mark.gpu


def test_noalloc_matches_device_entry_point(gpu_ctx):
    import torch
    tris = scenes.scene(2)
    rays = scenes.rays(2, count=1 << 21)
    gpu_ctx.set_tris(tris)
    mask = prt.T | prt.PRIMITIVE_ID | prt.UV
    size, _ = prt.layout(mask)
    stream = torch.cuda.Stream()
    sp = stream.cuda_stream
    d_rays = torch.from_numpy(rays).cuda()
    for n in (len(rays), 300001):  # the second call smaller: the scratch keeps the first call's contents
        ref = torch.zeros(n * size, dtype=torch.uint8, device="cuda")
        got = torch.full((n * size,), 0xAB, dtype=torch.uint8, device="cuda")
        torch.cuda.synchronize()
        print(f"noalloc: {n} rays", flush=True)
        gpu_ctx.nearest_hits_device(d_rays.data_ptr(), n, mask, ref.data_ptr(), sp)
        gpu_ctx.nearest_hits_device_noalloc(d_rays.data_ptr(), n, mask, got.data_ptr(), sp)
        gpu_ctx.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(ref, got), f"no-alloc records differ from the device entry point at n={n}"
        rec = got.cpu().numpy().view(prt.hit_dtype(mask))
        idx = np.sort(np.random.default_rng(n).choice(n, 4096, replace=False))
        want = oracle.trace(tris, rays[idx], oracle.LOWEST_ID)
        assert rec["t"][idx].tobytes() == want["t"].tobytes()
        hit = want["t"] < np.inf
        assert (rec["primitive_id"][idx][hit] == want["primitive_id"][hit]).all()


def test_noalloc_argument_errors(gpu_ctx):
    """Bad layouts and null pointers are refused before anything is launched."""
    gpu_ctx.set_tris(scenes.scene(0))
    L = prt.lib()
    mask = prt.T | prt.PRIMITIVE_ID
    size, off = prt.layout(mask)
    offs = np.asarray(off, np.int32)
    rc = L.prt_mi355x_nearest_hits_device_noalloc(gpu_ctx.h, None, 10, mask, None, size,
                                                  offs.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == prt.ERR_ARG
    rc = L.prt_mi355x_nearest_hits_device_noalloc(gpu_ctx.h, ctypes.c_void_p(16), 10, mask, ctypes.c_void_p(16),
                                                  size + 4, offs.ctypes.data_as(ctypes.c_void_p), None)
    assert rc == prt.ERR_LAYOUT
