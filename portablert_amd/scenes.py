"""Deterministic synthetic workloads (scenes and rays) for the nearest-hit path.

Thin ctypes view of ``lib/libprt_scenes.so`` (csrc/scenes.cpp). Workload numbers:
0 soup + pinhole, 1 displaced icosphere (+ bounce rays), 2 room with incoherent rays,
3 batch of 64 soups, 4 room with training / inference ray streams, 5 the validation example's
1024 x 1024 bunny camera and 6 the bunny benchmark's 4096 x 4096 camera, both on a closed
5,120-triangle stand-in for the bunny mesh (the mesh itself has no licence file).
"""
import ctypes
import hashlib
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libprt_scenes.so")

SOUP, SPHERE, ROOM, BATCH, LEARNED, BUNNY, BUNNY_BENCH = 0, 1, 2, 3, 4, 5, 6

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (python -c 'import __graft_entry__ as g; g.build()')")
        L = ctypes.CDLL(LIB)
        vp, u64, ci = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        L.prt_workload_scenes.argtypes = [ci]
        L.prt_scene_tris.restype = u64
        L.prt_scene_tris.argtypes = [ci, ci]
        L.prt_scene_generate.argtypes = [ci, ci, vp]
        L.prt_rays_count.restype = u64
        L.prt_rays_count.argtypes = [ci, ci, ci]
        L.prt_rays_generate.argtypes = [ci, ci, ci, u64, u64, vp]
        L.prt_bounce_rays.argtypes = [vp, u64, vp, vp, vp, vp, u64, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def n_scenes(workload):
    return lib().prt_workload_scenes(workload)


def scene(workload, index=0):
    """Triangles of a scene, float32 array of shape (n, 9)."""
    n = lib().prt_scene_tris(workload, index)
    out = np.empty((n, 9), np.float32)
    if lib().prt_scene_generate(workload, index, _p(out)):
        raise ValueError(f"no scene {index} in workload {workload}")
    return out


def n_rays(workload, index=0, stream=0):
    return lib().prt_rays_count(workload, index, stream)


def rays(workload, index=0, stream=0, first=0, count=None):
    """Rays [first, first+count) of a ray set, float32 array of shape (count, 6)."""
    total = n_rays(workload, index, stream)
    if count is None:
        count = total - first
    out = np.empty((count, 6), np.float32)
    if lib().prt_rays_generate(workload, index, stream, first, count, _p(out)):
        raise ValueError("ray range out of bounds")
    return out


def ray_sample(workload, idx, index=0, stream=0):
    """Rays at arbitrary indices (each generated independently)."""
    idx = np.asarray(idx, dtype=np.uint64)
    out = np.empty((len(idx), 6), np.float32)
    for k, i in enumerate(idx):
        out[k] = rays(workload, index, stream, int(i), 1)[0]
    return out


def bounce_rays(tris, prim_rays, t, prim_id, p):
    """4 cosine-weighted bounce rays per finite primary hit (workload 1)."""
    tris = np.ascontiguousarray(tris, np.float32)
    prim_rays = np.ascontiguousarray(prim_rays, np.float32)
    t = np.ascontiguousarray(t, np.float32)
    prim_id = np.ascontiguousarray(prim_id, np.uint32)
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    n = len(prim_rays)
    out = np.empty((4 * n, 6), np.float32)
    n_out = ctypes.c_uint64(0)
    lib().prt_bounce_rays(_p(tris), len(tris), _p(prim_rays), _p(t), _p(prim_id), _p(p), n, _p(out),
                          ctypes.byref(n_out))
    return out[: n_out.value]


def sha256(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_obj(path):
    """Triangles of a Wavefront OBJ file (faces with 3 vertices only; others skipped), in file
    order, as the reference validation example reads the bunny (examples/validation/main.cpp:
    139-164). Negative (relative) indices are supported."""
    verts, tris = [], []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if not parts:
                continue
            if parts[0] == "v":
                verts.append([float(x) for x in parts[1:4]])
            elif parts[0] == "f" and len(parts) == 4:
                idx = []
                for tok in parts[1:]:
                    i = int(tok.split("/")[0])
                    idx.append(i - 1 if i > 0 else len(verts) + i)
                tris.append(verts[idx[0]] + verts[idx[1]] + verts[idx[2]])
    return np.asarray(tris, np.float32).reshape(-1, 9)
