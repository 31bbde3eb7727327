"""CPU oracle for the nearest-hit path -- TEST INFRASTRUCTURE ONLY.

Python view of ``liboracle.so`` (a C restatement of the reference CPU backend, see oracle.h)
and, where it was built, of ``_ref/libprtref.so`` (the reference CPU backend compiled from its
own sources). Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import this package, and only as the checker or the timed CPU baseline.

Tie modes: ``DFS`` reproduces the reference exactly (first candidate in depth-first order wins
at equal t, bvh.hpp:247 and :255-256); ``LOWEST_ID`` resolves equal t to the lowest primitive id.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libprtref.so")

DFS = 0
LOWEST_ID = 1
WATERTIGHT = 0x100  # OR into a mode: the watertight leaf test instead of Moller-Trumbore

# FullHitReg layout (reference hitreg.hpp:29-44), 32 bytes
HIT_DTYPE = np.dtype([("u", "<f4"), ("v", "<f4"), ("t", "<f4"), ("primitive_id", "<u4"),
                      ("valid", "u1"), ("pad", "u1", 3), ("px", "<f4"), ("py", "<f4"), ("pz", "<f4")])
assert HIT_DTYPE.itemsize == 32

_lib = None
_ref = None


def build():
    """Compile liboracle.so (and the reference shim when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        L.orc_build.restype = vp
        L.orc_build.argtypes = [vp, sz]
        L.orc_free.argtypes = [vp]
        L.orc_node_count.restype = sz
        L.orc_node_count.argtypes = [vp]
        L.orc_node.argtypes = [vp, sz, vp, vp, vp, vp, vp]
        L.orc_trace.argtypes = [vp, vp, sz, ctypes.c_int, ctypes.c_int, vp]
        L.orc_brute.argtypes = [vp, sz, vp, sz, ctypes.c_int, vp]
        L.orc_slab.argtypes = [vp, vp, vp]
        L.orc_mt.argtypes = [vp, vp, vp, vp, vp]
        L.orc_wt.argtypes = [vp, vp, vp, vp, vp]
        L.orc_brute_mode.argtypes = [vp, sz, vp, sz, ctypes.c_int, ctypes.c_int, vp]
        L.orc_fp64.argtypes = [vp, sz, vp, vp, sz, vp, vp, vp]
        _lib = L
    return _lib


def have_ref():
    return os.path.exists(REF_LIB)


def ref():
    global _ref
    if _ref is None:
        if not have_ref():
            raise RuntimeError("reference CPU backend not built (needs the reference sources)")
        R = ctypes.CDLL(REF_LIB)
        vp, u64 = ctypes.c_void_p, ctypes.c_uint64
        R.ref_trace_full.argtypes = [vp, u64, vp, u64, vp]
        R.ref_time_trace.argtypes = [vp, u64, vp, u64, ctypes.c_int]
        R.ref_time_trace.restype = ctypes.c_double
        R.ref_hitreg_layout.argtypes = [ctypes.c_int, vp]
        _ref = R
    return _ref


def _f32(a, cols):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a.reshape(-1, cols)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class BVH2:
    """The reference's binned-SAH BVH2 (bvh.hpp:58-193), built on the CPU."""

    def __init__(self, tris):
        self.tris = _f32(tris, 9)
        self.h = lib().orc_build(_p(self.tris), len(self.tris))
        if not self.h:
            raise MemoryError("orc_build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def node_count(self):
        return lib().orc_node_count(self.h)

    def node(self, i):
        leaf, tri, l, r = (ctypes.c_uint32() for _ in range(4))
        b = np.zeros(6, np.float32)
        lib().orc_node(self.h, i, ctypes.byref(leaf), ctypes.byref(tri), ctypes.byref(l),
                       ctypes.byref(r), _p(b))
        return leaf.value, tri.value, l.value, r.value, b

    def trace(self, rays, mode=LOWEST_ID, threads=0):
        rays = _f32(rays, 6)
        out = np.zeros(len(rays), HIT_DTYPE)
        rc = lib().orc_trace(self.h, _p(rays), len(rays), mode, threads, _p(out))
        if rc:
            raise RuntimeError("orc_trace failed")
        return out


def trace(tris, rays, mode=LOWEST_ID, threads=0):
    return BVH2(tris).trace(rays, mode, threads)


def brute(tris, rays, threads=0, mode=LOWEST_ID):
    """Minimum t over all triangles (own-box slab and leaf test), lowest id; mode may carry
    WATERTIGHT."""
    tris, rays = _f32(tris, 9), _f32(rays, 6)
    out = np.zeros(len(rays), HIT_DTYPE)
    lib().orc_brute_mode(_p(tris), len(tris), _p(rays), len(rays), mode, threads, _p(out))
    return out


def slab(ray, bmin, bmax):
    r, lo, hi = _f32(ray, 6), _f32(bmin, 3), _f32(bmax, 3)
    return bool(lib().orc_slab(_p(r), _p(lo), _p(hi)))


def mt(tri, ray):
    """(hit, t, u, v) of the reference leaf test on one pair."""
    tr, r = _f32(tri, 9), _f32(ray, 6)
    t, u, v = ctypes.c_float(np.nan), ctypes.c_float(np.nan), ctypes.c_float(np.nan)
    hit = lib().orc_mt(_p(tr), _p(r), ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return bool(hit), t.value, u.value, v.value


def wt(tri, ray):
    """(hit, t, u, v) of the watertight leaf test on one pair."""
    tr, r = _f32(tri, 9), _f32(ray, 6)
    t, u, v = ctypes.c_float(np.nan), ctypes.c_float(np.nan), ctypes.c_float(np.nan)
    hit = lib().orc_wt(_p(tr), _p(r), ctypes.byref(t), ctypes.byref(u), ctypes.byref(v))
    return bool(hit), t.value, u.value, v.value


def fp64(tris, rays, ids):
    """Double-precision Moller-Trumbore for (ray i, triangle ids[i])."""
    tris, rays = _f32(tris, 9), _f32(rays, 6)
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    n = len(rays)
    t, u, v = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().orc_fp64(_p(tris), len(tris), _p(rays), _p(ids), n, _p(t), _p(u), _p(v))
    return t, u, v


def ref_trace(tris, rays):
    """The compiled reference CPU backend, all tags (FullHitReg records)."""
    tris, rays = _f32(tris, 9), _f32(rays, 6)
    out = np.zeros(len(rays), HIT_DTYPE)
    ref().ref_trace_full(_p(tris), len(tris), _p(rays), len(rays), _p(out))
    return out


def ref_time(tris, rays, reps=3):
    """Seconds of the compiled reference's nearest_hits<t, primitive_id> over the rays (its own
    CPU backend and thread split), best of reps calls after set_tris."""
    tris, rays = _f32(tris, 9), _f32(rays, 6)
    s = ref().ref_time_trace(_p(tris), len(tris), _p(rays), len(rays), int(reps))
    if s < 0:
        raise RuntimeError("reference trace returned a short record array")
    return s


def ref_layout(mask):
    out = np.zeros(9, np.int32)
    if ref().ref_hitreg_layout(mask, _p(out)):
        raise ValueError(mask)
    return out
