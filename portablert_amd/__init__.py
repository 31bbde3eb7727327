"""portableRT MI355X backend: Python view of the C ABI (include/prt_mi355x.h).

Mirrors the reference's operator interface (include/portableRT/backend.hpp,
nearest_hits_impl.hpp): ``available_backends()``, ``select_backend(b)``, ``b.set_tris(tris)``,
``nearest_hits(rays, *tags)``, ``b.device_name()``, ``b.is_available()``. Hit records come back
as numpy structured arrays with the exact byte layout of ``HitReg<Tags...>``
(hitreg.hpp:29-61). Everything runs in the gfx950 HIP library; if it is not built the import
of the native layer raises, there is no fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib", "libprt_mi355x.so")

# filter tags (hitreg.hpp:48-54) as field-mask bits
UV, T, PRIMITIVE_ID, P, VALID = 1, 2, 4, 8, 16
ALL_TAGS = (UV, T, PRIMITIVE_ID, P, VALID)

OK, ERR_ARG, ERR_NODEV, ERR_HIP, ERR_OOM, ERR_STATE, ERR_LAYOUT, ERR_INPUT = range(8)

OPT = dict(node_format=1, cull=2, leaf_size=3, block=4, waves_per_cu=5, lds_stack=6, chunk=7, leaf_phase=8,
           count_stats=9, time_kernel=10, alloc_limit=11, margin_rel_log2=12, builder=13, refill=14,
           collapse=15, leaf_test=16, seeds=17)
STAT = dict(build_ms=1, trace_ms=2, nodes=3, depth=4, node_visits=5, tri_tests=6, leaf_visits=7,
            node_bytes=8, api_ms=9, n_tris=10, node_phases=11, tri_phases=12, culled_pops=13, refit_ms=21, ncull_tris=22)

_lib = None


class PrtError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def lib():
    """The native library; raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built: run python portablert_amd/build.py")
        # PyTorch ships its own HIP runtime under the same SONAME (libamdhip64.so.7) but links it
        # by file name, so loading this library first would put two HIP runtimes in the process
        # and torch would then see no GPU. Loading torch first makes this library bind to the
        # runtime torch already loaded: one runtime per process.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB)
        vp, u64, ci, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, ctypes.c_uint32
        L.prt_mi355x_probe.argtypes = [ctypes.POINTER(ci), ctypes.c_char_p, ctypes.c_size_t]
        L.prt_mi355x_create.argtypes = [ci, ctypes.POINTER(vp)]
        L.prt_mi355x_destroy.argtypes = [vp]
        L.prt_mi355x_destroy.restype = None
        L.prt_mi355x_last_error.argtypes = [vp]
        L.prt_mi355x_last_error.restype = ctypes.c_char_p
        L.prt_mi355x_device_name.argtypes = [vp, ctypes.c_char_p, ctypes.c_size_t]
        L.prt_mi355x_set_tris.argtypes = [vp, vp, u64]
        L.prt_mi355x_set_tris_device.argtypes = [vp, vp, u64]
        L.prt_mi355x_refit.argtypes = [vp, vp, u64]
        L.prt_mi355x_refit_device.argtypes = [vp, vp, u64]
        L.prt_mi355x_nearest_hits.argtypes = [vp, vp, u64, u32, vp, u32, vp]
        L.prt_mi355x_nearest_hits_device.argtypes = [vp, vp, u64, u32, vp, u32, vp, vp]
        L.prt_mi355x_nearest_hits_device_noalloc.argtypes = [vp, vp, u64, u32, vp, u32, vp, vp]
        L.prt_mi355x_synchronize.argtypes = [vp]
        L.prt_mi355x_set_option.argtypes = [vp, ci, ctypes.c_int64]
        L.prt_mi355x_get_option.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_int64)]
        L.prt_mi355x_get_stat.argtypes = [vp, ci, ctypes.POINTER(ctypes.c_double)]
        L.prt_mi355x_layout.argtypes = [u32, vp]
        L.prt_mi355x_export.argtypes = [vp, ci, vp, ctypes.POINTER(u64)]
        _lib = L
    return _lib


def mask_of(tags):
    m = 0
    for t in tags:
        m |= int(t)
    return m


def layout(mask):
    """(size, offsets u v t id valid px py pz) of HitReg for a field mask."""
    out = np.zeros(9, np.int32)
    if lib().prt_mi355x_layout(mask, out.ctypes.data_as(ctypes.c_void_p)):
        raise ValueError(mask)
    return int(out[0]), [int(x) for x in out[1:]]


def hit_dtype(mask):
    """numpy dtype with the byte layout of HitReg for `mask` (absent fields omitted)."""
    size, off = layout(mask)
    names, formats, offsets = [], [], []
    for name, o, f in zip(("u", "v", "t", "primitive_id", "valid", "px", "py", "pz"), off,
                          ("<f4", "<f4", "<f4", "<u4", "u1", "<f4", "<f4", "<f4")):
        if o >= 0:
            names.append(name)
            formats.append(f)
            offsets.append(o)
    return np.dtype(dict(names=names, formats=formats, offsets=offsets, itemsize=size))


def probe():
    """(number of gfx950 devices, name of the first)."""
    n = ctypes.c_int(0)
    buf = ctypes.create_string_buffer(256)
    lib().prt_mi355x_probe(ctypes.byref(n), buf, 256)
    return n.value, buf.value.decode()


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if isinstance(a, np.ndarray) else ctypes.c_void_p(a)


class Context:
    """One device context of the C ABI: stream, BVH, staging buffers."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        rc = lib().prt_mi355x_create(device, ctypes.byref(h))
        if rc:
            raise PrtError(rc, f"no usable gfx950 device {device}")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().prt_mi355x_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: module globals may already be gone
            pass

    def _check(self, rc):
        if rc:
            raise PrtError(rc, lib().prt_mi355x_last_error(self.h).decode())

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        self._check(lib().prt_mi355x_device_name(self.h, buf, 256))
        return buf.value.decode()

    def set_option(self, name, value):
        self._check(lib().prt_mi355x_set_option(self.h, OPT[name], int(value)))

    def get_option(self, name):
        v = ctypes.c_int64()
        self._check(lib().prt_mi355x_get_option(self.h, OPT[name], ctypes.byref(v)))
        return v.value

    def stat(self, name):
        v = ctypes.c_double()
        self._check(lib().prt_mi355x_get_stat(self.h, STAT[name], ctypes.byref(v)))
        return v.value

    def set_tris(self, tris):
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self._check(lib().prt_mi355x_set_tris(self.h, _p(tris), len(tris)))

    def set_tris_device(self, ptr, n):
        self._check(lib().prt_mi355x_set_tris_device(self.h, ctypes.c_void_p(ptr), n))

    def refit(self, tris):
        """The same triangles at new positions: keeps the tree, recomputes its boxes."""
        tris = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        self._check(lib().prt_mi355x_refit(self.h, _p(tris), len(tris)))

    def refit_device(self, ptr, n):
        self._check(lib().prt_mi355x_refit_device(self.h, ctypes.c_void_p(ptr), n))

    def nearest_hits(self, rays, mask=mask_of(ALL_TAGS)):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        size, off = layout(mask)
        out = np.zeros(len(rays), hit_dtype(mask))
        offs = np.asarray(off, np.int32)
        self._check(lib().prt_mi355x_nearest_hits(self.h, _p(rays), len(rays), mask, _p(out), size, _p(offs)))
        return out

    def nearest_hits_device(self, rays_ptr, n, mask, out_ptr, stream=None):
        size, off = layout(mask)
        offs = np.asarray(off, np.int32)
        self._check(lib().prt_mi355x_nearest_hits_device(self.h, ctypes.c_void_p(rays_ptr), n, mask,
                                                         ctypes.c_void_p(out_ptr), size, _p(offs),
                                                         ctypes.c_void_p(stream or 0)))

    def nearest_hits_device_noalloc(self, rays_ptr, n, mask, out_ptr, stream=None):
        """nearest_hits_device without any allocation (a fixed device scratch array, shared by
        every call on the GPU)."""
        size, off = layout(mask)
        offs = np.asarray(off, np.int32)
        self._check(lib().prt_mi355x_nearest_hits_device_noalloc(self.h, ctypes.c_void_p(rays_ptr), n, mask,
                                                                 ctypes.c_void_p(out_ptr), size, _p(offs),
                                                                 ctypes.c_void_p(stream or 0)))

    def synchronize(self):
        self._check(lib().prt_mi355x_synchronize(self.h))

    def export(self, which):
        """Host copy of the device BVH: 'nodes_q', 'nodes_f' or 'tris' (raw bytes)."""
        w = {"nodes_q": 0, "nodes_f": 1, "tris": 2}[which]
        n = ctypes.c_uint64()
        self._check(lib().prt_mi355x_export(self.h, w, None, ctypes.byref(n)))
        out = np.zeros(n.value, np.uint8)
        if n.value:
            self._check(lib().prt_mi355x_export(self.h, w, _p(out), ctypes.byref(n)))
        return out


# ------------------------------------------------------------------ reference-shaped API

class Backend:
    """backend.hpp:12-29: name(), set_tris, is_available, init, shutdown, device_name."""

    def __init__(self, name):
        self._name = name

    def name(self):
        return self._name


class MI355XBackend(Backend):
    def __init__(self, device=0):
        super().__init__("MI355X")
        self.device = device
        self.ctx = None

    def is_available(self):
        try:
            n, _ = probe()
        except (OSError, RuntimeError):
            return False
        return n > self.device

    def init(self):
        if self.ctx is None:
            self.ctx = Context(self.device)

    def shutdown(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None

    def device_name(self):
        self.init()
        return self.ctx.device_name()

    def set_tris(self, tris):
        self.init()
        self.ctx.set_tris(tris)

    def nearest_hits(self, rays, *tags):
        self.init()
        return self.ctx.nearest_hits(rays, mask_of(tags if tags else ALL_TAGS))


_all = []
_available = []
_selected = None


def _register(b):
    """RegisterBackend (backend.hpp:85-95): first available backend becomes selected."""
    global _selected
    if b.is_available():
        if _selected is None:
            select_backend(b)
        _available.append(b)
    _all.append(b)


def all_backends():
    if not _all:
        _register(MI355XBackend())
    return list(_all)


def available_backends():
    all_backends()
    return list(_available)


def select_backend(b):
    """select_backend (src/backend.cpp:50-56): shutdown the old, init the new."""
    global _selected
    if _selected is not None:
        _selected.shutdown()
    _selected = b
    b.init()


def selected_backend():
    all_backends()
    return _selected


def nearest_hits(rays, *tags):
    """Free nearest_hits (nearest_hits_impl.hpp:28-36): dispatch to the selected backend;
    no tags means all tags (backend.hpp:77-79)."""
    b = selected_backend()
    if b is None:
        raise RuntimeError("Unknown backend")
    return b.nearest_hits(rays, *tags)


# ------------------------------------------------------------------ learned t_max seeding

MLP_IN, MLP_HIDDEN, CLUSTERS = 192, 256, 4096
MLP_OUT = CLUSTERS + 1
LAYER_SHAPES = [(MLP_HIDDEN, MLP_IN), (MLP_HIDDEN, MLP_HIDDEN), (MLP_HIDDEN, MLP_HIDDEN), (MLP_OUT, MLP_HIDDEN)]
STAT.update(mlp_ms=14, seeded_ms=15, train_ms=16, train_loss=17, clusters=18, seed_top1=19, seed_top4=20)


def _bind_learned(L):
    if getattr(L, "_learned_bound", False):
        return L
    vp, u64, ci = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    L.prt_mi355x_learned_setup.argtypes = [vp, u64]
    L.prt_mi355x_learned_train.argtypes = [vp, vp, u64, ci, ci, ctypes.c_float]
    L.prt_mi355x_nearest_hits_learned_device.argtypes = [vp, vp, u64, ctypes.c_uint32, vp, ctypes.c_uint32, vp, vp]
    L.prt_mi355x_learned_eval.argtypes = [vp, vp, u64, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.prt_mi355x_learned_clusters.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32), vp]
    L.prt_mi355x_mlp_params.argtypes = [vp, ci, ci, vp, vp]
    L.prt_mi355x_mlp_grad.argtypes = [vp, vp, u64, vp, ctypes.POINTER(ctypes.c_float), vp]
    L.prt_mi355x_mlp_grads.argtypes = [vp, ci, vp, vp]
    L.prt_mi355x_mlp_top4.argtypes = [vp, vp, u64, vp, vp]
    L._learned_bound = True
    return L


def _learned_lib():
    return _bind_learned(lib())


def learned_setup(ctx, seed=1):
    ctx._check(_learned_lib().prt_mi355x_learned_setup(ctx.h, seed))


def learned_train(ctx, rays_ptr, n, epochs=1, batch=8192, lr=1e-3):
    ctx._check(_learned_lib().prt_mi355x_learned_train(ctx.h, ctypes.c_void_p(rays_ptr), n, epochs, batch, lr))


def nearest_hits_learned_device(ctx, rays_ptr, n, mask, out_ptr, stream=None):
    size, off = layout(mask)
    offs = np.asarray(off, np.int32)
    ctx._check(_learned_lib().prt_mi355x_nearest_hits_learned_device(
        ctx.h, ctypes.c_void_p(rays_ptr), n, mask, ctypes.c_void_p(out_ptr), size, _p(offs), ctypes.c_void_p(stream or 0)))


def learned_eval(ctx, rays_ptr, n):
    t1, t4 = ctypes.c_double(), ctypes.c_double()
    ctx._check(_learned_lib().prt_mi355x_learned_eval(ctx.h, ctypes.c_void_p(rays_ptr), n, ctypes.byref(t1),
                                                      ctypes.byref(t4)))
    return t1.value, t4.value


def learned_clusters(ctx, n_tris):
    count = ctypes.c_uint32()
    tc = np.zeros(n_tris, np.uint32)
    ctx._check(_learned_lib().prt_mi355x_learned_clusters(ctx.h, ctypes.byref(count), _p(tc)))
    return count.value, tc


def mlp_params(ctx, layer):
    o, i = LAYER_SHAPES[layer]
    w, b = np.zeros((o, i), np.float32), np.zeros(o, np.float32)
    ctx._check(_learned_lib().prt_mi355x_mlp_params(ctx.h, layer, 1, _p(w), _p(b)))
    return w, b


def set_mlp_params(ctx, layer, w, b):
    w = np.ascontiguousarray(w, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    ctx._check(_learned_lib().prt_mi355x_mlp_params(ctx.h, layer, 0, _p(w), _p(b)))


def mlp_grad(ctx, rays, labels):
    """Loss, encoded inputs (uint16 bf16 bits [n][192]) and per-layer (gW^T, gb) of one batch."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    labels = np.ascontiguousarray(labels, np.uint32)
    n = len(rays)
    loss = ctypes.c_float()
    x0 = np.zeros((n, MLP_IN), np.uint16)
    ctx._check(_learned_lib().prt_mi355x_mlp_grad(ctx.h, _p(rays), n, _p(labels), ctypes.byref(loss), _p(x0)))
    grads = []
    for layer in range(4):
        o, i = LAYER_SHAPES[layer]
        gw, gb = np.zeros((o, i), np.float32), np.zeros(o, np.float32)
        ctx._check(_learned_lib().prt_mi355x_mlp_grads(ctx.h, layer, _p(gw), _p(gb)))
        grads.append((gw, gb))
    return loss.value, x0, grads


def mlp_top4(ctx, rays, with_logits=False):
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
    n = len(rays)
    top = np.zeros((n, 4), np.uint32)
    logits = np.zeros((n, MLP_OUT), np.float32) if with_logits else None
    ctx._check(_learned_lib().prt_mi355x_mlp_top4(ctx.h, _p(rays), n, _p(top),
                                                   _p(logits) if with_logits else None))
    return (top, logits) if with_logits else top


# ------------------------------------------------------------------ batched multi-scene tracing

BATCH_BUCKETS = 4  # scenes by triangle count: < 4096, < 16384, < 65536, more


def batch_bucket(tris):
    return 0 if tris < 4096 else 1 if tris < 16384 else 2 if tris < 65536 else 3


def _bind_batch(L):
    if getattr(L, "_batch_bound", False):
        return L
    vp, u32, ci = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    L.prt_mi355x_batch_setup.argtypes = [vp, u32, vp, vp, u32]
    L.prt_mi355x_batch_run.argtypes = [vp, vp, vp, vp, vp]
    L.prt_mi355x_batch_capture.argtypes = [vp, vp, vp, vp]
    L.prt_mi355x_batch_replay.argtypes = [vp, vp]
    L.prt_mi355x_batch_check.argtypes = [vp]
    L.prt_mi355x_batch_get_params.argtypes = [vp, ci, vp]
    L.prt_mi355x_batch_set_params.argtypes = [vp, ci, vp]
    L.prt_mi355x_batch_tune.argtypes = [vp, vp, vp, vp, ci, dp, dp]
    L.prt_mi355x_batch_tiles.argtypes = [u32, vp, vp, vp, vp, ctypes.POINTER(ctypes.c_uint64)]
    L._batch_bound = True
    return L


def _batch_lib():
    return _bind_batch(lib())


def batch_tiles(tri_counts, ray_counts, chunk=(64, 64, 64, 64)):
    """Work tiles [n][4] = (first ray lo, hi, count, scene) of the batch traversal (no device)."""
    tc = np.ascontiguousarray(tri_counts, np.uint64)
    rc = np.ascontiguousarray(ray_counts, np.uint64)
    ch = np.ascontiguousarray(chunk, np.int32)
    n = ctypes.c_uint64(0)
    L = _batch_lib()
    rc_ = L.prt_mi355x_batch_tiles(len(tc), _p(tc), _p(rc), _p(ch), None, ctypes.byref(n))
    if rc_:
        raise PrtError(rc_, "batch_tiles: bad arguments")
    out = np.zeros((n.value, 4), np.uint32)
    rc_ = L.prt_mi355x_batch_tiles(len(tc), _p(tc), _p(rc), _p(ch), _p(out), ctypes.byref(n))
    if rc_:
        raise PrtError(rc_, "batch_tiles: bad arguments")
    return out


class Batch:
    """Many scenes built and traced by one set of launches on a Context (device buffers)."""

    def __init__(self, ctx, tri_counts, ray_counts, mask):
        self.ctx = ctx
        self.tri_counts = np.ascontiguousarray(tri_counts, np.uint64)
        self.ray_counts = np.ascontiguousarray(ray_counts, np.uint64)
        self.mask = mask
        ctx._check(_batch_lib().prt_mi355x_batch_setup(ctx.h, len(self.tri_counts), _p(self.tri_counts),
                                                       _p(self.ray_counts), mask))

    def run(self, tris_ptr, rays_ptr, out_ptr, stream=None):
        self.ctx._check(_batch_lib().prt_mi355x_batch_run(self.ctx.h, ctypes.c_void_p(tris_ptr),
                                                          ctypes.c_void_p(rays_ptr), ctypes.c_void_p(out_ptr),
                                                          ctypes.c_void_p(stream or 0)))

    def capture(self, tris_ptr, rays_ptr, out_ptr):
        self.ctx._check(_batch_lib().prt_mi355x_batch_capture(self.ctx.h, ctypes.c_void_p(tris_ptr),
                                                              ctypes.c_void_p(rays_ptr), ctypes.c_void_p(out_ptr)))

    def replay(self, stream=None):
        self.ctx._check(_batch_lib().prt_mi355x_batch_replay(self.ctx.h, ctypes.c_void_p(stream or 0)))

    def check(self):
        self.ctx._check(_batch_lib().prt_mi355x_batch_check(self.ctx.h))

    def params(self, bucket):
        v = np.zeros(3, np.int32)
        rc = _batch_lib().prt_mi355x_batch_get_params(self.ctx.h, bucket, _p(v))
        if rc:
            raise PrtError(rc, "batch_get_params")
        return tuple(int(x) for x in v)

    def set_params(self, bucket, values):
        v = np.ascontiguousarray(values, np.int32)
        self.ctx._check(_batch_lib().prt_mi355x_batch_set_params(self.ctx.h, bucket, _p(v)))

    def tune(self, tris_ptr, rays_ptr, out_ptr, reps=3):
        before, after = ctypes.c_double(), ctypes.c_double()
        self.ctx._check(_batch_lib().prt_mi355x_batch_tune(self.ctx.h, ctypes.c_void_p(tris_ptr),
                                                           ctypes.c_void_p(rays_ptr), ctypes.c_void_p(out_ptr), reps,
                                                           ctypes.byref(before), ctypes.byref(after)))
        return before.value, after.value
