"""Comparison helpers shared by the tests (bitwise, field by field)."""
import numpy as np

FIELDS = ("t", "u", "v", "primitive_id", "valid", "px", "py", "pz")


def bits(a):
    a = np.ascontiguousarray(a)
    if a.dtype == np.float32:
        return a.view(np.uint32)
    return a.astype(np.int64)


def as_full(raw):
    """uint8[n, 32] FullHitReg bytes -> structured array."""
    import oracle
    return np.ascontiguousarray(raw).view(oracle.HIT_DTYPE).reshape(-1)


def diff_records(got, want, fields=FIELDS, ids=True):
    """Dict field -> number of rays whose bits differ. u, v and primitive_id are compared only
    where want is a hit (the reference leaves them indeterminate on a miss). With ids=False
    (comparing against the reference's depth-first tie order) u and v are compared only where
    both picked the same triangle; t and p must match everywhere."""
    want_valid = np.asarray(want["t"]) < np.inf
    same_tri = np.ones(len(want), bool)
    if not ids and "primitive_id" in got.dtype.names and "primitive_id" in want.dtype.names:
        same_tri = np.asarray(got["primitive_id"]) == np.asarray(want["primitive_id"])
    out = {}
    for f in fields:
        if f not in got.dtype.names or f not in want.dtype.names:
            continue
        if f == "primitive_id" and not ids:
            continue
        g, w = bits(got[f]), bits(want[f])
        if f == "valid":
            g, w = g != 0, w != 0
        bad = g != w
        if f in ("u", "v", "primitive_id"):
            bad &= want_valid & same_tri
        n = int(bad.sum())
        if n:
            out[f] = n
    return out


def assert_same(got, want, fields=FIELDS, ids=True, what=""):
    d = diff_records(got, want, fields, ids)
    if d:
        bad = np.zeros(len(want), bool)
        for f in d:
            g, w = bits(got[f]), bits(want[f])
            bad |= (g != 0) != (w != 0) if f == "valid" else g != w
        i = int(np.flatnonzero(bad)[0])
        raise AssertionError(f"{what}: differing fields {d}; first ray {i}: got {got[i]} want {want[i]}")
