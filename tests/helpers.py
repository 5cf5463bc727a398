"""Shared helpers of the test suite (numpy/torch glue and mesh invariants)."""
import numpy as np


def ulp_diff(a, b):
    """Elementwise distance in units in the last place between float32 arrays."""
    a = np.ascontiguousarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.ascontiguousarray(b, np.float32).view(np.int32).astype(np.int64)
    # map the sign-magnitude encoding onto a monotonic integer line
    a = np.where(a < 0, -(a & 0x7FFFFFFF), a)
    b = np.where(b < 0, -(b & 0x7FFFFFFF), b)
    return np.abs(a - b)


def edge_use_counts(tris):
    """Histogram {uses: number of undirected edges} of a triangle list (T,3)."""
    t = np.asarray(tris, np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    key = e[:, 0] * (1 << 32) + e[:, 1]
    _, cnt = np.unique(key, return_counts=True)
    vals, num = np.unique(cnt, return_counts=True)
    return dict(zip(vals.tolist(), num.tolist()))


def euler_characteristic(nv, tris):
    t = np.asarray(tris, np.int64)
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    e.sort(axis=1)
    ne = np.unique(e[:, 0] * (1 << 32) + e[:, 1]).size
    used = np.unique(t).size
    return used - ne + len(t), used == nv


def oriented_consistently(tris):
    """Every directed edge appears at most once (consistent winding)."""
    t = np.asarray(tris, np.int64)
    d = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    key = d[:, 0] * (1 << 32) + d[:, 1]
    return np.unique(key).size == key.size


def to_np(t):
    return t.detach().cpu().numpy()
