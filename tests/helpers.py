"""Shared helpers of the tests: synthetic frames and the pyramid on both paths."""
import numpy as np

from fusion_amd.api import Scene

K640 = (525.0, 525.0, 319.5, 239.5)


def scene_frames(kind=0, size=(4.0, 4.0, 3.0), seed=0, frames=(0,), cols=640, rows=480, K=K640, noise=True):
    sc = Scene(kind, size, seed)
    scale = cols / 640.0
    Kc = (K[0] * scale, K[1] * scale, (cols - 1) / 2.0, (rows - 1) / 2.0) if cols != 640 else K
    out = [sc.render(f, cols, rows, Kc, 5000.0, noise) for f in frames]
    return sc, Kc, out


def f32(x):
    return float(np.float32(x))


def same_bits(a, b):
    """bit-identical float arrays (NaNs compared as NaN, any payload)"""
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a[~na].view(np.uint32), b[~nb].view(np.uint32))


def mismatch(a, b):
    a = np.asarray(a, dtype=np.float32)
    b = np.asarray(b, dtype=np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    bad = (na != nb) | (~na & ~nb & (a.view(np.uint32) != b.view(np.uint32)))
    return int(bad.sum()), bad.size


def rows_set(xyz):
    return set(map(tuple, np.asarray(xyz).reshape(-1, 3).tolist()))
