"""CPU tests of the map-snapshot checks (fu_map_check_file, the host pass of fu_map_load).

A small map image is written in the snapshot format (magic, version, the map parameters, the
16 counters, block bytes, allocated and used counts, the used entries as (index, key, ptr,
next), the block keys, the pool, an FNV-1a-64 checksum), with a hash table of 8 buckets so
that excess chains form.  The valid file passes; each structural corruption, written with a
correct checksum so that only the structure checks can catch it, is refused, as are a wrong
checksum, a truncated file and other map parameters.  No GPU is used.
"""
import ctypes as C
import struct

import numpy as np
import pytest

from fusion_amd import native as N

NB, NX, NK = 8, 16, 32
VOX, TRUNC = 0.01, float(np.float32(0.01) * np.float32(3.0))
BB = 512 * 8  # f32 sdf + u8 weight + rgb
C_ALLOC, C_EXCESS = 0, 1


def _hash(k):
    x, y, z = k
    return ((((x * 73856093) & 0xFFFFFFFF) ^ ((y * 19349669) & 0xFFFFFFFF) ^ ((z * 83492791) & 0xFFFFFFFF))
            & 0xFFFFFFFF) % NB


def _pack(k):
    return sum(((v & 0x1FFFFF) << s) for v, s in zip(k, (42, 21, 0)))


def _fnv(data):
    h = 1469598103934665603
    for b in data:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def _image(n=20, seed=0):
    """entries (index -> [key, ptr, next]), block keys, counters of a map of n blocks"""
    rng = np.random.default_rng(seed)
    coords = set()
    while len(coords) < n:
        coords.add(tuple(int(v) for v in rng.integers(-50, 50, 3)))
    ent, keys, excess = {}, [], 0
    for j, c in enumerate(sorted(coords)):
        key = _pack(c)
        b = _hash(c)
        keys.append(key)
        if b not in ent:
            ent[b] = [key, j * 512, 0]
            continue
        i = b
        while ent[i][2]:
            i = NB + ent[i][2] - 1
        ent[NB + excess] = [key, j * 512, 0]
        ent[i][2] = excess + 1
        excess += 1
    ctr = [0] * 16
    ctr[C_ALLOC], ctr[C_EXCESS] = n, excess
    assert excess >= 4  # chains of several entries
    return ent, keys, ctr


def _write(path, ent, keys, ctr, n_alloc=None, checksum_ok=True, truncate=0, params=None):
    nb, nx, nk = params or (NB, NX, NK)
    n_alloc = len(keys) if n_alloc is None else n_alloc
    b = bytearray(b"FUSNMAP\0")
    b += struct.pack("<IIIIII", 1, nb, nx, nk, 0, 255)
    b += struct.pack("<ffff", VOX, TRUNC, 0.1, 3.0)
    b += struct.pack("<16I", *ctr)
    b += struct.pack("<III", BB, n_alloc, len(ent))
    for i in sorted(ent):
        key, ptr, nxt = ent[i]
        b += struct.pack("<IQii", i, key, ptr, nxt)
    b += struct.pack("<%dQ" % len(keys), *keys)
    b += np.random.default_rng(1).integers(0, 256, n_alloc * BB, dtype=np.uint8).tobytes()
    h = _fnv(bytes(b)) ^ (0 if checksum_ok else 1)
    b += struct.pack("<Q", h)
    with open(path, "wb") as f:
        f.write(bytes(b[:len(b) - truncate]))


def _check(path, params=None):
    nb, nx, nk = params or (NB, NX, NK)
    p = N.MapParams()
    N.lib().fu_map_params_default(C.byref(p))
    p.num_buckets, p.num_excess, p.num_blocks = nb, nx, nk
    p.voxel_size, p.trunc_dist, p.depth_min, p.depth_max = VOX, TRUNC, 0.1, 3.0
    p.voxel_format, p.max_weight = 0, 255
    st = N.lib().fu_map_check_file(C.byref(p), str(path).encode())
    return st, N.lib().fu_last_error().decode()


def _chain(ent):
    """a bucket whose chain has at least two excess entries: [head index, excess indices...]"""
    for b in range(NB):
        if b in ent:
            c, i = [b], b
            while ent[i][2]:
                i = NB + ent[i][2] - 1
                c.append(i)
            if len(c) >= 3:
                return c
    raise AssertionError("no chain of three")


def test_valid_snapshot_passes(tmp_path):
    ent, keys, ctr = _image()
    _write(tmp_path / "m", ent, keys, ctr)
    assert _check(tmp_path / "m")[0] == 0


def _mutations():
    def ptr_range(ent, keys, ctr):
        ent[0 if 0 in ent else min(ent)][1] = len(keys) * 512

    def ptr_align(ent, keys, ctr):
        ent[min(ent)][1] += 8

    def next_range(ent, keys, ctr):
        i = max(i for i in ent if i >= NB)
        ent[i][2] = ctr[C_EXCESS] + 1

    def cycle(ent, keys, ctr):
        c = _chain(ent)
        ent[c[-1]][2] = c[1] - NB + 1

    def shared(ent, keys, ctr):
        c = _chain(ent)
        other = next(b for b in range(NB) if b in ent and b != c[0] and ent[b][2] == 0)
        ent[other][2] = c[2] - NB + 1

    def wrong_bucket(ent, keys, ctr):
        heads = [b for b in range(NB) if b in ent]
        a, b = heads[0], heads[1]
        ent[a][0], ent[b][0] = ent[b][0], ent[a][0]
        ja, jb = ent[a][1] // 512, ent[b][1] // 512
        keys[ja], keys[jb] = keys[jb], keys[ja]

    def key_mismatch(ent, keys, ctr):
        keys[3] = _pack((999, 0, 0))

    def duplicate(ent, keys, ctr):
        c = _chain(ent)
        key = ent[c[0]][0]
        ent[c[1]][0] = key
        keys[ent[c[1]][1] // 512] = key

    def excess_counter(ent, keys, ctr):
        ctr[C_EXCESS] -= 2

    def alloc_counter(ent, keys, ctr):
        ctr[C_ALLOC] += 1

    def orphan(ent, keys, ctr):
        i = max(i for i in ent if i >= NB)
        ent[i + 1] = [_pack((700, 1, 2)), -1, 0]
        ctr[C_EXCESS] += 1

    return [ptr_range, ptr_align, next_range, cycle, shared, wrong_bucket, key_mismatch, duplicate, excess_counter,
            alloc_counter, orphan]


@pytest.mark.parametrize("mut", _mutations(), ids=lambda f: f.__name__)
def test_structural_corruption_refused(tmp_path, mut):
    ent, keys, ctr = _image()
    mut(ent, keys, ctr)
    _write(tmp_path / "m", ent, keys, ctr)
    st, msg = _check(tmp_path / "m")
    assert st != 0 and "map snapshot" in msg, msg


def test_checksum_truncation_and_parameters_refused(tmp_path):
    ent, keys, ctr = _image()
    _write(tmp_path / "a", ent, keys, ctr, checksum_ok=False)
    assert "checksum" in _check(tmp_path / "a")[1]
    _write(tmp_path / "b", ent, keys, ctr, truncate=100)
    assert _check(tmp_path / "b")[0] != 0
    _write(tmp_path / "c", ent, keys, ctr)
    st, msg = _check(tmp_path / "c", params=(NB, NX, NK * 2))
    assert st != 0 and "other parameters" in msg
