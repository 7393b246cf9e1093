"""The oracle's keypoint map (KeyMap, Mapping/DeviceMap.cu:203-296; ResetKeyPoints /
InsertKeyPoints / CollectKeyPoints, Mapping/FuseMap.cu:366-472) and adjacency matrix
(Tracking/KeyPoints.cu:12-57), pinned by independent numpy statements of the same rules:

* the bucket of a key: the signed remainder of the wrapped products' xor of its grid cell
  (the float quotients pos / 0.01 truncated), made non-negative;
* insertion one key at a time: a valid key of the same cell in the bucket gives its position
  and slot; otherwise the first free slot takes the key; a full bucket drops it;
* collection: the valid keys in slot order, the first of them not copied (offset 0);
* the adjacency matrix from its closed form in FP64 (acos, exp of the summed differences;
  4-D dots with w = 1 for the direction; the clamp turning NaN into -1; 0 for pairs closer
  than 1 cm and for NaN scores).
"""
import numpy as np

from oracle import binding as O


def _hash(cell, max_keys):
    x, y, z = (np.int64(c) for c in cell)
    h = (((x * 73856093) & 0xFFFFFFFF) ^ ((y * 19349669) & 0xFFFFFFFF) ^ ((z * 83492791) & 0xFFFFFFFF))
    h = int(np.uint32(h).view(np.int32))
    r = int(np.fmod(h, max_keys))
    return r + max_keys if r < 0 else r


def _cell(pos, grid=0.01):
    return tuple(int(np.trunc(np.float32(p) / np.float32(grid))) for p in pos)


def _keys(rng, n, spread=2.0):
    k = np.zeros(n, O.SURF)
    k["valid"] = 1
    k["pos"] = rng.uniform(-spread, spread, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    k["normal"][:, :3] = nrm
    k["normal"][:, 3] = rng.uniform(-1, 1, n)
    k["descriptor"] = rng.normal(size=(n, 64)).astype(np.float32)
    return k


def _insert_numpy(slots, keys, index, max_keys, nb):
    for i in range(len(keys)):
        c = _cell(keys[i]["pos"])
        b = _hash(c, max_keys)
        first, found = -1, -1
        for j in range(nb):
            s = b * nb + j
            if not slots[s]["valid"] and first == -1:
                first = s
            if slots[s]["valid"] and _cell(slots[s]["pos"]) == c:
                found = s
                break
        if found >= 0:
            index[i] = found
            keys[i]["pos"] = slots[found]["pos"]
        elif first >= 0:
            index[i] = first
            slots[first] = keys[i]


def test_insert_same_cell_full_bucket_and_collect():
    rng = np.random.default_rng(0)
    for max_keys, nb, n in ((97, 2, 300), (100000, 5, 2000), (1, 3, 10)):
        km = O.KeyMap(max_keys=max_keys, n_buckets=nb)
        keys = _keys(rng, n)
        keys[5]["pos"] = keys[3]["pos"] + np.float32(1e-4)  # mostly the same 1 cm cell as key 3
        keys[7]["pos"] = -np.abs(keys[7]["pos"])  # negative cell coordinates
        idx = np.full(n, -7, np.int32)
        k_or, i_or = keys.copy(), idx.copy()
        km.insert(k_or, i_or)
        slots = np.zeros(max_keys * nb, O.SURF)
        k_np, i_np = keys.copy(), idx.copy()
        _insert_numpy(slots, k_np, i_np, max_keys, nb)
        got = km.read()
        assert np.array_equal(got.view(np.uint8), slots.view(np.uint8)), (max_keys, nb)
        assert np.array_equal(i_or, i_np) and np.array_equal(k_or.view(np.uint8), k_np.view(np.uint8))
        if max_keys == 1:
            assert (i_or == -7).sum() == n - nb  # one bucket of nb slots: the rest dropped
        # a second insert of the same keys finds them all: nothing changes in the map
        again_k, again_i = keys.copy(), np.full(n, -7, np.int32)
        km.insert(again_k, again_i)
        assert np.array_equal(km.read().view(np.uint8), got.view(np.uint8))
        placed = i_or >= 0
        assert np.array_equal(again_i[placed], i_or[placed])
        # collection in slot order; offset 0 is not copied
        valid = got[got["valid"] == 1]
        out, cnt = km.collect(len(valid) + 4)
        assert cnt == len(valid)
        assert np.all(out[:1].view(np.uint8) == 0)
        assert np.array_equal(out[1:cnt].view(np.uint8), valid[1:].view(np.uint8))
        km.reset()
        assert km.read()["valid"].sum() == 0 and km.collect(4)[1] == 0


def _adjacency_numpy(f, m, dist):
    n = len(f)
    out = np.zeros((n, n))
    P = lambda k: k["pos"].astype(np.float64)
    N4 = lambda k: k["normal"].astype(np.float64)
    # clamp (KeyPoints.cu:6-10) as written: a NaN fails both comparisons and becomes -1
    c = lambda a: -1.0 if not a > -1.0 else min(a, 1.0)
    for y in range(n):
        for x in range(n):
            if x == y:
                out[y, x] = np.exp(-np.float64(dist[x]))
                continue
            df, dm = P(f[x]) - P(f[y]), P(m[x]) - P(m[y])
            d00, d01 = np.linalg.norm(df), np.linalg.norm(dm)
            if d00 <= 1e-2 or d01 <= 1e-2:
                continue
            e10 = np.append(df / d00, 1.0)
            e11 = np.append(dm / d01, 1.0)
            a0 = np.arccos(c(N4(f[x]) @ N4(f[y])))
            b0 = np.arccos(c(e10 @ N4(f[x])))
            g0 = np.arccos(c(e10 @ N4(f[y])))
            a1 = np.arccos(c(N4(m[x]) @ N4(m[y])))
            b1 = np.arccos(c(e11 @ N4(m[x])))
            g1 = np.arccos(c(e11 @ N4(m[y])))
            s = np.exp(-(abs(d00 - d01) + abs(a0 - a1) + abs(b0 - b1) + abs(g0 - g1)))
            out[y, x] = 0.0 if np.isnan(s) else s
    return out


def test_adjacency_closed_form():
    rng = np.random.default_rng(1)
    n = 60
    f = _keys(rng, n, 1.0)
    # the map keys: the frame keys moved rigidly, so consistent pairs score high
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = f.copy()
    m["pos"] = (f["pos"] @ R.T + [0.1, -0.2, 0.05]).astype(np.float32)
    m["normal"][:, :3] = (f["normal"][:, :3] @ R.T).astype(np.float32)
    m["pos"][10:20] += rng.normal(0, 0.2, (10, 3)).astype(np.float32)  # inconsistent pairs
    f["pos"][30] = f["pos"][31] + np.float32(0.004)  # a pair closer than 1 cm
    f["normal"][40, 0] = np.nan  # the clamp turns a NaN dot into -1
    f["pos"][50, 1] = np.nan  # NaN scores become 0
    dist = rng.uniform(0, 2, n).astype(np.float32)
    got = O.adjacency(f, m, dist)
    want = _adjacency_numpy(f, m, dist)
    assert np.abs(got - want).max() < 2e-5, np.abs(got - want).max()
    assert got[30, 31] == 0 and got[31, 30] == 0 and np.all(got[50, np.arange(n) != 50] == 0)
    assert np.all(got[40, :40] > 0)
    assert got[0, 1] > 0.5 and got[12, 1] < got[2, 1]
