"""The keypoint map (KeyMap insert / collect / reset, Mapping/DeviceMap.cu:203-296,
Mapping/FuseMap.cu:366-472) and the adjacency matrix (Tracking/KeyPoints.cu:12-57) on the GPU
against the oracle, with keypoints supplied by the test (positions, normals, descriptors; no
detector): the map's slots, the inserted keys' positions and indices and the collected keys
bit-exact; the adjacency scores within 2e-6 (acosf and expf of the device's
math library against the host's)."""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.test_oracle_keymap import _keys

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("max_keys,nb,n", [(97, 2, 300), (100000, 5, 5000), (1, 3, 10), (1000, 5, 20000),
                                           (262147, 5, 5000)])
def test_keymap_insert_collect_reset(gpu, max_keys, nb, n):
    """(262147, 5): 1,310,735 slots, 1,281 groups of the collection, so its prefix pass gives
    lanes 0-639 runs of two group counts, lane 640 one and the rest none; the whole slot array
    (377 MB) is compared only for the smaller maps"""
    rng = np.random.default_rng(max_keys + n)
    whole = max_keys * nb <= 600_000
    g = A.KeyMap(max_keys=max_keys, n_buckets=nb)
    o = O.KeyMap(max_keys=max_keys, n_buckets=nb)
    for rep in range(3):  # repeated inserts: new keys, keys of occupied cells, full buckets
        keys = _keys(rng, n, 0.5 if rep else 2.0)
        keys[1::7]["pos"] = keys[0::7][: len(keys[1::7])]["pos"]  # keys sharing a cell
        idx = np.full(n, -3, np.int32)
        kd, id_ = A.surf_to_device(keys), torch.from_numpy(idx.copy()).cuda()
        g.InsertKeyPoints(kd, id_)
        ko, io = keys.copy(), idx.copy()
        o.insert(ko, io)
        assert np.array_equal(id_.cpu().numpy(), io), rep
        assert np.array_equal(A.surf_from_device(kd).view(np.uint8), ko.view(np.uint8)), rep
        if whole:
            assert np.array_equal(g.read().view(np.uint8), o.read().view(np.uint8)), rep
        cap = 3 * n
        out_g, ng = g.CollectKeyPoints(cap)
        out_o, no = o.collect(cap)
        assert ng == no
        m = min(ng, cap)
        assert np.array_equal(A.surf_from_device(out_g)[1:m].view(np.uint8), out_o[1:m].view(np.uint8))
    g.ResetKeyPoints()
    o.reset()
    if whole:
        assert np.array_equal(g.read().view(np.uint8), o.read().view(np.uint8))
    assert g.CollectKeyPoints(4)[1] == 0


def test_adjacency_matrix(gpu):
    rng = np.random.default_rng(5)
    for n in (1, 37, 300):
        f = _keys(rng, n, 1.0)
        m = f.copy()
        m["pos"] += rng.normal(0, 0.02, (n, 3)).astype(np.float32)
        if n > 10:
            f["pos"][3] = f["pos"][4] + np.float32(0.003)
            f["pos"][7, 2] = np.nan
            f["normal"][9, 1] = np.nan
        dist = rng.uniform(0, 3, n).astype(np.float32)
        got = A.BuildAdjacencyMatrix(A.surf_to_device(f), A.surf_to_device(m), torch.from_numpy(dist).cuda())
        want = O.adjacency(f, m, dist)
        got = got.cpu().numpy()
        assert np.array_equal(got == 0, want == 0)
        assert np.abs(got - want).max() <= 2e-6, np.abs(got - want).max()
