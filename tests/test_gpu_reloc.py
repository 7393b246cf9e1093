"""Detector-free relocalisation on the GPU against the oracle: keypoint lifting
(Frame::ExtractKeyPoints' test and back-projection, Core/Frame.cc:93-200) bit-exact on a
rendered frame; the 2-NN matcher (the role of knnMatch, k = 2, L2, Tracking.cc:436) with
indices exact and distances bit-exact, across train chunks, ties and tiny inputs; and
Tracker::Relocalise's default path (Tracking.cc:411-467) recovering a camera pose from keypoints
supplied by the test (descriptors and positions; no detector)."""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.helpers import scene_frames
from tests.test_gpu_pyramid import _pyr_oracle

pytestmark = pytest.mark.gpu


def test_lift_keypoints_bit_exact(gpu):
    sc, K, fr = scene_frames(frames=(0,))
    _, d, v, n = _pyr_oracle(fr[0][0], K)
    depth, nmap = d[0], n[0]
    rng = np.random.default_rng(11)
    rows, cols = depth.shape
    xy = np.concatenate([rng.uniform(0, cols, (3000, 1)), rng.uniform(0, rows, (3000, 1))], 1)
    xy = np.concatenate([xy, [[1.0, 100.0], [478.9, 100.0], [479.0, 100.0], [600.0, 100.0], [300.0, 478.5],
                              [300.0, 479.0], [np.nan, 10.0], [12.0, np.inf]]]).astype(np.float32)
    p, nr, ok = A.lift_keypoints(torch.from_numpy(xy).cuda(), torch.from_numpy(depth).cuda(),
                                 torch.from_numpy(nmap).cuda(), K)
    po, no, oko = O.lift_keypoints(xy, depth, nmap, K)
    ok = ok.cpu().numpy()
    assert np.array_equal(ok, oko) and ok.sum() > 500
    assert np.array_equal(p.cpu().numpy().view(np.uint32)[ok == 1], po.view(np.uint32)[ok == 1])
    assert np.array_equal(nr.cpu().numpy().view(np.uint32)[ok == 1], no.view(np.uint32)[ok == 1])
    # the border cases: x <= 1, x >= rows - 1 (as written), y >= rows - 1, NaN, inf
    assert not ok[[-8, -6, -5, -3, -2, -1]].any()


@pytest.mark.parametrize("nq,nt", [(700, 20000), (5, 1), (3, 0), (300, 8192), (1, 8193)])
def test_match_knn2_matches_oracle(gpu, nq, nt):
    rng = np.random.default_rng(nq + nt)
    t = rng.normal(size=(max(nt, 1), 64)).astype(np.float32)[:nt]
    if nt > 8200:  # a duplicate in the next chunk: the lower index wins the tie
        t[8192 + 5] = t[5]
    src = rng.integers(0, max(nt, 1), nq)
    q = (t[src] + rng.normal(0, 0.05, (nq, 64))).astype(np.float32) if nt else rng.normal(size=(nq, 64)).astype(
        np.float32)
    if nt > 8200:
        q[0] = t[5]
    idx, dist = A.match_knn2(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda() if nt else None)
    io, do = O.match_knn2(q, t)
    assert np.array_equal(idx.cpu().numpy(), io)
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), do.view(np.uint32))
    if nt > 8200:
        assert io[0, 0] == 5 and io[0, 1] == 8197 and do[0, 0] == 0


def test_relocalise_recovers_the_pose(gpu):
    rng = np.random.default_rng(12)
    nk = 2000
    keys = np.zeros(nk, A.SURF_DTYPE)
    keys["valid"] = 1
    keys["pos"] = rng.uniform([-2, -2, 0.5], [2, 2, 4.0], (nk, 3)).astype(np.float32)
    nrm = rng.normal(size=(nk, 3))
    keys["normal"][:, :3] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    keys["descriptor"] = rng.normal(size=(nk, 64)).astype(np.float32)
    km = A.KeyMap(max_keys=20011, n_buckets=5)
    kd = A.surf_to_device(keys)
    index = torch.full((nk,), -1, dtype=torch.int32, device="cuda")
    km.InsertKeyPoints(kd, index)
    stored = A.surf_from_device(kd)  # positions as the map holds them (same-cell keys merged)
    assert (index.cpu().numpy() >= 0).all()
    # a camera at a known pose sees 600 of the keys (1 mm noise, perturbed descriptors) and
    # 250 keypoints with no counterpart in the map
    th = 0.4
    R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    t = np.array([0.3, -0.2, 0.1])
    pose = np.eye(4)
    pose[:3, :3], pose[:3, 3] = R, t
    seen = rng.choice(nk, 600, replace=False)
    world = stored["pos"][seen].astype(np.float64)
    cam = (world - t) @ R + rng.normal(0, 0.001, (600, 3))
    desc = stored["descriptor"][seen] + rng.normal(0, 0.05, (600, 64)).astype(np.float32)
    cam = np.concatenate([cam, rng.uniform([-1, -1, 0.5], [1, 1, 3], (250, 3))])
    desc = np.concatenate([desc, rng.normal(size=(250, 64)).astype(np.float32)])
    got, npairs = km.relocalise(torch.from_numpy(desc).cuda(), cam, iterations=100, seed=7)
    assert got is not None and npairs >= 600
    assert np.abs(got[:3, 3] - t).max() < 3e-3
    assert np.abs(got[:3, :3] - R).max() < 3e-3
    # the pairs that pass the ratio test are the ones the oracle's matcher gives
    out, n = km.CollectKeyPoints(nk + 8)
    rec = A.surf_from_device(out)[:n]
    rec[0] = np.zeros(1, A.SURF_DTYPE)  # CollectKeyPoints never writes the first record
    io, do = O.match_knn2(desc, rec["descriptor"])
    assert npairs == int((do[:, 0].astype(np.float64) < 0.9 * do[:, 1].astype(np.float64)).sum())
    # too few keypoints: refused
    assert km.relocalise(torch.from_numpy(desc[:2]).cuda(), cam[:2])[0] is None
    km.close()
