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


@pytest.mark.parametrize("nq,nt", [(700, 20000), (5, 1), (3, 0), (300, 8192), (1, 8193), (4000, 40000)])
def test_match_knn2_matches_oracle(gpu, nq, nt):
    """the train set is cut into chunks whose length depends on the sizes (128 records for
    (700, 20000) and (1, 8193), 256 for (4000, 40000)); ties are planted across chunks, and
    across chunks that different lanes of the merge hold (chunk c goes to lane c % 64)"""
    rng = np.random.default_rng(nq + nt)
    t = rng.normal(size=(max(nt, 1), 64)).astype(np.float32)[:nt]
    if nt > 8200:  # duplicates later in the train set: the lower index wins the tie
        t[8192 + 5] = t[5]
        t[300] = t[9001] = t[7]
    src = rng.integers(0, max(nt, 1), nq)
    q = (t[src] + rng.normal(0, 0.05, (nq, 64))).astype(np.float32) if nt else rng.normal(size=(nq, 64)).astype(
        np.float32)
    if nt > 8200:
        q[0] = t[5]
        if nq > 1:
            q[1] = t[7]
    idx, dist = A.match_knn2(torch.from_numpy(q).cuda(), torch.from_numpy(t).cuda() if nt else None)
    io, do = O.match_knn2(q, t)
    assert np.array_equal(idx.cpu().numpy(), io)
    assert np.array_equal(dist.cpu().numpy().view(np.uint32), do.view(np.uint32))
    if nt > 8200:
        assert io[0, 0] == 5 and io[0, 1] == 8197 and do[0, 0] == 0
        if nq > 1:
            assert io[1, 0] == 7 and io[1, 1] == 300 and do[1, 1] == 0


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


# ------------------------------------------------------------------ graph matching
def _sel_matrix(rng, n, exclusive=0.05):
    C = rng.random((n, n)).astype(np.float32)
    C[rng.random((n, n)) < exclusive] = np.float32(0.002)
    return C


def _same_selection(C, n_lists, max_sel, score, offset=0):
    """offset: the matrix starts this many floats into its allocation (a view not 16-byte aligned)"""
    n = C.shape[0]
    dev = torch.zeros(n * n + offset, dtype=torch.float32, device="cuda")
    dev[offset:] = torch.from_numpy(C.ravel()).cuda()
    g = A.select_subgraphs(dev[offset:].view(n, n), n_lists, max_sel, score)
    o = O.graph_select(C, n_lists, max_sel, score)
    assert np.array_equal(g[2].view(np.uint32), o[2].view(np.uint32))  # rank sums, bit for bit
    assert np.array_equal(g[3], o[3])
    assert np.array_equal(g[0], o[0])
    for a, b in zip(g[1], o[1]):
        assert np.array_equal(a, b)
    return o


def test_select_subgraphs_matches_oracle(gpu):
    from tests.test_oracle_graph import _hand_matrix
    o = _same_selection(_hand_matrix(), 10, 400, 0.1)
    assert list(o[1][0]) == [2, 3, 4, 5]
    rng = np.random.default_rng(21)
    for n, n_lists, max_sel, score in ((1, 10, 400, 0.1), (2, 10, 400, 0.1), (3, 64, 400, 0.1), (50, 10, 400, 0.1),
                                       (700, 10, 400, 0.1), (700, 12, 1024, 0.05), (700, 10, 1, 0.1)):
        print("select", n, n_lists, max_sel, score, flush=True)
        _same_selection(_sel_matrix(rng, n), n_lists, max_sel, score)
    _same_selection(_sel_matrix(rng, 700), 10, 400, 0.1, offset=1)
    # few qualify: the scan runs over several 1024-pair chunks; duplicate columns tie; a NaN column
    C = _sel_matrix(rng, 5000)
    C[:, 17] = C[:, 3]
    C[40, 99] = np.nan
    print("select 5000", flush=True)
    o = _same_selection(C, 10, 400, 0.97)
    assert o[3][-1] == 99 and 400 > o[0][0] > 100
    # no pairs: every list empty
    z = A.select_subgraphs(torch.zeros((0, 0), device="cuda"), 10, 400, 0.1)
    assert not z[0].any()


def _graph_keys(rng, nk=1500):
    keys = np.zeros(nk, A.SURF_DTYPE)
    keys["valid"] = 1
    keys["pos"] = rng.uniform([-2, -2, 0.5], [2, 2, 4.0], (nk, 3)).astype(np.float32)
    nrm = rng.normal(size=(nk, 3))
    keys["normal"][:, :3] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    keys["descriptor"] = rng.normal(size=(nk, 64)).astype(np.float32)
    # 100 decoys: another place with nearly the same descriptor as key i (the ratio test fails)
    keys["pos"][nk - 100:] = rng.uniform([-2, -2, 0.5], [2, 2, 4.0], (100, 3)).astype(np.float32)
    keys["descriptor"][nk - 100:] = keys["descriptor"][:100] + rng.normal(0, 0.02, (100, 64)).astype(np.float32)
    return keys


def test_graph_candidates_match_the_oracle(gpu):
    rng = np.random.default_rng(31)
    keys = _graph_keys(rng)
    km = A.KeyMap(max_keys=20011, n_buckets=5)
    kd = A.surf_to_device(keys)
    km.InsertKeyPoints(kd, torch.full((len(keys),), -1, dtype=torch.int32, device="cuda"))
    out, nk = km.CollectKeyPoints(len(keys) + 8)
    rec = A.surf_from_device(out)[:nk]
    rec[0] = np.zeros(1, A.SURF_DTYPE)  # never written by CollectKeyPoints
    th = 0.5
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    t = np.array([-0.2, 0.1, 0.3])
    seen = rng.choice(np.arange(1, 700), 400, replace=False)
    seen = np.concatenate([seen, np.arange(1, 60)])  # includes keys that have decoys
    world = rec["pos"][seen].astype(np.float64)
    cam = ((world - t) @ R + rng.normal(0, 0.001, (len(seen), 3))).astype(np.float32)
    ncam = (rec["normal"][seen][:, :3].astype(np.float64) @ R).astype(np.float32)
    nrm = np.concatenate([ncam, np.zeros((len(seen), 1), np.float32)], 1)
    desc = (rec["descriptor"][seen] + rng.normal(0, 0.05, (len(seen), 64))).astype(np.float32)
    dd = torch.from_numpy(desc).cuda()
    for seed in (3, 4):
        poses, npairs = km.graph_candidates(dd, cam, nrm, iterations=100, seed=seed)
        # the same pairs from the oracle's list rules over the GPU's matches and matrix
        idx, dist = A.match_knn2(dd, torch.from_numpy(np.ascontiguousarray(rec["descriptor"])).cuda())
        fk, mk, d, q, nm = O.graph_inputs(idx.cpu().numpy(), dist.cpu().numpy(), rec, cam, nrm)
        assert npairs == nm and nm > len(seen)  # some queries kept both neighbours
        Cg = A.BuildAdjacencyMatrix(A.surf_to_device(fk), A.surf_to_device(mk), torch.from_numpy(d).cuda())
        nsel, lists, _, _ = O.graph_select(Cg.cpu().numpy(), 10, 400, 0.1)
        M, pairs = O.graph_lists(nsel, lists, 400, 5)
        assert M >= 1 and len(pairs) > 50
        want = []
        for c in range(M):
            T, _ = A.pose_estimate(fk["pos"][pairs].astype(np.float64), mk["pos"][pairs].astype(np.float64),
                                   iterations=100, seed=seed + c)
            if T is not None:
                want.append(np.linalg.inv(T))
        assert len(poses) == len(want)
        for g, w in zip(poses, want):
            assert np.abs(g - w).max() < 1e-9
            assert np.abs(g[:3, 3] - t).max() < 3e-3 and np.abs(g[:3, :3] - R).max() < 3e-3
    # too few matches: no candidates
    assert len(km.graph_candidates(dd[:2], cam[:2], nrm[:2])[0]) == 0
    km.close()


def _lost_gpu(K, fr, empty, mp):
    s = A.System(A.system_params(cols=320, rows=240, K=K, map=mp))
    st = []
    for f in list(fr[:6]) + [empty] * 10 + [fr[6]]:
        s.GrabImage(*A.to_device(*f))
        st.append(s.result()["track_status"])
    assert st == [4] + [1] * 5 + [2] * 10 + [5]
    return s


def test_validate_poses_match_the_oracle(gpu):
    from tests.test_oracle_graph import _lost_oracle, _perturb
    from tests.helpers import scene_frames
    sc, K, fr = scene_frames(frames=range(0, 9), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    mp = dict(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
              trunc_dist=float(np.float32(0.01) * np.float32(3.0)), max_new_blocks=1 << 16)
    ref = A.System(A.system_params(cols=320, rows=240, K=K, map=mp))
    for f in fr[:7]:
        ref.GrabImage(*A.to_device(*f))
    truth = ref.result()["pose"]
    ref.close()
    far = np.eye(4)
    far[:3, 3] = [0, 0, 40.0]
    cands = [far, _perturb(truth, [0.25, -0.2, 0.15], 4.0, [1, 2, 0]), _perturb(truth, [0.012, -0.01, 0.01], 1.5, [0, 1, 1]),
             _perturb(truth, [-0.01, 0.0, 0.01], 1.0, [1, 0, 0])]
    s = _lost_gpu(K, fr, empty, mp)
    o = _lost_oracle(K, fr, empty)
    ok, acc, chosen = s.validate_poses(cands)
    ook, oacc, ochosen, opose = o.validate_poses(cands)
    assert ok == bool(ook) and list(acc) == list(oacc) == [-1, 0, 1, 1] and chosen == ochosen == 2
    r = s.result()
    assert r["track_status"] == 6 and r["fused"] == 0
    assert np.abs(r["pose"] - opose).max() < 5e-4
    assert np.abs(r["pose"][:3, 3] - truth[:3, 3]).max() < 1e-3
    for f in fr[7:9]:
        s.GrabImage(*A.to_device(*f))
        st, op = o.grab(*f)
        g = s.result()
        assert g["track_status"] == st == 1 and g["fused"] == 0
        assert np.abs(g["pose"] - op).max() < 2e-3
    with pytest.raises(A.N.FusionError):  # no longer lost
        s.validate_poses(cands)
    s.close()


def test_validate_poses_twice_on_one_system(gpu):
    """ValidatePose keeps its ICP tracker from one call to the next: an attempt whose every
    candidate fails (one skipped, one rejected by ICP) and then one that succeeds give, on the
    second call, the accepted flags, choice and relocalised pose of a fresh lost system, bit
    for bit, and the next frame tracks alike"""
    from tests.test_oracle_graph import _perturb
    sc, K, fr = scene_frames(frames=range(0, 8), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    mp = dict(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
              trunc_dist=float(np.float32(0.01) * np.float32(3.0)), max_new_blocks=1 << 16)
    ref = A.System(A.system_params(cols=320, rows=240, K=K, map=mp))
    for f in fr[:7]:
        ref.GrabImage(*A.to_device(*f))
    truth = ref.result()["pose"]
    ref.close()
    far = np.eye(4)
    far[:3, 3] = [0, 0, 40.0]
    bad = [far, _perturb(truth, [0.25, -0.2, 0.15], 4.0, [1, 2, 0])]
    good = [_perturb(truth, [0.012, -0.01, 0.01], 1.5, [0, 1, 1]), _perturb(truth, [-0.01, 0.0, 0.01], 1.0, [1, 0, 0])]
    s = _lost_gpu(K, fr, empty, mp)
    ok, acc, chosen = s.validate_poses(bad)
    assert not ok and list(acc) == [-1, 0] and chosen == -1
    assert s.result()["track_status"] == 5  # still lost
    ok, acc, chosen = s.validate_poses(good)
    fresh = _lost_gpu(K, fr, empty, mp)
    ok2, acc2, chosen2 = fresh.validate_poses(good)
    assert ok and ok2 and list(acc) == list(acc2) == [1, 1] and chosen == chosen2 == 0
    r, r2 = s.result(), fresh.result()
    assert r["track_status"] == r2["track_status"] == 6
    assert np.array_equal(r["pose"], r2["pose"])
    s.GrabImage(*A.to_device(*fr[7]))
    fresh.GrabImage(*A.to_device(*fr[7]))
    g, g2 = s.result(), fresh.result()
    assert g["track_status"] == g2["track_status"] == 1 and np.array_equal(g["pose"], g2["pose"])
    s.close()
    fresh.close()


def test_relocalise_graph_end_to_end(gpu):
    """a lost system relocalised from a key map built on an earlier frame: keypoints lifted
    from frame 3, put in the map at frame 3's pose; frame 6's keypoints observe the same points
    (projected with frame 6's tracked pose) with perturbed descriptors"""
    from tests.helpers import scene_frames
    sc, K, fr = scene_frames(frames=range(0, 9), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    mp = dict(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
              trunc_dist=float(np.float32(0.01) * np.float32(3.0)), max_new_blocks=1 << 16)
    ref = A.System(A.system_params(cols=320, rows=240, K=K, map=mp))
    poses = []
    for f in fr[:7]:
        ref.GrabImage(*A.to_device(*f))
        poses.append(ref.result()["pose"])
    ref.close()
    rng = np.random.default_rng(41)
    _, d3, _, n3 = _pyr_oracle(fr[3][0], K)
    # the bounds test as written refuses x >= rows - 1, most of a landscape frame's right half
    xy = np.stack([rng.uniform(4, 316, 3000), rng.uniform(4, 236, 3000)], 1).astype(np.float32)
    p3, nr3, ok3 = O.lift_keypoints(xy, d3[0], n3[0], K)
    p3, nr3 = p3[ok3 == 1].astype(np.float64), nr3[ok3 == 1].astype(np.float64)
    P3 = poses[3]
    keys = np.zeros(len(p3), A.SURF_DTYPE)
    keys["valid"] = 1
    keys["pos"] = (p3 @ P3[:3, :3].T + P3[:3, 3]).astype(np.float32)
    keys["normal"][:, :3] = (nr3[:, :3] @ P3[:3, :3].T).astype(np.float32)
    keys["normal"][:, 3] = nr3[:, 3]  # the 4-D dots of the consistency scores see w as well
    keys["descriptor"] = rng.normal(size=(len(p3), 64)).astype(np.float32)
    km = A.KeyMap(max_keys=20011, n_buckets=5)
    km.InsertKeyPoints(A.surf_to_device(keys), torch.full((len(keys),), -1, dtype=torch.int32, device="cuda"))
    # frame 6 sees the same world points where they project inside its image
    P6 = poses[6]
    c6 = (keys["pos"].astype(np.float64) - P6[:3, 3]) @ P6[:3, :3]
    u = c6[:, 0] / c6[:, 2] * K[0] + K[2]
    v = c6[:, 1] / c6[:, 2] * K[1] + K[3]
    inside = (c6[:, 2] > 0.3) & (u > 3) & (u < 316) & (v > 3) & (v < 236)
    _, d6, _, n6 = _pyr_oracle(fr[6][0], K)
    xy6 = np.stack([u[inside], v[inside]], 1).astype(np.float32)
    p6, nr6, ok6 = O.lift_keypoints(xy6, d6[0], n6[0], K)
    sel = ok6 == 1
    desc = (keys["descriptor"][inside][sel] + rng.normal(0, 0.05, (sel.sum(), 64))).astype(np.float32)
    assert sel.sum() > 300
    s = _lost_gpu(K, fr, empty, mp)
    ok, nc = s.relocalise_graph(km, torch.from_numpy(desc).cuda(), p6[sel], nr6[sel], iterations=100, seed=5)
    assert ok and nc >= 1
    r = s.result()
    assert r["track_status"] == 6
    assert np.abs(r["pose"][:3, 3] - P6[:3, 3]).max() < 2e-3 and np.abs(r["pose"][:3, :3] - P6[:3, :3]).max() < 2e-3
    for f in fr[7:9]:
        s.GrabImage(*A.to_device(*f))
        assert s.result()["track_status"] == 1
    s.close()
    km.close()


def test_validate_poses_with_the_photometric_term_and_in_a_batch(gpu):
    """ValidatePose runs ICP only (Tracking.cc:657, 681) also for a system that tracks with the
    photometric term, whose intensity pyramid is swapped in with the relocalised frame: GPU =
    oracle in choice and pose, and RGB+ICP tracking resumes.  Two systems driven through one
    batch and relocalised alike continue bit for bit alike in the batch."""
    from tests.test_oracle_graph import _perturb
    from tests.helpers import scene_frames
    sc, K, fr = scene_frames(frames=range(0, 9), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    mp = dict(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
              trunc_dist=float(np.float32(0.01) * np.float32(3.0)), max_new_blocks=1 << 16)
    seq = list(fr[:6]) + [empty] * 10 + [fr[6]]
    want = [4] + [1] * 5 + [2] * 10 + [5]
    # photometric term
    s = A.System(A.system_params(cols=320, rows=240, K=K, map=mp, use_rgb=1))
    o = O.System(cols=320, rows=240, K=K, num_buckets=mp["num_buckets"], num_excess=mp["num_excess"],
                 num_blocks=mp["num_blocks"], voxel_size=0.01, use_rgb=True)
    st, ost, poses = [], [], []
    for f in seq:
        s.GrabImage(*A.to_device(*f))
        r = s.result()
        st.append(r["track_status"])
        poses.append(r["pose"])
        ost.append(o.grab(*f)[0])
    assert st == ost == want
    near = _perturb(poses[5], [0.012, -0.01, 0.01], 1.5, [0, 1, 1])
    far = np.eye(4)
    far[:3, 3] = [0, 0, 40.0]
    ok, acc, chosen = s.validate_poses([far, near])
    ook, oacc, ochosen, opose = o.validate_poses([far, near])
    assert ok and ook == 1 and list(acc) == list(oacc) == [-1, 1] and chosen == ochosen == 1
    assert np.abs(s.result()["pose"] - opose).max() < 5e-4
    for f in fr[7:9]:
        s.GrabImage(*A.to_device(*f))
        g = s.result()
        ost_, op = o.grab(*f)
        assert g["track_status"] == ost_ == 1 and g["rgb_count"] > 0
        assert np.abs(g["pose"] - op).max() < 2e-3
    s.close()
    # two systems in one batch
    pair = [A.System(A.system_params(cols=320, rows=240, K=K, map=mp)) for _ in range(2)]
    bt = A.Batch(2)
    for x in pair:
        bt.attach(x)
    for f in seq:
        d, c = A.to_device(*f)
        bt.grab(pair, [d, d], [c, c])
    assert [x.result()["track_status"] for x in pair] == [5, 5]
    got = [x.validate_poses([far, near]) for x in pair]
    assert all(g[0] for g in got) and all(list(g[1]) == [-1, 1] for g in got)
    r0, r1 = pair[0].result(), pair[1].result()
    assert r0["track_status"] == r1["track_status"] == 6 and np.array_equal(r0["pose"], r1["pose"])
    for f in fr[7:9]:
        d, c = A.to_device(*f)
        bt.grab(pair, [d, d], [c, c])
        r0, r1 = pair[0].result(), pair[1].result()
        assert r0["track_status"] == r1["track_status"] == 1 and np.array_equal(r0["pose"], r1["pose"])
    for x in pair:
        bt.detach(x)
        x.close()


def test_lost_system_through_a_snapshot(gpu, tmp_path):
    """a lost session saved and loaded stays lost; validation needs a frame given after the
    load, and then relocalises it exactly as the original session does"""
    from tests.test_oracle_graph import _perturb
    from tests.helpers import scene_frames
    sc, K, fr = scene_frames(frames=range(0, 8), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    mp = dict(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
              trunc_dist=float(np.float32(0.01) * np.float32(3.0)), max_new_blocks=1 << 16)
    ref = A.System(A.system_params(cols=320, rows=240, K=K, map=mp))
    for f in fr[:7]:
        ref.GrabImage(*A.to_device(*f))
    truth = ref.result()["pose"]  # frame 6 as tracking places it
    ref.close()
    s = _lost_gpu(K, fr, empty, mp)  # frames 0..5, ten empty frames, frame 6 (lost)
    path = str(tmp_path / "lost.fusess")
    s.save(path)
    t = A.System(A.system_params(cols=320, rows=240, K=K, map=mp))
    t.load(path)
    cand = [_perturb(truth, [0.012, -0.01, 0.01], 1.5, [0, 1, 1])]
    with pytest.raises(A.N.FusionError):  # no frame since the load
        t.validate_poses(cand)
    t.GrabImage(*A.to_device(*fr[6]))
    assert t.result()["track_status"] == 5
    a = s.validate_poses(cand)
    b = t.validate_poses(cand)
    assert a[0] and b[0] and list(a[1]) == list(b[1]) and a[2] == b[2]
    assert np.array_equal(s.result()["pose"], t.result()["pose"])
    s.GrabImage(*A.to_device(*fr[7]))
    t.GrabImage(*A.to_device(*fr[7]))
    assert s.result()["track_status"] == t.result()["track_status"] == 1
    assert np.array_equal(s.result()["pose"], t.result()["pose"])
    s.close()
    t.close()
