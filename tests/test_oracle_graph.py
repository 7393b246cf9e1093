"""Graph-matching relocalisation on the CPU: the oracle's sub-graph selection
(Tracker::GenerateGraph, Tracking/Tracking.cc:515-596), its list rule (:591-629), its match
list (Relocalise with graph matching on, :432-453, 485-505) and Tracker::ValidatePose
(:636-682) on a lost oracle system.

* selection: a seven-pair matrix worked by hand (a five-pair clique, a pair nearly exclusive
  with one clique member that scores higher with the head, an outlier); column sums that tie
  exactly, so the lower index goes first; the cap on selected pairs; lists past the last pair;
  a NaN rank sorted last; and, on random matrices, the definitions themselves checked pair by
  pair (selected = the first qualifying pairs in rank order; a pair is refined away exactly
  when a later pair of its list is a witness);
* list rule: the first list of three or more selected pairs, called min(5, such lists) times,
  skipped when it keeps three or fewer refined pairs;
* match list: the ratio test's winners once, the others with both neighbours, the invalid
  map record dropped;
* ValidatePose: on a system lost after ten empty frames, a far pose is skipped for visibility,
  a pose 35 cm off is rejected by ICP, and a pose 2 cm / 1.5 degrees off the pose that
  tracking gives the frame on the same model is refined to within 1 mm of it; the next frames
  track.
"""
import numpy as np

from oracle import binding as O
from tests.helpers import scene_frames


def _hand_matrix():
    C = np.full((7, 7), 0.05, np.float32)
    for a in range(5):
        for b in range(5):
            C[a, b] = 0.8
    np.fill_diagonal(C, 0.9)
    C[0, 5] = C[5, 0] = 0.95  # pair 5 scores higher with the head than pair 1 does
    C[1, 5] = C[5, 1] = 0.001  # and is nearly exclusive with pair 1
    C[5, 6] = C[6, 5] = 0.0
    return C


def test_select_known_answer():
    C = _hand_matrix()
    nsel, lists, rank, order = O.graph_select(C, 10, 400, 0.1)
    # column sums: 5.10, 4.151, 4.2 (x3), 2.001, 1.15; the three 4.2 sums are equal, so index order
    assert np.allclose(rank, [5.10, 4.151, 4.2, 4.2, 4.2, 2.001, 1.15], atol=1e-6)
    assert rank[2] == rank[3] == rank[4]
    assert list(order) == [0, 2, 3, 4, 1, 5, 6]
    # list 0: head 0, then 2, 3, 4, 1, 5 (6 scores 0.05 with the head); pair 1 is refined away
    # (5 comes later, is nearly exclusive with it and scores 0.95 > 0.8 with the head); the head
    # is never kept
    assert list(nsel) == [6, 4, 3, 2, 1, 1, 1, 0, 0, 0]
    assert list(lists[0]) == [2, 3, 4, 5]
    assert list(lists[1]) == [3, 4, 1]  # head 2: 5 and 6 score 0.05 with it
    assert list(lists[2]) == [4, 1]
    assert all(len(x) == 0 for x in lists[3:])
    # at most three selected per list
    nsel3, lists3, _, _ = O.graph_select(C, 10, 3, 0.1)
    assert nsel3[0] == 3 and list(lists3[0]) == [2, 3]
    # the list rule: lists 0, 1, 2 qualify, all three calls take list 0
    M, pairs = O.graph_lists(nsel, lists, 400, 5)
    assert M == 3 and list(pairs) == [2, 3, 4, 5]
    M, pairs = O.graph_lists(nsel, lists, 400, 2)
    assert M == 2
    # the first qualifying list with three refined pairs: no call
    M, pairs = O.graph_lists(np.array([2, 4, 3]), [np.array([]), np.array([3, 4, 1]), np.array([1, 2])], 400, 5)
    assert M == 0
    M, _ = O.graph_lists(np.array([2, 1, 0]), [np.array([])] * 3, 400, 5)
    assert M == 0


def test_select_nan_rank_and_small_sizes():
    C = _hand_matrix()
    C[3, 4] = np.nan  # column 4's sum is NaN: it ranks last
    _, _, rank, order = O.graph_select(C, 3, 400, 0.1)
    assert np.isnan(rank[4]) and order[-1] == 4
    for n in (0, 1, 2):
        nsel, lists, _, order = O.graph_select(np.ones((n, n), np.float32), 4, 400, 0.1)
        assert list(nsel) == [min(n, 2) - i if i < n else 0 for i in range(4)]
        assert all(len(x) == 0 for x in lists)


def test_select_definitions_on_random_matrices():
    rng = np.random.default_rng(8)
    for n, max_sel, score in ((40, 400, 0.1), (300, 60, 0.3), (200, 400, 0.05)):
        C = rng.random((n, n)).astype(np.float32)
        C[rng.random((n, n)) < 0.05] = rng.random() * 4e-3  # some nearly exclusive pairs
        nsel, lists, rank, order = O.graph_select(C, 10, max_sel, score)
        r64 = C.astype(np.float64).sum(0).astype(np.float32)
        assert np.array_equal(rank, r64)
        assert list(order) == sorted(range(n), key=lambda x: (-float(rank[x]), x))
        for i in range(10):
            head = order[i]
            sel = [head] + [p for p in order[i + 1:] if C[head, p] > score][: max_sel - 1]
            assert nsel[i] == len(sel)
            if len(sel) < 3:
                assert len(lists[i]) == 0
                continue
            kept = []
            for k in range(1, len(sel)):
                a = sel[k]
                witness = any((C[a, b] < 5e-3 or C[b, a] < 5e-3) and C[head, b] > C[head, a] for b in sel[k + 1:])
                if not witness:
                    kept.append(a)
            assert list(lists[i]) == kept


def test_graph_match_list():
    keys = np.zeros(4, O.SURF)
    keys["valid"] = [0, 1, 1, 1]  # the first collected record is never written: invalid
    keys["pos"] = np.arange(12, dtype=np.float32).reshape(4, 3)
    idx = np.array([[1, 2], [2, 3], [0, 1], [3, -1]], np.int32)
    dist = np.array([[0.1, 0.5], [0.3, 0.31], [0.2, 0.9], [0.4, np.inf]], np.float32)
    fpos = np.arange(12, dtype=np.float32).reshape(4, 3) + 100
    fnrm = np.tile(np.float32([0, 0, -1, 0]), (4, 1))
    fk, mk, d, q, nm = O.graph_inputs(idx, dist, keys, fpos, fnrm)
    # query 0 passes the ratio test (one pair), 1 fails (two), 2 passes but its key is the
    # invalid record, 3 has one neighbour (passes against +inf)
    assert nm == 5
    assert list(q) == [0, 1, 1, 3]
    assert np.array_equal(mk["pos"], keys["pos"][[1, 2, 3, 3]])
    assert np.array_equal(d, np.float32([0.1, 0.3, 0.31, 0.4]))
    assert np.array_equal(fk["pos"], fpos[[0, 1, 1, 3]]) and np.all(fk["valid"] == 1)


def _lost_oracle(K, fr, empty):
    o = O.System(cols=320, rows=240, K=K, num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15,
                 voxel_size=0.01)
    st = [o.grab(*f)[0] for f in list(fr[:6]) + [empty] * 10 + [fr[6]]]
    assert st == [4] + [1] * 5 + [2] * 10 + [5]
    return o


def _perturb(P, t, deg, axis):
    a = np.radians(deg)
    w = np.asarray(axis, float) / np.linalg.norm(axis)
    Wx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    R = np.eye(3) + np.sin(a) * Wx + (1 - np.cos(a)) * Wx @ Wx
    Q = P.copy()
    Q[:3, :3] = R @ P[:3, :3]
    Q[:3, 3] += t
    return Q


def test_validate_pose_on_a_lost_system():
    sc, K, fr = scene_frames(frames=range(0, 9), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    o = _lost_oracle(K, fr, empty)
    # frame 6 as tracking places it on the same model (frames 0..5), without the ten lost frames
    ref = O.System(cols=320, rows=240, K=K, num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15,
                   voxel_size=0.01)
    truth = [ref.grab(*f) for f in fr[:7]][6][1]
    far = np.eye(4)
    far[:3, 3] = [0, 0, 40.0]  # 40 m away: nothing visible
    off = _perturb(truth, [0.25, -0.2, 0.15], 4.0, [1, 2, 0])
    near = _perturb(truth, [0.012, -0.01, 0.01], 1.5, [0, 1, 1])
    near2 = _perturb(truth, [-0.01, 0.0, 0.01], 1.0, [1, 0, 0])
    ok, acc, chosen, pose = o.validate_poses([far, off, near, near2])
    assert list(acc) == [-1, 0, 1, 1] and chosen == 2 and ok == 1
    assert np.abs(pose[:3, 3] - truth[:3, 3]).max() < 1e-3
    assert np.abs(pose[:3, :3] - truth[:3, :3]).max() < 1e-3
    # relocalised: the next frames track (the first is not swapped in, as after any relocalisation)
    for f in fr[7:9]:
        st, p = o.grab(*f)
        assert st == 1
    # not lost any more
    assert o.validate_poses([near])[0] == -1


def test_validate_pose_without_a_passing_candidate():
    sc, K, fr = scene_frames(frames=range(0, 7), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    o = _lost_oracle(K, fr, empty)
    far = np.eye(4)
    far[:3, 3] = [0, 0, 40.0]
    ok, acc, chosen, _ = o.validate_poses([far])
    assert ok == 0 and list(acc) == [-1] and chosen == -1
    ok, acc, chosen, _ = o.validate_poses([])
    assert ok == 0 and chosen == -1
    # still lost: the next frame is not tracked
    assert o.grab(*fr[6])[0] == 5
