"""The oracle's ICP sums (or_icp_sums) against an independent numpy restatement of
Tracking/Reduction.cu:227-286 at a moving pose (the true motion between frames 0 and 2 of
room 0 at 640x480).

Pins: the association's projection rule (float invz, float products, + 0.5 in double,
truncation; Reduction.cu:237-239), the acceptance tests, the row in the last camera frame and
the 28 products, through the exact correspondence count and the FP64 sums of the same FP32
products to 1e-12 of their absolute sums.  On these maps the all-double projection associates
other pixels (checked below); tests/test_oracle_planted.py shows that oracle variants with that
projection, or with the current frame's normal in the row, fail the comparison.
"""
import numpy as np

from oracle import binding as O
from tests.helpers import K640, scene_frames
from tests.icp_numpy import associate, sums64


def _pyr(d16, K):
    _, f0 = O.filter_depth(d16, 5000.0, 3.0)
    v = O.compute_vmap(f0, *K, 3.0)
    return v, O.compute_nmap(v)


def _moving_pair():
    sc, K, fr = scene_frames(frames=(0, 2))
    v0, n0 = _pyr(fr[0][0], K)
    v1, n1 = _pyr(fr[1][0], K)
    gt = np.linalg.inv(sc.pose(0)) @ sc.pose(2)  # frame 2 in frame 0's camera
    return K, v0, n0, v1, n1, gt


def _close(s, ref, absum):
    return np.all(np.abs(s - ref) <= 1e-12 * np.maximum(absum, 1.0))


def test_icp_sums_match_independent_fp64_at_a_moving_pose():
    K, v0, n0, v1, n1, gt = _moving_pair()
    _, s64 = O.icp_sums(v1, v0, n1, n0, gt, np.eye(4), K)
    ok, j, row = associate(v1, n1, v0, n0, gt, np.eye(4), K)
    s_np, absum = sums64(ok, row)
    assert s64[28] == s_np[28] > 200000  # the correspondence count, exactly
    assert _close(s64, s_np, absum), np.abs(s64 - s_np).max()
    # the same with a last pose that is not the identity (both poses moved by one rigid motion)
    W = np.array([[0.96, -0.28, 0.0, 0.4], [0.28, 0.96, 0.0, -0.3], [0.0, 0.0, 1.0, 0.2], [0, 0, 0, 1.0]])
    _, s64w = O.icp_sums(v1, v0, n1, n0, W @ gt, W, K)
    okw, _, roww = associate(v1, n1, v0, n0, W @ gt, W, K)
    s_w, absw = sums64(okw, roww)
    assert s64w[28] == s_w[28] and _close(s64w, s_w, absw)


def test_projection_rule_decides_associations_on_these_maps():
    """the reference's float projection and an all-double one land some pixels on different
    last-frame pixels here, so the sums above tell the two rules apart"""
    K, v0, n0, v1, n1, gt = _moving_pair()
    ok, j, row = associate(v1, n1, v0, n0, gt, np.eye(4), K)
    okd, jd, rowd = associate(v1, n1, v0, n0, gt, np.eye(4), K, rule="double")
    assert int((j != jd).sum()) > 0
    s_ref, absum = sums64(ok, row)
    s_dbl, _ = sums64(okd, rowd)
    assert not _close(s_dbl, s_ref, absum)
