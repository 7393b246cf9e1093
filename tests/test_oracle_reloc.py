"""The relocalisation pieces on the CPU: the oracle's keypoint lifting (Frame::ExtractKeyPoints'
test and back-projection, Core/Frame.cc:93-200) and 2-NN matcher (knnMatch, k = 2, L2) against
independent numpy statements, and the library's host-side PoseEstimate (Optimization/Solver.cc)
against the closed form on exact data.

* lifting: numpy float32 element-wise operations (single IEEE operations) restate InterpDepth,
  InterpNormal and the back-projection; a fronto-parallel plane gives the closed form
  (x - cx) * z / fx; the bounds test as written rejects x >= rows - 1 on a wide image;
* matching: distances in float64 rank the train descriptors of well-separated data; equal
  descriptors tie and the lower index wins;
* PoseEstimate: noise-free inliers of a known rigid motion plus outliers give that motion
  (Kabsch by numpy's SVD on the inliers) and exactly the inlier set; fewer than three pairs,
  collinear data and a mirror image are refused.
"""
import numpy as np

from fusion_amd import api as A
from oracle import binding as O

f32 = np.float32


def _lift_numpy(xy, depth, nmap, K, cutoff):
    rows, cols = depth.shape
    fx, fy, cx, cy = (f32(k) for k in K)
    out_p = np.full((len(xy), 3), np.nan, f32)
    out_n = np.full((len(xy), 4), np.nan, f32)
    ok = np.zeros(len(xy), np.uint8)
    for i, (x, y) in enumerate(xy.astype(f32)):
        if x <= 1 or y <= 1 or y >= cols - 1 or x >= rows - 1:  # as written
            continue
        if np.isnan(x) or np.isnan(y) or x >= cols - 1 or y >= rows - 1:  # would leave the image
            continue
        ax, ay = f32(x - np.floor(x)), f32(y - np.floor(y))
        x0, x1, y0, y1 = int(np.floor(x)), int(np.ceil(x)), int(np.floor(y)), int(np.ceil(y))
        d = [depth[y0, x0], depth[y1, x0], depth[y0, x1], depth[y1, x1]]
        if any(np.isnan(v) or float(v) < 0.3 or v > f32(cutoff) for v in d):
            continue
        d00, d10, d01, d11 = d
        e0 = f32(d01 * ax) + f32(d00 * f32(1 - ax))
        e1 = f32(d11 * ax) + f32(d10 * f32(1 - ax))
        fin = f32(f32(f32(1 - ay) * e0) + f32(ay * e1))
        if not float(abs(f32(fin - d00))) <= 0.005:
            continue
        dp = fin
        if not (float(dp) > 0.3 and dp < f32(cutoff)):
            continue
        n00, n10, n01, n11 = nmap[y0, x0], nmap[y1, x0], nmap[y0, x1], nmap[y1, x1]
        g0 = (n01 * ax).astype(f32) + (n00 * f32(1 - ax)).astype(f32)
        g1 = (n11 * ax).astype(f32) + (n10 * f32(1 - ax)).astype(f32)
        fn = (g0 * f32(1 - ay)).astype(f32) + (g1 * ay).astype(f32)
        df = (fn - n00).astype(f32)
        s = f32(0)
        for c in range(4):
            s = f32(s + f32(df[c] * df[c]))
        if not float(np.sqrt(s, dtype=f32)) <= 0.1:
            continue
        out_p[i] = [f32(f32(dp * f32(x - cx)) / fx), f32(f32(dp * f32(y - cy)) / fy), dp]
        out_n[i] = fn
        ok[i] = 1
    return out_p, out_n, ok


def _tilted(rows, cols, K):
    """depth of a tilted plane and its (constant) normal map"""
    yy, xx = np.mgrid[0:rows, 0:cols].astype(np.float64)
    z = 1.2 + 0.0015 * xx + 0.0008 * yy
    depth = z.astype(f32)
    n = np.array([-0.3, -0.15, -1.0, 0.0])
    n[:3] /= np.linalg.norm(n[:3])
    nmap = np.broadcast_to(n.astype(f32), (rows, cols, 4)).copy()
    return depth, nmap


def test_lift_matches_numpy_and_the_plane():
    rng = np.random.default_rng(3)
    rows, cols = 120, 160
    K = (131.25, 131.25, 79.5, 59.5)
    depth, nmap = _tilted(rows, cols, K)
    depth[40:44, 60:64] = np.nan  # holes reject their neighbourhood
    depth[80, 100] = 3.5  # beyond the cut-off
    nmap[10:14, 10:14, 0] += 0.5  # a normal step larger than 0.1
    xy = np.concatenate([rng.uniform(0, 160, (400, 2)), [[61.5, 41.2], [100.2, 80.3], [11.5, 11.5], [150.0, 50.0],
                                                          [50.0, 119.5], [118.95, 30.0], [1.0, 50.0]]]).astype(f32)
    p, n, ok = O.lift_keypoints(xy, depth, nmap, K)
    pn, nn, okn = _lift_numpy(xy, depth, nmap, K, 3.0)
    assert np.array_equal(ok, okn)
    assert np.array_equal(p[ok == 1].view(np.uint32), pn[ok == 1].view(np.uint32))
    assert np.array_equal(n[ok == 1].view(np.uint32), nn[ok == 1].view(np.uint32))
    assert np.all(np.isnan(p[ok == 0]))
    assert ok.sum() > 50
    # the bounds test as written: x against the height (119), y against the width (159);
    # y = 119.5 (< width, but its neighbours leave the 120-row image) is refused as well
    assert ok[-4] == 0 and ok[-1] == 0 and ok[-3] == 0  # x = 150 >= rows - 1; x = 1; y = 119.5
    assert ok[-2] == 1  # x = 118.95 < rows - 1 on the plane
    # on the plane the back-projection is the closed form
    sel = ok == 1
    zz = p[sel, 2].astype(np.float64)
    assert np.allclose(p[sel, 0], (xy[sel, 0] - K[2]) * zz / K[0], rtol=1e-6, atol=1e-7)
    assert np.allclose(p[sel, 1], (xy[sel, 1] - K[3]) * zz / K[1], rtol=1e-6, atol=1e-7)


def test_match_ranks_by_distance_and_breaks_ties_low():
    rng = np.random.default_rng(4)
    t = rng.normal(size=(300, 64)).astype(f32)
    q = (t[rng.integers(0, 300, 40)] + rng.normal(0, 0.05, (40, 64))).astype(f32)
    idx, dist = O.match_knn2(q, t)
    d64 = np.sqrt(((q[:, None, :].astype(np.float64) - t[None, :, :]) ** 2).sum(-1))
    order = np.argsort(d64, axis=1, kind="stable")[:, :2]
    assert np.array_equal(idx, order)
    assert np.allclose(dist, np.take_along_axis(d64, order, 1), rtol=1e-6)
    # duplicates: the lower index first, then the next
    t2 = np.concatenate([t[:5], t[:5]])
    i2, d2 = O.match_knn2(t[:5], t2)
    assert np.array_equal(i2, np.stack([np.arange(5), np.arange(5) + 5], 1)) and np.all(d2 == 0)
    # one train descriptor: the second is missing
    i3, d3 = O.match_knn2(q[:3], t[:1])
    assert np.all(i3[:, 1] == -1) and np.all(np.isinf(d3[:, 1]))


def _rigid(rng):
    w = rng.normal(size=3)
    w /= np.linalg.norm(w)
    a = rng.uniform(0.1, 1.0)
    Wx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    R = np.eye(3) + np.sin(a) * Wx + (1 - np.cos(a)) * Wx @ Wx
    return R, rng.uniform(-1, 1, 3)


def _kabsch(src, ref):
    sm, rm = src.mean(0), ref.mean(0)
    U, _, Vt = np.linalg.svd((src - sm).T @ (ref - rm))
    R = U @ Vt
    return R, sm - R @ rm


def test_pose_estimate_recovers_the_motion_and_its_inliers():
    rng = np.random.default_rng(5)
    for trial in range(6):
        R, t = _rigid(rng)
        n = 80
        ref = rng.uniform(-2, 2, (n, 3))
        src = ref @ R.T + t
        out = rng.random(n) < 0.3
        src[out] += rng.uniform(0.5, 1.5, (out.sum(), 3)) * rng.choice([-1, 1], (out.sum(), 3))
        T, inl = A.pose_estimate(src, ref, iterations=100, seed=trial)
        assert T is not None
        assert np.array_equal(inl, ~out)
        Rk, tk = _kabsch(src[~out], ref[~out])
        assert np.abs(T[:3, :3] - Rk).max() < 1e-9 and np.abs(T[:3, 3] - tk).max() < 1e-9
        assert np.abs(T[:3, :3] - R).max() < 1e-9 and np.abs(T[:3, 3] - t).max() < 1e-9
        # the same seed, the same result
        T2, _ = A.pose_estimate(src, ref, iterations=100, seed=trial)
        assert np.array_equal(T, T2)


def test_pose_estimate_refusals():
    rng = np.random.default_rng(6)
    assert A.pose_estimate(np.zeros((2, 3)), np.zeros((2, 3)))[0] is None
    line = np.outer(np.linspace(0, 1, 20), [1.0, 2.0, 0.5])
    assert A.pose_estimate(line, line + 0.1)[0] is None  # every sample collinear
    ref = rng.uniform(-1, 1, (30, 3))
    mirror = ref * np.array([1.0, 1.0, -1.0])  # a reflection, not a rotation
    assert A.pose_estimate(mirror, ref)[0] is None
