"""Frame sources and trajectory tools (fusion_amd/ingest.py) on the CPU: the TUM list pairing
of MainTum.cc:100-134, lossless PNG round trips, the apt0 naming of MainTum.cc:74-93, the
TUM trajectory format and the absolute trajectory error."""
import os

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from fusion_amd import ingest as I


def test_tum_lists_pair_by_line_after_three_header_lines(tmp_path):
    d = tmp_path / "seq"
    d.mkdir()
    (d / "depth.txt").write_text("# depth maps\n# file: x\n# timestamp filename\n"
                                 "1.000 depth/a.png\n2.000 depth/b.png\n3.500 depth/c.png\n4.000 depth/d.png\n")
    # colour stamps differ and there is one line fewer: pairing is by line index
    (d / "rgb.txt").write_text("# color images\n# file: x\n# timestamp filename\n"
                               "0.990 rgb/A.png\n2.100   rgb/B.png\n3.400\trgb/C.png\n")
    items = I.tum_lists(str(d))
    root = str(d) + "/"
    assert items == [(1.0, root + "depth/a.png", root + "rgb/A.png"), (2.0, root + "depth/b.png", root + "rgb/B.png"),
                     (3.5, root + "depth/c.png", root + "rgb/C.png")]


def _frames(n=3, rows=48, cols=64):
    rng = np.random.default_rng(0)
    out = []
    for _ in range(n):
        d = rng.integers(0, 65536, size=(rows, cols)).astype(np.uint16)
        d[0, 0], d[1, 1] = 0, 65535
        c = rng.integers(0, 256, size=(rows, cols, 3)).astype(np.uint8)
        out.append((d, c))
    return out


def test_tum_sequence_round_trip_is_lossless(tmp_path):
    fr = _frames()
    stamps = [10.0 + 0.033 * i for i in range(len(fr))]
    I.write_tum_sequence(str(tmp_path / "seq"), fr, stamps)
    src = I.TumSource(str(tmp_path / "seq"))
    assert len(src) == 3 and (src.rows, src.cols) == (48, 64) and src.K == I.TUM_FR1["K"]
    for i, (d, c) in enumerate(src):
        assert d.dtype == np.uint16 and np.array_equal(d, fr[i][0])
        assert c.dtype == np.uint8 and np.array_equal(c, fr[i][1])
        assert abs(src.timestamp(i) - stamps[i]) < 1e-6


def test_apt0_layout(tmp_path):
    from PIL import Image
    fr = _frames(n=2)
    yy, xx = np.mgrid[0:48, 0:64]
    smooth = np.stack([2 * xx + 40, 3 * yy + 20, xx + yy + 60], axis=2).astype(np.uint8)
    fr = [(d, smooth) for d, _ in fr]
    for i, (d, c) in enumerate(fr):
        I.write_depth_png(str(tmp_path / ("frame-%06d.depth.png" % i)), d)
        Image.fromarray(c, "RGB").save(str(tmp_path / ("frame-%06d.color.jpg" % i)), quality=95)
    I.write_depth_png(str(tmp_path / "frame-000003.depth.png"), fr[0][0])  # after a gap: not read
    src = I.Apt0Source(str(tmp_path))
    assert len(src) == 2 and src.depth_scale == 1000.0 and src.K == (583.0, 583.0, 320.0, 240.0)
    d, c = src.frame(1)
    assert np.array_equal(d, fr[1][0])
    # JPEG is lossy: a smooth image comes back within a few levels, in RGB order
    assert c.shape == fr[1][1].shape and np.abs(c.astype(int) - fr[1][1]).mean() < 2


def test_stream_source_checks_shapes():
    fr = _frames(n=2)
    it = iter(fr + [None])
    src = I.StreamSource(lambda: next(it), cols=64, rows=48)
    got = list(src)
    assert len(got) == 2 and np.array_equal(got[1][0], fr[1][0])
    bad = I.StreamSource(lambda: (np.zeros((4, 4), np.uint16), np.zeros((4, 4, 3), np.uint8)), cols=64, rows=48)
    with pytest.raises(ValueError):
        next(iter(bad))


def test_trajectory_round_trip_and_quaternions(tmp_path):
    rng = np.random.default_rng(3)
    poses, stamps = [], []
    for i in range(50):
        P = np.eye(4)
        P[:3, :3] = Rotation.from_rotvec(rng.normal(size=3) * (0.5 if i % 2 else 3.0)).as_matrix()
        P[:3, 3] = rng.normal(size=3)
        poses.append(P)
        stamps.append(1000.0 + i / 30.0)
    p = str(tmp_path / "traj.txt")
    I.write_tum_trajectory(p, stamps, poses)
    ts, got = I.read_tum_trajectory(p)
    assert np.allclose(ts, stamps, atol=1e-6)
    for a, b in zip(poses, got):
        assert np.abs(a - b).max() < 1e-7
    # the quaternion convention is (x, y, z, w), as scipy's
    line = open(p).readline().split()
    q = np.array([float(x) for x in line[4:8]])
    qs = Rotation.from_matrix(poses[0][:3, :3]).as_quat()
    assert np.allclose(q, qs if qs[3] >= 0 else -qs, atol=1e-8)


def test_ate_is_alignment_invariant_and_measures_noise():
    rng = np.random.default_rng(4)
    gt = []
    for i in range(100):
        P = np.eye(4)
        P[:3, 3] = [np.cos(i / 10), np.sin(i / 10), 0.01 * i]
        gt.append(P)
    T = np.eye(4)
    T[:3, :3] = Rotation.from_rotvec([0.3, -0.2, 0.9]).as_matrix()
    T[:3, 3] = [1.0, 2.0, -3.0]
    moved = [T @ P for P in gt]
    rmse, R, t = I.ate_rmse(moved, gt)
    assert rmse < 1e-12 and np.allclose(R, T[:3, :3].T, atol=1e-12)
    noisy = []
    off = rng.normal(size=(100, 3)) * 0.01
    off -= off.mean(0)
    for P, o in zip(gt, off):
        Q = P.copy()
        Q[:3, 3] += o
        noisy.append(Q)
    rmse, _, _ = I.ate_rmse(noisy, gt)
    assert rmse <= np.sqrt((off ** 2).sum(1).mean()) + 1e-12 and rmse > 0.5 * 0.01 * np.sqrt(3)


def test_prefetch_keeps_order(tmp_path):
    fr = _frames(n=7)
    I.write_tum_sequence(str(tmp_path / "seq"), fr, [float(i) for i in range(7)])
    src = I.TumSource(str(tmp_path / "seq"))
    got = list(I.prefetch(src, workers=3, depth=2))
    assert len(got) == 7 and all(np.array_equal(g[0], f[0]) and np.array_equal(g[1], f[1]) for g, f in zip(got, fr))
