"""CPU tests of the synthetic scene generator: determinism, agreement with its own analytic
SDF, noise model and trajectory speeds."""
import numpy as np

from fusion_amd.api import Scene

K = (525.0, 525.0, 319.5, 239.5)


def test_deterministic_and_seeded():
    a = Scene(0, (4.0, 4.0, 3.0), 0).render(5)
    b = Scene(0, (4.0, 4.0, 3.0), 0).render(5)
    c = Scene(0, (4.0, 4.0, 3.0), 1).render(5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert not np.array_equal(a[0], c[0])


def test_depth_lies_on_the_analytic_surface():
    sc = Scene(1, (12.0, 8.0, 3.0), 1)
    pose = sc.pose(7)
    d, _ = sc.render_pose(pose, noise=False)
    ys, xs = np.nonzero(d)
    sel = np.random.default_rng(0).choice(len(xs), 400, replace=False)
    for y, x in zip(ys[sel], xs[sel]):
        z = d[y, x] / 5000.0
        pc = np.array([(x - K[2]) / K[0] * z, (y - K[3]) / K[1] * z, z, 1.0])
        pw = pose @ pc
        # quantisation to 1/5000 m along the optical axis
        assert abs(sc.sdf(*pw[:3])) < 2e-4 * max(1.0, z)


def test_noise_model():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    clean, _ = sc.render_pose(sc.pose(3), noise=False)
    noisy, _ = sc.render_pose(sc.pose(3), noise=True, noise_seed=3)
    dropped = (noisy == 0) & (clean > 0)
    assert 0.007 < dropped.mean() < 0.013  # 1 % dropout
    m = (noisy > 0) & (clean > 0)
    z = clean[m] / 5000.0
    r = (noisy[m].astype(np.float64) - clean[m]) / 5000.0 / (0.0015 * z * z)
    assert abs(r.mean()) < 0.02 and 0.95 < r.std() < 1.05


def test_trajectory_speeds():
    for kind, size in ((0, (4.0, 4.0, 3.0)), (1, (12.0, 8.0, 3.0))):
        sc = Scene(kind, size, 2)
        P = [sc.pose(f) for f in range(0, 600)]
        v = [np.linalg.norm(P[i + 1][:3, 3] - P[i][:3, 3]) * 30 for i in range(len(P) - 1)]
        assert 0.27 < np.median(v) < 0.33
        w = []
        for i in range(len(P) - 1):
            R = P[i][:3, :3].T @ P[i + 1][:3, :3]
            w.append(np.degrees(np.arccos(np.clip((np.trace(R) - 1) / 2, -1, 1))) * 30)
        assert max(w) < 22.0  # heading rate bounded by 20 deg/s plus the slow pitch oscillation
        # camera stays inside the scene, away from the walls
        pos = np.array([p[:3, 3] for p in P])
        assert pos.min() > 0.3 and (pos[:, 0].max() < size[0] - 0.3) and (pos[:, 1].max() < size[1] - 0.3)
