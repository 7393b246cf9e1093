"""Independent checks of oracle functions that the pipeline tests alone do not pin.

* Integration (Mapping/FuseMap.cu:140-191): two frames of a fronto-parallel wall at different
  depths and colours, so the running average and the colour blend both depend on the 4-D
  weight w = n.normalised((p, 1)) = (p.z + 1) / sqrt(|p|^2 + 1); checked against that closed
  form in FP64 per voxel.
* Z-range tiles (Mapping/RenderScene.cu:54-179): the eight corners of every visible block
  projected in numpy, and the tiles that the reference's fragments cover taken from the
  closed form of createRenderingBlockList's loops (fragment columns run to ny, the lower right
  corner lacks the fragment offset), independently of the oracle's loops.
* The acceptance test's Euler sines (Tracking/Tracking.cc:398-408) against scipy's intrinsic
  XYZ Euler angles.
"""
import numpy as np
from scipy.spatial.transform import Rotation

from oracle import binding as O
from tests.helpers import K640, scene_frames

f32 = np.float32


# ----------------------------------------------------------------------- integration
def test_integration_of_two_walls_weight_and_colour():
    H, W = 60, 80
    Kq = (100.0, 100.0, 39.5, 29.5)
    Z1, Z2 = f32(1.0), f32(1.012)
    c1, c2 = np.array([200, 40, 120]), np.array([20, 180, 61])
    m = O.Map(num_buckets=1 << 14, num_excess=1 << 12, num_blocks=1 << 13)
    frames = []
    for Z, c in ((Z1, c1), (Z2, c2)):
        depth = np.full((H, W), Z, np.float32)
        n = O.compute_nmap(O.compute_vmap(depth, *Kq, 3.0))
        # a constant depth's normals are exactly (0, 0, 1, 1) away from the last row and column
        assert np.array_equal(n[:-1, :-1].reshape(-1, 4), np.tile(f32([0, 0, 1, 1]), ((H - 1) * (W - 1), 1)))
        frames.append((depth, np.broadcast_to(c.astype(np.uint8), (H, W, 3)).copy(), n))
    m.fuse_map_color(*frames[0], np.eye(4), Kq)
    first_blocks = set(map(tuple, m.export_blocks().tolist()))
    m.fuse_map_color(*frames[1], np.eye(4), Kq)
    blocks = m.export_blocks()
    sdf, wt, col = m.read_blocks(blocks)
    T = float(f32(0.01) * f32(3.0))
    checked = 0
    for bi, b in enumerate(blocks):
        if tuple(b) not in first_blocks:
            continue
        li = np.arange(512)
        vx = np.stack([b[0] * 8 + li % 8, b[1] * 8 + (li // 8) % 8, b[2] * 8 + li // 64], -1)
        p = (vx.astype(f32) * f32(0.01)).astype(np.float64)  # the voxel's float position
        u = Kq[0] * p[:, 0] / p[:, 2] + Kq[2]
        v = Kq[1] * p[:, 1] / p[:, 2] + Kq[3]
        s1, s2 = float(Z1) - p[:, 2], float(Z2) - p[:, 2]
        keep = (u >= 1) & (u < W - 2) & (v >= 1) & (v < H - 2) & (s1 >= -T + 1e-5)
        keep &= (np.abs(s1 - T) > 1e-5) & (np.abs(s2 - T) > 1e-5)
        if not keep.any():
            continue
        w = (p[:, 2] + 1.0) / np.sqrt((p ** 2).sum(1) + 1.0)
        e1, e2 = np.minimum(1.0, s1 / T), np.minimum(1.0, s2 / T)
        want_sdf = (e1 + w * e2) / (1.0 + w)
        blend = (w * 0.2)[:, None] * c2[None, :] + (1.0 - w * 0.2)[:, None] * c1[None, :]
        for k in np.nonzero(keep)[0]:
            assert wt[bi, k] == 2, (b, k)
            assert abs(sdf[bi, k] - want_sdf[k]) < 1e-5, (b, k, sdf[bi, k], want_sdf[k])
            exact = np.floor(blend[k])
            near = np.abs(blend[k] - np.round(blend[k])) < 1e-3
            got = col[bi, k].astype(np.float64)
            assert np.all((got == exact) | (near & (np.abs(got - np.round(blend[k])) <= 1))), (got, blend[k])
            checked += 1
    assert checked > 2000, checked


# ----------------------------------------------------------------------- z-range tiles
def _zrange_brute(blocks, pose, K, zc, zr, vox, dmin, dmax):
    R = np.asarray(pose, np.float64)
    Ri = R[:3, :3].T.astype(f32)
    t = R[:3, 3].astype(f32)
    fx, fy, cx, cy = (f32(k) for k in K)
    zmin = np.full((zr, zc), f32(100.0), np.float32)
    zmax = np.zeros((zr, zc), np.float32)
    total = 0
    corners = np.array([[(c & 1), (c >> 1) & 1, (c >> 2) & 1] for c in range(8)])
    for b in blocks:
        q = (b[None, :] + corners) * 8
        p = q.astype(f32) * f32(vox) - t
        pc = np.stack([Ri[i, 0] * p[:, 0] + Ri[i, 1] * p[:, 1] + Ri[i, 2] * p[:, 2] for i in range(3)], -1)
        use = ~(pc[:, 2].astype(np.float64) < 0.2)
        if not use.any():
            continue
        pc = pc[use]
        px = (fx * pc[:, 0] / pc[:, 2] + cx) / f32(8)
        py = (fy * pc[:, 1] / pc[:, 2] + cy) / f32(8)
        ulx, lrx = max(int(np.floor(px).min()), 0), min(int(np.ceil(px).max()), zc - 1)
        uly, lry = max(int(np.floor(py).min()), 0), min(int(np.ceil(py).max()), zr - 1)
        zlo, zhi = max(pc[:, 2].min(), f32(dmin)), pc[:, 2].max()
        if ulx > lrx or uly > lry or zhi < f32(dmin):
            continue
        nx = int(np.ceil(f32(lrx - ulx + 1) / f32(16)))
        ny = int(np.ceil(f32(lry - uly + 1) / f32(16)))
        total += nx * ny
        # fragment 0 covers 16 columns and rows from the upper left; a second fragment column
        # exists when ny >= 2 and then covers column ulx + 16 only (its lower right is
        # ulx + 16); the same for rows
        x1 = min(lrx, ulx + (16 if ny >= 2 else 15))
        y1 = min(lry, uly + 16)
        zmin[uly:y1 + 1, ulx:x1 + 1] = np.minimum(zmin[uly:y1 + 1, ulx:x1 + 1], zlo)
        zmax[uly:y1 + 1, ulx:x1 + 1] = np.maximum(zmax[uly:y1 + 1, ulx:x1 + 1], zhi)
    return zmin, zmax, total


def _close_range_map():
    """blocks of a tilted wall 0.25-0.9 m in front of the camera: blocks near the camera span
    more than 16 tiles, and the image border clips some of them to one fragment row"""
    H, W = 480, 640
    u = np.arange(W, dtype=np.float64)
    depth = np.broadcast_to((0.25 + 0.65 * u / W).astype(np.float32), (H, W)).copy()
    depth[:, :] += (0.05 * np.sin(np.arange(H) / 40.0)[:, None]).astype(np.float32)
    m = O.Map(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15)
    m.create_blocks(depth, np.eye(4), K640)
    return m


def test_zrange_tiles_brute_force():
    m = _close_range_map()
    pose = np.eye(4)
    pose[:3, 3] = (0.01, -0.02, -0.03)
    n = m.check_visibility(pose, 640, 480, K640)
    vis = m.export_visible()
    assert n == len(vis) > 100
    zmin, zmax, nf, ok = m.create_rendering_blocks(80, 60, pose, n, K640)
    bmin, bmax, total = _zrange_brute(vis, pose, K640, 80, 60, 0.01, m.depth_min, m.depth_max)
    assert ok and nf == total
    assert np.array_equal(zmin, bmin) and np.array_equal(zmax, bmax)
    # on the room scene, from a pose other than the one that allocated the blocks
    sc, K, fr = scene_frames(frames=(0,))
    raw, _ = O.filter_depth(fr[0][0], 5000.0, 3.0)
    m2 = O.Map(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15)
    m2.create_blocks(raw, np.eye(4), K)
    pose2 = np.linalg.inv(sc.pose(0)) @ sc.pose(6)
    n2 = m2.check_visibility(pose2, 640, 480, K)
    zmin2, zmax2, nf2, _ = m2.create_rendering_blocks(80, 60, pose2, n2, K)
    b2min, b2max, total2 = _zrange_brute(m2.export_visible(), pose2, K, 80, 60, 0.01, m2.depth_min, m2.depth_max)
    assert nf2 == total2 and np.array_equal(zmin2, b2min) and np.array_equal(zmax2, b2max)


# ----------------------------------------------------------------------- acceptance test
def test_euler_sines_of_the_acceptance_test_against_scipy():
    rng = np.random.default_rng(7)
    rots = list(Rotation.random(400, random_state=11))
    # small motions, as tracking sees them, and large ones up to 80 degrees about Y
    rots += [Rotation.from_rotvec(rng.normal(0, 0.05, 3)) for _ in range(400)]
    rots += [Rotation.from_euler("XYZ", [rng.uniform(-3, 3), rng.uniform(-1.4, 1.4), rng.uniform(-3, 3)])
             for _ in range(400)]
    for r in rots:
        p = np.eye(4)
        p[:3, :3] = r.as_matrix()
        p[:3, 3] = rng.normal(0, 0.1, 3)
        a, b, c = r.as_euler("XYZ")  # intrinsic: R = Rx(a) Ry(b) Rz(c)
        if abs(np.cos(b)) < 1e-6:
            continue
        want = np.linalg.norm(np.sin([a, b, c]))
        assert abs(O.euler_sine_norm(p) - want) < 1e-12, (r.as_matrix(), O.euler_sine_norm(p), want)
