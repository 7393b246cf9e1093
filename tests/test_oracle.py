"""CPU tests that pin the oracle (oracle/) against closed forms, textbook routines, brute
force on tiny inputs and invariants, chosen so that a dropped term, a wrong sign or index,
or a transposed operand fails one of them.  The reference ships no tests or golden vectors
for this path; the marching-cubes tables are pinned separately by their fingerprints."""
import numpy as np
import pytest
from scipy.linalg import expm

from fusion_amd.api import Scene
from oracle import binding as O

K = (525.0, 525.0, 319.5, 239.5)


# ----------------------------------------------------------------------- exp
def test_expf_accuracy_and_special_values():
    xs = np.concatenate([np.linspace(-103.0, 88.0, 200001), np.linspace(-1, 1, 20001)]).astype(np.float32)
    got = np.array([O.lib().or_expf(float(x)) for x in xs[::7]], dtype=np.float32)
    ref = np.exp(xs[::7].astype(np.float64))
    fin = ref > 1.2e-38  # normal range
    ulp = np.abs(got[fin].astype(np.float64) - ref[fin]) / np.spacing(ref[fin].astype(np.float32)).astype(np.float64)
    assert ulp.max() <= 2.0, ulp.max()
    assert O.lib().or_expf(0.0) == 1.0
    assert np.isnan(O.lib().or_expf(float("nan")))
    assert O.lib().or_expf(100.0) == float("inf") and O.lib().or_expf(-200.0) == 0.0


# ----------------------------------------------------------------------- pyramid
def test_filter_depth_constant_and_cutoff():
    d = np.full((24, 32), 7500, np.uint16)  # 1.5 m
    raw, filt = O.filter_depth(d, 5000.0, 3.0)
    assert np.all(raw == np.float32(7500) * np.float32(1.0 / 5000.0))
    assert np.allclose(filt, 1.5, rtol=0, atol=2e-7)
    d[3, 4] = 20000  # 4 m: beyond the cutoff, and range-weighted out of its neighbours
    raw, filt = O.filter_depth(d, 5000.0, 3.0)
    assert np.isnan(raw[3, 4]) and np.isnan(filt[3, 4])
    w = np.exp(-(1 / 32.0 + 2.5 ** 2 * 2))  # the 4 m pixel's weight at (3, 5)
    assert abs(filt[3, 5] - 1.5) < 2.5 * w * 1.01


def test_filter_depth_brute_force_small():
    """5x5 bilateral with the reference's window [p-2, min(p+3, n-1)) (the last row and
    column never contribute) and weights exp(-(dp^2/32 + dz^2 * 2)), in float64"""
    rng = np.random.default_rng(1)
    d = rng.integers(4000, 9000, size=(9, 11)).astype(np.uint16)
    d[2, 3] = 0
    raw, filt = O.filter_depth(d, 5000.0, 3.0)
    z = d.astype(np.float64) / 5000.0
    H, W = d.shape
    for y in range(H):
        for x in range(W):
            ys = range(max(y - 2, 0), min(y + 3, H - 1))
            xs = range(max(x - 2, 0), min(x + 3, W - 1))
            w = np.array([[np.exp(-(((x - cx) ** 2 + (y - cy) ** 2) / 32.0 + (z[y, x] - z[cy, cx]) ** 2 * 2.0))
                           for cx in xs] for cy in ys])
            v = np.array([[z[cy, cx] for cx in xs] for cy in ys])
            assert abs(filt[y, x] - (w * v).sum() / w.sum()) < 3e-6


def test_pyrdown_interior_is_binomial_and_border_weights_mirror():
    rng = np.random.default_rng(2)
    src = rng.uniform(0.5, 2.5, size=(24, 32)).astype(np.float32)
    dst = O.pyrdown_gauss(src)
    g = np.array([1, 4, 6, 4, 1], np.float64)
    k = np.outer(g, g)
    for y in range(1, 11):
        for x in range(1, 15):
            win = src[2 * y - 2:2 * y + 3, 2 * x - 2:2 * x + 3].astype(np.float64)
            assert abs(dst[y, x] - (win * k).sum() / 256.0) < 2e-6
    # right border: window columns [2x-2, W-1) with weights indexed from the far end, so
    # the centre column 2x = W-2 gets weight 1 and its left neighbours 4 and 6
    one = np.zeros((24, 32), np.float32)
    one[:, 30] = 1.0
    dst = O.pyrdown_gauss(one)
    assert abs(dst[5, 15] - 1.0 / 11.0) < 1e-7
    # NaNs are skipped and the weights renormalised
    src2 = np.ones((24, 32), np.float32)
    src2[10, 10] = np.nan
    assert np.allclose(O.pyrdown_gauss(src2), 1.0)


def test_vmap_pinhole_and_thresholds():
    d = np.full((6, 8), 1.25, np.float32)
    d[0, 0] = 0.1  # 0.1f > 0.1 in double: valid
    d[0, 1] = np.float32(np.nextafter(np.float32(0.1), np.float32(0)))  # below: invalid
    d[0, 2] = 3.0  # not < cutoff: invalid
    d[0, 3] = np.nan
    v = O.compute_vmap(d, 500.0, 400.0, 3.5, 2.5, 3.0)
    assert not np.isnan(v[0, 0, 0]) and np.all(np.isnan(v[0, 1:4, 0]))
    yy, xx = np.mgrid[0:6, 0:8]
    ok = ~np.isnan(v[..., 0])
    assert np.allclose(v[..., 0][ok], (d * (xx - 3.5) / 500.0)[ok], rtol=1e-6)
    assert np.allclose(v[..., 1][ok], (d * (yy - 2.5) / 400.0)[ok], rtol=1e-6)
    assert np.all(v[..., 3][ok] == 1.0)


def test_nmap_of_a_tilted_plane():
    # plane z = 2 + 0.3 x_cam, observed by a pinhole camera: the normal is (-0.3, 0, 1)/|.|
    # pointing away from the camera
    H, W, f = 40, 50, 100.0
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    a = (xx - 25) / f
    z = (2.0 / (1.0 - 0.3 * a)).astype(np.float32)
    v = O.compute_vmap(z, f, f, 25.0, 20.0, 10.0)
    n = O.compute_nmap(v)
    exp = np.array([-0.3, 0.0, 1.0]) / np.hypot(0.3, 1.0)
    inner = n[:-1, :-1, :3].reshape(-1, 3)
    assert np.abs(inner - exp).max() < 2e-3
    assert np.all(np.isnan(n[-1, :, 0])) and np.all(np.isnan(n[:, -1, 0]))


def test_resize_map_means_and_4d_normalisation():
    v = np.arange(4 * 6 * 4, dtype=np.float32).reshape(4, 6, 4)
    n = np.zeros((4, 6, 4), np.float32)
    n[..., 2] = 1.0
    n[..., 3] = 1.0
    vd, nd = O.resize_map(v, n)
    assert np.array_equal(vd, v.reshape(2, 2, 3, 2, 4).mean(axis=(1, 3)))
    # w = 1 takes part in the normalisation: (0, 0, 1, 1) / sqrt(2)
    assert np.allclose(nd, [0, 0, 1 / np.sqrt(2), 1 / np.sqrt(2)], atol=1e-7)
    v[1, 1, 0] = np.nan
    vd, _ = O.resize_map(v, n)
    assert np.all(np.isnan(vd[0, 0])) and not np.isnan(vd[0, 1, 0])


# ----------------------------------------------------------------------- hash and allocation
def test_hash_wraps_and_reduces_unsigned():
    for (x, y, z, nb) in [(1, 2, 3, 1000003), (-1, 0, 0, 1 << 21), (-5, 7, -9, 999983), (100000, -100000, 1, 12345)]:
        h = ((x * 73856093) ^ (y * 19349669) ^ (z * 83492791)) & 0xFFFFFFFF
        assert O.lib().or_hash(x, y, z, nb) == h % nb
    # a signed remainder would have been negative here and shifted into another bucket
    assert O.lib().or_hash(-1, 0, 0, 1000000) != ((-73856093) % 1000000)


def _single_ray_map(depth_m, vox=0.01, tv=3.0, **kw):
    m = O.Map(voxel_size=vox, trunc_dist=float(np.float32(vox) * np.float32(tv)), **kw)
    d = np.full((1, 1), np.nan, np.float32)
    d[0, 0] = depth_m
    # principal point at the pixel: the ray is the optical axis
    m.create_blocks(d, np.eye(4), (500.0, 500.0, 0.0, 0.0))
    return m


def test_allocation_along_the_optical_axis():
    """samples at most half a voxel apart cover every voxel of [z - T/2, z + T/2]"""
    for z in (0.5, 1.237, 2.9):
        m = _single_ray_map(z)
        T = np.float32(0.01) * np.float32(3.0)
        lo = int(np.floor((z - T / 2) / 0.01 + 1e-6))
        hi = int(np.floor(min(3.0, z + T / 2) / 0.01 - 1e-6))
        want = {(0, 0, v // 8) for v in range(lo, hi + 1)}
        got = set(map(tuple, m.export_blocks().tolist()))
        assert got == want, (z, got, want)


def test_allocation_independent_of_table_shape():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    d16, _ = sc.render(2, 160, 120, (131.25, 131.25, 79.5, 59.5))
    raw, _ = O.filter_depth(d16, 5000.0, 3.0)
    Kq = (131.25, 131.25, 79.5, 59.5)
    sets, cnt = [], []
    for nb, nx in ((1, 1 << 14), (7, 1 << 14), (1 << 16, 1 << 10)):
        m = O.Map(num_buckets=nb, num_excess=nx, num_blocks=1 << 14)
        m.create_blocks(raw, np.eye(4), Kq)
        b = m.export_blocks()
        sets.append(set(map(tuple, b.tolist())))
        cnt.append(m.counters())
        assert len(sets[-1]) == len(b)  # no duplicates
    assert sets[0] == sets[1] == sets[2]
    n = len(sets[0])
    assert cnt[0]["n_excess"] == n - 1 and cnt[0]["n_allocated"] == n
    buckets = {O.lib().or_hash(*p, 7) for p in sets[1]}
    assert cnt[1]["n_excess"] == n - len(buckets)


def test_heap_and_excess_exhaustion():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    d16, _ = sc.render(2, 160, 120, (131.25, 131.25, 79.5, 59.5))
    raw, _ = O.filter_depth(d16, 5000.0, 3.0)
    Kq = (131.25, 131.25, 79.5, 59.5)
    full = O.Map(num_buckets=1 << 16, num_excess=1 << 14, num_blocks=1 << 14)
    full.create_blocks(raw, np.eye(4), Kq)
    allb = set(map(tuple, full.export_blocks().tolist()))
    small = O.Map(num_buckets=1 << 16, num_excess=1 << 14, num_blocks=50)
    small.create_blocks(raw, np.eye(4), Kq)
    got = set(map(tuple, small.export_blocks().tolist()))
    assert len(got) == 50 and got <= allb and small.counters()["heap_full"] > 0
    ex = O.Map(num_buckets=1, num_excess=10, num_blocks=1 << 14)
    ex.create_blocks(raw, np.eye(4), Kq)
    assert len(ex.export_blocks()) == 11 and ex.counters()["excess_full"] > 0


def test_visibility_brute_force():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    d16, _ = sc.render(0)
    raw, _ = O.filter_depth(d16, 5000.0, 3.0)
    m = O.Map(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15)
    m.create_blocks(raw, np.eye(4), K)
    P = sc.pose(0)
    pose = np.linalg.inv(P) @ sc.pose(25)
    n = m.check_visibility(pose, 640, 480, K)
    vis = set(map(tuple, m.export_visible().tolist()))
    assert n == len(vis) > 0
    Rinv = pose[:3, :3].T
    t = pose[:3, 3]
    sure_in, sure_out = 0, 0
    for b in map(tuple, m.export_blocks().tolist()):
        c = np.array([[b[0] + (i & 1), b[1] + ((i >> 1) & 1), b[2] + ((i >> 2) & 1)] for i in range(8)]) * 0.08
        p = (c - t) @ Rinv.T
        u = K[0] * p[:, 0] / p[:, 2] + K[2]
        v = K[1] * p[:, 1] / p[:, 2] + K[3]
        inside = (p[:, 2] > 1e-3) & (u >= 0) & (v >= 0) & (u < 640) & (v < 480) & (p[:, 2] >= 0.1) & (p[:, 2] <= 3.0)
        margin = np.min(np.abs(np.concatenate([u, u - 640, v, v - 480, p[:, 2] - 0.1, p[:, 2] - 3.0])))
        if margin < 1e-3:
            continue  # too close to a boundary for float32 vs float64
        assert (b in vis) == bool(inside.any())
        sure_in += b in vis
        sure_out += b not in vis
    assert sure_in > 100 and sure_out > 100


# ----------------------------------------------------------------------- integration
def test_integration_of_a_fronto_parallel_wall():
    """constant depth Z, identity pose: sdf = min(1, (Z - z)/T) for voxels with z <= Z + T,
    weight = number of frames, colour of a never-updated voxel = the first frame's"""
    H, W = 60, 80
    Kq = (100.0, 100.0, 39.5, 29.5)
    Z = np.float32(1.0)
    depth = np.full((H, W), Z, np.float32)
    rgb = np.full((H, W, 3), 90, np.uint8)
    v = O.compute_vmap(depth, *Kq, 3.0)
    n = O.compute_nmap(v)
    m = O.Map(num_buckets=1 << 14, num_excess=1 << 12, num_blocks=1 << 12)
    for frame in range(3):
        m.fuse_map_color(depth, rgb, n, np.eye(4), Kq)
    blocks = m.export_blocks()
    sdf, w, col = m.read_blocks(blocks)
    T = np.float32(0.01) * np.float32(3.0)
    checked = 0
    for bi, b in enumerate(blocks):
        for li in range(512):
            x, y, z = b[0] * 8 + li % 8, b[1] * 8 + (li // 8) % 8, b[2] * 8 + li // 64
            p = np.array([x, y, z], np.float64) * 0.01
            u, vv = Kq[0] * p[0] / p[2] + Kq[2], Kq[1] * p[1] / p[2] + Kq[3]
            if not (1 <= u < W - 2 and 1 <= vv < H - 2):
                continue  # stay clear of the image border and the NaN normal column
            if abs(Z - p[2] + T) < 1e-4:
                continue  # on the truncation boundary: float32 vs float64
            if Z - p[2] < -T:
                assert w[bi, li] == 0
                continue
            checked += 1
            assert w[bi, li] == 3
            assert abs(sdf[bi, li] - min(1.0, (Z - p[2]) / T)) < 2e-5
    assert checked > 1000


def test_weight_cap():
    H, W = 20, 20
    Kq = (50.0, 50.0, 9.5, 9.5)
    depth = np.full((H, W), 1.0, np.float32)
    n = O.compute_nmap(O.compute_vmap(depth, *Kq, 3.0))
    m = O.Map(num_buckets=1 << 12, num_excess=1 << 10, num_blocks=1 << 10, max_weight=7)
    for _ in range(12):
        m.fuse_map_color(depth, np.zeros((H, W, 3), np.uint8), n, np.eye(4), Kq)
    _, w, _ = m.read_blocks(m.export_blocks())
    assert w.max() == 7


# ----------------------------------------------------------------------- ICP
def test_solve6_and_se3_exp_closed_forms():
    rng = np.random.default_rng(3)
    for _ in range(20):
        M = rng.normal(size=(6, 6))
        A = M @ M.T + 0.1 * np.eye(6)
        b = rng.normal(size=6)
        assert np.allclose(O.solve6(A, b), np.linalg.solve(A, b), rtol=1e-10, atol=1e-12)
    for scale in (1e-9, 1e-4, 0.05, 1.0):
        xi = rng.normal(size=6) * scale
        up, w = xi[:3], xi[3:]
        twist = np.zeros((4, 4))
        twist[:3, :3] = [[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]]
        twist[:3, 3] = up  # tangent order (upsilon, omega): translation first
        assert np.allclose(O.se3_exp(xi), expm(twist), atol=1e-12)


def _pyr(d16, Kq):
    _, f0 = O.filter_depth(d16, 5000.0, 3.0)
    ds = [f0, O.pyrdown_gauss(f0)]
    ds.append(O.pyrdown_gauss(ds[1]))
    vs, ns = [], []
    for l in range(3):
        s = np.float32(1 << l)
        v = O.compute_vmap(ds[l], *(np.float32(k) / s for k in Kq), 3.0)
        vs.append(v)
        ns.append(O.compute_nmap(v))
    return vs, ns


def test_icp_recovers_a_known_motion():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    P = sc.pose(10)
    d0, _ = sc.render_pose(P)
    v0, n0 = _pyr(d0, K)
    dT = O.se3_exp(np.array([0.012, -0.006, 0.009, 0.004, -0.008, 0.006]))
    d1, _ = sc.render_pose(P @ dT)
    v1, n1 = _pyr(d1, K)
    st, pose, err, cnt = O.compute_se3(v1, n1, v0, n0, K, np.eye(4))
    assert st == 1 and cnt > 100000
    assert np.abs(pose - dT).max() < 5e-4
    # zero motion: the identity, exactly up to rounding
    st, pose, _, _ = O.compute_se3(v0, n0, v0, n0, K, np.eye(4))
    assert st == 1 and np.abs(pose - np.eye(4)).max() < 1e-6


def test_icp_sums_fp32_close_to_fp64():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    d0, _ = sc.render(0)
    d1, _ = sc.render(2)
    v0, n0 = _pyr(d0, K)
    v1, n1 = _pyr(d1, K)
    s32, s64 = O.icp_sums(v1[0], v0[0], n1[0], n0[0], np.eye(4), np.eye(4), K)
    assert s32[28] == s64[28] > 100000
    assert np.all(np.abs(s32 - s64) <= 1e-3 * np.abs(s64).max())


# ----------------------------------------------------------------------- raycast and mesh
def test_raycast_lies_on_the_scene_surface():
    sc = Scene(0, (4.0, 4.0, 3.0), 0)
    P0 = sc.pose(0)
    m = O.Map(num_buckets=1 << 17, num_excess=1 << 16, num_blocks=1 << 16, voxel_size=0.01)
    for f in (0, 2, 4):
        d16, rgb = sc.render(f, noise=False)
        raw, filt = O.filter_depth(d16, 5000.0, 3.0)
        n = O.compute_nmap(O.compute_vmap(filt, *K, 3.0))
        m.fuse_map_color(raw, rgb, n, np.linalg.inv(P0) @ sc.pose(f), K)
    pose = np.linalg.inv(P0) @ sc.pose(4)
    nv = m.counters()["n_visible"]
    zmin, zmax, nf, ok = m.create_rendering_blocks(80, 60, pose, nv, K)
    assert ok and nf >= nv * 0.9
    v, n = m.raycast(640, 480, zmin, zmax, pose, K)
    valid = ~np.isnan(v[..., 0])
    assert valid.mean() > 0.5
    pts = v[valid][:, :3].astype(np.float64)
    world = (P0 @ pose @ np.c_[pts, np.ones(len(pts))].T).T[:, :3]
    sel = np.random.default_rng(0).choice(len(world), 2000, replace=False)
    dist = np.abs([sc.sdf(*p) for p in world[sel]])
    # the reference's half-step refinement and truncation-vs-floor lookups leave the hit
    # up to about a voxel off the surface; normals point towards the camera
    assert np.median(dist) < 0.01 and np.percentile(dist, 90) < 0.02
    nn = n[valid][:, :3]
    assert np.median(np.sum(nn * pts, axis=1) / np.linalg.norm(pts, axis=1)) < -0.5


def test_marching_cubes_sphere_is_a_closed_surface():
    m = O.Map(num_buckets=1 << 14, num_excess=1 << 12, num_blocks=1 << 12, voxel_size=0.01)
    c, r = (0.123, -0.207, 0.5), 0.13
    m.set_sphere(c, r)
    vert, norm, col, meta, total = m.mesh()
    assert total == len(vert) > 500
    d = np.linalg.norm(vert.reshape(-1, 3) - np.array(c), axis=1)
    assert np.abs(d - r).max() < 0.01 * 0.6
    # closed 2-manifold of genus 0: V - E + F = 2 after merging shared vertices
    key = np.round(vert.reshape(-1, 3) / 1e-6).astype(np.int64)
    _, idx = np.unique(key, axis=0, return_inverse=True)
    tris = idx.reshape(-1, 3)
    edges = np.sort(np.concatenate([tris[:, [0, 1]], tris[:, [1, 2]], tris[:, [2, 0]]]), axis=1)
    ue, counts = np.unique(edges, axis=0, return_counts=True)
    assert np.all(counts == 2)
    assert idx.max() + 1 - len(ue) + len(tris) == 2
    # outward normals (the sdf gradient) and the voxel colour
    nrm = norm.reshape(-1, 3)
    radial = (vert.reshape(-1, 3) - np.array(c)) / d[:, None]
    assert np.min(np.sum(nrm * radial, axis=1)) > 0.9
    assert np.all(col.reshape(-1, 3) == [100, 150, 200])


# ----------------------------------------------------------------------- photometric term
def test_intensity_is_the_truncated_luma_of_every_colour():
    """all 2^24 colours against exact rational arithmetic; where the exact luma is an integer
    (63 colours) the double expression may land just below it, and there the oracle must
    equal numpy's uncontracted double evaluation in the same order"""
    v = np.arange(1 << 24, dtype=np.uint32)
    rgb = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], axis=1).astype(np.uint8).reshape(4096, 4096, 3)
    out = O.image_to_intensity(rgb).ravel().astype(np.int64)
    c = rgb.reshape(-1, 3).astype(np.int64)
    num = 2125 * c[:, 1] + 7154 * c[:, 0] + 721 * c[:, 2]  # 0.2125 G + 0.7154 R + 0.0721 B (channel 0 = R)
    exact = num // 10000
    tie = (num % 10000) == 0
    assert np.array_equal(out[~tie], exact[~tie])
    cf = c[tie].astype(np.float64)
    dbl = (0.2125 * cf[:, 1] + 0.7154 * cf[:, 0]) + 0.0721 * cf[:, 2]
    assert np.array_equal(out[tie], np.trunc(dbl).astype(np.int64))
    assert ((out[tie] == exact[tie]) | (out[tie] == exact[tie] - 1)).all()
    assert int(tie.sum()) > 0 and int((out[tie] != exact[tie]).sum()) == 63


def test_pyrdown_u8_is_the_truncated_float_pyramid():
    """interior: the binomial 5x5 mean (scipy), truncated; everywhere, including the mirrored
    border weights: the float pyramid (pinned above) of the same bytes, truncated"""
    from scipy.ndimage import correlate
    rng = np.random.default_rng(4)
    src = rng.integers(0, 256, size=(36, 52), dtype=np.uint8)
    dst = O.pyrdown_u8(src)
    assert dst.shape == (18, 26) and dst.dtype == np.uint8
    assert np.array_equal(dst, np.trunc(O.pyrdown_gauss(src.astype(np.float32))).astype(np.uint8))
    g = np.array([1, 4, 6, 4, 1], np.float64)
    full = correlate(src.astype(np.float64), np.outer(g, g) / 256.0, mode="constant")
    ys, xs = np.mgrid[1:17, 1:25]
    ref = np.floor(full[2 * ys, 2 * xs] + 1e-9).astype(np.uint8)
    assert np.array_equal(dst[1:17, 1:25], ref)
    assert np.array_equal(O.pyrdown_u8(np.full((20, 20), 77, np.uint8)), np.full((10, 10), 77, np.uint8))


def test_derivative_image_sobel_interior_and_clipped_border():
    from scipy.ndimage import correlate
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, size=(17, 23), dtype=np.uint8)
    dx, dy = O.derivative_image(img)
    assert dx.dtype == np.int16 and dy.dtype == np.int16
    kx = np.array([[-1, 0, 1], [-2, 0, 2], [-1, 0, 1]], np.float64)  # +x to the right
    ix = correlate(img.astype(np.float64), kx, mode="constant")
    iy = correlate(img.astype(np.float64), kx.T, mode="constant")
    assert np.array_equal(dx[1:-1, 1:-1], ix[1:-1, 1:-1].astype(np.int16))
    assert np.array_equal(dy[1:-1, 1:-1], iy[1:-1, 1:-1].astype(np.int16))
    # a ramp a x + b y: 8a, 8b inside
    yy, xx = np.mgrid[0:12, 0:16]
    rdx, rdy = O.derivative_image((3 * xx + 5 * yy).astype(np.uint8))
    assert (rdx[1:-1, 1:-1] == 24).all() and (rdy[1:-1, 1:-1] == 40).all()
    # top-left corner: the 2x2 window takes kernel entries 8, 7, 6, 5 (counted down in window
    # order), i.e. dx = -v00 + v10 - 2 v11 and dy = -v00 - 2 v01 - v10 (Pyrdown.cu:226-233)
    v = img.astype(np.int64)
    assert dx[0, 0] == -v[0, 0] + v[1, 0] - 2 * v[1, 1]
    assert dy[0, 0] == -v[0, 0] - 2 * v[0, 1] - v[1, 0]


def _plane_frame(t, cols=640, rows=480, Kq=K, Z=1.5):
    """a fronto-parallel plane z = Z (world) with a smooth grey texture, seen by a camera at
    translation t (no rotation): depth u16 at 5000/m and the colour image"""
    v, u = np.mgrid[0:rows, 0:cols].astype(np.float64)
    X = t[0] + (Z - t[2]) * (u - Kq[2]) / Kq[0]
    Y = t[1] + (Z - t[2]) * (v - Kq[3]) / Kq[1]
    inten = 128.0 + 50.0 * np.sin(2 * np.pi * X / 0.15) * np.cos(2 * np.pi * Y / 0.2)
    g = np.clip(np.floor(inten + 0.5), 1, 255).astype(np.uint8)
    d = np.full((rows, cols), int(round((Z - t[2]) * 5000)), np.uint16)
    return d, np.repeat(g[:, :, None], 3, axis=2)


def _plane_frame_pose(T, cols=640, rows=480, Kq=K, Z=1.5):
    """the textured plane z = Z seen by a camera at any pose T (ray-plane intersection)"""
    v, u = np.mgrid[0:rows, 0:cols].astype(np.float64)
    dc = np.stack([(u - Kq[2]) / Kq[0], (v - Kq[3]) / Kq[1], np.ones_like(u)], -1)
    d = dc @ T[:3, :3].T
    o = T[:3, 3]
    s = (Z - o[2]) / d[..., 2]  # the camera-frame ray has z = 1, so s is the depth
    X = o[0] + s * d[..., 0]
    Y = o[1] + s * d[..., 1]
    inten = 128.0 + 50.0 * np.sin(2 * np.pi * X / 0.15) * np.cos(2 * np.pi * Y / 0.2)
    g = np.clip(np.floor(inten + 0.5), 1, 255).astype(np.uint8)
    return np.floor(s * 5000 + 0.5).astype(np.uint16), np.repeat(g[:, :, None], 3, axis=2)


def _images(rgb):
    im = [O.image_to_intensity(rgb)]
    im += [O.pyrdown_u8(im[0])]
    im += [O.pyrdown_u8(im[1])]
    der = [O.derivative_image(i) for i in im]
    return im, [d[0] for d in der], [d[1] for d in der]


def test_rgb_residuals_of_a_frame_against_itself():
    """identity pose and the same frame twice: every eligible pixel matches itself with zero
    difference, so sigma = 0, every weight is 1 and b = 0"""
    d, rgb = _plane_frame((0.0, 0.0, 0.0))
    vs, ns = _pyr(d, K)
    im, dx, dy = _images(rgb)
    (cnt, sq), sigma, s32, s64 = O.rgb_sums(im[0], im[0], vs[0], vs[0], dx[0], dy[0], np.eye(4), np.eye(4), K)
    rows, cols = im[0].shape
    yy, xx = np.mgrid[0:rows, 0:cols]
    inner = (xx >= 5) & (xx < cols - 5) & (yy >= 5) & (yy < rows - 5)
    win = np.ones_like(inner)
    for a in range(-2, 2):  # the 4x4 window [y-2, y+2) x [x-2, x+2)
        for b in range(-2, 2):
            win &= np.roll(np.roll(im[0], -a, 0), -b, 1) > 0
    grad = (dx[0] != 0) | (dy[0] != 0)
    assert cnt == int((inner & win & grad & ~np.isnan(vs[0][..., 0])).sum()) > 200000
    assert sq == 0 and sigma == 0.0
    assert s64[28] == cnt
    A, b = O.unpack_sums(s64)
    assert np.all(b == 0) and np.all(np.linalg.eigvalsh(A) >= -1e-6)


def test_rgb_term_recovers_in_plane_motion_that_icp_cannot():
    """a textured plane: point-to-plane ICP leaves the in-plane translation free (its normal
    matrix is singular there), the photometric term pins it; the blended tracker must recover
    a known sideways motion.  The residual pairs pixels by nearest-pixel projection, so every
    pose within half a level-0 pixel of the truth (1.43 mm here) has the same residuals: the
    bound is 0.6 pixel."""
    t = np.array([0.012, -0.009, 0.0])
    d0, c0 = _plane_frame((0.0, 0.0, 0.0))
    d1, c1 = _plane_frame(t)
    v0, n0 = _pyr(d0, K)
    v1, n1 = _pyr(d1, K)
    i0, _, _ = _images(c0)
    i1, dx1, dy1 = _images(c1)
    s32, s64 = O.icp_sums(v1[0], v0[0], n1[0], n0[0], np.eye(4), np.eye(4), K)
    A, _ = O.unpack_sums(s64)
    ev = np.linalg.eigvalsh(A)
    assert ev[2] < 1e-6 * ev[-1]  # tx, ty and the roll are unobservable by ICP
    st, pose, err, cnt, rcnt = O.compute_se3_rgbd(v1, n1, v0, n0, i1, i0, dx1, dy1, K, np.eye(4))
    px = 1.5 / K[0]
    assert st == 1 and rcnt > 100
    assert np.abs(pose[:3, 3] - t).max() < 0.6 * px, pose[:3, 3]
    assert np.abs(pose[:3, :3] - np.eye(3)).max() < 2e-3
    # the sign of the photometric rows matters: the reversed motion is recovered as well
    st, pose, _, _, _ = O.compute_se3_rgbd(v0, n0, v1, n1, i0, i1, *_images(c0)[1:], K, np.eye(4))
    assert st == 1 and np.abs(pose[:3, 3] + t).max() < 0.6 * px, pose[:3, 3]


def test_rgb_rows_alone_recover_a_6dof_motion():
    """Gauss-Newton on the photometric sums alone (numpy solve, the oracle's exponential),
    coarse to fine: every column of the Jacobian rows and the sign of the residual take part"""
    xi = np.array([0.006, -0.004, 0.005, np.radians(0.6), np.radians(-0.5), np.radians(0.8)])
    T = O.se3_exp(xi)
    d0, c0 = _plane_frame_pose(np.eye(4))
    d1, c1 = _plane_frame_pose(T)
    v0, _ = _pyr(d0, K)
    v1, _ = _pyr(d1, K)
    i0, _, _ = _images(c0)
    i1, dx1, dy1 = _images(c1)
    last, delta, nxt = np.eye(4), np.eye(4), np.eye(4)
    for l, iters in ((2, 10), (1, 5), (0, 4)):
        Kl = [float(np.float32(k) / np.float32(1 << l)) for k in K]
        for _ in range(iters):
            _, _, _, s64 = O.rgb_sums(i1[l], i0[l], v1[l], v0[l], dx1[l], dy1[l], nxt, last, Kl)
            A, b = O.unpack_sums(s64)
            delta = O.se3_exp(np.linalg.solve(A, b)) @ delta
            nxt = last @ np.linalg.inv(delta)
    assert np.abs(nxt[:3, 3] - T[:3, 3]).max() < 5e-4
    assert np.degrees(np.abs(nxt[:3, :3] - T[:3, :3]).max()) < 0.03


def test_rgb_sums_fp32_close_to_fp64_and_failure_on_no_texture():
    d0, c0 = _plane_frame((0.0, 0.0, 0.0))
    d1, c1 = _plane_frame((0.01, 0.0, 0.0))
    v0, _ = _pyr(d0, K)
    v1, _ = _pyr(d1, K)
    i0, _, _ = _images(c0)
    i1, dx1, dy1 = _images(c1)
    (cnt, sq), sigma, s32, s64 = O.rgb_sums(i1[0], i0[0], v1[0], v0[0], dx1[0], dy1[0], np.eye(4), np.eye(4), K)
    assert cnt > 100000 and sq > 0 and sigma == np.float32(np.sqrt(np.float32(sq) / np.float32(cnt)))
    assert np.all(np.abs(s32 - s64) <= 1e-3 * np.abs(s64).max())
    # a flat grey image has no gradient: no residuals, RGBStep fails the frame
    flat = np.full_like(c1, 90)
    fi, fdx, fdy = _images(flat)
    v0n, n0n = _pyr(d0, K)
    st, pose, _, _, rcnt = O.compute_se3_rgbd(v0n, n0n, v0n, n0n, fi, fi, fdx, fdy, K, np.eye(4))
    assert st == 2 and rcnt == 0 and np.array_equal(pose, np.eye(4))


def test_sobel_scale_float_division_equals_double_rounding():
    """the GPU divides the Sobel responses by 9 in float; the reference computes
    (float)((double)s / 9.0): the two agree for every short s"""
    s = np.arange(-32768, 32768, dtype=np.int64)
    ref = (s.astype(np.float64) / 9.0).astype(np.float32)
    gpu = s.astype(np.float32) / np.float32(9.0)
    assert np.array_equal(ref.view(np.uint32), gpu.view(np.uint32))


def test_squared_norm_thresholds_of_the_gpu_association():
    """the GPU compares squared norms with these bounds instead of taking square roots:
    sqrtf(s) < 0.6f <=> s < T1, sqrtf(d) <= 0.1f <=> d <= T2, (double)sqrtf(e) < 0.1 <=> e <= T3
    (square roots in float are correctly rounded and monotonic, as numpy's are)"""
    T1 = np.float32(float.fromhex("0x1.70a3d8p-2"))
    T2 = np.float32(float.fromhex("0x1.47ae16p-7"))
    T3 = np.float32(float.fromhex("0x1.47ae12p-7"))
    up = lambda x: np.nextafter(x, np.float32(1))
    dn = lambda x: np.nextafter(x, np.float32(0))
    assert np.sqrt(T1) >= np.float32(0.6) and np.sqrt(dn(T1)) < np.float32(0.6)
    assert np.sqrt(T2) <= np.float32(0.1) and np.sqrt(up(T2)) > np.float32(0.1)
    assert float(np.sqrt(T3)) < 0.1 and float(np.sqrt(up(T3))) >= 0.1
    # and on random values around and away from the bounds
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 1, 100000).astype(np.float32),
                        (T1.view(np.int32) + np.arange(-500, 500)).astype(np.int32).view(np.float32),
                        (T2.view(np.int32) + np.arange(-500, 500)).astype(np.int32).view(np.float32),
                        (T3.view(np.int32) + np.arange(-500, 500)).astype(np.int32).view(np.float32)])
    r = np.sqrt(x)
    assert np.array_equal(r < np.float32(0.6), x < T1)
    assert np.array_equal(r <= np.float32(0.1), x <= T2)
    assert np.array_equal(r.astype(np.float64) < 0.1, x <= T3)
