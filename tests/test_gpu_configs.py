"""The benchmark configurations as GPU parity tests against the CPU oracle.

small scene (System::GrabImage, Core/System.cc:92-138): 30 frames of 640x480 of the 4x4x3 m
    room at 10 mm voxels (truncation 3 voxels, 2^18 buckets, 200k blocks), ICP 10/5/4, a raycast
    every frame and a marching-cubes mesh at the end.
single stream at scale: the 12x8x3 m multi-room scene at 5 mm voxels (truncation 8 voxels,
    2^21 buckets and 2M blocks) in the three voxel formats.
batched ICP: 256 frame pairs of uneven size (320x240, 640x480, 1280x720 drawn uniformly) from
    the small scene with seeded relative poses of up to 5 cm and 3 degrees.

The GPU runs the pipeline on its own.  The oracle is driven from the GPU's state so that its
result can be compared exactly: each frame's tracking from the same maps and last pose as the
GPU's (pose within 1e-5 rad and 1e-5 m of the oracle's FP64 result), the map fused and raycast
with the GPU's tracked pose (block set, voxels, model maps and mesh bit-exact).
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.helpers import K640, rows_set, same_bits

pytestmark = pytest.mark.gpu


def _rot_err(Ra, Rb):
    return float(np.arccos(np.clip((np.trace(Ra.T @ Rb) - 1) / 2, -1, 1)))


def _pyramid(d16, K):
    raw, f0 = O.filter_depth(d16, 5000.0, 3.0)
    ds = [f0, O.pyrdown_gauss(f0)]
    ds.append(O.pyrdown_gauss(ds[1]))
    vs, ns = [], []
    for l in range(3):
        s = np.float32(1 << l)
        v = O.compute_vmap(ds[l], *(np.float32(k) / s for k in K), 3.0)
        vs.append(v)
        ns.append(O.compute_nmap(v))
    return raw, vs, ns


def _model_pyramid(v0, n0):
    vs, ns = [v0], [n0]
    for _ in range(2):
        v, n = O.resize_map(vs[-1], ns[-1])
        vs.append(v)
        ns.append(n)
    return vs, ns


def _run_and_check(sc, frames, map_params, oracle_map, K, zc, zr, twin=None):
    """the GPU system over `frames`; per frame the oracle's tracking from the GPU's inputs and
    its fusion and raycast with the GPU's pose.  Returns the GPU system and the poses."""
    s = A.System(A.system_params(map=map_params))
    last_pose = np.eye(4)
    model = None
    poses = []
    for i, f in enumerate(frames):
        d16, rgb = sc.render(f)
        raw, vs, ns = _pyramid(d16, K)
        if i > 0:
            mv, mn = _model_pyramid(*model)
            st_o, pose_o, _, _ = O.compute_se3(vs, ns, mv, mn, K, last_pose)
        s.GrabImage(*A.to_device(d16, rgb))
        r = s.result()
        if i == 0:
            assert r["track_status"] == 4 and np.array_equal(r["pose"], np.eye(4))
        else:
            assert r["track_status"] == st_o, (f, r["track_status"], st_o)
            assert _rot_err(r["pose"][:3, :3], pose_o[:3, :3]) <= 1e-5, f
            assert np.linalg.norm(r["pose"][:3, 3] - pose_o[:3, 3]) <= 1e-5, f
        if r["track_status"] not in (1, 4):
            continue
        pose = r["pose"]
        poses.append(pose)
        last_pose = pose
        nv = oracle_map.fuse_map_color(raw, rgb, ns[0], pose, K)
        if twin is not None:
            twin.fuse_map_color(raw, rgb, ns[0], pose, K)
        zmin, zmax, nf, ok = oracle_map.create_rendering_blocks(zc, zr, pose, nv, K)
        assert r["n_visible"] == nv and r["n_fragments"] == nf, (f, r["n_visible"], nv, r["n_fragments"], nf)
        vo, no = oracle_map.raycast(640, 480, zmin, zmax, pose, K) if ok else (vs[0], ns[0])
        vg = s.view("model_vmap", 0).cpu().numpy()
        ng = s.view("model_nmap", 0).cpu().numpy()
        assert same_bits(vg, vo) and same_bits(ng, no), f
        model = (vo, no)
    return s, poses


def _compare_voxels(g, o, blocks):
    sg, wg, cg = g.read_blocks(blocks)
    so, wo, co = o.read_blocks(blocks)
    assert np.array_equal(wg, wo) and np.array_equal(cg, co)
    assert same_bits(sg, so)


@pytest.mark.timeout(600)
def test_small_scene_30_frames_and_mesh(gpu):
    print("small scene: 30 frames + mesh", flush=True)
    vox = 0.01
    mp = dict(num_buckets=1 << 18, num_excess=1 << 17, num_blocks=200_000, voxel_size=vox,
              trunc_dist=float(np.float32(vox) * np.float32(3.0)), max_new_blocks=1 << 16)
    sc = A.Scene(0, (4.0, 4.0, 3.0), 0)
    o = O.Map(num_buckets=mp["num_buckets"], num_excess=mp["num_excess"], num_blocks=mp["num_blocks"], voxel_size=vox,
              trunc_dist=mp["trunc_dist"])
    s, poses = _run_and_check(sc, range(30), mp, o, K640, 80, 60)
    assert len(poses) >= 27
    g = s.mapping()
    ob = o.export_blocks()
    assert rows_set(g.export_blocks()) == rows_set(ob) and len(ob) > 2000
    _compare_voxels(g, o, ob)
    vg, ng, cg, mg = g.MeshScene(max_triangles=2_000_000, with_meta=True)
    vo, no, co, mo, total = o.mesh()
    assert mg.shape[0] == total == mo.shape[0] > 20000
    key = lambda m: np.lexsort(m.T[::-1])
    kg, ko = key(mg.cpu().numpy()), key(mo)
    assert np.array_equal(mg.cpu().numpy()[kg], mo[ko])  # topology: (voxel, cube index, triangle) set
    assert same_bits(vg.cpu().numpy()[kg], vo[ko]) and same_bits(ng.cpu().numpy()[kg], no[ko])
    assert np.array_equal(cg.cpu().numpy()[kg], co[ko])


@pytest.mark.timeout(900)
def test_multiroom_at_scale_parity_and_fp64_error(gpu):
    """2^21 buckets and 2M blocks at 5 mm: tracking, block set, voxels and raycasts exact for
    six frames of the multi-room scene; then the same frames at the tracked poses in each voxel
    format against the FP64 re-integration (bar: FP32 forms within 1e-5 of the truncation
    distance, FP16 within 2^-8; weights exact)"""
    print("multi-room 12x8x3 m at 5 mm, 2^21 buckets, 2M blocks", flush=True)
    vox = 0.005
    mp = dict(num_buckets=1 << 21, num_excess=1 << 21, num_blocks=2_000_000, voxel_size=vox,
              trunc_dist=float(np.float32(vox) * np.float32(8.0)), max_new_blocks=1 << 18)
    sc = A.Scene(1, (12.0, 8.0, 3.0), 0)
    frames = range(0, 12, 2)
    o = O.Map(num_buckets=mp["num_buckets"], num_excess=mp["num_excess"], num_blocks=mp["num_blocks"], voxel_size=vox,
              trunc_dist=mp["trunc_dist"])
    twin = O.Map(num_buckets=mp["num_buckets"], num_excess=mp["num_excess"], num_blocks=mp["num_blocks"],
                 voxel_size=vox, trunc_dist=mp["trunc_dist"], precision=1)
    s, poses = _run_and_check(sc, frames, mp, o, K640, 80, 60, twin=twin)
    assert len(poses) == len(frames)
    g = s.mapping()
    ob = o.export_blocks()
    assert rows_set(g.export_blocks()) == rows_set(ob)
    _compare_voxels(g, o, ob)
    s.close()
    del g, o
    s64 = twin.read_blocks_f64(ob)
    _, w64, _ = twin.read_blocks(ob)
    m = ~np.isnan(s64)
    print("%d blocks, %d observed voxels" % (len(ob), int(m.sum())), flush=True)
    for fmt, bound in ((0, 1e-5), (1, 1e-5), (2, 2.0 ** -8)):
        gm = A.Mapping(**mp, voxel_format=fmt)
        for f, pose in zip(frames, poses):
            d16, rgb = sc.render(f)
            raw, vs, ns = _pyramid(d16, K640)
            gm.FuseMapColor(torch.from_numpy(raw).cuda(), torch.from_numpy(rgb).cuda(),
                            torch.from_numpy(ns[0]).cuda(), pose, K640)
        assert rows_set(gm.export_blocks()) == rows_set(ob)
        sg, wg, _ = gm.read_blocks(ob)
        assert np.array_equal(wg, w64)
        assert np.array_equal(np.isnan(sg), ~m)
        err = np.abs(sg[m].astype(np.float64) - s64[m])
        print("format %d: max |sdf - sdf64| = %.3g T, mean %.3g T, 99.9%% %.3g T" % (
            fmt, err.max(), err.mean(), np.percentile(err, 99.9)), flush=True)
        assert err.max() <= bound
        gm.close()


def _se3(rng, tmax=0.05, amax=np.deg2rad(3.0)):
    """a rigid motion with translation and rotation angle drawn uniformly up to the bounds"""
    t = rng.normal(size=3)
    t *= rng.uniform(0, tmax) / np.linalg.norm(t)
    w = rng.normal(size=3)
    w /= np.linalg.norm(w)
    a = rng.uniform(0, amax)
    Wx = np.array([[0, -w[2], w[1]], [w[2], 0, -w[0]], [-w[1], w[0], 0]])
    T = np.eye(4)
    T[:3, :3] = np.eye(3) + np.sin(a) * Wx + (1 - np.cos(a)) * Wx @ Wx
    T[:3, 3] = t
    return T


@pytest.mark.timeout(900)
def test_batched_icp_256_uneven_pairs(gpu):
    """every pair's batched ComputeSE3 equals the tracker alone bit for bit; every ICP step of
    the oracle's iteration, rerun on the GPU from the same pose, gives an update
    x = (upsilon, omega) within 1e-5 m / 1e-5 rad of the oracle's (the bar); the final pose
    after the 19 iterations within 1e-5 rad / 1e-5 m of the FP64 oracle for at least 95 % of the
    pairs and within 1e-4 for all.  The final pose compounds the steps: where a pair's
    association set shifts with the pose in the last digits, the GPU's and the oracle's
    iterations part by up to a few 1e-5 while each step agrees to 1e-8 (pairs 84 and 127)"""
    rng = np.random.default_rng(2024)
    sizes = [((320, 240), (640, 480), (1280, 720))[k] for k in rng.integers(0, 3, 256)]
    sc = A.Scene(0, (4.0, 4.0, 3.0), 0)
    cu = lambda xs: [torch.from_numpy(x).cuda() for x in xs]
    pairs, trackers = [], []
    for i, (cols, rows) in enumerate(sizes):
        s = cols / 640.0
        K = (525.0 * s, 525.0 * s, (cols - 1) / 2.0, (rows - 1) / 2.0)
        P = sc.pose(int(rng.integers(0, 200)))
        dT = _se3(rng)
        d0, _ = sc.render_pose(P, cols, rows, K, noise=True, noise_seed=2 * i)
        d1, _ = sc.render_pose(P @ dT, cols, rows, K, noise=True, noise_seed=2 * i + 1)
        _, v0, n0 = _pyramid(d0, K)
        _, v1, n1 = _pyramid(d1, K)
        pairs.append((K, v0, n0, v1, n1, dT))
        t = A.Tracker(cols, rows)
        t.set_last_pose(np.eye(4))
        trackers.append(t)
    print("256 pairs rendered", flush=True)
    dev = [(cu(p[3]), cu(p[4]), cu(p[1]), cu(p[2])) for p in pairs]
    b = A.IcpBatch(256)
    res = b.ComputeSE3(trackers, [d[0] for d in dev], [d[1] for d in dev], [d[2] for d in dev], [d[3] for d in dev],
                       [p[0] for p in pairs])
    worst_pose, worst_step, steps, ok, beyond = 0.0, 0.0, 0, 0, []
    for i, (p, d, r) in enumerate(zip(pairs, dev, res)):
        K, v0, n0, v1, n1, dT = p
        alone = A.Tracker(v1[0].shape[1], v1[0].shape[0])
        ra = alone.ComputeSE3(*d, K)
        assert np.array_equal(ra["pose"], r["pose"]) and ra["status"] == r["status"], i
        st, pose_o, trace = O.compute_se3_trace(v1, n1, v0, n0, K, np.eye(4))
        assert r["status"] == st, (i, r["status"], st)
        ok += st == 1
        e = max(_rot_err(r["pose"][:3, :3], pose_o[:3, :3]), np.linalg.norm(r["pose"][:3, 3] - pose_o[:3, 3]))
        worst_pose = max(worst_pose, e)
        assert e <= 1e-4, (i, e)
        if e > 1e-5:
            beyond.append((i, float(e)))
        for level, pose_c, x_o in trace:
            s = np.float32(1 << level)
            Kl = tuple(float(np.float32(k) / s) for k in K)
            A_g, b_g, _ = A.ICPStep(d[0][level], d[2][level], d[1][level], d[3][level], pose_c, np.eye(4), Kl)
            x_g = np.linalg.solve(A_g, b_g)
            de = max(np.abs(x_g[:3] - x_o[:3]).max(), np.abs(x_g[3:] - x_o[3:]).max())
            worst_step = max(worst_step, de)
            steps += 1
            assert de <= 1e-5, (i, level, de)
    print("256 pairs: %d tracked, %d steps; worst pose error %.3g, worst step update error %.3g; "
          "final pose beyond 1e-5 (pair, error): %s" % (ok, steps, worst_pose, worst_step, beyond), flush=True)
    assert len(beyond) <= 0.05 * len(pairs), beyond
    assert ok >= 128
