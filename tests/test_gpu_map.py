"""GPU parity of allocation, visibility and TSDF integration (Mapping/DeviceMap.cu,
Mapping/FuseMap.cu) against the CPU oracle on identical inputs.

Bar: the allocated and visible block sets are identical as sets (pool order differs);
the FP32 voxel form (sdf f32 + weight) is bit-identical per voxel, colour identical.
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.helpers import rows_set, scene_frames

pytestmark = pytest.mark.gpu


def _inputs(d16, rgb, K):
    raw, f0 = O.filter_depth(d16, 5000.0, 3.0)
    v = O.compute_vmap(f0, *K, 3.0)
    n = O.compute_nmap(v)
    return raw, n


def _fuse_both(params, frames, poses, K, fmt=0):
    g = A.Mapping(**params, voxel_format=fmt)
    o = O.Map(num_buckets=params["num_buckets"], num_excess=params["num_excess"], num_blocks=params["num_blocks"],
              voxel_size=params["voxel_size"], trunc_dist=params["trunc_dist"], max_weight=params.get("max_weight", 255))
    for (d16, rgb), pose in zip(frames, poses):
        raw, n = _inputs(d16, rgb, K)
        nv_g = g.FuseMapColor(torch.from_numpy(raw).cuda(), torch.from_numpy(rgb).cuda(), torch.from_numpy(n).cuda(),
                              pose, K)
        nv_o = o.fuse_map_color(raw, rgb, n, pose, K)
        assert nv_g == nv_o, (nv_g, nv_o)
    return g, o


def _params(voxel=0.01, buckets=1 << 17, blocks=1 << 16, excess=1 << 16):
    return dict(num_buckets=buckets, num_excess=excess, num_blocks=blocks, voxel_size=voxel,
                trunc_dist=float(np.float32(voxel) * np.float32(3.0)), max_new_blocks=1 << 16)


def _compare_maps(g, o, exact_sdf=True, tol=None):
    gb, ob = g.export_blocks(), o.export_blocks()
    assert rows_set(gb) == rows_set(ob), (len(gb), len(ob))
    assert rows_set(g.export_visible()) == rows_set(o.export_visible())
    sg, wg, cg = g.read_blocks(ob)
    so, wo, co = o.read_blocks(ob)
    assert np.array_equal(wg, wo), int((wg != wo).sum())
    assert np.array_equal(cg, co), int((cg != co).sum())
    assert np.array_equal(np.isnan(sg), np.isnan(so))
    m = ~np.isnan(so)
    if exact_sdf:
        assert np.array_equal(sg[m].view(np.uint32), so[m].view(np.uint32)), int((sg[m] != so[m]).sum())
    else:
        assert np.max(np.abs(sg[m] - so[m])) <= tol
    return len(ob)


def test_fuse_single_frame_identity(gpu):
    sc, K, frames = scene_frames(frames=(0,))
    print("fuse single frame", flush=True)
    g, o = _fuse_both(_params(), frames, [np.eye(4)], K)
    n = _compare_maps(g, o)
    assert n > 500


def test_fuse_sequence_moving_pose(gpu):
    """five frames fused at their ground-truth poses relative to frame 0"""
    sc, K, frames = scene_frames(frames=(0, 3, 6, 9, 12))
    P0 = sc.pose(0)
    poses = [np.linalg.inv(P0) @ sc.pose(f) for f in (0, 3, 6, 9, 12)]
    g, o = _fuse_both(_params(), frames, poses, K)
    _compare_maps(g, o)


def test_fuse_5mm_1280(gpu):
    sc, K, frames = scene_frames(frames=(0, 4), cols=1280, rows=720)
    P0 = sc.pose(0)
    poses = [np.eye(4), np.linalg.inv(P0) @ sc.pose(4)]
    g, o = _fuse_both(_params(voxel=0.005, buckets=1 << 19, blocks=1 << 18, excess=1 << 18), frames, poses, K)
    _compare_maps(g, o)


def test_weight16_format_matches(gpu):
    sc, K, frames = scene_frames(frames=(0, 2))
    P0 = sc.pose(0)
    poses = [np.eye(4), np.linalg.inv(P0) @ sc.pose(2)]
    g, o = _fuse_both(_params(), frames, poses, K, fmt=1)
    _compare_maps(g, o)


def test_fp16_sdf_within_tolerance(gpu):
    """FP16 storage: |sdf - FP32 oracle| <= 2^-9 of the truncation distance after 3 frames
    (fp16 spacing is 2^-11 on [0.5, 1); weights and colours are exact)"""
    sc, K, frames = scene_frames(frames=(0, 2, 4))
    P0 = sc.pose(0)
    poses = [np.linalg.inv(P0) @ sc.pose(f) for f in (0, 2, 4)]
    g, o = _fuse_both(_params(), frames, poses, K, fmt=2)
    _compare_maps(g, o, exact_sdf=False, tol=2.0 ** -9)


def test_all_in_one_bucket(gpu):
    """every block collides (one bucket): the excess chains carry the whole set"""
    sc, K, frames = scene_frames(frames=(0,), cols=320, rows=240)
    p = _params(buckets=1, excess=1 << 15, blocks=1 << 15)
    g, o = _fuse_both(p, frames, [np.eye(4)], K)
    n = _compare_maps(g, o)
    st = g.stats()
    assert st["n_excess"] == n - 1


def test_heap_exhaustion(gpu):
    """a pool smaller than the frame's block set: exactly `capacity` blocks, all from the
    oracle's set, and the overflow flag raised"""
    sc, K, frames = scene_frames(frames=(0,))
    p = _params(blocks=300)
    g = A.Mapping(**p)
    d16, rgb = frames[0]
    raw, n = _inputs(d16, rgb, K)
    g.FuseMapColor(torch.from_numpy(raw).cuda(), torch.from_numpy(rgb).cuda(), torch.from_numpy(n).cuda(),
                   np.eye(4), K)
    full = O.Map(**{k: v for k, v in _params().items() if k != "max_new_blocks"})
    full.fuse_map_color(raw, rgb, n, np.eye(4), K)
    gb = rows_set(g.export_blocks())
    assert len(gb) == 300
    assert gb <= rows_set(full.export_blocks())
    assert g.stats()["flags"] & 1


def test_reset_clears(gpu):
    sc, K, frames = scene_frames(frames=(0,), cols=320, rows=240)
    g, o = _fuse_both(_params(), frames, [np.eye(4)], K)
    g.Reset()
    assert len(g.export_blocks()) == 0
    st = g.stats()
    assert st["n_allocated"] == 0 and st["flags"] == 0


def test_empty_frame(gpu):
    K = (525.0, 525.0, 319.5, 239.5)
    d16 = np.zeros((480, 640), np.uint16)
    rgb = np.zeros((480, 640, 3), np.uint8)
    g, o = _fuse_both(_params(), [(d16, rgb)], [np.eye(4)], K)
    assert len(g.export_blocks()) == 0 and len(o.export_blocks()) == 0


@pytest.mark.parametrize("fmt,bound", [(0, 1e-3), (1, 1e-3), (2, 2.0 ** -8)])
def test_precision_vs_fp64_twin(gpu, fmt, bound):
    """Per-voxel TSDF of each storage format against an FP64 re-integration of the same
    frames (the oracle's FP64 twin: same allocation, the update arithmetic in double), 5 mm
    voxels, truncation 8 voxels, 8 frames with ground-truth poses.  Bar: |sdf - sdf64| within
    1e-3 of the truncation distance for the FP32 forms; FP16 within 2^-8 (storage rounding,
    spacing 2^-11 near 1).  Weights must match the FP64 twin's exactly."""
    sc, K, frames = scene_frames(frames=range(0, 16, 2))
    P0 = sc.pose(0)
    poses = [np.linalg.inv(P0) @ sc.pose(f) for f in range(0, 16, 2)]
    vox = 0.005
    p = _params(voxel=vox, buckets=1 << 19, blocks=300_000, excess=1 << 18)
    p["trunc_dist"] = float(np.float32(vox) * np.float32(8.0))
    g = A.Mapping(**p, voxel_format=fmt)
    o = O.Map(num_buckets=p["num_buckets"], num_excess=p["num_excess"], num_blocks=p["num_blocks"], voxel_size=vox,
              trunc_dist=p["trunc_dist"], precision=1)
    for (d16, rgb), pose in zip(frames, poses):
        raw, n = _inputs(d16, rgb, K)
        g.FuseMapColor(torch.from_numpy(raw).cuda(), torch.from_numpy(rgb).cuda(), torch.from_numpy(n).cuda(), pose, K)
        o.fuse_map_color(raw, rgb, n, pose, K)
    ob = o.export_blocks()
    assert rows_set(g.export_blocks()) == rows_set(ob)
    sg, wg, _ = g.read_blocks(ob)
    s64 = o.read_blocks_f64(ob)
    _, wo, _ = o.read_blocks(ob)
    assert np.array_equal(wg, wo)
    m = ~np.isnan(s64)
    assert np.array_equal(np.isnan(sg), ~m)
    err = np.abs(sg[m].astype(np.float64) - s64[m])
    print("format %d: %d voxels, max |sdf - sdf64| = %.3g T, mean %.3g T, 99.9%% %.3g T" % (
        fmt, int(m.sum()), err.max(), err.mean(), np.percentile(err, 99.9)), flush=True)
    assert err.max() <= bound


def test_range_specialised_arithmetic_equals_ieee(gpu):
    """the integration kernel's shared-reciprocal divisions and range-specialised square root
    give the compiler's IEEE results inside their ranges (2^28 random cases of each), and the
    one-instruction float/double -> int conversions follow their rule for any input, and the
    bilateral weights' exp equals the general one bit for bit"""
    import ctypes as C
    bad = (C.c_uint32 * 7)()
    A.check(A.N.lib().fu_selftest_arith(12345, 1 << 28, bad, None), "fu_selftest_arith")
    assert list(bad) == [0] * 7, list(bad)
