"""GPU parity of the z-range tiles, the raycast (Mapping/RenderScene.cu) and marching cubes
(Mapping/MeshScene.cu) against the CPU oracle on the same fused map.

Bar: tile maps, vertex and normal maps bit-identical; the mesh identical as a set of
triangles keyed by (voxel, cube index, triangle number), with bit-identical vertices,
normals and colours per triangle.
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.helpers import mismatch, same_bits, scene_frames
from tests.test_gpu_map import _fuse_both, _params

pytestmark = pytest.mark.gpu


def _map_pair(voxel=0.01, frames=(0, 3, 6), cols=640, rows=480, fmt=0):
    sc, K, fr = scene_frames(frames=frames, cols=cols, rows=rows)
    P0 = sc.pose(frames[0])
    poses = [np.linalg.inv(P0) @ sc.pose(f) for f in frames]
    g, o = _fuse_both(_params(voxel=voxel, buckets=1 << 18, blocks=1 << 17, excess=1 << 17), fr, poses, K, fmt=fmt)
    return sc, K, g, o, poses


@pytest.mark.parametrize("voxel", [0.01, 0.005])
def test_raycast_bit_exact(gpu, voxel):
    print("raycast parity voxel %g" % voxel, flush=True)
    sc, K, g, o, poses = _map_pair(voxel=voxel)
    pose = poses[-1]
    nv = g.stats()["n_visible"]
    zg = g.CreateRenderingBlocks(80, 60, pose, nv, K)
    zo = o.create_rendering_blocks(80, 60, pose, nv, K)
    assert zg[2] == zo[2] and zg[3] == zo[3]
    assert same_bits(zg[0].cpu().numpy(), zo[0]), mismatch(zg[0].cpu().numpy(), zo[0])
    assert same_bits(zg[1].cpu().numpy(), zo[1])
    vg, ng = g.Raycast(640, 480, zg[0], zg[1], pose, K)
    vo, no = o.raycast(640, 480, zo[0], zo[1], pose, K)
    vg, ng = vg.cpu().numpy(), ng.cpu().numpy()
    assert same_bits(vg, vo), mismatch(vg, vo)
    assert same_bits(ng, no), mismatch(ng, no)
    assert (~np.isnan(vo[..., 0])).mean() > 0.5


def test_raycast_other_view_and_sizes(gpu):
    """a pose away from the fused ones, and a 320x240 image with its own tile map"""
    sc, K, g, o, poses = _map_pair()
    P0 = sc.pose(0)
    pose = np.linalg.inv(P0) @ sc.pose(20)
    nv = g.CheckBlockVisibility(pose, 640, 480, K, 3.0, 0.1)
    assert nv == o.check_visibility(pose, 640, 480, K, 3.0, 0.1)
    Ks = (262.5, 262.5, 159.5, 119.5)
    zg = g.CreateRenderingBlocks(40, 30, pose, nv, Ks)
    zo = o.create_rendering_blocks(40, 30, pose, nv, Ks)
    assert same_bits(zg[0].cpu().numpy(), zo[0]) and same_bits(zg[1].cpu().numpy(), zo[1])
    vg, ng = g.Raycast(320, 240, zg[0], zg[1], pose, Ks)
    vo, no = o.raycast(320, 240, zo[0], zo[1], pose, Ks)
    assert same_bits(vg.cpu().numpy(), vo) and same_bits(ng.cpu().numpy(), no)


def test_raycast_f16_close(gpu):
    sc, K, g, o, poses = _map_pair(fmt=2)
    pose = poses[-1]
    nv = g.stats()["n_visible"]
    zg = g.CreateRenderingBlocks(80, 60, pose, nv, K)
    vg, _ = g.Raycast(640, 480, zg[0], zg[1], pose, K)
    zo = o.create_rendering_blocks(80, 60, pose, nv, K)
    vo, _ = o.raycast(640, 480, zo[0], zo[1], pose, K)
    vg = vg.cpu().numpy()
    both = ~np.isnan(vg[..., 0]) & ~np.isnan(vo[..., 0])
    iou = both.sum() / max(1, (~np.isnan(vg[..., 0]) | ~np.isnan(vo[..., 0])).sum())
    assert iou > 0.99
    assert np.median(np.abs(vg[both][:, 2] - vo[both][:, 2])) < 1e-3


def _mesh_dict(vert, norm, col, meta):
    d = {}
    for i in range(meta.shape[0]):
        d[tuple(meta[i].tolist())] = (vert[i], norm[i], col[i])
    return d


def test_mesh_bit_exact(gpu):
    print("marching cubes parity", flush=True)
    sc, K, g, o, poses = _map_pair(frames=(0, 3))
    vg, ng, cg, mg = g.MeshScene(max_triangles=1_000_000, with_meta=True)
    vo, no, co, mo, total = o.mesh()
    assert mg.shape[0] == total == mo.shape[0]
    assert total > 1000
    dg = _mesh_dict(vg.cpu().numpy(), ng.cpu().numpy(), cg.cpu().numpy(), mg.cpu().numpy())
    do = _mesh_dict(vo, no, co, mo)
    assert set(dg) == set(do)
    bad = 0
    for k, (v, n, c) in do.items():
        gv, gn, gc = dg[k]
        if not (same_bits(gv, v) and same_bits(gn, n) and np.array_equal(gc, c)):
            bad += 1
    assert bad == 0, bad


def test_mesh_capacity_clamp(gpu):
    sc, K, g, o, poses = _map_pair(frames=(0,))
    _, _, _, mo, total = o.mesh()
    vg, ng, cg, mg = g.MeshScene(max_triangles=100, with_meta=True)
    assert mg.shape[0] == 100 and total > 100
    assert g.stats()["flags"] & 32
