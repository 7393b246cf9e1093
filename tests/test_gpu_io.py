"""On-disk formats: the map and session snapshots (System::WriteMapToDisk/ReadMapFromDisk,
Core/System.cc:195-271) and the PLY writer (System::WriteMeshToDisk, :140-193).

Bar: a loaded map is the saved map (same blocks in the same pool slots, bit-identical voxels),
a session resumed from a snapshot continues bit-identically to one that never left the GPU,
damaged or mismatched files are refused, and the PLY text is the reference's format.
"""
import os

import numpy as np
import pytest
import torch

from fusion_amd import api as A
from tests.helpers import same_bits, scene_frames

pytestmark = pytest.mark.gpu

VOX = 0.01
MAP = dict(num_buckets=1 << 17, num_excess=1 << 16, num_blocks=120_000, voxel_size=VOX,
           trunc_dist=float(np.float32(VOX) * np.float32(3.0)), max_new_blocks=1 << 15)


def _system(use_rgb=0, **kw):
    return A.System(A.system_params(map=dict(MAP, **kw), use_rgb=use_rgb))


def _blocks(m, sort=False):
    xyz = m.export_blocks()  # pool order
    if sort:  # blocks allocated by different runs take pool slots in atomic order
        xyz = xyz[np.lexsort(xyz.T[::-1])]
    sdf, w, rgb = m.read_blocks(xyz)
    return xyz, sdf, w, rgb


@pytest.mark.parametrize("use_rgb", [0, 1])
def test_session_snapshot_resumes_bit_identically(gpu, tmp_path, use_rgb):
    print("session snapshot, use_rgb", use_rgb, flush=True)
    _, _, frames = scene_frames(frames=range(8))
    dev = [A.to_device(*f) for f in frames]
    a = _system(use_rgb)
    for f in dev[:5]:
        a.GrabImage(*f)
    path = str(tmp_path / "session.fusn")
    a.save(path)
    b = _system(use_rgb)
    b.load(path)
    ra, rb = [], []
    for f in dev[5:]:
        a.GrabImage(*f)
        ra.append(a.result())
        b.GrabImage(*f)
        rb.append(b.result())
    for x, y in zip(ra, rb):
        assert x["track_status"] == y["track_status"] and x["frame_index"] == y["frame_index"]
        assert np.array_equal(x["pose"], y["pose"])
        assert x["n_allocated"] == y["n_allocated"] and x["n_visible"] == y["n_visible"]
    assert all(x["track_status"] == 1 for x in ra)
    xa, sa, wa, ca = _blocks(a.mapping(), sort=True)
    xb, sb, wb, cb = _blocks(b.mapping(), sort=True)
    assert np.array_equal(xa, xb)
    assert same_bits(sa, sb) and np.array_equal(wa, wb) and np.array_equal(ca, cb)
    assert same_bits(a.view("model_vmap").cpu().numpy(), b.view("model_vmap").cpu().numpy())
    if use_rgb:
        for l in range(3):
            assert np.array_equal(a.view("last_image", l).cpu().numpy(), b.view("last_image", l).cpu().numpy())
    a.close()
    b.close()


def test_map_snapshot_round_trip_and_refusals(gpu, tmp_path):
    _, _, frames = scene_frames(frames=range(3))
    s = _system()
    for f in frames:
        s.GrabImage(*A.to_device(*f))
    m = s.mapping()
    path = str(tmp_path / "map.fusn")
    m.save(path)
    ref = _blocks(m)
    st = m.stats()
    m2 = A.Mapping(**MAP)
    m2.load(path)
    got = _blocks(m2)
    assert np.array_equal(ref[0], got[0]) and same_bits(ref[1], got[1])  # the same blocks in the same pool slots
    assert np.array_equal(ref[2], got[2]) and np.array_equal(ref[3], got[3])
    st2 = m2.stats()
    assert (st2["n_allocated"], st2["n_excess"], st2["flags"]) == (st["n_allocated"], st["n_excess"], st["flags"])
    # the loaded map meshes to the same triangles
    v1 = m.MeshScene(max_triangles=1_000_000)[0].reshape(-1, 9).cpu().numpy()
    v2 = m2.MeshScene(max_triangles=1_000_000)[0].reshape(-1, 9).cpu().numpy()
    assert same_bits(v1, v2)
    # other parameters: refused, map untouched
    m3 = A.Mapping(**dict(MAP, num_buckets=1 << 16))
    with pytest.raises(A.N.FusionError):
        m3.load(path)
    # damaged files: refused, and the map they were loaded into is left as it was
    raw = open(path, "rb").read()
    bad = str(tmp_path / "bad.fusn")

    def refused(data):
        open(bad, "wb").write(bytes(data))
        with pytest.raises(A.N.FusionError):
            m2.load(bad)
        now = _blocks(m2)
        assert np.array_equal(ref[0], now[0]) and same_bits(ref[1], now[1]) and np.array_equal(ref[2], now[2])

    refused(raw[: len(raw) // 2])
    flip = bytearray(raw)
    flip[len(raw) - 100] ^= 0x40
    refused(flip)
    refused(b"not a snapshot at all")
    # a structural error under a correct checksum: the first entry's block index past the
    # allocated blocks (entries start after the 124-byte header; 20 bytes each)
    struct_bad = bytearray(raw[:-8])
    n_alloc = int.from_bytes(struct_bad[116:120], "little")
    struct_bad[124 + 12:124 + 16] = (n_alloc * 512).to_bytes(4, "little")
    h = 1469598103934665603
    for b in struct_bad:
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    refused(struct_bad + h.to_bytes(8, "little"))
    with pytest.raises(A.N.FusionError):
        s.load(path)  # a map snapshot is not a session snapshot
    s.close()


def test_refused_session_load_leaves_the_session_unchanged(gpu, tmp_path):
    _, _, frames = scene_frames(frames=range(7))
    dev = [A.to_device(*f) for f in frames]
    a, b = _system(), _system()
    for f in dev[:4]:
        a.GrabImage(*f)
        b.GrabImage(*f)
    path = str(tmp_path / "s.fusn")
    a.save(path)
    raw = open(path, "rb").read()
    bad = str(tmp_path / "bad.fusn")
    for data in (raw[: len(raw) - 4096], raw[: len(raw) // 3], raw[:-1] + bytes([raw[-1] ^ 1])):
        open(bad, "wb").write(data)
        with pytest.raises(A.N.FusionError):
            b.load(bad)
    for f in dev[4:]:
        a.GrabImage(*f)
        b.GrabImage(*f)
        ra, rb = a.result(), b.result()
        assert ra["track_status"] == rb["track_status"] == 1 and np.array_equal(ra["pose"], rb["pose"])
    xa, sa, wa, ca = _blocks(a.mapping(), sort=True)
    xb, sb, wb, cb = _blocks(b.mapping(), sort=True)
    assert np.array_equal(xa, xb) and same_bits(sa, sb) and np.array_equal(wa, wb)
    a.close()
    b.close()


def _ply_header(n_vert, n_face, fmt):
    # System::WriteMeshToDisk (Core/System.cc:151-166)
    return ["ply", "format %s 1.0" % fmt, "element vertex %d" % n_vert, "property float x", "property float y",
            "property float z", "property float nx", "property float ny", "property float nz",
            "property uchar red", "property uchar green", "property uchar blue", "element face %d" % n_face,
            "property list uchar uint vertex_indices", "end_header"]


def test_ply_writer_formats(gpu, tmp_path):
    _, _, frames = scene_frames(frames=(0,), cols=320, rows=240)
    s = A.System(A.system_params(cols=320, rows=240, K=(262.5, 262.5, 159.5, 119.5), map=MAP))
    s.GrabImage(*A.to_device(*frames[0]))
    v, n, c, _ = s.mapping().MeshScene(max_triangles=400_000)
    nt = v.shape[0]
    assert nt > 1000
    vh = v.reshape(-1, 3).cpu().numpy()
    nh = n.reshape(-1, 3).cpu().numpy()
    ch = c.reshape(-1, 3).cpu().numpy()
    p = str(tmp_path / "scene.ply")
    A.WriteMeshToDisk(p, v, n, c)
    lines = open(p).read().split("\n")
    hdr = _ply_header(3 * nt, nt, "ascii")
    assert lines[: len(hdr)] == hdr
    body = lines[len(hdr):]
    for i in (0, 1, 7, 3 * nt // 2, 3 * nt - 1):  # std::ostream's default float output is %g
        want = " ".join(["%g" % float(x) for x in vh[i]] + ["%g" % float(x) for x in nh[i]] +
                        ["%d" % int(x) for x in ch[i]])
        assert body[i] == want, (i, body[i], want)
    vals = np.array([list(map(float, l.split()[:6])) for l in body[: 3 * nt]], dtype=np.float64)
    assert np.allclose(vals[:, :3], vh, rtol=1e-5, atol=1e-6)  # 6 significant digits
    faces = body[3 * nt: 4 * nt]
    assert faces[0] == "3 0 1 2" and faces[-1] == "3 %d %d %d" % (3 * nt - 3, 3 * nt - 2, 3 * nt - 1)
    assert body[4 * nt:] == [""]
    # binary: the same elements, bit-exact
    pb = str(tmp_path / "scene_bin.ply")
    A.WriteMeshToDisk(pb, v, n, c, binary=True)
    raw = open(pb, "rb").read()
    hb = "\n".join(_ply_header(3 * nt, nt, "binary_little_endian")) + "\n"
    assert raw[: len(hb)] == hb.encode()
    vt = np.dtype([("p", "<f4", 3), ("n", "<f4", 3), ("c", "u1", 3)])
    verts = np.frombuffer(raw, dtype=vt, count=3 * nt, offset=len(hb))
    assert same_bits(verts["p"], vh) and same_bits(verts["n"], nh) and np.array_equal(verts["c"], ch)
    ft = np.dtype([("k", "u1"), ("i", "<u4", 3)])
    fcs = np.frombuffer(raw, dtype=ft, count=nt, offset=len(hb) + 27 * 3 * nt)
    assert (fcs["k"] == 3).all() and np.array_equal(fcs["i"], np.arange(3 * nt, dtype=np.uint32).reshape(nt, 3))
    assert len(raw) == len(hb) + 27 * 3 * nt + 13 * nt
    s.close()


def test_served_session_evicted_and_resumed(gpu, tmp_path):
    """eviction = fu_system_save + fu_server_close; resume = fu_server_open + fu_system_load:
    the resumed session's frames complete with the poses of a session never evicted"""
    _, K, frames = scene_frames(frames=range(7))
    dev = [A.to_device(*f) for f in frames]
    p = A.system_params(map=MAP)
    ref = _system()
    ref_poses = []
    for f in dev:
        ref.GrabImage(*f)
        ref_poses.append(ref.result()["pose"])
    sv = A.Server(max_batch=4, max_inflight=2, lanes=1)
    sid = sv.open(p)
    for k, f in enumerate(dev[:4]):
        sv.submit(sid, *f, tag=k)
    done = sv.drain()
    path = str(tmp_path / "evicted.fusn")
    sv.system(sid).save(path)
    sv.close_session(sid)
    sid2 = sv.open(A.system_params(map=MAP))
    sv.system(sid2).load(path)
    for k, f in enumerate(dev[4:], start=4):
        sv.submit(sid2, *f, tag=k)
    done += sv.drain()
    assert [c["tag"] for c in done] == list(range(7))
    for c in done:
        assert np.array_equal(c["pose"], ref_poses[c["tag"]]), c["tag"]
    sv.close()
    ref.close()


def test_tum_replay_matches_in_memory(gpu, tmp_path):
    """a sequence written in the TUM layout and read back through TumSource tracks exactly as
    the in-memory frames do; its trajectory file and ATE against the scene's ground truth"""
    from fusion_amd import ingest as I
    src = I.SceneSource(frames=6)
    frames = list(src)
    stamps = [src.timestamp(i) for i in range(len(src))]
    I.write_tum_sequence(str(tmp_path / "seq"), frames, stamps)
    tum = I.TumSource(str(tmp_path / "seq"), K=src.K, depth_scale=5000.0)
    assert len(tum) == 6
    a = _system()
    b = _system()
    pa, pb = [], []
    for f, g in zip(frames, tum):
        a.GrabImage(*A.to_device(*f))
        pa.append(a.result()["pose"])
        b.GrabImage(*A.to_device(*g))
        pb.append(b.result()["pose"])
    assert all(np.array_equal(x, y) for x, y in zip(pa, pb))
    traj = str(tmp_path / "est.txt")
    I.write_tum_trajectory(traj, stamps, pb)
    ts, back = I.read_tum_trajectory(traj)
    assert len(back) == 6 and max(np.abs(x - y).max() for x, y in zip(back, pb)) < 1e-7
    gt = [np.linalg.inv(src.ground_truth(0)) @ src.ground_truth(i) for i in range(6)]
    rmse, _, _ = I.ate_rmse(pb, gt)
    assert rmse < 0.02, rmse
    a.close()
    b.close()
