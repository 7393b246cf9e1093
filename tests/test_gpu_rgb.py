"""The photometric term on the GPU against the CPU oracle: intensity pyramid and Sobel
derivatives (Tracking/Pyrdown.cu:127-249), RGBStep (Tracking/Reduction.cu:377-627) and
ComputeSE3 with icpOnly = false (Tracking/Tracking.cc:296-409), alone, in the per-frame
pipeline and in batches.

Bar: images, pyramids, derivatives and the residual pass's integer sums are bit-exact; the
29 photometric sums are within FP32 summation error of the oracle's FP64 accumulation of the
same FP32 rows; tracked poses are within 1e-5 rad and 1e-5 m of the oracle's.
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from oracle import binding as O
from tests.helpers import scene_frames
from tests.test_gpu_pyramid import _pyr_oracle
from tests.test_oracle import _plane_frame_pose

pytestmark = pytest.mark.gpu

K = (525.0, 525.0, 319.5, 239.5)
VOX = 0.01
MAP = dict(num_buckets=1 << 18, num_excess=1 << 17, num_blocks=200_000, voxel_size=VOX,
           trunc_dist=float(np.float32(VOX) * np.float32(3.0)), max_new_blocks=1 << 16)


def _rot_err(Ra, Rb):
    return float(np.arccos(np.clip((np.trace(Ra.T @ Rb) - 1) / 2, -1, 1)))


def _cu(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _images_oracle(rgb):
    im = [O.image_to_intensity(rgb)]
    for _ in range(2):
        im.append(O.pyrdown_u8(im[-1]))
    der = [O.derivative_image(i) for i in im]
    return im, [d[0] for d in der], [d[1] for d in der]


def _images_gpu(rgb_dev):
    im = [A.ImageToIntensity(rgb_dev)]
    for _ in range(2):
        im.append(A.PyrDownGaussU8(im[-1]))
    der = [A.ComputeDerivativeImage(i) for i in im]
    return im, [d[0] for d in der], [d[1] for d in der]


def test_intensity_of_every_colour(gpu):
    v = np.arange(1 << 24, dtype=np.uint32)
    rgb = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], axis=1).astype(np.uint8).reshape(4096, 4096, 3)
    g = A.ImageToIntensity(_cu(rgb)).cpu().numpy()
    assert np.array_equal(g, O.image_to_intensity(rgb))


@pytest.mark.parametrize("cols,rows", [(640, 480), (1280, 720), (322, 242), (37, 23)])
def test_image_pyramid_and_derivatives_bit_exact(gpu, cols, rows):
    print("images %dx%d" % (cols, rows), flush=True)
    rng = np.random.default_rng(cols)
    if cols % 4 == 0 and rows % 4 == 0:
        _, _, fr = scene_frames(frames=(3,), cols=cols, rows=rows)
        rgb = fr[0][1]
    else:
        rgb = rng.integers(0, 256, size=(rows, cols, 3), dtype=np.uint8)
    og, odx, ody = _images_oracle(rgb)
    gg, gdx, gdy = _images_gpu(_cu(rgb))
    for l in range(3):
        assert np.array_equal(gg[l].cpu().numpy(), og[l]), l
        assert np.array_equal(gdx[l].cpu().numpy(), odx[l]), l
        assert np.array_equal(gdy[l].cpu().numpy(), ody[l]), l


def _plane_pair(xi):
    T = O.se3_exp(np.asarray(xi, np.float64))
    d0, c0 = _plane_frame_pose(np.eye(4))
    d1, c1 = _plane_frame_pose(T)
    return T, (d0, c0), (d1, c1)


def _scene_pair(f0=0, f1=2, cols=640, rows=480):
    sc, Kc, fr = scene_frames(frames=(f0, f1), cols=cols, rows=rows)
    return Kc, fr[0], fr[1], np.linalg.inv(sc.pose(f0)) @ sc.pose(f1)


@pytest.mark.parametrize("case", ["plane", "scene"])
def test_rgb_step_matches_oracle(gpu, case):
    if case == "plane":
        T, (d0, c0), (d1, c1) = _plane_pair([0.006, -0.004, 0.005, 0.01, -0.008, 0.012])
        Kc = K
    else:
        Kc, (d0, c0), (d1, c1), T = _scene_pair()
    _, _, v0, _ = _pyr_oracle(d0, Kc)
    _, _, v1, _ = _pyr_oracle(d1, Kc)
    i0, _, _ = _images_oracle(c0)
    i1, dx1, dy1 = _images_oracle(c1)
    # a pose guess part-way to the truth, as inside the iterations
    guess = O.se3_exp(0.5 * np.array([T[0, 3], T[1, 3], T[2, 3], 0, 0, 0]))
    for l in range(3):
        Kl = [float(np.float32(k) / np.float32(1 << l)) for k in Kc]
        (cnt, sq), sigma, s32, s64 = O.rgb_sums(i1[l], i0[l], v1[l], v0[l], dx1[l], dy1[l], guess, np.eye(4), Kl)
        A_g, b_g, r_g, (gc, gq) = A.RGBStep(_cu(i1[l]), _cu(i0[l]), _cu(v1[l]), _cu(v0[l]), _cu(dx1[l]),
                                            _cu(dy1[l]), guess, np.eye(4), Kl)
        assert (gc, gq) == (cnt, sq) and cnt > 100, (l, gc, cnt)
        assert r_g[1] == s64[28] == cnt
        A_o, b_o = O.unpack_sums(s64)
        scale = np.abs(A_o).max()
        assert np.abs(A_g - A_o).max() <= 2e-6 * scale, (l, np.abs(A_g - A_o).max() / scale)
        assert np.abs(b_g - b_o).max() <= 2e-6 * max(np.abs(b_o).max(), 1e-30), l
        assert abs(r_g[0] - s64[27]) <= 2e-6 * s64[27]


def _tracker_inputs(pair0, pair1, Kc):
    d0, c0 = pair0
    d1, c1 = pair1
    _, _, v0, n0 = _pyr_oracle(d0, Kc)
    _, _, v1, n1 = _pyr_oracle(d1, Kc)
    i0, _, _ = _images_oracle(c0)
    i1, dx1, dy1 = _images_oracle(c1)
    return (v0, n0, i0), (v1, n1, i1, dx1, dy1)


@pytest.mark.parametrize("case", ["plane", "scene"])
def test_compute_se3_rgbd_vs_oracle(gpu, case):
    if case == "plane":
        T, p0, p1 = _plane_pair([0.012, -0.009, 0.0, 0.0, 0.0, 0.0])
        Kc = K
    else:
        Kc, p0, p1, T = _scene_pair()
    (v0, n0, i0), (v1, n1, i1, dx1, dy1) = _tracker_inputs(p0, p1, Kc)
    st, pose_o, err_o, cnt_o, rcnt_o = O.compute_se3_rgbd(v1, n1, v0, n0, i1, i0, dx1, dy1, Kc, np.eye(4))
    t = A.Tracker(640, 480)
    t.set_last_pose(np.eye(4))
    cu = lambda xs: [_cu(x) for x in xs]
    r = t.ComputeSE3RGBD(cu(v1), cu(n1), cu(v0), cu(n0), cu(i1), cu(i0), cu(dx1), cu(dy1), Kc)
    assert r["status"] == st == 1, (r["status"], st)
    # the last iteration's counts come from poses that differ by rounding: a handful of
    # pixels near an edge can change sides (exact on identical inputs: test_rgb_step_matches_oracle)
    assert abs(r["rgb_count"] - rcnt_o) <= 1e-3 * rcnt_o and abs(r["count"] - cnt_o) <= 1e-3 * cnt_o
    assert _rot_err(r["pose"][:3, :3], pose_o[:3, :3]) <= 1e-5
    assert np.linalg.norm(r["pose"][:3, 3] - pose_o[:3, 3]) <= 1e-5
    if case == "plane":  # the in-plane motion ICP alone cannot see (0.6 pixel resolution)
        assert np.abs(r["pose"][:3, 3] - T[:3, 3]).max() < 0.6 * 1.5 / K[0]


def _run(frames, **kw):
    s = A.System(A.system_params(map=MAP, use_rgb=1, **kw))
    out = []
    for d16, rgb in frames:
        s.GrabImage(*A.to_device(d16, rgb))
        out.append(s.result())
    return s, out


def test_pipeline_with_rgb_matches_oracle(gpu):
    print("rgb pipeline 8 frames vs oracle", flush=True)
    sc, Kc, frames = scene_frames(frames=range(8))
    s, res = _run(frames)
    o = O.System(num_buckets=MAP["num_buckets"], num_excess=MAP["num_excess"], num_blocks=MAP["num_blocks"],
                 voxel_size=VOX, use_rgb=True)
    for i, (d16, rgb) in enumerate(frames):
        st, pose = o.grab(d16, rgb)
        r = res[i]
        assert r["track_status"] == st, (i, r["track_status"], st)
        assert _rot_err(r["pose"][:3, :3], pose[:3, :3]) <= 1e-5 * (i + 1), i
        assert np.linalg.norm(r["pose"][:3, 3] - pose[:3, 3]) <= 1e-5 * (i + 1), i
        if i > 0:
            # the model maps were raycast from poses that differ by up to the bound above, so
            # projections near pixel edges and silhouettes can land on other pixels: the
            # counts agree to 0.1 % (exact on identical inputs: test_rgb_step_matches_oracle)
            assert abs(r["rgb_count"] - o.rgb_count()) <= 1e-3 * o.rgb_count(), (i, r["rgb_count"], o.rgb_count())
    assert all(r["track_status"] in (1, 4) for r in res)
    for l in range(3):
        assert np.array_equal(s.view("image", l).cpu().numpy(), o.image(0, l)), l
        assert np.array_equal(s.view("last_image", l).cpu().numpy(), o.image(1, l)), l
        assert np.array_equal(s.view("dIdx", l).cpu().numpy(), o.image(2, l)), l
        assert np.array_equal(s.view("dIdy", l).cpu().numpy(), o.image(3, l)), l
    # and the sequence is tracked: frame-to-model tracking of this scene drifts by about
    # 2.5 cm over 8 frames on both the ICP-only and the blended path (the raycast model is
    # offset by about half a voxel from the surface), so the bound catches divergence only
    gt = np.linalg.inv(sc.pose(0)) @ sc.pose(7)
    assert np.linalg.norm(res[-1]["pose"][:3, 3] - gt[:3, 3]) < 0.04


def test_rgb_modes_agree(gpu):
    """device-resident, hipGraph-replayed and plain control give the same poses with the
    photometric term (plain solves on the host: last-ulp FP64 differences only)"""
    sc, Kc, frames = scene_frames(frames=range(5))
    _, dev = _run(frames)
    _, gra = _run(frames, use_graph=True)
    _, pla = _run(frames, plain=True)
    for a, b, c in zip(dev, gra, pla):
        assert a["track_status"] == b["track_status"] == c["track_status"]
        assert np.array_equal(a["pose"], b["pose"])
        assert np.abs(a["pose"] - c["pose"]).max() < 1e-9
        assert a["rgb_count"] == b["rgb_count"] == c["rgb_count"]


def test_textureless_frame_fails_and_keeps_last_images(gpu):
    """a frame of one flat colour has no intensity gradient: RGBStep finds no residuals and
    fails the frame (Tracking.cc:341-346); the last frame's images stay those of frame 0"""
    sc, Kc, frames = scene_frames(frames=(0, 1))
    flat = (frames[1][0], np.full_like(frames[1][1], 90))
    s, res = _run([frames[0], flat])
    assert res[1]["track_status"] == 2 and res[1]["rgb_count"] == 0 and res[1]["fused"] == 0
    assert np.array_equal(res[1]["pose"], res[0]["pose"])
    im0, _, _ = _images_oracle(frames[0][1])
    for l in range(3):
        assert np.array_equal(s.view("last_image", l).cpu().numpy(), im0[l]), l
    # and tracking resumes against frame 0 on the next good frame
    s.GrabImage(*A.to_device(*frames[1]))
    assert s.result()["track_status"] == 1


def test_rgb_argument_errors(gpu):
    with pytest.raises(A.N.FusionError):
        A.System(A.system_params(cols=2560, rows=2048, map=MAP, use_rgb=1))  # over 2^22 pixels
    s = A.System(A.system_params(map=MAP))
    with pytest.raises(A.N.FusionError):
        s.view("image", 0)  # ICP-only systems have no intensity images
    b = A.Batch(2)
    b.attach(s)
    s2 = A.System(A.system_params(map=MAP, use_rgb=1))
    with pytest.raises(A.N.FusionError):
        b.attach(s2)  # one tracking mode per batch context
    b.close()
