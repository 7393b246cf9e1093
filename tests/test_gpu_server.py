"""The continuous-batching server: every session's results are the same alone and in the mix
(bit-identical poses, maps and model maps), admission by the voxel-block budget, and the
submit/poll interface.
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from tests.helpers import same_bits
from tests.test_gpu_batch import MAP, SESSIONS, _frames, _state, _system

pytestmark = pytest.mark.gpu


def _alone(frames, use_rgb=0):
    res, state = [], []
    for i, (K, fr) in enumerate(frames):
        s = _system(i, K, use_rgb)
        r = []
        for d, c in fr:
            s.GrabImage(d, c)
            r.append(s.result())
        res.append(r)
        state.append(_state(s))
        s.close()
    return res, state


def _params(i, K, use_rgb=0):
    kind, size, seed, cols, rows, start, skip = SESSIONS[i]
    return A.system_params(cols=cols, rows=rows, K=K, map=MAP, use_rgb=use_rgb)


@pytest.mark.parametrize("closed_loop,use_rgb,host,graph", [(False, 0, False, False), (True, 0, False, False),
                                                            (False, 1, False, False), (False, 0, True, False),
                                                            (True, 0, False, True), (False, 1, True, True)])
def test_server_same_alone_and_in_mix(gpu, closed_loop, use_rgb, host, graph):
    """host: frames in pinned host memory, uploaded inside the steps that take them; graph:
    steps whose membership repeats run as captured graphs"""
    frames = _frames()
    alone_res, alone_state = _alone(frames, use_rgb)
    sv = A.Server(max_batch=3, max_inflight=2, lanes=2, use_graph=graph)
    sids = [sv.open(_params(i, K, use_rgb)) for i, (K, _) in enumerate(frames)]
    trace = []
    rng = np.random.default_rng(5)
    for i, (K, fr) in enumerate(frames):
        t = float(rng.uniform(0, 5))
        for f, (d, c) in enumerate(fr):
            if host:
                d, c = d.cpu().pin_memory(), c.cpu().pin_memory()
            trace.append((sids[i], d, c, t, i * 1000 + f))
            t += float(rng.exponential(4.0))  # about 250 frames/s per session, jittered
    torch.cuda.synchronize()
    print("trace", len(trace), "frames, closed loop", closed_loop, flush=True)
    done = sv.run_trace(trace, closed_loop=closed_loop)
    assert len(done) == len(trace)
    captures, replays = sv.graph_counts()
    print("graphs captured", captures, "replayed", replays, flush=True)
    if not graph:
        assert captures == replays == 0
    elif closed_loop:  # a lane's sessions step together: its membership repeats
        assert replays > 0
    seen = {}
    for c in done:
        i, f = divmod(c["tag"], 1000)
        assert c["session"] == sids[i]
        assert f == seen.get(i, -1) + 1, "frames of a session complete in order"
        seen[i] = f
        ra = alone_res[i][f]
        assert c["track_status"] == ra["track_status"], (i, f)
        assert np.array_equal(c["pose"], ra["pose"]), (i, f)
        assert c["n_visible"] == ra["n_visible"], (i, f)
        assert c["t_done_ms"] >= c["t_submit_ms"]
    for i, sid in enumerate(sids):
        xa, sa, wa, ca, va, na = alone_state[i]
        xb, sb, wb, cb, vb, nb = _state(sv.system(sid))
        assert np.array_equal(xa, xb), i
        assert same_bits(sa, sb) and np.array_equal(wa, wb) and np.array_equal(ca, cb), i
        assert same_bits(va, vb) and same_bits(na, nb), i
    sv.close()


def test_server_budget_admission(gpu):
    p = A.system_params(map=MAP)
    one = A.N.lib().fu_map_bytes(A.C.byref(p.map))
    sv = A.Server(block_budget=2 * one + one // 2)
    a = sv.open(p)
    b = sv.open(p)
    with pytest.raises(A.BudgetExceeded):
        sv.open(p)
    sv.close_session(a)
    c = sv.open(p)  # admitted once the budget frees up
    assert c not in (a, b)
    sv.close()


def test_server_submit_poll(gpu):
    frames = _frames()
    K, fr = frames[0]
    sv = A.Server(max_batch=4, max_inflight=2, lanes=1)
    sid = sv.open(_params(0, K))
    for f, (d, c) in enumerate(fr):
        sv.submit(sid, d, c, tag=f)
    assert sv.pending() == len(fr)
    with pytest.raises(A.N.FusionError):
        sv.close_session(sid)  # frames still queued
    done = sv.drain()
    assert [c["tag"] for c in done] == list(range(len(fr)))
    assert sv.pending() == 0
    r = sv.system(sid).result()
    assert np.array_equal(r["pose"], done[-1]["pose"])
    sv.close_session(sid)
    sv.close()
    torch.cuda.synchronize()


def test_server_mesh_request(gpu):
    """a mesh request queued between a session's frames extracts the map as it is after the
    frames before it: the same triangles (as a set) as fu_mesh_scene on a system that ran
    those frames alone"""
    frames = _frames()
    K, fr = frames[1]
    cut = 3
    s = _system(1, K)
    for d, c in fr[:cut]:
        s.GrabImage(d, c)
    v_a, _, c_a, _ = s.mapping().MeshScene(max_triangles=1_000_000)
    v_a = v_a.reshape(-1, 9).cpu().numpy()
    s.close()
    sv = A.Server(max_batch=4, max_inflight=2, lanes=1)
    sid = sv.open(_params(1, K))
    cap = 1_000_000
    vb = torch.empty((cap, 9), dtype=torch.float32, device="cuda")
    nb = torch.empty_like(vb)
    cb = torch.empty((cap, 9), dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    for f, (d, c) in enumerate(fr):
        if f == cut:
            sv.submit_mesh(sid, vb, nb, cb, cap, tag=999)
        sv.submit(sid, d, c, tag=f)
    done = sv.drain()
    kinds = [(c["kind"], c["tag"]) for c in done]
    assert kinds == [(0, f) for f in range(cut)] + [(1, 999)] + [(0, f) for f in range(cut, len(fr))]
    m = done[cut]
    assert m["n_triangles"] == m["n_triangles_total"] == v_a.shape[0] > 0
    v_b = vb[: m["n_triangles"]].cpu().numpy()
    key = lambda v: np.unique(v.view(np.uint32), axis=0)
    assert np.array_equal(key(v_a), key(v_b))
    sv.close_session(sid)
    sv.close()


def test_served_session_relocalised_through_its_system(gpu):
    """a served session that is lost is relocalised on its system handle once its submitted
    frames have completed (fu_server_system + fu_system_validate_poses), then served on, with
    the poses of a system run alone on the same frames, bit for bit"""
    from tests.helpers import scene_frames
    from tests.test_oracle_graph import _perturb
    sc, K, fr = scene_frames(frames=range(0, 9), cols=320, rows=240)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    mp = dict(num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15, voxel_size=0.01,
              trunc_dist=float(np.float32(0.01) * np.float32(3.0)), max_new_blocks=1 << 16)
    params = lambda: A.system_params(cols=320, rows=240, K=K, map=mp)
    ref = A.System(params())
    for f in fr[:7]:
        ref.GrabImage(*A.to_device(*f))
    cand = [_perturb(ref.result()["pose"], [0.012, -0.01, 0.01], 1.5, [0, 1, 1])]
    ref.close()
    seq = list(fr[:6]) + [empty] * 10 + [fr[6]]
    dev = [A.to_device(*f) for f in seq + list(fr[7:9])]
    # alone
    s = A.System(params())
    alone = []
    for i, (d, c) in enumerate(dev):
        if i == len(seq):
            assert s.validate_poses(cand)[0]
            alone.append(s.result())
        s.GrabImage(d, c)
        alone.append(s.result())
    s.close()
    # served, next to a second session that tracks throughout
    sv = A.Server(max_batch=4, max_inflight=2, lanes=1)
    sid = sv.open(params())
    other = sv.open(params())
    served = []
    for i, (d, c) in enumerate(dev):
        if i == len(seq):
            sys_view = sv.system(sid)
            ok, acc, chosen = sys_view.validate_poses(cand)
            assert ok and list(acc) == [1] and chosen == 0
            served.append(sys_view.result())
        sv.submit(sid, d, c, tag=i)
        sv.submit(other, *dev[min(i, 5)], tag=1000 + i)
        done = [x for x in sv.drain() if x["session"] == sid]
        assert len(done) == 1
        served.append(done[0])
    assert [r["track_status"] for r in served] == [r["track_status"] for r in alone]
    assert [r["track_status"] for r in served][-4:] == [5, 6, 1, 1]
    for a, b in zip(alone, served):
        assert np.array_equal(a["pose"], b["pose"])
    sv.close_session(sid)
    sv.close_session(other)
    sv.close()
    torch.cuda.synchronize()
