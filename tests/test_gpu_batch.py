"""Batched frames of many systems (fu_batch_grab) against the same systems run alone.

Each session's work-group decomposition is the same alone and in a batch, so every result
is required to be bit-identical: poses, tracking status, model maps, block sets and voxels.
The batch composition changes from step to step (sessions join late, skip a step, differ in
frame size), as it does under continuous batching.
"""
import numpy as np
import pytest
import torch

from fusion_amd import api as A
from tests.helpers import rows_set, same_bits, scene_frames

pytestmark = pytest.mark.gpu

VOX = 0.01
MAP = dict(num_buckets=1 << 17, num_excess=1 << 16, num_blocks=120_000, voxel_size=VOX,
           trunc_dist=float(np.float32(VOX) * np.float32(3.0)), max_new_blocks=1 << 15)

# (scene kind, size, seed, cols, rows, first step, steps skipped)
SESSIONS = [
    (0, (4.0, 4.0, 3.0), 0, 640, 480, 0, ()),
    (0, (3.0, 3.0, 2.5), 3, 320, 240, 0, (3,)),
    (0, (5.0, 4.0, 3.0), 7, 640, 480, 2, ()),
    (1, (8.0, 6.0, 3.0), 2, 1280, 720, 1, (4,)),
    # partial 16x16 tiles at every level (202x154, 101x77): the ragged edges of the batched
    # pyramid, resize and input-pack kernels
    (0, (4.0, 4.0, 3.0), 5, 404, 308, 1, ()),
]
STEPS = 6


def _frames():
    out = []
    for kind, size, seed, cols, rows, start, skip in SESSIONS:
        n = STEPS - start - len(skip)
        _, K, fr = scene_frames(kind, size, seed, frames=range(n), cols=cols, rows=rows)
        out.append((K, [A.to_device(*f) for f in fr]))
    return out


def _system(i, K, use_rgb=0):
    kind, size, seed, cols, rows, start, skip = SESSIONS[i]
    return A.System(A.system_params(cols=cols, rows=rows, K=K, map=MAP, use_rgb=use_rgb))


def _state(s):
    m = s.mapping()
    xyz = m.export_blocks()
    order = np.lexsort(xyz.T[::-1])
    xyz = xyz[order]
    sdf, w, rgb = m.read_blocks(xyz)
    vm = s.view("model_vmap").cpu().numpy()
    nm = s.view("model_nmap").cpu().numpy()
    if s.params.use_rgb:  # the last tracked frame's intensity pyramid
        vm = [vm] + [s.view("last_image", l).cpu().numpy().astype(np.float32) for l in range(3)]
        vm = np.concatenate([v.ravel() for v in vm])
    return xyz, sdf, w, rgb, vm, nm


def _run_batched(frames, use_rgb):
    """the sessions' frames through one batch context, with the step membership of SESSIONS"""
    print("batched", flush=True)
    b = A.Batch(8)
    systems = [_system(i, K, use_rgb) for i, (K, _) in enumerate(frames)]
    for s in systems:
        b.attach(s)
    nxt = [0] * len(systems)
    batch_res = [[] for _ in systems]
    for step in range(STEPS):
        members = [i for i, (_, _, _, _, _, start, skip) in enumerate(SESSIONS) if step >= start and step not in skip]
        print("step", step, "sessions", members, flush=True)
        b.grab([systems[i] for i in members], [frames[i][1][nxt[i]][0] for i in members],
               [frames[i][1][nxt[i]][1] for i in members])
        for i in members:
            batch_res[i].append(systems[i].result())
            nxt[i] += 1
    for i in range(len(systems)):
        assert nxt[i] == len(frames[i][1])
    return batch_res, systems, b


@pytest.mark.parametrize("use_rgb", [0, 1])
def test_batch_equals_alone(gpu, use_rgb):
    frames = _frames()
    alone_res, alone_state = [], []
    for i, (K, fr) in enumerate(frames):
        print("alone session", i, "use_rgb", use_rgb, flush=True)
        s = _system(i, K, use_rgb)
        res = []
        for d, c in fr:
            s.GrabImage(d, c)
            res.append(s.result())
        alone_res.append(res)
        alone_state.append(_state(s))
        s.close()
    torch.cuda.synchronize()

    batch_res, systems, b = _run_batched(frames, use_rgb)
    for i, s in enumerate(systems):
        for f, (ra, rb) in enumerate(zip(alone_res[i], batch_res[i])):
            assert ra["track_status"] == rb["track_status"], (i, f)
            assert np.array_equal(ra["pose"], rb["pose"]), (i, f, ra["pose"], rb["pose"])
            assert ra["n_visible"] == rb["n_visible"] and ra["n_allocated"] == rb["n_allocated"], (i, f)
            assert ra["rgb_count"] == rb["rgb_count"], (i, f)
        xa, sa, wa, ca, va, na = alone_state[i]
        xb, sb, wb, cb, vb, nb = _state(s)
        assert np.array_equal(xa, xb), i
        assert same_bits(sa, sb) and np.array_equal(wa, wb) and np.array_equal(ca, cb), i
        assert same_bits(va, vb) and same_bits(na, nb), i
        assert alone_res[i][-1]["track_status"] in (1, 4), i  # the sequences do track
    b.close()


# launch geometries far from the default: other wave layouts, every XCD-ordered stage, odd
# work-group counts (a grid that is not a multiple of 8 exercises the XCD mapping's remainder)
TUNINGS = [
    dict(raycast_wave=1, xcd_order=0xFF, integrate_groups=3001, visible_groups=777, alloc_groups=259),
    dict(raycast_wave=2, xcd_order=(1 << 2) | (1 << 7), integrate_groups=65536, visible_groups=20000,
         alloc_groups=9000),
]


@pytest.mark.parametrize("use_rgb", [0, 1])
def test_launch_geometry_does_not_change_results(gpu, use_rgb):
    frames = _frames()
    base = A.default_tuning()
    runs = []
    try:
        for t in [base] + TUNINGS:
            print("tuning", t, "use_rgb", use_rgb, flush=True)
            A.set_tuning(**t)
            res, systems, b = _run_batched(frames, use_rgb)
            runs.append((res, [_state(s) for s in systems]))
            for s in systems:
                s.close()
            b.close()
    finally:
        A.set_tuning(**base)
    (r0, s0) = runs[0]
    for k, (r1, s1) in enumerate(runs[1:]):
        for i in range(len(SESSIONS)):
            for f, (ra, rb) in enumerate(zip(r0[i], r1[i])):
                assert ra["track_status"] == rb["track_status"], (k, i, f)
                assert np.array_equal(ra["pose"], rb["pose"]), (k, i, f)
                assert ra["n_visible"] == rb["n_visible"] and ra["n_allocated"] == rb["n_allocated"], (k, i, f)
                assert ra["rgb_count"] == rb["rgb_count"], (k, i, f)
            xa, sa, wa, ca, va, na = s0[i]
            xb, sb, wb, cb, vb, nb = s1[i]
            assert np.array_equal(xa, xb), (k, i)
            assert same_bits(sa, sb) and np.array_equal(wa, wb) and np.array_equal(ca, cb), (k, i)
            assert same_bits(va, vb) and same_bits(na, nb), (k, i)


# step memberships of the captured-step test (indices into SESSIONS): a step's graph is
# captured the second time its membership is met and replayed afterwards
GRAPH_SESSIONS = (0, 1, 4)
GRAPH_STEPS = [(0, 1, 4), (0, 1, 4), (0, 1, 4), (0, 1, 4), (0, 4), (0, 1, 4), (0, 4), (0, 4)]


def _run_schedule(frames, use_rgb, use_graph, retune=None):
    b = A.Batch(8)
    if use_graph:
        b.set_graph(True)
    systems = {i: _system(i, frames[i][0], use_rgb) for i in GRAPH_SESSIONS}
    for i in GRAPH_SESSIONS:
        b.attach(systems[i])
    nxt = dict.fromkeys(GRAPH_SESSIONS, 0)
    res = {i: [] for i in GRAPH_SESSIONS}
    counts = []
    for step, members in enumerate(GRAPH_STEPS):
        if retune is not None and step == retune[0]:
            A.set_tuning(**retune[1])
        b.grab([systems[i] for i in members], [frames[i][1][nxt[i]][0] for i in members],
               [frames[i][1][nxt[i]][1] for i in members])
        for i in members:
            res[i].append(systems[i].result())
            nxt[i] += 1
        counts.append(b.graph_counts())
    state = {i: _state(systems[i]) for i in GRAPH_SESSIONS}
    for x in systems.values():
        x.close()
    b.close()
    return res, state, counts


@pytest.mark.parametrize("use_rgb", [0, 1])
def test_captured_steps_equal_direct_launches(gpu, use_rgb):
    """fu_batch_set_graph: every step's results and the final maps bit-identical to direct
    launches; the first step of a membership runs directly, the second captures, later ones
    replay; frames passed at fresh device addresses every step (the captured part does not read
    them); a change of launch geometry drops the graphs"""
    frames = {}
    for i in GRAPH_SESSIONS:
        kind, size, seed, cols, rows, _, _ = SESSIONS[i]
        n = sum(i in m for m in GRAPH_STEPS)
        _, K, fr = scene_frames(kind, size, seed, frames=range(n), cols=cols, rows=rows)
        frames[i] = (K, [A.to_device(*f) for f in fr])
    print("direct launches, use_rgb", use_rgb, flush=True)
    r0, s0, c0 = _run_schedule(frames, use_rgb, False)
    assert c0[-1] == (0, 0)
    print("captured steps", flush=True)
    r1, s1, c1 = _run_schedule(frames, use_rgb, True)
    # (captures, graph launches) after each step: step 0 holds first frames (direct), (0,1,4)
    # is captured at step 2, (0,4) met at step 4 and captured at step 6
    assert c1 == [(0, 0), (0, 0), (1, 1), (1, 2), (1, 2), (1, 3), (2, 4), (2, 5)], c1
    base = A.default_tuning()
    try:
        print("captured steps, launch geometry changed at step 3", flush=True)
        r2, s2, c2 = _run_schedule(frames, use_rgb, True, retune=(3, TUNINGS[0]))
    finally:
        A.set_tuning(**base)
    # the change drops the graph of (0,1,4) and the memberships met once: (0,1,4) runs directly
    # at step 3 and is captured again at step 5, (0,4) is captured at step 6
    assert c2 == [(0, 0), (0, 0), (1, 1), (1, 1), (1, 1), (2, 2), (3, 3), (3, 4)], c2
    for r, st in ((r1, s1), (r2, s2)):
        for i in GRAPH_SESSIONS:
            for f, (ra, rb) in enumerate(zip(r0[i], r[i])):
                assert ra["track_status"] == rb["track_status"], (i, f)
                assert np.array_equal(ra["pose"], rb["pose"]), (i, f)
                assert ra["n_visible"] == rb["n_visible"] and ra["n_allocated"] == rb["n_allocated"], (i, f)
                assert ra["rgb_count"] == rb["rgb_count"], (i, f)
            for a, b in zip(s0[i], st[i]):
                assert same_bits(a, b) if a.dtype.kind == "f" else np.array_equal(a, b), i
    assert r0[0][-1]["track_status"] == 1


def test_batch_argument_errors(gpu):
    K = (525.0, 525.0, 319.5, 239.5)
    b = A.Batch(2)
    s1 = A.System(A.system_params(map=MAP))
    s2 = A.System(A.system_params(map=MAP))
    s3 = A.System(A.system_params(map=dict(MAP, voxel_format=2)))
    b.attach(s1)
    with pytest.raises(A.N.FusionError):
        b.attach(s1)  # already attached
    with pytest.raises(A.N.FusionError):
        b.attach(s3)  # other voxel format
    d = torch.zeros((480, 640), dtype=torch.int16, device="cuda")
    c = torch.zeros((480, 640, 3), dtype=torch.uint8, device="cuda")
    with pytest.raises(A.N.FusionError):
        b.grab([s1, s2], [d, d], [c, c])  # s2 not attached
    b.attach(s2)
    with pytest.raises(A.N.FusionError):
        b.grab([s1, s1], [d, d], [c, c])  # twice in one step
    b.grab([s1, s2], [d, d], [c, c])  # empty frames: nothing to fuse, still valid
    r = s1.result()
    assert r["track_status"] == 4 and r["frame_index"] == 1
    b.detach(s1)
    s1.close()
    s2.close()  # destroying an attached system detaches it
    s3.close()
    b.close()
    del K
