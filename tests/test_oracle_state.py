"""The tracker's state machine in the oracle's System::GrabImage (Tracking/Tracking.cc:59-112,
Core/System.cc:92-111), against the sequence read off the reference's code by hand:

* frame 0: state 1, InitTracking, fused (status FIRST);
* ten frames that fail tracking: status FAILED, nothing fused; the tenth makes
  noMissedFrames = 10 > 9, so the next frame's state is -1 (lost);
* lost frames: Relocalise fails without a keypoint map: status LOST, nothing happens;
* after a relocalisation pose is given: that frame is swapped and raycast, not fused
  (status RELOC, num_frames_after_reloc = 0); the next frame is tracked but GrabFrame returns
  false (lastState was -1): nothing fused, not swapped; then frames are swapped and raycast,
  and fused again once num_frames_after_reloc reaches 10, i.e. from the tenth frame after
  the relocalisation on.
"""
import numpy as np

from oracle import binding as O
from tests.helpers import scene_frames

K320 = (262.5, 262.5, 159.5, 119.5)


def _weight_sum(o):
    b = o.map.export_blocks()
    return int(o.map.read_blocks(b)[1].astype(np.int64).sum()) if len(b) else 0


def test_lost_after_ten_misses_and_relocalisation():
    sc, K, fr = scene_frames(frames=range(0, 16), cols=320, rows=240, K=(525.0, 525.0, 319.5, 239.5))
    o = O.System(cols=320, rows=240, K=K, num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    st, p0 = o.grab(*fr[0])
    assert st == 4
    w = _weight_sum(o)
    assert w > 0
    for k in range(10):
        st, p = o.grab(*empty)
        assert st == 2 and np.array_equal(p, p0), k
    for k in range(3):  # lost: even a good frame is not tracked
        st, p = o.grab(*fr[1])
        assert st == 5 and np.array_equal(p, p0), k
    assert _weight_sum(o) == w
    # relocalisation at the first frame's pose; the next frames continue from frame 1
    assert o.relocalise(p0) == 0
    assert o.relocalise(p0) == 0  # still lost until the next frame
    st, p = o.grab(*fr[1])
    assert st == 6 and np.array_equal(p, p0)
    assert o.relocalise(p0) == -1  # no longer lost
    assert _weight_sum(o) == w
    seen = []
    for k in range(2, 14):
        st, p = o.grab(*fr[k])
        seen.append((st, _weight_sum(o)))
    # tracked (or rejected by the final error test: while fusion waits the model does not
    # grow); the map grows only from the tenth frame after the relocalisation on (frames
    # 2..10 are frames 1..9 after it: num_frames_after_reloc 1..9)
    assert all(s in (1, 3) for s, _ in seen) and sum(s == 1 for s, _ in seen) >= 9, seen
    assert seen[9][0] == 1
    assert all(ws == w for _, ws in seen[:9]), seen
    assert seen[9][1] > w and seen[10][1] > seen[9][1], seen


def test_missed_count_resets_on_a_tracked_frame():
    sc, K, fr = scene_frames(frames=range(0, 4), cols=320, rows=240, K=(525.0, 525.0, 319.5, 239.5))
    o = O.System(cols=320, rows=240, K=K, num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15)
    empty = (np.zeros((240, 320), np.uint16), np.zeros((240, 320, 3), np.uint8))
    assert o.grab(*fr[0])[0] == 4
    for _ in range(9):
        assert o.grab(*empty)[0] == 2
    assert o.grab(*fr[1])[0] == 1  # nine misses: not lost; a tracked frame resets the count
    for _ in range(9):
        assert o.grab(*empty)[0] == 2
    assert o.grab(*fr[2])[0] == 1
    assert o.relocalise(np.eye(4)) == -1


def test_mapping_disabled_freezes_the_model():
    """Tracker::mappingDisabled (System.cc:108-126): with mapping off, frames are tracked and
    rendered against the model, which does not change (the same blocks, the same voxels);
    with mapping on again the map grows from the next frame"""
    sc, K, fr = scene_frames(frames=range(0, 9), cols=320, rows=240, K=(525.0, 525.0, 319.5, 239.5))
    o = O.System(cols=320, rows=240, K=K, num_buckets=1 << 16, num_excess=1 << 15, num_blocks=1 << 15)
    for f in fr[:3]:
        assert o.grab(*f)[0] in (1, 4)
    b0 = o.map.export_blocks()
    v0 = o.map.read_blocks(b0)
    o.set_mapping(False)
    for f in fr[3:6]:
        assert o.grab(*f)[0] == 1
    b1 = o.map.export_blocks()
    assert np.array_equal(b0, b1)
    v1 = o.map.read_blocks(b1)
    assert all(np.array_equal(x, y, equal_nan=True) if x.dtype.kind == "f" else np.array_equal(x, y)
               for x, y in zip(v0, v1))
    o.set_mapping(True)
    assert o.grab(*fr[6])[0] == 1
    assert len(o.map.export_blocks()) > len(b0)
