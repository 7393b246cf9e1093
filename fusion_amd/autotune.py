"""Autotuner of the launch geometry of the batched stages (fu_tuning).

The knobs (include/fusion_amd.h, fu_tuning) change how the batched kernels are laid out on
the GPU, never what they compute: the pixel layout of a ray-march wave, whether consecutive
work-groups of ICP / integration / raycast run on one XCD (one L2), and the work-group counts
of the grid-stride stages.  The tuner measures candidates on a live workload and keeps the
fastest, one knob at a time (coordinate descent in the order of SPACE), each knob judged by
the HIP-event time of the stage it affects.

Candidates of a knob are measured round-robin, `reps` rounds, so that slow drift of the
workload (maps growing as frames are fused) does not favour one of them; a candidate replaces
the current value only if its median is faster by more than `margin`.  Results are cached in a
JSON file keyed by GPU architecture and workload (frame size, sessions per step, voxel format,
tracking mode).  No reference counterpart: the reference's launch shapes are compile-time
constants.
"""
import json
import os

import numpy as np

# (knob, stage whose time judges it, candidates).  "xcd_order:<stage>" is one bit of the
# xcd_order mask.
SPACE = [
    ("raycast_wave", "raycast", [0, 1, 2]),
    ("xcd_order:raycast", "raycast", [0, 1]),
    ("xcd_order:integrate", "integrate", [0, 1]),
    ("integrate_groups", "integrate", [4096, 8192, 16384, 32768, 65536]),
    ("xcd_order:icp", "icp", [0, 1]),
    ("visible_groups", "visible", [2048, 4096, 8192, 16384]),
    ("alloc_groups", "alloc", [512, 1024, 2048, 4096, 8192]),
]
STAGE_BIT = {"resize": 0, "pyramid": 1, "icp": 2, "alloc": 3, "visible": 4, "integrate": 5, "zrange": 6,
             "raycast": 7}
DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")


def get_knob(tuning, knob):
    if knob.startswith("xcd_order:"):
        return (tuning["xcd_order"] >> STAGE_BIT[knob.split(":", 1)[1]]) & 1
    return tuning[knob]


def with_knob(tuning, knob, value):
    t = dict(tuning)
    if knob.startswith("xcd_order:"):
        bit = 1 << STAGE_BIT[knob.split(":", 1)[1]]
        t["xcd_order"] = (t["xcd_order"] | bit) if value else (t["xcd_order"] & ~bit)
    else:
        t[knob] = int(value)
    return t


def search(measure, base, space=SPACE, reps=3, margin=0.01):
    """Coordinate descent over `space` from the tuning dict `base`.  measure(tuning) runs one
    round of the workload with that tuning and returns {stage: ms per step}.  Returns (best
    tuning, report); the report lists every knob's candidates with their median times."""
    cur = dict(base)
    report = []
    for knob, stage, cands in space:
        now = get_knob(cur, knob)
        if now not in cands:
            cands = [now] + list(cands)
        times = {c: [] for c in cands}
        for _ in range(reps):
            for c in cands:
                times[c].append(float(measure(with_knob(cur, knob, c))[stage]))
        med = {c: float(np.median(v)) for c, v in times.items()}
        best = min(cands, key=lambda c: (med[c], c != now))
        if med[best] < med[now] * (1.0 - margin):
            cur = with_knob(cur, knob, best)
        else:
            best = now
        report.append(dict(knob=knob, stage=stage, ms={str(c): round(m, 4) for c, m in med.items()}, chosen=best))
    return cur, report


def palindrome(n):
    """frame indices 0, 1, .., n-1, n-2, .., 1, 0, 1, ..: a replay in which every step moves
    the camera by one frame"""
    k = 0
    while True:
        if n <= 1:
            yield 0
            continue
        period = 2 * (n - 1)
        j = k % period
        yield j if j < n else period - j
        k += 1


def tune_server(A, server, sids, frames, reps=3, space=SPACE, margin=0.01):
    """Tune on a running server: `sids` are open sessions, frames[j] the device frames (depth,
    rgb) of session j, replayed as a palindrome (one frame of every session per measurement,
    with the server's per-stage profiling on).  Applies the best tuning and returns (tuning,
    report)."""
    order = palindrome(min(len(f) for f in frames))

    def measure(t):
        A.set_tuning(**t)
        f = next(order)
        server.set_profiling(True)
        server.run_trace([(sid, frames[j][f][0], frames[j][f][1], 0.0, 7_000_000 + j) for j, sid in enumerate(sids)])
        st = server.stage_times()
        server.set_profiling(False)
        steps = max(1.0, st["steps"])
        return {k: v / steps for k, v in st["ms"].items()}

    base = A.get_tuning()
    try:
        best, report = search(measure, base, space, reps, margin)
    except BaseException:
        A.set_tuning(**base)
        raise
    A.set_tuning(**best)
    return best, report


def cache_key(arch, cols, rows, per_step, voxel_format, use_rgb):
    return "%s|%dx%d|%d|fmt%d|rgb%d" % (arch.split(":")[0], cols, rows, per_step, voxel_format, use_rgb)


def device_arch(torch):
    return torch.cuda.get_device_properties(0).gcnArchName


def load(key, path=DEFAULT_CACHE):
    """the cached tuning of `key`, or None"""
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        return None
    e = doc.get("entries", {}).get(key)
    return dict(e["tuning"]) if e else None


def save(key, tuning, report=None, path=DEFAULT_CACHE):
    try:
        with open(path) as f:
            doc = json.load(f)
    except (OSError, ValueError):
        doc = {"version": 1, "entries": {}}
    doc.setdefault("entries", {})[key] = {"tuning": dict(tuning), "report": report}
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(doc, f, indent=1, sort_keys=True)
    os.replace(tmp, path)
