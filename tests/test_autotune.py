"""The launch-geometry autotuner's search and cache, and the library's tuning setter (host
code only; the GPU test that geometry never changes a result is in test_gpu_batch.py)."""
import ctypes as C
import itertools

import numpy as np
import pytest

from fusion_amd import autotune as T
from fusion_amd import native as N

BASE = dict(raycast_wave=0, xcd_order=0, integrate_groups=16384, visible_groups=8192, alloc_groups=2048)


def _fake(optimum, noise=0.0, seed=0):
    """a workload whose stage times are smallest at `optimum` (separable per knob), with
    multiplicative noise; counts the measurements"""
    rng = np.random.default_rng(seed)
    calls = []

    def measure(t):
        calls.append(dict(t))
        ms = {s: 1.0 for s in T.STAGE_BIT}
        for knob, stage, cands in T.SPACE:
            v, o = T.get_knob(t, knob), T.get_knob(optimum, knob)
            ms[stage] += 0.0 if v == o else 0.05 * (1 + abs(np.log2(v + 1.0) - np.log2(o + 1.0)))
        return {s: m * (1.0 + noise * rng.standard_normal()) for s, m in ms.items()}

    return measure, calls


def test_search_finds_a_separable_optimum_through_noise():
    opt = dict(raycast_wave=1, xcd_order=(1 << 7) | (1 << 2), integrate_groups=32768, visible_groups=4096,
               alloc_groups=1024)
    measure, calls = _fake(opt, noise=0.005)
    best, report = T.search(measure, BASE, reps=5)
    assert best == opt
    assert [r["knob"] for r in report] == [k for k, _, _ in T.SPACE]
    assert len(calls) == 5 * sum(len(c) for _, _, c in T.SPACE)


def test_search_keeps_the_current_value_within_the_margin():
    # every candidate within 0.5 %: nothing changes
    def flat(t):
        return {s: 1.0 + 0.004 * (T.get_knob(t, "raycast_wave") + T.get_knob(t, "alloc_groups") / 8192.0)
                for s in T.STAGE_BIT}

    best, _ = T.search(flat, dict(BASE, raycast_wave=2), reps=2, margin=0.01)
    assert best == dict(BASE, raycast_wave=2)


def test_search_adds_an_off_list_current_value():
    measure, calls = _fake(dict(BASE, alloc_groups=777))
    space = [("alloc_groups", "alloc", [512, 1024])]
    best, report = T.search(measure, dict(BASE, alloc_groups=777), space=space, reps=1)
    assert best["alloc_groups"] == 777 and set(report[0]["ms"]) == {"777", "512", "1024"}


def test_xcd_bits_are_independent():
    t = dict(BASE)
    for stage in ("icp", "integrate", "raycast"):
        t = T.with_knob(t, "xcd_order:" + stage, 1)
    assert t["xcd_order"] == (1 << 2) | (1 << 5) | (1 << 7)
    t = T.with_knob(t, "xcd_order:integrate", 0)
    assert t["xcd_order"] == (1 << 2) | (1 << 7)
    assert T.get_knob(t, "xcd_order:raycast") == 1 and T.get_knob(t, "xcd_order:integrate") == 0


def test_palindrome_moves_one_frame_per_step():
    seq = list(itertools.islice(T.palindrome(4), 14))
    assert seq == [0, 1, 2, 3, 2, 1, 0, 1, 2, 3, 2, 1, 0, 1]
    assert list(itertools.islice(T.palindrome(1), 3)) == [0, 0, 0]


def test_cache_round_trip(tmp_path):
    path = str(tmp_path / "tuning.json")
    key = T.cache_key("gfx950:sramecc+:xnack-", 640, 480, 32, 0, 0)
    assert key == "gfx950|640x480|32|fmt0|rgb0"
    assert T.load(key, path) is None
    T.save(key, dict(BASE, raycast_wave=1), [{"knob": "raycast_wave"}], path=path)
    T.save("other", BASE, None, path=path)
    assert T.load(key, path) == dict(BASE, raycast_wave=1)
    assert T.load("other", path) == BASE


def test_committed_cache_entries_are_valid_tunings():
    import json
    import os
    if not os.path.exists(T.DEFAULT_CACHE):
        pytest.skip("no committed tuning cache")
    doc = json.load(open(T.DEFAULT_CACHE))
    lib = N.lib()
    for key, e in doc["entries"].items():
        t = N.Tuning(**e["tuning"])
        assert lib.fu_set_tuning(C.byref(t)) == 0, key
    d = N.Tuning()
    lib.fu_tuning_default(C.byref(d))
    assert lib.fu_set_tuning(C.byref(d)) == 0


def test_library_tuning_setter_validates():
    lib = N.lib()
    d = N.Tuning()
    lib.fu_tuning_default(C.byref(d))
    assert (d.raycast_wave, d.xcd_order, d.integrate_groups, d.visible_groups, d.alloc_groups) == \
        (1, 0, 16384, 8192, 2048)
    try:
        for field, bad in [("raycast_wave", 3), ("raycast_wave", -1), ("xcd_order", 1 << 8),
                           ("integrate_groups", 255), ("visible_groups", 262145), ("alloc_groups", 0)]:
            t = N.Tuning.from_buffer_copy(d)
            setattr(t, field, bad)
            assert lib.fu_set_tuning(C.byref(t)) == -1, field
        t = N.Tuning.from_buffer_copy(d)
        t.raycast_wave, t.xcd_order, t.integrate_groups = 2, 0xA4, 262144
        assert lib.fu_set_tuning(C.byref(t)) == 0
        g = N.Tuning()
        assert lib.fu_get_tuning(C.byref(g)) == 0
        assert (g.raycast_wave, g.xcd_order, g.integrate_groups) == (2, 0xA4, 262144)
    finally:
        assert lib.fu_set_tuning(C.byref(d)) == 0
