"""CPU tests: the marching-cubes tables against the reference's fingerprints, and the C
ABI library's exports and argument checking (no GPU needed, no compute calls)."""
import ctypes as C
import hashlib
import os
import re

import numpy as np
import pytest

from fusion_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# sha256(','.join(decimal ints))[:16] of Mapping/Constant.h's edgeTable (:4-37),
# triangleTable (:39-296) and vertexTable (:298-555)
FP_EDGE, FP_TRI, FP_VERT = "07256d860bcfcb99", "446bbde4d59001da", "d245c01cc5e47b59"


def _fp(v):
    return hashlib.sha256(",".join(str(int(x)) for x in v).encode()).hexdigest()[:16]


def _tables():
    e = (C.c_int32 * 256)()
    t = (C.c_int32 * 4096)()
    v = (C.c_int32 * 256)()
    assert N.lib().fu_mc_tables(e, t, v) == 0
    return np.array(e[:]), np.array(t[:]).reshape(256, 16), np.array(v[:])


def test_mc_table_fingerprints():
    e, t, v = _tables()
    assert _fp(e) == FP_EDGE
    assert _fp(t.ravel()) == FP_TRI
    assert _fp(v) == FP_VERT


def test_mc_tables_self_consistent():
    e, t, v = _tables()
    assert v.sum() == 2460
    for c in range(256):
        tri = t[c][t[c] >= 0]
        assert len(tri) == v[c] and len(tri) % 3 == 0 and len(tri) <= 15
        for edge in tri:
            assert e[c] & (1 << edge)
        # complementary cases cut the same edges
        assert e[c] == e[255 - c]


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fu_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = C.CDLL(N.LIB_PATH)
    names = _declared("fusion_amd.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding declares a signature for each
    assert set(names) <= set(N.SIGNATURES), set(names) - set(N.SIGNATURES)


def test_scene_symbols_exported():
    lib = C.CDLL(N.SCENE_PATH)
    names = _declared("fusion_scene.h")
    assert names and all(hasattr(lib, n) for n in names)
    assert set(names) <= set(N.SCENE_SIGNATURES)


def test_argument_errors_without_device():
    lib = N.lib()
    p = N.MapParams()
    lib.fu_map_params_default(C.byref(p))
    assert p.num_buckets == 1 << 21 and abs(p.voxel_size - 0.005) < 1e-9 and p.max_weight == 255
    h = C.c_void_p()
    bad = N.MapParams.from_buffer_copy(p)
    bad.num_buckets = 0
    assert lib.fu_map_create(C.byref(bad), C.byref(h)) == -1
    assert b"positive" in lib.fu_last_error()
    bad = N.MapParams.from_buffer_copy(p)
    bad.max_weight = 300  # u8 weights cap at 255
    assert lib.fu_map_create(C.byref(bad), C.byref(h)) == -1
    bad = N.MapParams.from_buffer_copy(p)
    bad.voxel_format = 7
    assert lib.fu_map_create(C.byref(bad), C.byref(h)) == -1
    assert lib.fu_filter_depth(None, None, None, 640, 480, 5000.0, 3.0, None) == -1
    assert lib.fu_compute_nmap(None, None, 0, 0, None) == -1
    sp = N.SystemParams()
    lib.fu_system_params_default(C.byref(sp))
    assert (sp.iters[0], sp.iters[1], sp.iters[2]) == (10, 5, 4)
    sp.cols = 30
    assert lib.fu_system_create(C.byref(sp), C.byref(h)) == -1


def test_map_bytes_scales_with_format():
    lib = N.lib()
    p = N.MapParams()
    lib.fu_map_params_default(C.byref(p))
    p.num_blocks = 1000
    sizes = []
    for fmt in (0, 1, 2):
        p.voxel_format = fmt
        sizes.append(lib.fu_map_bytes(C.byref(p)))
    # per block: 4096 B (f32+u8+rgb), 4608 B (f32+u16+rgb), 3072 B (f16+u8+rgb)
    assert sizes[1] - sizes[0] == 1000 * 512 and sizes[0] - sizes[2] == 1000 * 1024


def test_api_rejects_non_contiguous_tensors():
    """the library reads images as dense row-major arrays: a strided view is refused before any
    call instead of being read with the wrong layout"""
    import torch
    from fusion_amd import api as A
    t = torch.zeros((4, 6, 4), dtype=torch.float32)
    assert A._addr(t) == t.data_ptr()
    with pytest.raises(N.FusionError):
        A._addr(t.transpose(0, 1))
    with pytest.raises(N.FusionError):
        A._p(t[:, ::2])


def test_relocalisation_argument_errors_without_device():
    """the relocalisation entry points refuse bad arguments before any device call"""
    lib = N.lib()
    I = C.c_int
    # sub-graph selection: null buffers, too many pairs, lists too long
    assert lib.fu_select_subgraphs(None, 10, 10, 400, 0.1, None, None, None, None, None, None) == -1
    assert b"null" in lib.fu_last_error()
    assert lib.fu_select_subgraphs(None, 46341, 10, 400, 0.1, None, None, None, None, None, None) == -1
    assert b"46340" in lib.fu_last_error()
    assert lib.fu_select_subgraphs(None, 10, 10, 1025, 0.1, None, None, None, None, None, None) == -1
    assert lib.fu_select_subgraphs(None, 10, 0, 400, 0.1, None, None, None, None, None, None) == -1
    poses = (C.c_double * 80)()
    npose, npairs, ok, chosen = I(), I(), I(), I()
    assert lib.fu_relocalise_graph_candidates(None, None, None, None, 5, 100, 0, poses, C.byref(npose),
                                              C.byref(npairs), None) == -1
    assert lib.fu_relocalise_keys(None, None, None, 5, 100, 0, poses, C.byref(npairs), C.byref(ok), None) == -1
    it = (I * 3)(3, 3, 2)
    assert lib.fu_system_validate_poses(None, poses, 1, it, 1e-3, None, C.byref(chosen), C.byref(ok), None) == -1
    assert lib.fu_system_relocalise_graph(None, None, None, None, None, 5, 100, 0, None, C.byref(ok), None) == -1
    assert lib.fu_match_knn2(None, -1, None, 0, None, None, None) == -1
    assert lib.fu_lift_keypoints(None, -1, None, None, 640, 480, 525.0, 525.0, 319.5, 239.5, 3.0, None, None, None,
                                 None) == -1
