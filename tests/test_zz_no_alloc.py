"""fu_icp_step_noalloc (ICPStep with its state in a fixed device array, for callers that cannot
allocate) against fu_icp_step through the C ABI: identical residual, A and b on two calls in a
row, the second on a smaller frame, then again after the allocating entry point has run."""
import ctypes as C

import numpy as np
import pytest
import torch

from fusion_amd import api as A
from fusion_amd import native as N
from tests.helpers import scene_frames
from tests.test_gpu_pyramid import _pyr_oracle

pytestmark = pytest.mark.gpu


def _maps(cols, rows):
    sc, K, fr = scene_frames(frames=(0, 2), cols=cols, rows=rows)
    _, _, v0, n0 = _pyr_oracle(fr[0][0], K)
    _, _, v1, n1 = _pyr_oracle(fr[1][0], K)
    gt = np.linalg.inv(sc.pose(0)) @ sc.pose(2)
    cu = lambda a: torch.from_numpy(a).cuda()
    return K, cu(v1[0]), cu(v0[0]), cu(n1[0]), cu(n0[0]), gt


def _step(fn, K, nv, lv, nn, ln, pose_curr):
    rows, cols, _ = nv.shape
    pc, pl = A.pose_f(pose_curr), A.pose_f(np.eye(4))
    res = (C.c_float * 2)()
    Am = (C.c_double * 36)()
    b = (C.c_double * 6)()
    N.check(fn(A._p(nv), A._p(lv), A._p(nn), A._p(ln), cols, rows, pc.R, pc.t, pl.R, pl.Rinv, pl.t, K[0], K[1], K[2],
               K[3], res, Am, b, A._s()), "icp step")
    return np.array(res[:], np.float32), np.array(Am[:]), np.array(b[:])


def test_noalloc_icp_step_equals_icp_step(gpu):
    lib = N.lib()
    big = _maps(640, 480)
    small = _maps(320, 240)
    print("noalloc icp step 640x480 then 320x240", flush=True)
    for maps in (big, small, big):
        K, nv, lv, nn, ln, gt = maps
        got = _step(lib.fu_icp_step_noalloc, K, nv, lv, nn, ln, gt)
        ref = _step(lib.fu_icp_step, K, nv, lv, nn, ln, gt)
        assert got[0][1] > 1000
        for x, y in zip(got, ref):
            assert np.array_equal(x.view(np.uint8), y.view(np.uint8))


def test_noalloc_refuses_bad_sizes(gpu):
    K, nv, lv, nn, ln, gt = _maps(320, 240)
    pc, pl = A.pose_f(gt), A.pose_f(np.eye(4))
    res = (C.c_float * 2)()
    Am = (C.c_double * 36)()
    b = (C.c_double * 6)()
    for cols, rows in ((0, 240), (1 << 16, 1 << 15)):
        st = N.lib().fu_icp_step_noalloc(A._p(nv), A._p(lv), A._p(nn), A._p(ln), cols, rows, pc.R, pc.t, pl.R,
                                         pl.Rinv, pl.t, K[0], K[1], K[2], K[3], res, Am, b, A._s())
        assert st != 0 and b"noalloc" in N.lib().fu_last_error()
