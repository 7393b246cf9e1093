"""The C++ drop-in layer (include/fusion_amd_compat.hpp, libfusion_amd_cxx.so): a System-shaped
test program written against the reference's Tracking/Reduction.h and Mapping/RenderScene.h
API (tests/cxx/test_compat.cpp) runs six frames, meshes the map and uses the keypoint map; each
of its calls is checked bit for bit against the same call through the C ABI inside the program.
Here its tracked poses are compared with the library's reference-style system (host solve
per ICP iteration) on the same frames."""
import os
import subprocess

import numpy as np
import pytest

from fusion_amd import api as A
from tests.helpers import scene_frames

pytestmark = pytest.mark.gpu

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cxx", "_build", "test_compat")


def test_cxx_dropin_driver(gpu):
    assert os.path.exists(EXE), "the C++ test driver is built by __graft_entry__.build()"
    print("C++ drop-in driver, 6 frames", flush=True)
    r = subprocess.run([EXE, "6"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "compat ok" in r.stdout
    checks = int(r.stdout.strip().splitlines()[-1].split()[-1])
    assert checks > 100
    lines = [l.split() for l in r.stdout.splitlines() if l.startswith("frame ")]
    assert len(lines) == 6
    _, _, frames = scene_frames(frames=range(6))
    s = A.System(A.system_params(plain=True, map=dict(num_buckets=1 << 17, num_excess=1 << 16, num_blocks=1 << 16,
                                                      voxel_size=0.01,
                                                      trunc_dist=float(np.float32(0.01) * np.float32(3.0)))))
    for l, f in zip(lines, frames):
        s.GrabImage(*A.to_device(*f))
        res = s.result()
        status = int(l[3])
        pose = np.array([float(v) for v in l[5:21]]).reshape(4, 4)
        assert res["track_status"] == status, (l[1], res["track_status"], status)
        assert np.abs(res["pose"] - pose).max() < 1e-9, (l[1], np.abs(res["pose"] - pose).max())
    s.close()
