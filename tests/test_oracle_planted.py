"""Each pinning test catches the mistake it is there for.

oracle/oracle.cpp builds, with -DOR_PLANT=k, a variant holding one plausible mistake (the
projection of the ICP association in double, the current frame's normal in the ICP row, a 3-D
dot for the integration weight, old and new colours swapped in the blend, fragment columns
counted to nx instead of ny, the X Euler sine from a transposed element).  Each variant is
built here with the oracle's own flags and the test that pins that function is run against it
in a child process; it must fail with an assertion.
"""
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

PLANTS = {
    1: "tests/test_oracle_icp.py::test_icp_sums_match_independent_fp64_at_a_moving_pose",
    2: "tests/test_oracle_icp.py::test_icp_sums_match_independent_fp64_at_a_moving_pose",
    3: "tests/test_oracle_pins.py::test_integration_of_two_walls_weight_and_colour",
    4: "tests/test_oracle_pins.py::test_integration_of_two_walls_weight_and_colour",
    5: "tests/test_oracle_pins.py::test_zrange_tiles_brute_force",
    6: "tests/test_oracle_pins.py::test_euler_sines_of_the_acceptance_test_against_scipy",
}


@pytest.fixture(scope="module")
def variants():
    with tempfile.TemporaryDirectory() as d:
        def build(k):
            out = os.path.join(d, "liboracle_plant%d.so" % k)
            subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
                            "-DOR_PLANT=%d" % k, "-shared", "oracle.cpp", "mc_tables.cpp", "-o", out],
                           cwd=ORACLE, check=True, capture_output=True)
            return k, out

        with ThreadPoolExecutor(max_workers=6) as ex:
            yield dict(ex.map(build, PLANTS))


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_pinning_test_fails_on_planted_mistake(variants, k):
    env = dict(os.environ, FU_ORACLE_LIB=variants[k], OMP_NUM_THREADS="4")
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider", PLANTS[k]],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 1 and "1 failed" in out and "AssertionError" in out, out[-3000:]
