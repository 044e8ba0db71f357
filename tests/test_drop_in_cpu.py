"""The drop-in EuclideanDistance class with no HIP device: the reference's
CPU branch (EuclideanDistance.cu:82-85, 107-117), run by the library's own
host loop, bit-equal to the oracle from the object, a copy, a moved object
and a copy at a reused address; FT_REQUIRE_DEVICE=1 makes init throw."""
import os
import shutil

import numpy as np
import pytest

import fluidtissue_amd as fa
from tests.drop_in import run_link_test

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


@pytest.fixture(scope="module")
def no_device():
    if fa.device_count() > 0:
        pytest.skip("a HIP device is present: the device path is tested under -m gpu")


@pytest.mark.parametrize("n", [12, 20])
def test_cpu_branch_matches_oracle(tmp_path, no_device, n):
    got, ref = run_link_test(tmp_path, copies=True, n=n)
    assert "no HIP device, running the CPU branch" in got["stderr"]
    for col in got["phi_cols"]:
        assert np.array_equal(col.view(np.uint64), ref.view(np.uint64))


def test_require_device_makes_init_throw(tmp_path, no_device):
    env = dict(os.environ, FT_REQUIRE_DEVICE="1")
    run = run_link_test(tmp_path, env=env, check=False)
    assert run.returncode != 0
    assert "FT_REQUIRE_DEVICE" in run.stderr
