"""CPU build of the point -> triangle stage's per-element steps (the device
kernels k_p2t_count, k_p2t_fill and k_point_area call the same functions):
a valid interface gives the ascending lists and ordered area sums, and a
vertex out of range, a full list or a list entry that is not a triangle is
reported without a write out of bounds (AddressSanitizer build)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_p2t_index_checks_cpu_build(tmp_path):
    exe = tmp_path / "p2t_guard_test"
    subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=all", "-I" + os.path.join(ROOT, "fluidtissue_amd", "csrc"),
                    os.path.join(ROOT, "tests", "cpp", "p2t_guard_test.cpp"), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "ok"
