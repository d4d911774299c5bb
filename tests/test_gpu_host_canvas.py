"""Runs the C++ known-answer tests of the Path/Paint/Canvas front end
(tests/cpp/hip_canvas_test.cc, built by __graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hip_canvas_test")


def test_cpp_front_end_known_answers():
    assert os.path.exists(BIN), "tests/cpp/hip_canvas_test is not built"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
