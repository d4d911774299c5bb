"""Address and undefined-behaviour sanitizers over the CPU code: the oracle
(rasteriser, stroker, renderer, tile bins, blur layers, StackBlur), the scene
generator and the host build of the device stroker, run over seeded slices of
the benchmark workloads (tests/sanitize/sanitize_check.cc, built by the
Makefile beside it with clang's -fsanitize=address,undefined on host code
only).  The run aborts on the first sanitizer report; it also checks that
the threaded oracle render equals the single-threaded one and that the host
stroker's outlines equal the oracle's bit for bit.  A second build runs the
C++ front end's SKP reader over a real SKP and corrupted copies of it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DIR = os.path.join(HERE, "sanitize")
CLANG = "/opt/rocm/llvm/bin/clang++"
HIPCC = "/opt/rocm/bin/hipcc"


def test_sanitized_cpu_code():
    if not (os.path.exists(CLANG) and os.path.exists(HIPCC)) or shutil.which("make") is None:
        pytest.skip("clang++/hipcc/make not available")
    subprocess.run(["make", "-s", "-j8", "sanitize_check"], cwd=DIR, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(DIR, "sanitize_check")], cwd=DIR, env=env,
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitized run clean" in r.stdout


def test_sanitized_picture_reader():
    """The SKP reader, playback, display-list replay and scene packing of
    the C++ front end under the sanitizers, over tiger.skp and 400 seeded
    corruptions and truncations of it (tests/sanitize/picture_check.cc)."""
    if not os.path.exists(CLANG) or shutil.which("make") is None:
        pytest.skip("clang++/make not available")
    lib = os.path.join(HERE, "..", "skity_amd", "libskity_hip.so")
    if not os.path.exists(lib):
        pytest.skip("libskity_hip.so is not built")
    subprocess.run(["make", "-s", "-j8", "picture_check"], cwd=DIR, check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    skp = os.path.join(HERE, "golden", "skp", "tiger.skp")
    r = subprocess.run([os.path.join(DIR, "picture_check"), skp], cwd=DIR, env=env,
                       capture_output=True, text=True, timeout=600)
    print(r.stdout[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "picture run clean" in r.stdout
