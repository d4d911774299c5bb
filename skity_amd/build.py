"""Builds the native libraries in-tree.

  skity_amd/libskity_hip.so    gfx950 kernels + C ABI (hipcc)
  skity_amd/libskity_scene.so  seeded scene generator (g++)
  skity_amd/libskity_hip_canvas.so  C++ Path/Paint/Canvas front end, SKP reader,
                                    display lists (g++)
  tests/cpp/hip_canvas_test    C++ known-answer tests of the front end (g++)
  oracle/liboracle.so          CPU oracle, test infrastructure (make)
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-fvisibility=hidden",
    # every float operation rounds separately, as in the reference build
    "-ffp-contract=off",
    "-Wno-unused-value", "-Wno-unused-result",
    # packed fp32 instructions are not used on this device
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, cwd=cwd, check=True)


# translation units of libskity_hip.so; the wide walks compile beside the
# main unit
HIP_UNITS = ["skity_hip.hip", "hipr_walk_wide.hip"]


def build_hip():
    out = os.path.join(HERE, "libskity_hip.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for unit in HIP_UNITS:
        obj = os.path.join(objdir, unit.replace(".hip", ".o"))
        cmd = [HIPCC] + HIP_FLAGS + ["-c", "-o", obj, os.path.join(HERE, "csrc", unit)]
        print("+", " ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])
    run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out + ".tmp"] + objs)
    os.replace(out + ".tmp", out)


def build_scene():
    src = os.path.join(HERE, "csrc", "scene_gen.cpp")
    out = os.path.join(HERE, "libskity_scene.so")
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fvisibility=hidden", "-pthread",
         "-o", out + ".tmp", src])
    os.replace(out + ".tmp", out)


def build_host():
    srcs = [os.path.join(HERE, "host", f) for f in ("hip_canvas.cc", "hip_picture.cc")]
    out = os.path.join(HERE, "libskity_hip_canvas.so")
    run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", out + ".tmp"] + srcs +
        ["-L" + HERE, "-lskity_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    test_src = os.path.join(ROOT, "tests", "cpp", "hip_canvas_test.cc")
    test_out = os.path.join(ROOT, "tests", "cpp", "hip_canvas_test")
    run(["g++", "-O2", "-std=c++17", "-o", test_out + ".tmp", test_src, "-L" + HERE,
         "-lskity_hip_canvas", "-lskity_hip", "-Wl,-rpath,$ORIGIN/../../skity_amd"])
    os.replace(test_out + ".tmp", test_out)


def build_oracle():
    run(["make", "-s", "-B", "liboracle.so"], cwd=os.path.join(ROOT, "oracle"))


def build_all():
    build_oracle()
    build_scene()
    build_hip()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "hip": build_hip, "scene": build_scene,
         "oracle": build_oracle, "host": build_host}[w]()
