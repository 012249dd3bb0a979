"""Builds the native libraries in-tree.

fledge_amd/libfledge_hip.so    HIP kernels + C ABI (hipcc, gfx950)
fledge_amd/libfledge_scene.so  host workload generators (g++)
fledge_amd/libfledge_host.so   C++ drop-ins over the C ABI (LBVHAccel, ...)

Device code is compiled with -ffp-contract=off (no implicit FMA: the kernels
reproduce the reference's operation order bit for bit) and with packed FP32
instructions disabled.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")

HIP_SOURCES = ["ctx.hip", "halton.hip", "lbvh.hip", "trace.hip", "volume.hip", "adjoint.hip", "volpath.hip", "film_io.hip",
               "scene_accel.hip", "ply_io.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: %s\n%s" % (" ".join(cmd), r.stdout))
    return r.stdout


def build_hip(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    hipcc = _hipcc()
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([hipcc, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, srcs))
    out = os.path.join(HERE, "libfledge_hip.so")
    tmp = out + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, out)
    return out


def build_scene():
    out = os.path.join(HERE, "libfledge_scene.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", tmp,
          os.path.join(CSRC, "scene.cpp"), "-lpthread"])
    os.replace(tmp, out)
    return out


def build_host():
    src = os.path.join(CSRC, "host", "fledge_gpu.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(HERE, "libfledge_host.so")
    tmp = out + ".tmp"
    # hipcc, as HIP: the host mirrors compile the device headers' traversal and
    # trackers with the kernels' floating-point flags
    flags = [f for f in HIP_FLAGS if f != "-fvisibility=hidden"]
    _run([_hipcc(), *flags, "-x", "hip", "-shared", "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(CSRC, "host"), "-o", tmp, src, "-L", HERE, "-lfledge_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, out)
    return out


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def build_all(oracle=True):
    build_scene()
    build_hip()
    build_host()
    if oracle:
        build_oracle()


if __name__ == "__main__":
    build_all(oracle="--no-oracle" not in sys.argv)
    print("built")
