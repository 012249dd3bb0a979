"""The C++ drop-ins (LBVHAccel, LBVHBuilder, GridVolume, GPUVolIntegrator,
pmr resources) through the compiled harness tests/cpp/dropin_test.cpp."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "cpp", "build", "dropin_test")


def _binary():
    if not os.path.exists(BIN):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    return BIN


def _run(args, timeout):
    r = subprocess.run([_binary(), *args], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True, timeout=timeout)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "ALL PASSED" in r.stdout


def test_dropins_host_side():
    """GridVolume::tr/sample (host mirror of the device trackers) bit-equal to
    the oracle's OpenVDBVolume::tr/sample; error paths; MakeTriangleMesh on
    ascii and binary PLY files; ParseSceneXML on a scene file in the
    reference's format and on the files it refuses.  No GPU call."""
    _run(["--cpu"], 120)


def test_traversal_stack_policies_host():
    """The device traversal with its private stack and with the LDS ring stack
    (rings of 1, 2 and 8 entries: the spill path is exercised) gives the same
    hits, t and node visits as the oracle, run on the CPU."""
    exe = os.path.join(HERE, "cpp", "build", "stack_test")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp")], check=True)
    r = subprocess.run([exe], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0 and "ALL PASSED" in r.stdout, r.stdout


def test_host_library_exports_dropins():
    out = subprocess.run(["nm", "-DC", os.path.join(os.path.dirname(HERE), "fledge_amd",
                                                    "libfledge_host.so")],
                         stdout=subprocess.PIPE, text=True, check=True).stdout
    for sym in ("fledge::LBVHAccel::intersect", "fledge::experimental::LBVHBuilder::build",
                "fledge::GridVolume::tr", "fledge::GridVolume::sample",
                "fledge::GPUVolIntegrator::render", "fledge::gpu_memory_resource::do_allocate",
                "fledge::managed_memory_resource::do_allocate",
                "fledge::async_memory_resource::do_allocate", "fledge::GlobalManagedMemoryResource",
                "fledge::SceneAccel::intersect", "fledge::MakeTriangleMesh", "fledge::HIPRender::render",
                "fledge::ParseSceneXML"):
        assert sym in out, sym


@pytest.mark.gpu
def test_dropins_on_gpu():
    """bvh_test-style: LBVHBuilder host mirror, GPU batch and oracle agree bit
    for bit on 20000 rays; LBVHAccel's intersect contract; SceneAccel over PLY
    meshes (host mirror and GPU) against the oracle's Scene::intersect;
    GPUVolIntegrator's film against the oracle; HIPRender (the RenderBase in
    the OptiX stub's place) on a volume-only and a mesh + medium scene;
    memory resources."""
    _run([], 300)
