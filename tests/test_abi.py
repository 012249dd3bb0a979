"""The C ABI: every function include/fledge_hip.h and include/fledge_scene.h
declare is exported by the built libraries and bound by the ctypes layer.
Loads the libraries only; no compute call (runs without a GPU)."""
import ctypes
import os
import re
import subprocess

from fledge_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, pattern):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(pattern, text)))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", lib], stdout=subprocess.PIPE, text=True,
                         check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_hip_header_symbols_exported_and_bound():
    names = _declared("fledge_hip.h", r"FLG_API\s+[\w\s\*]+?\b(flg_\w+)\s*\(")
    assert len(names) >= 30, names
    exported = _exported(N.HIP_LIB)
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = ctypes.CDLL(N.HIP_LIB)
    for n in names:
        assert getattr(lib, n) is not None
    unbound = [n for n in names if n not in N.declared_symbols()]
    assert not unbound, unbound
    # nothing beyond the header leaks out of the library
    extra = sorted(s for s in exported if s.startswith("flg_") and s not in names)
    assert not extra, extra


def test_scene_header_symbols_exported():
    names = _declared("fledge_scene.h", r"\b(flg_scene_\w+)\s*\(")
    assert len(names) >= 7, names
    exported = _exported(N.SCENE_LIB)
    assert not [n for n in names if n not in exported]


def test_status_strings_without_gpu():
    lib = N.hip()
    assert lib.flg_status_string(0).decode()
    for st in range(1, 6):
        assert lib.flg_status_string(st).decode()


def test_diff_tape_bytes_without_gpu():
    """flg_volpath_diff_tape_bytes: width*height*spp * 4 * (max_depth + 3);
    -1 beyond 63 bounces or for an invalid configuration."""
    lib = N.hip()

    def tb(w, h, spp, depth):
        c = N.RenderCfg()
        c.width, c.height, c.spp, c.max_depth = w, h, spp, depth
        return lib.flg_volpath_diff_tape_bytes(ctypes.byref(c))

    assert tb(256, 256, 16, 32) == 256 * 256 * 16 * 4 * 35
    assert tb(3, 5, 7, 1) == 3 * 5 * 7 * 4 * 4
    assert tb(8, 8, 1, 63) == 64 * 4 * 66
    assert tb(8, 8, 1, 64) == -1
    assert tb(0, 8, 1, 8) == -1 and tb(8, 8, 1, 0) == -1
    assert tb(1 << 16, 1 << 15, 1, 8) == -1  # 2^31 paths
    assert lib.flg_volpath_diff_tape_bytes(None) == -1
