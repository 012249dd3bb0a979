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
