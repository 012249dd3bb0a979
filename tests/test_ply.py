"""PLY ingestion (flg_ply_load, MakeTriangleMesh's reader, src/plymesh.cpp:
19-86) on files written here from the PLY format description: every format
and scalar type, optional normals and texture coordinates, skipped elements
and properties, and the error cases.  No GPU."""
import os
import struct

import numpy as np
import pytest

from fledge_amd import _native as N
from fledge_amd import api


def _sphere():
    v, ind = api.gen_octasphere(3, 0.05, seed=7, base_freq=2.0)
    return v.reshape(-1, 3).astype(np.float32), ind.reshape(-1, 3).astype(np.int32)


def _write(path, fmt, elems, comments=("written by test_ply",)):
    """elems: list of (name, [(kind, type, name)], rows); kind "s" scalar,
    ("l", count_type) list; rows: lists of values (a list value is a list)."""
    codes = {"char": "b", "int8": "b", "uchar": "B", "uint8": "B", "short": "h", "int16": "h",
             "ushort": "H", "uint16": "H", "int": "i", "int32": "i", "uint": "I", "uint32": "I",
             "float": "f", "float32": "f", "double": "d", "float64": "d"}
    hdr = ["ply", "format %s 1.0" % fmt] + ["comment " + c for c in comments]
    for name, props, rows in elems:
        hdr.append("element %s %d" % (name, len(rows)))
        for kind, t, pname in props:
            hdr.append("property list %s %s %s" % (kind[1], t, pname) if kind != "s" else
                       "property %s %s" % (t, pname))
    hdr.append("end_header")
    out = ("\n".join(hdr) + "\n").encode()
    end = "<" if fmt == "binary_little_endian" else ">"
    for name, props, rows in elems:
        for row in rows:
            if fmt == "ascii":
                toks = []
                for (kind, t, _), val in zip(props, row):
                    vals = [len(val)] + list(val) if kind != "s" else [val]
                    for j, x in enumerate(vals):
                        toks.append(repr(float(x)) if (t in ("float", "float32", "double", "float64") and
                                                        (kind == "s" or j > 0)) else str(int(x)))
                out += (" ".join(toks) + "\n").encode()
            else:
                for (kind, t, _), val in zip(props, row):
                    if kind == "s":
                        out += struct.pack(end + codes[t], val)
                    else:
                        out += struct.pack(end + codes[kind[1]], len(val))
                        out += struct.pack(end + codes[t] * len(val), *val)
    with open(path, "wb") as f:
        f.write(out)


def _mesh_elems(v, ind, vtype="float", itype="int", ctype="uchar", normals=None, uv=None, uvnames=("u", "v")):
    props = [("s", vtype, "x"), ("s", vtype, "y"), ("s", vtype, "z"), ("s", "uchar", "red")]
    rows = [[float(a), float(b), float(c), k % 256] for k, (a, b, c) in enumerate(v)]
    if normals is not None:
        props += [("s", "double", "nx"), ("s", "double", "ny"), ("s", "double", "nz")]
        rows = [r + [float(x) for x in n] for r, n in zip(rows, normals)]
    if uv is not None:
        props += [("s", "float", uvnames[0]), ("s", "float", uvnames[1])]
        rows = [r + [float(x) for x in t] for r, t in zip(rows, uv)]
    faces = [[list(map(int, f)), 7] for f in ind]
    return [("vertex", props, rows),
            ("face", [(("l", ctype), itype, "vertex_indices"), ("s", "int", "material")], faces)]


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
def test_ply_formats_round_trip(tmp_path, fmt):
    v, ind = _sphere()
    p = str(tmp_path / "m.ply")
    elems = _mesh_elems(v, ind)
    elems.insert(1, ("edge", [("s", "int", "vertex1"), ("s", "int", "vertex2")], [[0, 1], [1, 2]]))
    _write(p, fmt, elems)
    gv, gi, gn, guv = api.load_ply(p)
    assert np.array_equal(gv.view(np.uint32), v.view(np.uint32))  # float text round-trips exactly
    assert np.array_equal(gi, ind)
    assert gn is None and guv is None


@pytest.mark.parametrize("vtype,itype,ctype", [("double", "uint", "uchar"), ("float32", "ushort", "int"),
                                               ("float64", "int32", "uint8")])
def test_ply_scalar_types_convert(tmp_path, vtype, itype, ctype):
    """double positions round to the nearest float; any integer index type."""
    v, ind = _sphere()
    v64 = v.astype(np.float64) * (1 + 1e-9)  # not float-representable
    p = str(tmp_path / "m.ply")
    _write(p, "binary_little_endian", _mesh_elems(v64, ind, vtype=vtype, itype=itype, ctype=ctype))
    gv, gi, _, _ = api.load_ply(p)
    want = v64.astype(np.float32) if vtype in ("double", "float64") else v64.astype(np.float32)
    assert np.array_equal(gv.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(gi, ind)


@pytest.mark.parametrize("uvnames", [("u", "v"), ("s", "t"), ("texture_u", "texture_v"), ("texture_s", "texture_t")])
def test_ply_normals_and_texcoords(tmp_path, uvnames):
    v, ind = _sphere()
    n = v / np.linalg.norm(v, axis=1, keepdims=True)
    uv = np.stack([v[:, 0] * 0.5 + 0.5, v[:, 1] * 0.5 + 0.5], axis=1).astype(np.float32)
    p = str(tmp_path / "m.ply")
    _write(p, "ascii", _mesh_elems(v, ind, normals=n, uv=uv, uvnames=uvnames))
    gv, gi, gn, guv = api.load_ply(p)
    assert np.array_equal(gn, n.astype(np.float64).astype(np.float32))
    assert np.array_equal(guv.view(np.uint32), uv.view(np.uint32))


def _expect_error(path, code=1):
    with pytest.raises(N.FlgError) as e:
        api.load_ply(path)
    assert "(%d)" % code in str(e.value)


def test_ply_errors(tmp_path):
    v, ind = _sphere()
    _expect_error(str(tmp_path / "missing.ply"))
    quad = str(tmp_path / "quad.ply")
    _write(quad, "ascii", [("vertex", [("s", "float", "x"), ("s", "float", "y"), ("s", "float", "z")],
                            [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]]),
                           ("face", [(("l", "uchar"), "int", "vertex_indices")], [[[0, 1, 2, 3]]])])
    _expect_error(quad)
    bad_index = str(tmp_path / "bad_index.ply")
    ind2 = ind.copy()
    ind2[5, 1] = len(v)
    _write(bad_index, "binary_little_endian", _mesh_elems(v, ind2))
    _expect_error(bad_index)
    truncated = str(tmp_path / "trunc.ply")
    _write(truncated, "binary_big_endian", _mesh_elems(v, ind))
    data = open(truncated, "rb").read()
    open(truncated, "wb").write(data[:-5])
    _expect_error(truncated)
    no_faces = str(tmp_path / "nofaces.ply")
    _write(no_faces, "ascii", [("vertex", [("s", "float", "x"), ("s", "float", "y"), ("s", "float", "z")],
                                [[0, 0, 0]])])
    _expect_error(no_faces)
    bad_magic = str(tmp_path / "magic.ply")
    open(bad_magic, "wb").write(b"plx\nformat ascii 1.0\nend_header\n")
    _expect_error(bad_magic)


def test_ply_vertices_without_faces(tmp_path):
    """A face element of count 0 is a mesh of no triangles; no face element
    at all is an error (MakeTriangleMesh fails without faces)."""
    v, _ = _sphere()
    p = str(tmp_path / "nofaces.ply")
    elems = _mesh_elems(v, np.zeros((0, 3), np.int32))
    _write(p, "binary_little_endian", elems)
    gv, gi, _, _ = api.load_ply(p)
    assert gv.shape == v.shape and (gi is None or gi.size == 0)


def test_ply_hostile_counts_and_values(tmp_path):
    """Counts the body cannot hold fail before buffers are sized from them; an
    index given as a float outside int64 (or NaN) is an error, not undefined
    behaviour; a subnormal ascii coordinate is a value, not a range error."""
    huge = str(tmp_path / "huge.ply")
    open(huge, "wb").write(b"ply\nformat binary_little_endian 1.0\nelement vertex 2000000000\n"
                           b"property float x\nproperty float y\nproperty float z\n"
                           b"element face 1\nproperty list uchar int vertex_indices\nend_header\n" + bytes(64))
    _expect_error(huge)
    for bad in ("1e300", "nan", "-1e19"):
        p = str(tmp_path / "fidx.ply")
        open(p, "w").write("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                           "property float z\nelement face 1\nproperty list uchar double vertex_indices\n"
                           "end_header\n0 0 0\n1 0 0\n0 1 0\n3 0 1 %s\n" % bad)
        _expect_error(p)
    sub = str(tmp_path / "subnormal.ply")
    open(sub, "w").write("ply\nformat ascii 1.0\nelement vertex 3\nproperty float x\nproperty float y\n"
                         "property float z\nelement face 1\nproperty list uchar int vertex_indices\n"
                         "end_header\n1e-40 0 0\n1 0 0\n0 1 0\n3 0 1 2\n")
    gv, gi, _, _ = api.load_ply(sub)
    assert gv[0, 0] == np.float32(1e-40) and gv[0, 0] != 0
    assert gi.reshape(-1).tolist() == [0, 1, 2]
    over = str(tmp_path / "overflow.ply")
    open(over, "w").write(open(sub).read().replace("1e-40", "1e39"))
    _expect_error(over)
    # a list of some other element whose count (near 2^32) the body cannot
    # hold: refused at the count, before any item is read
    import time
    for fmt, cnt in (("binary_little_endian", struct.pack("<I", 4000000000)), ("ascii", b"4000000000")):
        lst = str(tmp_path / "biglist.ply")
        hdr = ("ply\nformat %s 1.0\nelement vertex 3\nproperty float x\nproperty float y\nproperty float z\n"
               "element face 1\nproperty list uchar int vertex_indices\nelement extra 1\n"
               "property list uint int values\nend_header\n" % fmt).encode()
        if fmt == "ascii":
            body = b"0 0 0\n1 0 0\n0 1 0\n3 0 1 2\n" + cnt + b" 1 2 3\n"
        else:
            body = struct.pack("<9f", 0, 0, 0, 1, 0, 0, 0, 1, 0) + struct.pack("<B3i", 3, 0, 1, 2) + cnt + bytes(12)
        open(lst, "wb").write(hdr + body)
        t0 = time.perf_counter()
        _expect_error(lst)
        assert time.perf_counter() - t0 < 1.0
