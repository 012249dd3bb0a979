"""Film output (Film::saveBuffer, ".exr"): flg_film_save_exr writes an
OpenEXR single-part scanline file; read back here with a parser written from
the published file layout (magic 20000630, version 2, attribute list
name/type/size/value ended by a null byte, offset table of one uint64 per
scanline block, blocks of y, byte count and per-channel rows in channel-list
order).  No GPU needed."""
import ctypes
import os
import struct

import numpy as np

from fledge_amd import _native as N


def _read_exr(path):
    data = open(path, "rb").read()
    assert struct.unpack_from("<i", data, 0)[0] == 20000630
    version = struct.unpack_from("<i", data, 4)[0]
    assert version & 0xFF == 2 and (version >> 8) == 0  # single-part scanline, short names
    pos, attrs = 8, {}
    while data[pos] != 0:
        end = data.index(b"\0", pos)
        name = data[pos:end].decode()
        tend = data.index(b"\0", end + 1)
        typ = data[end + 1:tend].decode()
        size = struct.unpack_from("<i", data, tend + 1)[0]
        attrs[name] = (typ, data[tend + 5:tend + 5 + size])
        pos = tend + 5 + size
    pos += 1
    chans, raw = [], attrs["channels"][1]
    p = 0
    while raw[p] != 0:
        end = raw.index(b"\0", p)
        ptype, plin, xs, ys = struct.unpack_from("<iB3xii", raw, end + 1)
        chans.append((raw[p:end].decode(), ptype, plin, xs, ys))
        p = end + 1 + 16
    assert p == len(raw) - 1
    x0, y0, x1, y1 = struct.unpack("<4i", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    offsets = struct.unpack_from("<%dQ" % h, data, pos)
    img = {c[0]: np.zeros((h, w), np.float32) for c in chans}
    for k, off in enumerate(offsets):
        y, nbytes = struct.unpack_from("<ii", data, off)
        assert y == y0 + k and nbytes == w * 4 * len(chans)
        for j, c in enumerate(chans):
            img[c[0]][k] = np.frombuffer(data, np.float32, w, off + 8 + j * w * 4)
    return attrs, chans, img, (x0, y0, x1, y1), len(data), offsets


def test_exr_layout_and_pixels(tmp_path):
    lib = N.hip()
    W, H = 7, 5
    rgb = np.random.default_rng(0).normal(size=(H, W, 3)).astype(np.float32)
    path = str(tmp_path / "film.exr")
    assert lib.flg_film_save_exr(path.encode(), W, H, rgb.ctypes.data) == 0
    attrs, chans, img, win, size, offsets = _read_exr(path)
    for name, typ in [("channels", "chlist"), ("compression", "compression"), ("dataWindow", "box2i"),
                      ("displayWindow", "box2i"), ("lineOrder", "lineOrder"), ("pixelAspectRatio", "float"),
                      ("screenWindowCenter", "v2f"), ("screenWindowWidth", "float")]:
        assert attrs[name][0] == typ, name
    assert [c[0] for c in chans] == sorted(c[0] for c in chans) == ["B", "G", "R"]
    assert all(c[1] == 2 and c[3] == 1 and c[4] == 1 for c in chans)  # FLOAT, full resolution
    assert attrs["compression"][1] == b"\x00" and attrs["lineOrder"][1] == b"\x00"
    assert win == (0, 0, W - 1, H - 1) and attrs["displayWindow"][1] == attrs["dataWindow"][1]
    assert struct.unpack("<f", attrs["pixelAspectRatio"][1])[0] == 1.0
    # the file's first scanline is the film's last row (the reference's
    # negative-stride write); values bit for bit
    for c, k in (("R", 0), ("G", 1), ("B", 2)):
        assert np.array_equal(img[c].view(np.uint32), rgb[::-1, :, k].view(np.uint32))
    assert size == offsets[-1] + 8 + W * 12  # nothing after the last block
    assert list(offsets) == sorted(offsets)


def test_exr_bad_arguments(tmp_path):
    lib = N.hip()
    rgb = np.zeros(12, np.float32)
    assert lib.flg_film_save_exr(None, 2, 2, rgb.ctypes.data) != 0
    assert lib.flg_film_save_exr(str(tmp_path / "a.exr").encode(), 0, 2, rgb.ctypes.data) != 0
    assert lib.flg_film_save_exr(str(tmp_path / "no_dir" / "a.exr").encode(), 2, 2, rgb.ctypes.data) != 0
    assert not os.path.exists(tmp_path / "no_dir")
