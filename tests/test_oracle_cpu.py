"""CPU tests: the oracle pinned against the reference's own known answers and
fixtures, closed forms and brute force (no GPU needed)."""
import ctypes
import math

import numpy as np
import pytest

import oracle_lib as orc
from scene_cases import reference_rank, scene_meshes, scene_rays
from fledge_amd import api

L = orc.lib()
P = ctypes.c_void_p
f32 = ctypes.c_float
i32 = ctypes.c_int32
L.orc_karras.argtypes = [P, i32, P, P, P, P, P]
L.orc_dist1d_sample.restype = f32
L.orc_dist1d_sample.argtypes = [P, i32, f32, P, P]
L.orc_hg_sample.restype = f32
L.orc_hg_sample.argtypes = [P, f32, f32, f32, P]
L.orc_hg_p.restype = f32
L.orc_hg_p.argtypes = [P, P, f32]
L.orc_pluecker.restype = i32
L.orc_pluecker.argtypes = [P, P, P, P, P, i32, P, P]
L.orc_bvh_test_rays.argtypes = [ctypes.c_int64, P]
L.orc_watertight_rays.argtypes = [ctypes.c_int64, P]
L.orc_offset_ray_origin.argtypes = [P, P, P, P]


def f32a(*x):
    return np.array(x, np.float32)


def _ulp_error(got, exact):
    """|got - exact| in units of the float spacing at exact (float64 math)."""
    e32 = exact.astype(np.float32)
    spacing = np.spacing(np.abs(e32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact) / spacing


# ------------------------------------------------------- transcendentals ---
def test_logf_within_one_ulp_of_log():
    """The path's logf (fdlibm e_logf method) against float64 log: < 1 ulp over
    the whole positive range, exact special values."""
    bits = np.concatenate([np.arange(1, 0x7F800000, 4099, dtype=np.uint32),
                           np.arange(0x3F700000, 0x3F880000, 13, dtype=np.uint32)])
    x = bits.view(np.float32)
    got = orc.logf(x)
    err = _ulp_error(got, np.log(x.astype(np.float64)))
    assert err.max() < 1.0, (float(x[err.argmax()]), float(err.max()))
    sp = orc.logf(np.array([1.0, 0.0, -0.0, np.inf, -1.0, np.nan], np.float32))
    assert sp[0] == 0.0 and sp[1] == -np.inf and sp[2] == -np.inf and sp[3] == np.inf
    assert np.isnan(sp[4]) and np.isnan(sp[5])


def test_sincosf_within_one_ulp():
    """sin/cos (msun double kernels, pi/2 reduction) against float64 sin/cos:
    < 1 ulp on the arguments the path uses (|x| < 8) and beyond."""
    x = np.concatenate([np.linspace(-8.0, 8.0, 400_001, dtype=np.float32),
                        np.linspace(-1e4, 1e4, 100_001, dtype=np.float32),
                        np.array([0.0, -0.0, 1e-30, np.float32(np.pi / 2)], np.float32)])
    s, c = orc.sincosf(x)
    xd = x.astype(np.float64)
    es, ec = _ulp_error(s, np.sin(xd)), _ulp_error(c, np.cos(xd))
    # near a zero of sin/cos the float spacing of the result collapses; bound
    # the absolute error there instead
    tiny_s, tiny_c = np.abs(np.sin(xd)) < 1e-6, np.abs(np.cos(xd)) < 1e-6
    assert es[~tiny_s].max() < 1.0 and ec[~tiny_c].max() < 1.0
    assert np.abs(s[tiny_s] - np.sin(xd[tiny_s])).max(initial=0) < 6e-8
    assert np.abs(c[tiny_c] - np.cos(xd[tiny_c])).max(initial=0) < 6e-8
    s0, c0 = orc.sincosf(np.array([0.0, -0.0], np.float32))
    assert s0[0] == 0.0 and s0[1] == 0.0 and c0[0] == 1.0 and c0[1] == 1.0


# ---------------------------------------------------------------- PCG32 ---
def test_pcg32_demo_kat():
    # pcg32-demo output for pcg32_srandom(42, 54)
    want = [0xa15c02b7, 0x7b47f409, 0xba1d3330, 0x83d2f293, 0xbfa4784b, 0xcbed606e]
    out = np.zeros(6, np.uint32)
    L.orc_pcg32(42, 54, 6, out.ctypes.data)
    assert out.tolist() == want
    from fledge_amd import _native as N
    out2 = np.zeros(6, np.uint32)
    N.scene().flg_scene_pcg32(42, 54, 6, out2.ctypes.data)
    assert out2.tolist() == want


def test_soup_generator_matches_oracle():
    v, i = api.gen_soup(5000)
    v2 = np.zeros_like(v)
    i2 = np.zeros_like(i)
    L.orc_gen_soup(42, 54, 5000, v2.ctypes.data, i2.ctypes.data)
    assert np.array_equal(v.view(np.uint32), v2.view(np.uint32)) and np.array_equal(i, i2)
    c = v.reshape(-1, 3, 3).mean(axis=1)
    assert c.min() > -0.02 and c.max() < 1.02


# --------------------------------------------------------------- Morton ---
def _interleave(x, y, z):
    r = 0
    for b in range(10):
        r |= ((x >> b) & 1) << (3 * b) | ((y >> b) & 1) << (3 * b + 1) | ((z >> b) & 1) << (3 * b + 2)
    return r


def test_morton_kats():
    assert L.orc_left_shift3(1023) == 0x09249249
    assert L.orc_left_shift3(1024) == 0x09249249  # 1024 clamps to 1023
    assert L.orc_encode_morton3(1023.0, 0.0, 0.0) == 0x09249249
    assert L.orc_encode_morton3(0.0, 1023.0, 0.0) == 0x12492492
    assert L.orc_encode_morton3(0.0, 0.0, 1023.0) == 0x24924924
    assert L.orc_encode_morton3(1024.0, 1024.0, 1024.0) == 0x3FFFFFFF
    assert L.orc_encode_morton3(1023.9, 0.99, 2.5) == _interleave(1023, 0, 2)  # truncation
    rng = np.random.default_rng(0)
    for x, y, z in rng.integers(0, 1024, size=(2000, 3)):
        assert L.orc_encode_morton3(float(x), float(y), float(z)) == _interleave(int(x), int(y), int(z))


# ---------------------------------------------------------------- Karras ---
def _karras(codes):
    codes = np.ascontiguousarray(codes, np.uint32)
    n = codes.size
    out = [np.zeros(n - 1, np.int32) for _ in range(3)] + [np.zeros(n - 1, np.uint8) for _ in range(2)]
    L.orc_karras(codes.ctypes.data, n, *[a.ctypes.data for a in out])
    return out


def test_karras_paper_figure():
    # Karras, HPG 2012, Fig. 3: keys 00001 00010 00100 00101 10011 11000 11001 11110
    codes = [0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110]
    split, lo, hi, ll, rl = _karras(codes)
    assert split.tolist() == [3, 0, 2, 1, 4, 6, 5]
    assert list(zip(lo.tolist(), hi.tolist())) == [(0, 7), (0, 1), (2, 3), (0, 3), (4, 7), (5, 7), (5, 6)]
    assert ll.tolist() == [0, 1, 1, 0, 1, 0, 1]
    assert rl.tolist() == [0, 1, 1, 0, 0, 1, 1]


def _brute_tree(keys):
    """Binary radix tree by recursive splitting at the highest differing bit of
    the 62-bit key (code << 32 | index); returns {(lo, hi): split}."""
    n = len(keys)
    aug = [(int(k) << 32) | i for i, k in enumerate(keys)]
    out = {}

    def rec(lo, hi):
        if lo == hi:
            return
        x = aug[lo] ^ aug[hi]
        bit = x.bit_length() - 1
        s = lo
        while s < hi and not ((aug[s + 1] >> bit) & 1):
            s += 1
        out[(lo, hi)] = s
        rec(lo, s)
        rec(s + 1, hi)

    rec(0, n - 1)
    return out


@pytest.mark.parametrize("seed,n,bits", [(1, 50, 30), (2, 300, 30), (3, 200, 4), (4, 64, 1), (5, 2, 30)])
def test_karras_vs_brute_radix_tree(seed, n, bits):
    rng = np.random.default_rng(seed)
    codes = np.sort(rng.integers(0, 1 << bits, n).astype(np.uint32))
    split, lo, hi, ll, rl = _karras(codes)
    want = _brute_tree(codes)
    got = {(int(a), int(b)): int(s) for a, b, s in zip(lo, hi, split)}
    assert got == want
    assert (ll == (lo == split)).all() and (rl == (hi == split + 1)).all()


@pytest.mark.parametrize("n", [2, 3, 100, 5000])
def test_lbvh_invariants(n):
    v, i = api.gen_soup(n, initseq=9)
    b = orc.Lbvh(v, i).export()
    # sorted codes non-decreasing, order a permutation, ties in index order
    ms, order = b["morton_sorted"], b["order"]
    assert np.all(np.diff(ms.astype(np.int64)) >= 0)
    assert np.array_equal(np.sort(order), np.arange(n))
    eq = ms[1:] == ms[:-1]
    assert np.all(order[1:][eq] > order[:-1][eq])
    assert np.array_equal(b["morton"][order], ms)
    # topology: every internal node except the root has exactly one parent,
    # every leaf exactly one, DFS covers every leaf once
    ni = n - 1
    nchild = np.zeros(ni, np.int64)
    lchild = np.zeros(n, np.int64)
    for k in range(ni):
        g = b["split"][k]
        for side, leaf in ((0, b["left_leaf"][k]), (1, b["right_leaf"][k])):
            if leaf:
                lchild[g + side] += 1
                assert b["leaf_parent"][g + side] == k
            else:
                nchild[g + side] += 1
                assert b["parent"][g + side] == k
    assert nchild[0] == 0 and (nchild[1:] == 1).all() and (lchild == 1).all()
    # bounds: node = union of children; root = union of all vertices
    tri = v.reshape(-1, 3, 3)
    tlo, thi = tri.min(axis=1), tri.max(axis=1)
    nb = b["node_bound"]
    for k in range(ni):
        g = b["split"][k]
        parts = []
        for side, leaf in ((0, b["left_leaf"][k]), (1, b["right_leaf"][k])):
            if leaf:
                t = order[g + side]
                parts.append(np.concatenate([tlo[t], thi[t]]))
            else:
                parts.append(nb[g + side])
        want = np.concatenate([np.minimum(parts[0][:3], parts[1][:3]), np.maximum(parts[0][3:], parts[1][3:])])
        assert np.array_equal(nb[k], want)
    allv = v.reshape(-1, 3)
    assert np.array_equal(nb[0], np.concatenate([allv.min(axis=0), allv.max(axis=0)]))
    assert np.array_equal(b["bound"], nb[0])


# ------------------------------------------------------------- Pluecker ---
def _pl(a, b, c, o, d, compat=0):
    t = np.zeros(1, np.float32)
    ng = np.zeros(3, np.float32)
    h = L.orc_pluecker(*[np.asarray(x, np.float32).ctypes.data for x in (a, b, c, o, d)], compat,
                       t.ctypes.data, ng.ctypes.data)
    return bool(h), float(t[0]), ng


def test_pluecker_kats():
    a, b, c = f32a(-1, -1, 2), f32a(1, -1, 2), f32a(0, 1, 2)
    h, t, ng = _pl(a, b, c, f32a(0, 0, 0), f32a(0, 0, 1))
    assert h and t == 2.0 and np.array_equal(ng, f32a(0, 0, 1))
    h, t, ng = _pl(a, c, b, f32a(0, 0, 0), f32a(0, 0, 1))  # two-sided
    assert h and t == 2.0 and np.array_equal(ng, f32a(0, 0, -1))
    h, _, _ = _pl(a, b, c, f32a(0, 0, 0), f32a(0, 0, -1))  # behind the origin
    assert not h
    h, _, _ = _pl(a, b, c, f32a(5, 0, 0), f32a(0, 0, 1))  # outside
    assert not h
    h, t, _ = _pl(a, b, c, f32a(0, 0, 2), f32a(0, 0, 1))  # t == 0 accepted
    assert h and t == 0.0
    # compat: StableNormalize's x^2 + y^2 + z + z is negative for a -z normal
    h, _, _ = _pl(a, c, b, f32a(0, 0, 0), f32a(0, 0, 1), compat=1)
    assert not h
    h, t, _ = _pl(a, b, c, f32a(0, 0, 0), f32a(0, 0, 1), compat=1)
    assert h and t == 2.0


def test_pluecker_shared_edge_watertight():
    # two triangles sharing the edge (0,-1)-(0,1) at z = 1; rays through the edge hit one of them
    a, b, c = f32a(0, -1, 1), f32a(0, 1, 1), f32a(-1, 0, 1)
    e = f32a(1, 0, 1)
    for y in np.linspace(-0.9, 0.9, 181, dtype=np.float32):
        o = f32a(0, y, 0)
        d = f32a(0, 0, 1)
        h1, _, _ = _pl(a, b, c, o, d)
        h2, _, _ = _pl(b, a, e, o, d)
        assert h1 or h2


def test_trace_matches_brute_force():
    v, i = api.gen_soup(20000, initseq=31)
    o = orc.Lbvh(v, i)
    rng = np.random.default_rng(7)
    n = 2000
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = rng.random((n, 3), dtype=np.float32) * 1.2 - 0.1
    d = rng.normal(size=(n, 3)).astype(np.float32)
    rays[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 3] = np.inf
    ids, ts = o.trace(rays)
    bid, bts, ties = orc.brute_trace(v, i, rays)
    ok = ties == 0
    assert (ids >= 0).sum() > 100
    assert np.array_equal(ids[ok], bid[ok]) and np.array_equal(ts[ok], bts[ok])


def test_bvh_test_fixed_ray_fixture():
    """bvh_test.hpp:143-168: the fixed ray must hit MakeMeshedSphere(1000, 1000, 5)."""
    v, i = api.gen_uv_sphere(1000, 1000, 5.0)
    assert i.size // 3 == 999 * 2 * 998
    o = orc.Lbvh(v, i)
    ray = np.array([[0.345564, 0.422557, 0.259793, np.finfo(np.float32).max,
                     0.524748325347900390625, 0.714325249195098876953125, 0.46301043033599853515625, 0]],
                   np.float32)
    ids, ts = o.trace(ray)
    assert ids[0] >= 0
    p = ray[0, :3] + ray[0, 4:7] * ts[0]
    assert abs(np.linalg.norm(p) - 5.0) < 1e-2
    # the 500k-ray differential harness (bvh_test.hpp:107-141) rays: all start
    # inside the sphere, so all must hit with |p| close to 5
    n = 20000
    rays = np.zeros((n, 8), np.float32)
    L.orc_bvh_test_rays(n, rays.ctypes.data)
    ids, ts = o.trace(rays)
    assert (ids >= 0).all()
    pts = rays[:, :3] + rays[:, 4:7] * ts[:, None]
    assert np.abs(np.linalg.norm(pts, axis=1) - 5.0).max() < 2e-2


def test_watertight_fixture():
    """shape_tests.cpp:22-48: rays from inside MakeMeshedSphere(16, 16, 1) always hit."""
    v, i = api.gen_uv_sphere(16, 16, 1.0)
    o = orc.Lbvh(v, i)
    n = 1000
    rays = np.zeros((n, 8), np.float32)
    L.orc_watertight_rays(n, rays.ctypes.data)
    ids, _ = o.trace(rays)
    assert (ids >= 0).all()


def test_octasphere_closed_and_outward():
    for level in (0, 1, 4):
        v, i = api.gen_octasphere(level, 0.0)
        t = i.reshape(-1, 3)
        assert t.shape[0] == 8 * 4 ** level and v.size // 3 == 4 * 4 ** level + 2
        edges = {}
        for a, b, c in t:
            for x, y in ((a, b), (b, c), (c, a)):
                edges[(x, y)] = edges.get((x, y), 0) + 1
        assert all(n == 1 for n in edges.values())
        assert all((y, x) in edges for (x, y) in edges)  # every directed edge has its twin
        p = v.reshape(-1, 3)
        nrm = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
        assert (np.einsum("ij,ij->i", nrm, p[t].mean(axis=1)) > 0).all()


# ---------------------------------------------------------------- Halton ---
def test_primes_kat():
    assert [L.orc_prime(k) for k in range(10)] == [2, 3, 5, 7, 11, 13, 17, 19, 23, 29]
    assert L.orc_num_primes() == 23000  # pi(262144)


def test_radical_inverse_closed_forms():
    ri = L.orc_radical_inverse
    ri.restype = f32
    assert [ri(2, a) for a in (1, 2, 3)] == [0.5, 0.25, 0.75]
    assert [ri(3, a) for a in (1, 2, 3)] == pytest.approx([1 / 3, 2 / 3, 1 / 9], rel=1e-7)
    for base in (2, 3, 5, 7, 101, 7919):
        for a in (1, 17, 1234, 99991, 1990655):
            digits, x, scale, nd = 0.0, a, 1.0, 0
            while x:
                scale /= base
                digits += (x % base) * scale
                x //= base
                nd += 1
            # float product chain of 1/base per digit (sampler.h:66-71): ~1 ulp per digit
            assert abs(ri(base, a) - digits) <= (nd + 2) * 2.0 ** -24 * digits


def test_halton_crt_offsets():
    for px, py in [(0, 0), (1, 1), (5, 200), (127, 127), (128, 3), (1919, 1079)]:
        off = L.orc_halton_offset(1920, 1080, px, py)
        assert off < 31104
        rev2 = int(format(px % 128, "07b")[::-1], 2)
        m, rev3 = py % 128, 0
        for _ in range(5):
            rev3 = rev3 * 3 + m % 3
            m //= 3
        assert off % 128 == rev2 and off % 243 == rev3


def test_halton_histogram_reference():
    """sampler_tests.cpp:24-45: HaltonSampler(128, {5,5}), pixel (1,1), 128 dims x 128 spp."""
    out = np.zeros(128 * 128, np.float32)
    L.orc_halton_sequence(5, 5, 128, 1, 1, 128, out.ctypes.data)
    hist = np.bincount(np.minimum((out * 10).astype(int), 9), minlength=10) / out.size
    assert np.abs(hist - 0.1).max() < 1e-2


# -------------------------------------------------- sampling & lights -----
def test_dist1d_inverse_cdf_quirk():
    f = np.ones(4, np.float32)
    pdf = np.zeros(1, np.float32)
    off = np.zeros(1, np.int32)
    # cdf = [0, .25, .5, .75, 1]; r = mid - 1 sends u = 0.3 to bin 0 (correct: 1)
    r = L.orc_dist1d_sample(f.ctypes.data, 4, 0.3, pdf.ctypes.data, off.ctypes.data)
    assert off[0] == 0 and pdf[0] == 1.0 and abs(r - 0.3) < 1e-7
    L.orc_dist1d_sample(f.ctypes.data, 4, 0.6, pdf.ctypes.data, off.ctypes.data)
    assert off[0] == 2
    L.orc_dist1d_sample(f.ctypes.data, 4, 0.8, pdf.ctypes.data, off.ctypes.data)
    assert off[0] == 3


def test_hg_phase_normalised_and_sampled():
    g = 0.3
    wo = f32a(0.0, 0.0, 1.0)
    # pdf integrates to 1 over the sphere
    cos = np.linspace(-1, 1, 20001)
    vals = []
    for c in cos[::100]:
        wi = f32a(math.sqrt(max(0.0, 1 - c * c)), 0.0, c)
        vals.append(L.orc_hg_p(wi.ctypes.data, wo.ctypes.data, g))
    integ = np.trapezoid(np.array(vals), cos[::100]) * 2 * math.pi
    assert abs(integ - 1.0) < 1e-3
    # sampled cos(theta) follows the HG distribution (CDF check)
    rng = np.random.default_rng(2)
    us = rng.random((20000, 2)).astype(np.float32)
    wi = np.zeros(3, np.float32)
    cs = []
    for u, v in us:
        p = L.orc_hg_sample(wo.ctypes.data, float(u), float(v), g, wi.ctypes.data)
        assert p > 0
        cs.append(float(np.dot(wi, wo)))
    cs = np.sort(np.array(cs))
    # HG CDF in mu = dot(wi, wo) for the convention 1 + g^2 + 2 g mu
    mu = np.linspace(-1, 1, 41)
    pdf = lambda m: (1 - g * g) / (1 + g * g + 2 * g * m) ** 1.5 / 2
    grid = np.linspace(-1, 1, 40001)
    cdf = np.concatenate([[0], np.cumsum((pdf(grid[1:]) + pdf(grid[:-1])) / 2 * np.diff(grid))])
    emp = np.searchsorted(cs, mu) / cs.size
    assert np.abs(emp - np.interp(mu, grid, cdf)).max() < 0.015


def test_offset_ray_origin_kat():
    out = np.zeros(3, np.float32)
    p, n, d = f32a(1.0, 2.0, 3.0), f32a(0.0, 0.0, 1.0), f32a(0.0, 0.0, 1.0)
    L.orc_offset_ray_origin(p.ctypes.data, n.ctypes.data, d.ctypes.data, out.ctypes.data)
    want = (np.array([3.0], np.float32).view(np.int32) + 256).view(np.float32)[0]
    assert out[0] == 1.0 and out[1] == 2.0 and out[2] == want
    p = f32a(0.001, -0.5, 0.0)
    L.orc_offset_ray_origin(p.ctypes.data, n.ctypes.data, f32a(0.0, 0.0, -1.0).ctypes.data, out.ctypes.data)
    assert out[0] == np.float32(0.001) and out[2] == np.float32(0.0) - np.float32(1 / 65536)


# ---------------------------------------------------------------- volume ---
CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]


def _chords(W, H, spp=1, s=0):
    rays = orc.camera_rays(CAM0, W, H, spp, s)
    o, d = rays[:, :3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    with np.errstate(divide="ignore"):
        t0 = (0.0 - o) / d
        t1 = (1.0 - o) / d
    tn = np.minimum(t0, t1).max(axis=1)
    tf = np.maximum(t0, t1).min(axis=1)
    return np.maximum(tf - np.maximum(tn, 0), 0)


def _trilinear(grid, p):
    """Density at points p [m, 3] in the unit box: index space p * n - 0.5, zero outside."""
    n = grid.shape[0]
    g = p * n - 0.5
    f = np.floor(g)
    w = g - f
    i = f.astype(int)
    out = np.zeros(len(p))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                ii, jj, kk = i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz
                ok = (ii >= 0) & (ii < n) & (jj >= 0) & (jj < n) & (kk >= 0) & (kk < n)
                v = np.zeros(len(p))
                v[ok] = grid[kk[ok], jj[ok], ii[ok]]
                wt = (w[:, 0] if dx else 1 - w[:, 0]) * (w[:, 1] if dy else 1 - w[:, 1]) * \
                     (w[:, 2] if dz else 1 - w[:, 2])
                out += wt * v
    return out


def test_volume_homogeneous_transmittance_closed_form():
    """Albedo 0, environment only: pixel = Le * E[ratio-tracking T] = Le exp(-sigma_t tau), with
    tau the optical depth of the trilinear density along each camera ray (quadrature)."""
    W = H = 12
    spp = 256
    grid = np.ones((8, 8, 8), np.float32)
    rgb, _, _ = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [0.0, 2.0, 0.0], [0, 0, 0, 0, 0, 1, 1.0],
                                   CAM0, W, H, spp, 4)
    want = np.zeros(W * H)
    u = (np.arange(400) + 0.5) / 400
    for s in range(spp):
        rays = orc.camera_rays(CAM0, W, H, spp, s).astype(np.float64)
        o, d = rays[:, :3], rays[:, 4:7]
        with np.errstate(divide="ignore"):
            t0, t1 = (0.0 - o) / d, (1.0 - o) / d
        tn = np.maximum(np.minimum(t0, t1).max(axis=1), 0)
        tf = np.maximum(t0, t1).min(axis=1)
        ln = np.maximum(tf - tn, 0)
        pts = o[:, None, :] + d[:, None, :] * (tn[:, None] + ln[:, None] * u[None, :])[:, :, None]
        rho = _trilinear(grid.astype(np.float64), pts.reshape(-1, 3)).reshape(W * H, -1)
        want += np.exp(-2.0 * rho.mean(axis=1) * ln)
    want = (want / spp).reshape(H, W)
    got = rgb[:, :, 0]
    # average over pixels of equal expectation: the Monte Carlo error shrinks
    assert abs(got.mean() - want.mean()) < 5e-3
    assert np.abs(got - want).max() < 0.08


def test_volume_white_furnace_oracle():
    """Albedo 1, constant environment 0.5, no point light: every pixel ~ 0.5."""
    grid = np.full((8, 8, 8), 0.5, np.float32)
    rgb, _, _ = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [4.0, 0.0, 0.3], [0, 0, 0, 0, 0, 1, 0.5],
                                   CAM0, 24, 24, 128, 512)
    assert abs(rgb.mean() - 0.5) < 0.01


def test_volume_point_light_single_scatter_scale():
    """Doubling the point-light intensity (no environment) doubles every pixel exactly
    (the light enters the estimator linearly; same Halton numbers)."""
    grid = api.gen_fbm_grid(16, seed=4)
    a, _, _ = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], [1, 2, 2, -2, 50.0, 0, 0],
                                 CAM0, 16, 16, 4, 8)
    b, _, _ = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], [1, 2, 2, -2, 100.0, 0, 0],
                                 CAM0, 16, 16, 4, 8)
    assert a.max() > 0
    assert np.allclose(b, 2 * a, rtol=1e-6, atol=0)


def test_walk_stream_uniform_and_independent():
    """The ratio-tracking walks' counter-based stream: every draw an exact
    multiple of 2^-24 in (0, 1]; uniform (mean, chi-square over 256 bins),
    free flights -log(1 - u) with the exponential's mean and variance 1, and no
    correlation between consecutive draws of a walk, between walks of a path,
    or between neighbouring paths."""
    npath, nwalk, nstep = 4096, 4, 64
    p, w, n = np.meshgrid(np.arange(npath), np.arange(nwalk), np.arange(nstep), indexing="ij")
    x = np.stack([p.ravel(), w.ravel(), n.ravel()], axis=1).astype(np.uint32)
    c = orc.walk_draw(x).astype(np.float64)
    N = c.size
    assert c.min() > 0.0 and c.max() <= 1.0
    k = c * 2.0 ** 24
    assert np.array_equal(k, np.round(k)) and k.min() >= 1 and k.max() <= 2 ** 24
    u = 1.0 - c
    assert abs(u.mean() - 0.5) < 5 * math.sqrt(1 / 12 / N)
    hist = np.bincount(np.minimum((u * 256).astype(int), 255), minlength=256)
    e = N / 256
    chi2 = ((hist - e) ** 2 / e).sum()
    assert chi2 < 255 + 6 * math.sqrt(2 * 255)
    ff = -np.log(c)
    assert abs(ff.mean() - 1.0) < 5 / math.sqrt(N) and abs(ff.var() - 1.0) < 0.02
    cube = u.reshape(npath, nwalk, nstep) - 0.5
    tol = 5 / math.sqrt(N)
    for a, b in [(cube[:, :, 1:], cube[:, :, :-1]),      # consecutive steps
                 (cube[:, 1:, :], cube[:, :-1, :]),      # walks of one path
                 (cube[1:, :, :], cube[:-1, :, :])]:     # neighbouring paths
        r = (a * b).mean() / (1 / 12)
        assert abs(r) < tol, r
    # distinct (path, walk) streams do not repeat one another
    assert len(np.unique(c.reshape(npath * nwalk, nstep)[:, :8], axis=0)) == npath * nwalk


# ------------------------------------------------------------- film (f2) --
CAMF = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]
LIGHTSF = [1.0, 2.0, 2.0, -2.0, 50.0, 1.0, 0.5]


def test_halton_stateful_matches_reference_sampler_test():
    """The reference's own sampler test (sampler_tests.cpp:24-45) driven
    through setPixel / get1D / reset: its histogram bound holds, and reset()
    lands on the next sample's index at dimension 0 (the per-sample sequence)."""
    out = orc.halton_stateful(128, 5, 5, 1, 1, 128)
    hist = np.bincount(np.minimum((out.ravel() * 10).astype(int), 9), minlength=10) / out.size
    assert np.abs(hist - 0.1).max() < 1e-2
    ref = np.zeros(128 * 128, np.float32)
    L.orc_halton_sequence(5, 5, 128, 1, 1, 128, ref.ctypes.data)
    assert np.array_equal(out.ravel().view(np.uint32), ref.view(np.uint32))


def test_film_filter_is_mitchell_with_b_c_zero():
    """The film's filter (radius 2, B = C = 1/3 in integer arithmetic = 0):
    Mitchell-Netravali's cubic at B = C = 0, (12|x|^3 - 18|x|^2 + 6) / 6 on
    |x| <= 1 and the outer piece, which vanishes at B = C = 0; so the weights
    of the pixels around a sample form a partition of unity in each axis."""
    def mn(x, B, C):
        x = abs(x)
        if x > 1:
            return ((-B - 6 * C) * x ** 3 + (6 * B + 30 * C) * x ** 2 + (-12 * B - 48 * C) * x + (8 * B + 24 * C)) / 6
        return ((12 - 9 * B - 6 * C) * x ** 3 + (-18 + 12 * B + 6 * C) * x ** 2 + (6 - 2 * B)) / 6

    xs = np.linspace(-2.5, 2.5, 2001, dtype=np.float32)
    got = np.array([L.orc_mitchell1d(float(x)) for x in xs], np.float64)
    exp = np.array([mn(float(x), 0.0, 0.0) for x in xs])
    assert np.abs(got - exp).max() < 1e-6
    assert L.orc_mitchell1d(0.0) == 1.0 and L.orc_mitchell1d(1.0) == 0.0 and L.orc_mitchell1d(-1.5) == 0.0
    assert abs(L.orc_mitchell1d(0.5) - 0.5) < 1e-7
    # partition of unity over the integer lattice
    for f in np.linspace(0.0, 1.0, 101, dtype=np.float32):
        s = sum(L.orc_mitchell1d(float(f) - k) for k in (-1, 0, 1, 2))
        assert abs(s - 1.0) < 1e-6
    # with the intended B = C = 1/3 the weights would differ (the quirk matters)
    assert abs(mn(0.5, 1 / 3, 1 / 3) - 0.5) > 1e-2


def _film_case(grid, W, H, spp, depth=16, lights=None):
    return orc.volpath_render_film(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], LIGHTSF if lights is None else lights,
                                   CAMF, W, H, spp, depth)


def test_film_schedule_maps_to_per_sample_schedule():
    """Render-loop schedule (Q10): sample s >= 2 of a pixel is the per-sample
    schedule's sample s - 1 (same Halton index, dimensions from 0), bit for
    bit; sample 1 shares index 0 with the pixel-centre path and continues its
    dimensions, so it is not sample 0 of the per-sample schedule."""
    grid = (api.gen_fbm_grid(16, seed=3)).astype(np.float32)
    W, H, spp = 18, 11, 5
    rgb, tl, tu, lk = _film_case(grid, W, H, spp)
    cl, cu = orc.volpath_samples(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], LIGHTSF, CAMF, W, H, spp, 16)
    assert np.array_equal(tl[:, 1:].view(np.uint32), cl[:, 1:spp - 1].view(np.uint32))
    assert np.array_equal(tu[:, 1:].view(np.uint32), cu[:, 1:spp - 1].view(np.uint32))
    assert not np.array_equal(tu[:, 0], cu[:, 0])
    # film positions lie in their pixel
    px = np.arange(W * H) % W
    py = np.arange(W * H) // W
    assert np.all((tu[..., 0] >= px[:, None]) & (tu[..., 0] < px[:, None] + 1))
    assert np.all((tu[..., 1] >= py[:, None]) & (tu[..., 1] < py[:, None] + 1))


def test_film_resolve_matches_float64_splat():
    """The film equals sum(w L) / sum(w) over every sample within the filter's
    support, recomputed in float64 from the per-sample values with the closed
    form 2|d|^3 - 3|d|^2 + 1, d = sample - 0.5 - pixel (bounds, offset and
    normalisation of addSample / commitSamples)."""
    grid = (api.gen_fbm_grid(16, seed=4)).astype(np.float32)
    W, H, spp = 21, 19, 4
    # chromatic lights: every channel splats its own radiance
    rgb, tl, tu, lk = _film_case(grid, W, H, spp, lights=[1, 2, 2, -2, 50.0, 20.0, 5.0, 1, 0.3, 0.5, 0.7])
    f = lambda d: np.where(np.abs(d) < 1, 2 * np.abs(d) ** 3 - 3 * d * d + 1, 0.0)
    uv = tu.reshape(-1, 2).astype(np.float64)
    for c in range(3):
        Ls = tl[..., c].reshape(-1).astype(np.float64)
        out = np.zeros((H, W))
        for y in range(H):
            for x in range(W):
                w = f(uv[:, 0] - 0.5 - x) * f(uv[:, 1] - 0.5 - y)
                out[y, x] = (w * Ls).sum() / w.sum() if w.sum() > 0 else 0.0
        got = rgb[:, :, c].astype(np.float64)
        assert np.allclose(got, out, rtol=2e-6, atol=1e-7)
    assert not np.array_equal(rgb[:, :, 0], rgb[:, :, 1])


def test_film_constant_radiance_and_single_sample():
    """An empty medium under the environment light: every path returns the
    environment radiance, so every film pixel is that radiance; with one sample
    per pixel nothing is splatted and the film stays 0."""
    grid = np.zeros((8, 8, 8), np.float32)
    rgb, tl, tu, lk = _film_case(grid, 16, 12, 6)
    assert np.all(tl == np.float32(0.5)) and lk == 0
    assert np.abs(rgb - 0.5).max() <= 2e-7
    rgb1, tl1, _, _ = _film_case(grid, 16, 12, 1)
    assert tl1.size == 0 and np.all(rgb1 == 0)


# ------------------------------------------------- surfaces + media (f1) --
def _cube_mesh(lo=0.0, hi=1.0):
    """The box [lo, hi]^3 as 12 triangles wound counter-clockwise seen from
    outside (outward Pluecker normals)."""
    c = np.array([[x, y, z] for z in (lo, hi) for y in (lo, hi) for x in (lo, hi)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    tris = []
    for a, b, cc, d in quads:
        tris += [(a, b, cc), (a, cc, d)]
    return c.ravel(), np.array(tris, np.int32).ravel()


def test_cube_mesh_is_closed_and_outward():
    v, ind = _cube_mesh()
    P = v.reshape(-1, 3)[ind.reshape(-1, 3)].astype(np.float64)
    n = np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0])
    cen = P.mean(axis=1) - 0.5
    assert np.all((n * cen).sum(axis=1) > 0)  # every normal points away from the centre
    edges = {}
    for t in ind.reshape(-1, 3):
        for k in range(3):
            e = (int(t[k]), int(t[(k + 1) % 3]))
            edges[e] = edges.get(e, 0) + 1
    assert all(edges.get((b, a), 0) == 1 for (a, b) in edges)  # each edge once each way


def test_volpath_mesh_empty_medium_sees_the_environment():
    """No medium: every camera ray crosses the interfaces and escapes at
    bounce 0, so every pixel is the environment radiance (or 0 without one)."""
    v, ind = api.gen_octasphere(3, 0.05, seed=7, base_freq=2.0)
    b = orc.Lbvh((v * 0.5 + 0.5).astype(np.float32), ind)
    e = np.zeros((16, 16, 16), np.float32)
    rgb, lk = b.volpath_mesh(e, [6.4, 1.6, 0.3], [1, 2, 2, -2, 50.0, 1, 0.5], CAMF, 24, 20, 3, 16)
    assert lk == 0 and np.all(rgb == np.float32(0.5))
    rgb, _ = b.volpath_mesh(e, [6.4, 1.6, 0.3], [1, 2, 2, -2, 50.0, 0, 0.5], CAMF, 24, 20, 3, 16)
    assert np.all(rgb == 0)


def test_volpath_cube_interface_matches_svol_in_expectation():
    """A cube interface exactly around the grid: VolPathIntegrator (delta
    tracking across interfaces, NEE before HG, no roulette) and
    SVolIntegrator (primary transmittance, HG before NEE, roulette after 10
    bounces) estimate the same radiance; their image means agree within the
    Monte Carlo error."""
    v, ind = _cube_mesh()
    b = orc.Lbvh(v, ind)
    g = (api.gen_fbm_grid(16, seed=9)).astype(np.float32)
    W = H = 16
    spp = 256
    med = [6.4, 1.6, 0.3]
    rgb_vp, _ = b.volpath_mesh(g, med, LIGHTSF, CAMF, W, H, spp, 64)
    rgb_sv, _, _ = orc.volpath_render(g, [0, 0, 0, 1, 1, 1], med, LIGHTSF, CAMF, W, H, spp, 64)
    a, c = float(rgb_vp.mean()), float(rgb_sv.mean())
    assert abs(a - c) / c < 0.02, (a, c)
    # and not by accident: a scene where the two differ (twice the density in SVol)
    rgb_2, _, _ = orc.volpath_render(2 * g, [0, 0, 0, 1, 1, 1], med, LIGHTSF, CAMF, W, H, spp, 64)
    assert abs(float(rgb_2.mean()) - a) / a > 0.05


def _mesh_diff_scene(n=12, seed=5, level=3):
    v, ind = api.gen_octasphere(level, 0.05, seed=7, base_freq=2.0)
    b = orc.Lbvh((v * 0.5 + 0.5).astype(np.float32), ind)
    g = (0.1 + 0.8 * api.gen_fbm_grid(n, seed=seed)).astype(np.float32)
    return b, g


def test_diff_mesh_estimates_volpath_in_expectation():
    """The differentiable estimator over an interface mesh (fixed majorant,
    proposal = evaluated grid, ratio tracking without roulette through the
    interfaces) and the reference's VolPathIntegrator estimate the same
    radiance: means over the pixels whose rays meet the medium within the
    Monte Carlo error; a grid of twice the density differs."""
    b, g = _mesh_diff_scene(16, 6)
    W = H = 16
    spp = 128
    S, A, G = 16.0, 0.8, 0.3
    mu = S * float(g.max())
    _, img, _ = b.diff_render(g.astype(np.float64), g, [S, A, G, mu], LIGHTSF, CAMF, W, H, spp, 32)
    rgb, _ = b.volpath_mesh(g, [A * S, (1 - A) * S, G], LIGHTSF, CAMF, W, H, spp, 32)
    m = rgb[:, :, 0] != np.float32(0.5)  # the other pixels see the environment only
    assert 0.2 < m.mean() < 0.8
    a, c = float(img[m].mean()), float(rgb[:, :, 0][m].mean())
    assert abs(a - c) / c < 0.015, (a, c)
    rgb2, _ = b.volpath_mesh(2 * g, [A * S, (1 - A) * S, G], LIGHTSF, CAMF, W, H, spp, 32)
    assert abs(float(rgb2[:, :, 0][m].mean()) - a) / a > 0.05


def test_diff_mesh_gradient_central_differences():
    """The FP64 twin's path-replay gradient over an interface mesh against
    central differences of its own forward with the proposal grid held fixed,
    for 24 voxels with a non-negligible gradient and the albedo; F = sum(dl *
    image).  Steps of 1e-5 (the error of a central difference falls as the
    step squared: 3e-3 at 1e-3, 3e-7 at 1e-5) and, for the float albedo
    parameter, 2^-14 added exactly.  Shadow rays cross the interfaces, so a
    lookup missing from a track's derivative, or an event's derivative
    weighted by the wrong share of the radiance, fails here."""
    b, g = _mesh_diff_scene(12, 5)
    W = H = 16
    spp, depth = 2, 16
    S, A, G = 8.0, 0.8, 0.3
    mu = S * float(g.max())
    theta = g.astype(np.float64)
    rng = np.random.default_rng(3)
    dl = rng.normal(size=(H, W))
    lp, _, _ = b.diff_render(theta, g, [S, A, G, mu], LIGHTSF, CAMF, W, H, spp, depth)
    grad, ga = b.diff_grad(theta, g, [S, A, G, mu], LIGHTSF, CAMF, W, H, spp, depth, lp, dl)
    assert np.abs(grad).max() > 0

    def F(th, alb=A):
        _, img, _ = b.diff_render(th, g, [S, alb, G, mu], LIGHTSF, CAMF, W, H, spp, depth)
        return float((dl * img).sum())

    cand = np.flatnonzero(np.abs(grad.ravel()) > 0.05 * np.abs(grad).max())
    pick = rng.choice(cand, size=min(24, cand.size), replace=False)
    eps = 1e-5
    for v in pick:
        tp = theta.copy()
        tp.flat[v] += eps
        tm = theta.copy()
        tm.flat[v] -= eps
        fd = (F(tp) - F(tm)) / (2 * eps)
        assert abs(grad.flat[v] - fd) <= 1e-5 * abs(fd) + 1e-9, (int(v), float(grad.flat[v]), fd)
    a0, h = float(np.float32(A)), 2.0 ** -14
    fa = (F(theta, a0 + h) - F(theta, a0 - h)) / (2 * h)
    assert abs(ga - fa) <= 1e-6 * abs(fa) + 1e-9


@pytest.mark.parametrize("n,dup,extra", [(40, [(5, 2), (17, 2), (30, 11)], []),
                                         (150, [], [(7, 3), (90, 40), (140, 41)])])
def test_scene_trace_is_lexicographic_min_over_meshes(n, dup, extra):
    """Scene::intersect over NaiveBVHAccel (its tree, box tests and strict
    replacement in each leaf) equals, per ray, the closest of the meshes' own
    hits and, between meshes at the same distance, the one first in the
    reference's leaf order: scene order up to 64 meshes (duplicated meshes
    tie everywhere), the median splits' order beyond."""
    meshes = scene_meshes(n, 3, dup, extra)
    lb = [orc.Lbvh(v, i) for v, i in meshes]
    rays = np.concatenate([scene_rays(3000, 4), scene_rays(1000, 5, tmax=6.0)])
    mesh, ids, ts = orc.scene_trace(lb, rays)
    rank = reference_rank(meshes)
    per = [b.trace(rays) for b in lb]
    best_t = np.full(len(rays), np.inf, np.float32)
    best_m = np.full(len(rays), -1)
    for m, (pid, pt) in enumerate(per):
        hit = pid >= 0
        better = hit & ((pt < best_t) | ((pt == best_t) & (best_m >= 0) & (rank[m] < rank[np.maximum(best_m, 0)])))
        better |= hit & (best_m < 0)
        best_t = np.where(better, pt, best_t)
        best_m = np.where(better, m, best_m)
    assert np.array_equal(mesh, best_m)
    hit = best_m >= 0
    want_id = np.array([per[m][0][i] if m >= 0 else -1 for i, m in enumerate(best_m)])
    assert np.array_equal(ids, want_id)
    assert np.array_equal(ts[hit].view(np.uint32), best_t[hit].view(np.uint32))
    assert np.array_equal(ts[~hit], rays[~hit, 3])
    assert hit.mean() > 0.3
    for j, k in dup:  # the duplicates tie and the lower scene index wins
        assert not np.any(mesh == j)
    for j, k in extra:  # ties on the shared surface go to the lower rank
        first, second = (j, k) if rank[j] < rank[k] else (k, j)
        both = (per[j][0] >= 0) & (per[k][0] >= 0) & (per[j][1] == per[k][1])
        assert both.sum() > 0 and not np.any(mesh[both] == second)


def test_volpath_mesh_strided_equals_full_on_its_pixels():
    """The strided oracle render (the combined scene's CPU baseline sample)
    computes exactly the full render's pixels 0, s, 2s, ... and nothing else."""
    from fledge_amd import api
    v, ind = api.gen_octasphere(3, 0.05, seed=7, base_freq=2.0)
    v = (v * 0.5 + 0.5).astype(np.float32)
    grid = api.gen_fbm_grid(16, seed=5)
    m = orc.Lbvh(v, ind)
    args = (grid, [6.4, 1.6, 0.3], [1, 2, 2, -2, 50.0, 1, 0.5], [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0],
            24, 20, 2, 8)
    full, _ = m.volpath_mesh(*args, nthreads=2)
    part, _, pa = m.volpath_mesh_strided(*args, 7, nthreads=2)
    f, p = full.reshape(-1, 3), part.reshape(-1, 3)
    sel = np.arange(0, 24 * 20, 7)
    assert pa == sel.size * 2
    assert np.array_equal(f[sel].view(np.uint32), p[sel].view(np.uint32))
    rest = np.setdiff1d(np.arange(24 * 20), sel)
    assert not p[rest].any() and f[sel].any()
