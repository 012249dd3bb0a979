"""GPU parity of the volume-only integrator and of the device copies of the
scalar helpers (log, sin/cos, Halton) against the CPU oracle."""
import numpy as np
import pytest
import torch

import oracle_lib as orc
from fledge_amd import _native as N
from fledge_amd import api

pytestmark = pytest.mark.gpu

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]


def _debug_eval(ctx, kind, x, n=None, out_dtype=torch.float32):
    xin = torch.from_numpy(np.ascontiguousarray(x)).cuda()
    n = x.shape[0] if n is None else n
    out = torch.empty(n, dtype=out_dtype, device="cuda")
    ctx.check(ctx.lib.flg_debug_eval(ctx.h, kind, api._ptr(xin), api._ptr(out), n))
    return out.cpu().numpy()


def test_logf_bitexact(ctx):
    bits = np.concatenate([
        np.arange(1, 0x7F800000, 977, dtype=np.uint32),           # denormals .. max float
        np.arange(0x3F000000, 0x3F800001, 7, dtype=np.uint32),    # [0.5, 1]
        np.array([0, 0x80000000, 0x7F800000, 0xBF800000, 0x33800000, 0x3F7FFFFF], np.uint32),
    ])
    x = bits.view(np.float32)
    g = _debug_eval(ctx, 0, x)
    o = orc.logf(x)
    assert np.array_equal(g.view(np.uint32), o.view(np.uint32))


def test_sincos_bitexact(ctx):
    x = np.linspace(-4.0, 7.0, 2_000_003, dtype=np.float32)
    o_s, o_c = orc.sincosf(x)
    g_s, g_c = _debug_eval(ctx, 1, x), _debug_eval(ctx, 2, x)
    bad = np.nonzero((g_s.view(np.uint32) != o_s.view(np.uint32)) | (g_c.view(np.uint32) != o_c.view(np.uint32)))[0]
    assert bad.size == 0, [(float(x[k]), float(g_s[k]), float(o_s[k]), float(g_c[k]), float(o_c[k])) for k in bad[:5]]


def test_halton_bitexact(ctx):
    rng = np.random.default_rng(1)
    n = 200000
    idx = rng.integers(0, 31104 * 64, n).astype(np.uint32)
    idx[2000:60000] = rng.integers(0, 2 ** 32, 58000, dtype=np.uint64).astype(np.uint32)  # the whole index range
    idx[60000:60100] = np.uint32(2 ** 32 - 1) - np.arange(100, dtype=np.uint32)
    dim = rng.integers(0, 3000, n).astype(np.uint32)
    dim[:1000] = 0
    dim[1000:2000] = 1
    packed = np.stack([idx, dim], axis=1).ravel()
    g = _debug_eval(ctx, 3, packed, n=n)
    L = orc.lib()
    primes = [L.orc_prime(i) for i in range(3000)]
    exp = np.empty(n, np.float32)
    for k in range(n):
        a = int(idx[k])
        d = int(dim[k])
        a = a // 128 if d == 0 else (a // 243 if d == 1 else a)
        v = L.orc_radical_inverse(primes[d], a)
        exp[k] = min(np.float32(v), np.float32(0.99999994))
    assert np.array_equal(g.view(np.uint32), exp.view(np.uint32))


def test_walk_draw_bitexact(ctx):
    """The differentiable estimator's ratio-tracking stream (path, walk, n)
    against the oracle's restatement, over the full counter ranges."""
    rng = np.random.default_rng(4)
    n = 300000
    x = np.stack([rng.integers(0, 2 ** 31, n), rng.integers(0, 64, n), rng.integers(0, 4096, n)],
                 axis=1).astype(np.uint32)
    x[:4096] = np.stack([np.zeros(4096), np.zeros(4096), np.arange(4096)], axis=1).astype(np.uint32)
    x[4096:8192, 0] = 0xFFFFFFFF
    g = _debug_eval(ctx, 4, x.ravel(), n=n)
    assert np.array_equal(g.view(np.uint32), orc.walk_draw(x).view(np.uint32))


def _scene(ctx, grid, spp, W, H, max_depth, point=True, env=True, sigma=(6.4, 1.6), g=0.3):
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], sigma[0], sigma[1], g)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam, has_point=point, has_env=env)
    out, st = vol.render(sc, W, H, spp, max_depth, stats=True)
    ref, lk, paths = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [sigma[0], sigma[1], g],
                                        [1.0 if point else 0.0, 2.0, 2.0, -2.0, 50.0,
                                         1.0 if env else 0.0, 0.5], CAM0, W, H, spp, max_depth)
    return out.cpu().numpy(), ref, st, lk


@pytest.mark.parametrize("point,env", [(True, True), (False, True), (True, False), (False, False)])
def test_volpath_parity_small(ctx, point, env):
    grid = api.gen_fbm_grid(32, seed=3)
    print("volpath small point=%s env=%s" % (point, env), flush=True)
    out, ref, st, lk = _scene(ctx, grid, 8, 64, 64, 32, point, env)
    assert st["lookups"] == lk
    rel = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= 1e-4
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))


def test_volpath_parity_fbm256_crop(ctx):
    """The 256^3 fBm volume (sigma_t_max 8, albedo 0.8, g 0.3) at 128x128, 8 spp."""
    grid = api.gen_fbm_grid(256, seed=1)
    out, ref, st, lk = _scene(ctx, grid, 8, 128, 128, 32)
    rel = np.abs(out - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= 1e-4
    assert st["lookups"] == lk


def test_volume_white_furnace(ctx):
    """Albedo 1, constant environment L = 0.5, no point light: every pixel -> 0.5."""
    grid = np.full((16, 16, 16), 0.5, np.float32)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 4.0, 0.0, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam, has_point=False, has_env=True, env=0.5)
    out = vol.render(sc, 64, 64, 64, 512).cpu().numpy()
    assert abs(out.mean() - 0.5) < 0.02


@pytest.mark.parametrize("schedule", [1, 2, 3])
@pytest.mark.parametrize("W,H,spp", [(96, 80, 8), (37, 23, 3), (211, 197, 5)])
def test_volpath_schedules_identical(ctx, schedule, W, H, spp):
    """One lane per path, persistent lanes with path regeneration and the
    queue-based wavefront give the same image and lookup count, bit for bit,
    and all match the oracle.  The regenerating lanes claim paths 128 at a
    time per wave: 61,440 paths fill whole chunks, 2,553 leave most waves
    without one, and 207,835 send waves back for more with a partial last
    chunk."""
    grid = api.gen_fbm_grid(32, seed=8)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam)
    out, st = vol.render(sc, W, H, spp, 32, stats=True, schedule=schedule)
    ref, lk, _ = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], [1, 2, 2, -2, 50.0, 1, 0.5],
                                    CAM0, W, H, spp, 32)
    assert st["lookups"] == lk
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


@pytest.mark.parametrize("schedule", [1, 2, 3])
def test_volpath_film_matches_oracle(ctx, schedule):
    """film = 1: the reference's render loop (pixel-centre path, samples
    1..spp-1 with the render loop's Halton schedule) resolved through the
    Mitchell film, against the oracle's single-threaded restatement: equal up
    to the film's summation order (the GPU gathers in (qx, qy, sample) order;
    the oracle splats in the reference's tile order), same lookup count.  The
    image spans several 16 x 16 tiles and an odd number of tasks."""
    grid = api.gen_fbm_grid(32, seed=10)
    W, H, spp = 45, 37, 4
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam)
    out, st = vol.render(sc, W, H, spp, 32, stats=True, schedule=schedule, film=1)
    ref, _, _, lk = orc.volpath_render_film(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3],
                                            [1, 2, 2, -2, 50.0, 1, 0.5], CAM0, W, H, spp, 32)
    got = out.cpu().numpy()
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= 1e-5, rel.max()
    assert st["lookups"] == lk and st["paths"] == W * H * spp
    other, _ = vol.render(sc, W, H, spp, 32, stats=True, schedule=1 if schedule != 1 else 2, film=1)
    assert torch.equal(out.view(torch.int32), other.view(torch.int32))


@pytest.mark.parametrize("film,depth", [(0, 32), (1, 8), (0, 1)])
def test_volpath_wavefront_graph_replays(ctx, film, depth):
    """The wavefront's stage sequence replayed from its hipGraph: every replay,
    with eager work and synchronisations between them, equals the eager
    render bit for bit; after the grid changed (a new majorant, so a new
    graph under a new key) the replay equals the regenerating schedule."""
    grid = api.gen_fbm_grid(32, seed=12)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam)
    W, H, spp = 53, 41, 5
    eager = vol.render(sc, W, H, spp, depth, schedule=3, film=film).clone()
    for k in range(4):
        g = vol.render(sc, W, H, spp, depth, schedule=3, film=film, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(g.view(torch.int32), eager.view(torch.int32)), k
        vol.render(sc, 17, 13, 2, 4, schedule=2)  # eager work between the replays
        torch.cuda.synchronize()
    vol.update(torch.from_numpy(grid * np.float32(0.5)).cuda())
    want = vol.render(sc, W, H, spp, depth, schedule=2, film=film)
    got = vol.render(sc, W, H, spp, depth, schedule=3, film=film, use_graph=True)
    torch.cuda.synchronize()
    assert torch.equal(got.view(torch.int32), want.view(torch.int32))


def test_volpath_film_constant_radiance_and_one_sample(ctx):
    """An empty medium under the environment light gives the environment
    radiance in every film pixel; one sample per pixel leaves the film at 0."""
    grid = np.zeros((8, 8, 8), np.float32)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam)
    out = vol.render(sc, 33, 17, 6, 8, film=1).cpu().numpy()
    assert np.abs(out - 0.5).max() <= 2e-7
    one, st = vol.render(sc, 33, 17, 1, 8, film=1, stats=True)
    assert torch.count_nonzero(one) == 0 and st["paths"] == 0


def test_volpath_full_benchmark_frame(ctx):
    """The benchmark's volume frame at full size (the 256^3 fBm grid,
    sigma_s 6.4, sigma_a 1.6, g 0.3, both lights, 512 x 512, 64 spp, depth
    32: 16,777,216 path samples): every pixel bit-equal to the oracle, and the
    same number of density lookups."""
    grid = api.gen_fbm_grid(256, seed=1)
    print("full volume frame", flush=True)
    out, ref, st, lk = _scene(ctx, grid, 64, 512, 512, 32)
    assert st["lookups"] == lk and st["paths"] == 512 * 512 * 64
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    # the wavefront schedule at the same full size
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    wf, sw = vol.render(api.vol_scene(cam), 512, 512, 64, 32, stats=True, schedule=3)
    assert sw["lookups"] == lk
    assert np.array_equal(wf.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_volpath_halton_indices_beyond_2pow24(ctx):
    """700 samples per pixel at 128 x 128 (Halton stride 31,104): sample
    indices reach 21.8 M, past 2^24, where a float quotient would no longer
    be exact.  Every 37th pixel against the oracle's integer digit loop."""
    grid = api.gen_fbm_grid(16, seed=4)
    W = H = 128
    spp, depth = 700, 8
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    out = vol.render(api.vol_scene(cam), W, H, spp, depth).cpu().numpy().reshape(-1, 3)
    L = orc.lib()
    rgb = np.zeros((H, W, 3), np.float32)
    lk = np.zeros(1, np.uint64)
    pa = np.zeros(1, np.uint64)
    stride = 37
    L.orc_volpath_render_strided(orc._p(grid), 16, 16, 16, orc._p(np.array([0, 0, 0, 1, 1, 1], np.float32)),
                                 orc._p(np.array([6.4, 1.6, 0.3], np.float32)),
                                 orc._p(orc.lights11([1, 2, 2, -2, 50.0, 1, 0.5])),
                                 orc._p(np.array(CAM0, np.float32)), W, H, spp, depth, stride, orc._p(rgb), 8,
                                 orc._p(lk), orc._p(pa))
    sel = np.arange(0, W * H, stride)
    assert int(pa[0]) == sel.size * spp
    ref = rgb.reshape(-1, 3)[sel]
    assert np.array_equal(out[sel].view(np.uint32), ref.view(np.uint32))


LIGHTS_RGB = [1, 2, 2, -2, 50.0, 20.0, 5.0, 1, 0.3, 0.5, 0.7]


@pytest.mark.parametrize("schedule,film", [(1, 0), (2, 0), (3, 0), (1, 1), (2, 1), (3, 1)])
def test_volpath_chromatic_lights_match_oracle(ctx, schedule, film):
    """RGB lights (point (50, 20, 5), environment (0.3, 0.5, 0.7)): every
    channel equals the oracle's Vector3f restatement, bit for bit per sample
    schedule, to the film's summation order in film mode; same lookups."""
    print("volpath chromatic schedule %d film %d" % (schedule, film), flush=True)
    grid = api.gen_fbm_grid(32, seed=14)
    W, H, spp = 45, 37, 5
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam, intensity=LIGHTS_RGB[4:7], env=LIGHTS_RGB[8:11])
    out, st = vol.render(sc, W, H, spp, 32, stats=True, schedule=schedule, film=film)
    got = out.cpu().numpy()
    if film:
        ref, _, _, lk = orc.volpath_render_film(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], LIGHTS_RGB, CAM0, W, H,
                                                spp, 32)
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)
        assert rel.max() <= 1e-5, rel.max()
    else:
        ref, lk, _ = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], LIGHTS_RGB, CAM0, W, H, spp, 32)
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert st["lookups"] == lk
    assert not np.array_equal(got[..., 0], got[..., 1]) and not np.array_equal(got[..., 1], got[..., 2])


def test_volpath_channels_scale_exactly(ctx):
    """Lights whose channels differ by powers of two give channels that differ
    by exactly those powers (float operations commute with them)."""
    grid = api.gen_fbm_grid(32, seed=15)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam, intensity=(64.0, 32.0, 16.0), env=(0.5, 0.25, 0.125))
    for schedule in (1, 2, 3):
        got = vol.render(sc, 40, 32, 4, 32, schedule=schedule).cpu().numpy()
        assert got[..., 0].max() > 0
        assert np.array_equal(got[..., 0], got[..., 1] * np.float32(2))
        assert np.array_equal(got[..., 0], got[..., 2] * np.float32(4))
