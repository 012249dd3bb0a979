"""Surfaces + media (VolPathIntegrator over a closed interface mesh) on the
GPU against the oracle's restatement: per-pixel radiance and lookup counts
bit for bit on both schedules, the empty medium, and agreement in expectation
with the volume-only integrator when a cube interface bounds the grid."""
import numpy as np
import pytest
import torch

import oracle_lib as orc
from fledge_amd import _native as N
from fledge_amd import api

pytestmark = pytest.mark.gpu

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]
LIGHTS = [1, 2, 2, -2, 50.0, 1, 0.5]
MED = [6.4, 1.6, 0.3]


def _sphere(level=4):
    v, ind = api.gen_octasphere(level, 0.05, seed=7, base_freq=2.0)
    return (v * 0.5 + 0.5).astype(np.float32), ind


def _scene(ctx, verts, ind, grid, env=True):
    bvh = api.LBVH.build(ctx, torch.from_numpy(verts).cuda(), torch.from_numpy(ind).cuda())
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], MED[0], MED[1], MED[2])
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    return bvh, vol, api.vol_scene(cam, has_env=env)


@pytest.mark.parametrize("schedule", [1, 2])
def test_volpath_mesh_matches_oracle(ctx, schedule):
    v, ind = _sphere(4)
    grid = api.gen_fbm_grid(24, seed=5)
    W, H, spp, D = 48, 36, 4, 16
    bvh, vol, sc = _scene(ctx, v, ind, grid)
    out, st = vol.render_mesh(bvh, sc, W, H, spp, D, stats=True, schedule=schedule)
    ref, lk = orc.Lbvh(v, ind).volpath_mesh(grid, MED, LIGHTS, CAM0, W, H, spp, D)
    got = out.cpu().numpy()
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= 1e-4, rel.max()
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert st["lookups"] == lk and st["paths"] == W * H * spp


def test_volpath_mesh_empty_medium(ctx):
    v, ind = _sphere(3)
    bvh, vol, sc = _scene(ctx, v, ind, np.zeros((16, 16, 16), np.float32))
    out, st = vol.render_mesh(bvh, sc, 40, 30, 4, 16, stats=True)
    assert st["lookups"] == 0 and torch.all(out == 0.5)


def test_volpath_cube_interface_matches_svol_in_expectation(ctx):
    """A cube interface exactly around the grid: image means of the two GPU
    integrators within the Monte Carlo error."""
    c = np.array([[x, y, z] for z in (0.0, 1.0) for y in (0.0, 1.0) for x in (0.0, 1.0)], np.float32)
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    ind = np.array([t for a, b, cc, d in quads for t in ((a, b, cc), (a, cc, d))], np.int32).ravel()
    grid = api.gen_fbm_grid(32, seed=9)
    bvh, vol, sc = _scene(ctx, c.ravel(), ind, grid)
    a = float(vol.render_mesh(bvh, sc, 64, 64, 64, 64).mean())
    b = float(vol.render(sc, 64, 64, 64, 64).mean())
    assert abs(a - b) / b < 0.01, (a, b)


@pytest.mark.parametrize("schedule", [1, 2])
def test_volpath_mesh_film_matches_oracle(ctx, schedule):
    """film = 1 over the interface mesh: the reference's render loop (the
    pixel-centre path, samples 1..spp-1 on the loop's Halton schedule)
    resolved through the Mitchell film, against the oracle's single-threaded
    loop: equal up to the film's summation order, same lookup count, and the
    same bits from both schedules.  The image spans several 16 x 16 tiles."""
    v, ind = _sphere(4)
    grid = api.gen_fbm_grid(24, seed=6)
    W, H, spp, D = 37, 29, 4, 16
    bvh, vol, sc = _scene(ctx, v, ind, grid)
    print("volpath mesh film schedule %d" % schedule, flush=True)
    out, st = vol.render_mesh(bvh, sc, W, H, spp, D, stats=True, schedule=schedule, film=1)
    ref, lk = orc.Lbvh(v, ind).volpath_mesh(grid, MED, LIGHTS, CAM0, W, H, spp, D, film=1)
    got = out.cpu().numpy()
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)
    assert rel.max() <= 1e-5, rel.max()
    assert st["lookups"] == lk and st["paths"] == W * H * spp
    other, _ = vol.render_mesh(bvh, sc, W, H, spp, D, stats=True, schedule=3 - schedule, film=1)
    assert torch.equal(out.view(torch.int32), other.view(torch.int32))


def test_volpath_mesh_film_one_sample_is_empty(ctx):
    """One sample per pixel: the render loop's only sample is the unsplatted
    pixel centre, so the film stays 0."""
    v, ind = _sphere(3)
    bvh, vol, sc = _scene(ctx, v, ind, api.gen_fbm_grid(16, seed=2))
    out, st = vol.render_mesh(bvh, sc, 20, 12, 1, 8, stats=True, film=1)
    assert torch.count_nonzero(out) == 0 and st["paths"] == 0


LIGHTS_RGB = [1, 2, 2, -2, 50.0, 20.0, 5.0, 1, 0.3, 0.5, 0.7]


@pytest.mark.parametrize("schedule,film", [(1, 0), (2, 0), (2, 1)])
def test_volpath_mesh_chromatic_lights_match_oracle(ctx, schedule, film):
    """RGB lights over the interface mesh: every channel equals the oracle's
    (bit for bit per sample, to the film's summation order in film mode)."""
    print("volpath mesh chromatic schedule %d film %d" % (schedule, film), flush=True)
    v, ind = _sphere(4)
    grid = api.gen_fbm_grid(24, seed=9)
    W, H, spp, D = 40, 30, 4, 16
    bvh, vol, _ = _scene(ctx, v, ind, grid)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    sc = api.vol_scene(cam, intensity=LIGHTS_RGB[4:7], env=LIGHTS_RGB[8:11])
    out, st = vol.render_mesh(bvh, sc, W, H, spp, D, stats=True, schedule=schedule, film=film)
    ref, lk = orc.Lbvh(v, ind).volpath_mesh(grid, MED, LIGHTS_RGB, CAM0, W, H, spp, D, film=film)
    got = out.cpu().numpy()
    if film:
        rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)
        assert rel.max() <= 1e-5, rel.max()
    else:
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert st["lookups"] == lk
    assert not np.array_equal(got[..., 0], got[..., 1])
