"""flg_volpath_render_noalloc (the volume render with its scratch in the
library's fixed per-GPU array) through the C ABI: bit-equal to
flg_volpath_render on two calls in a row, the second smaller, on both
schedules and in film mode; frames beyond the fixed scratch and scenes whose
environment table is not yet built are refused before any launch."""
import ctypes

import numpy as np
import pytest
import torch

from fledge_amd import _native as N
from fledge_amd import api

pytestmark = pytest.mark.gpu

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]


def _cfg(w, h, spp, depth, schedule=0, film=0):
    c = N.RenderCfg()
    c.width, c.height, c.spp, c.max_depth = w, h, spp, depth
    c.schedule = schedule
    c.film = film
    return c


def _both(ctx, vol, scene, w, h, spp, depth, schedule=0, film=0):
    cfg = _cfg(w, h, spp, depth, schedule, film)
    a = torch.full((h, w, 3), -1.0, dtype=torch.float32, device="cuda")
    b = torch.full((h, w, 3), -2.0, dtype=torch.float32, device="cuda")
    ctx.check(ctx.lib.flg_volpath_render(ctx.h, vol.h, ctypes.byref(scene), ctypes.byref(cfg), api._ptr(a), None))
    ctx.check(ctx.lib.flg_volpath_render_noalloc(ctx.h, vol.h, ctypes.byref(scene), ctypes.byref(cfg),
                                                 api._ptr(b)))
    torch.cuda.synchronize()
    return a.cpu().numpy(), b.cpu().numpy()


def test_noalloc_render_matches_render(ctx):
    grid = api.gen_fbm_grid(24, seed=5)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    scene = api.vol_scene(cam)
    for schedule in (0, 1):
        # a frame, then a smaller one: the fixed scratch keeps the first
        # call's contents and the second must not see them
        for (w, h, spp) in ((96, 64, 16), (40, 24, 4)):
            print("noalloc case: schedule %d, %dx%d, %d spp" % (schedule, w, h, spp), flush=True)
            a, b = _both(ctx, vol, scene, w, h, spp, 16, schedule)
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
            assert a.max() > 0
    print("noalloc case: film", flush=True)
    for (w, h, spp) in ((64, 48, 9), (16, 16, 3)):
        a, b = _both(ctx, vol, scene, w, h, spp, 16, 0, 1)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    # chromatic lights; the environment colour (0.6, 0.0, 0.8) has norm 1, as
    # the grey 1/sqrt(3) would not, so its table is built by the first call
    print("noalloc case: chromatic", flush=True)
    rgb = api.vol_scene(cam, intensity=(50.0, 20.0, 5.0), env=(0.6, 0.0, 0.8))
    a, b = _both(ctx, vol, rgb, 48, 32, 8, 16)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    assert not np.array_equal(a[..., 0], a[..., 1])
    vol.close()


def test_noalloc_render_refuses_what_it_cannot_do(ctx):
    grid = api.gen_fbm_grid(16, seed=6)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], 6.4, 1.6, 0.3)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    out = torch.zeros((8, 8, 3), dtype=torch.float32, device="cuda")
    # environment table not yet built for this volume
    scene = api.vol_scene(cam, env=0.25)
    st = ctx.lib.flg_volpath_render_noalloc(ctx.h, vol.h, ctypes.byref(scene), ctypes.byref(_cfg(8, 8, 2, 8)),
                                            api._ptr(out))
    assert st == N.FLG_ERR_INVALID and b"environment" in ctx.lib.flg_last_error(ctx.h)
    # a frame beyond the fixed scratch (4096 x 4096 x 2 paths); the output is
    # never written, so a small buffer is safe here
    st = ctx.lib.flg_volpath_render_noalloc(ctx.h, vol.h, ctypes.byref(scene),
                                            ctypes.byref(_cfg(4096, 4096, 2, 8)), api._ptr(out))
    assert st == N.FLG_ERR_RANGE and b"scratch" in ctx.lib.flg_last_error(ctx.h)
    # without an environment light nothing needs building: it runs
    nopt = api.vol_scene(cam, has_env=False)
    a, b = _both(ctx, vol, nopt, 8, 8, 2, 8)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    vol.close()
