"""GPU differentiable forward and path-replay adjoint against the FP64 CPU
twin of the same estimator (analytic gradient) and central finite
differences (eps 1e-3) of that twin; gradients within 1e-3 relative."""
import numpy as np
import pytest
import torch

import oracle_lib as orc
from fledge_amd import _native as N
from fledge_amd import api, diff

pytestmark = pytest.mark.gpu

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]
LIGHTS = [1.0, 2.0, 2.0, -2.0, 50.0, 1.0, 0.5]
S, ALBEDO, G = 8.0, 0.8, 0.3



def _grey(lpath):
    """Per-path radiance of grey lights as float64: one channel, after
    checking that the three are equal."""
    a = lpath.cpu().numpy()
    assert np.array_equal(a[:, 0], a[:, 1]) and np.array_equal(a[:, 0], a[:, 2])
    return a[:, 0].astype(np.float64)

def _setup(ctx, grid, W, H, spp, depth, mu):
    dv = diff.DiffVolume(ctx, torch.from_numpy(grid).cuda(), g=G)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    scene = api.vol_scene(cam)
    img, lpath = diff.forward(dv, scene, W, H, spp, depth, S, ALBEDO, mu)
    return dv, scene, img, lpath


def _grid(n, seed):
    g = api.gen_fbm_grid(n, seed=seed)
    return (0.1 + 0.8 * g).astype(np.float32)  # keeps 1 - s rho / mu away from 0


def test_diff_forward_matches_fp64_twin(ctx):
    grid = _grid(24, 5)
    W = H = 48
    spp, depth = 4, 32
    mu = S * float(grid.max())
    dv, scene, img, lpath = _setup(ctx, grid, W, H, spp, depth, mu)
    lp64, img64 = orc.diff_render(grid.astype(np.float64), grid, [S, ALBEDO, G, mu], LIGHTS, CAM0,
                                  W, H, spp, depth)
    got = img.cpu().numpy()[:, :, 0]
    rel = np.abs(got - img64) / np.maximum(np.abs(img64), 1e-6)
    assert rel.max() < 1e-4
    # per path: float (GPU) against the FP64 twin, relative to the path's value
    # or, for paths far darker than average (products of near-zero ratio
    # tracking factors, where float and double round apart), to the mean path
    lp = _grey(lpath)
    scale = max(float(np.abs(lp64).mean()), 1e-6)
    rel_p = np.abs(lp - lp64) / np.maximum(np.abs(lp64), scale)
    print("per-path: max %.3g at %d (gpu %.9g, fp64 %.9g), mean path %.4g" % (
        rel_p.max(), int(rel_p.argmax()), lp[rel_p.argmax()], lp64[rel_p.argmax()], scale), flush=True)
    assert rel_p.max() < 1e-4


def test_diff_forward_estimates_the_plain_integrator(ctx):
    """Same scene through the reference delta-tracking integrator and the
    differentiable one: equal in expectation (image means within noise)."""
    grid = _grid(32, 6)
    W = H = 64
    spp = 64
    mu = S * float(grid.max())
    dv, scene, img, _ = _setup(ctx, grid, W, H, spp, 32, mu)
    vol = api.Volume(ctx, torch.from_numpy(grid).cuda(), [0, 0, 0, 1, 1, 1], ALBEDO * S,
                     (1 - ALBEDO) * S, G)
    ref = vol.render(scene, W, H, spp, 32)
    a, b = float(img.mean()), float(ref.mean())
    assert abs(a - b) / b < 0.02


def _adjoint_case(ctx, n, W, H, spp, depth, seed):
    grid = _grid(n, seed)
    mu = S * float(grid.max())
    dv, scene, img, lpath = _setup(ctx, grid, W, H, spp, depth, mu)
    tape = diff.new_tape(ctx, W, H, spp, depth)
    diff.forward(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lpath=lpath, tape=tape)
    rng = np.random.default_rng(seed)
    dl = rng.normal(size=(H, W, 3)).astype(np.float32)
    grad, galpha = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lpath,
                                torch.from_numpy(dl).cuda(), tape=tape)
    lp64, _ = orc.diff_render(grid.astype(np.float64), grid, [S, ALBEDO, G, mu], LIGHTS, CAM0,
                              W, H, spp, depth)
    g64, ga64 = orc.diff_grad(grid.astype(np.float64), grid, [S, ALBEDO, G, mu], LIGHTS, CAM0,
                              W, H, spp, depth, lp64, dl.astype(np.float64).sum(axis=2))
    return grid, mu, grad.cpu().numpy().astype(np.float64), float(galpha.item()), g64, ga64, dl


def test_adjoint_matches_fp64_gradient(ctx):
    grid, mu, g, ga, g64, ga64, dl = _adjoint_case(ctx, 24, 48, 48, 4, 32, 7)
    assert np.abs(g64).max() > 0
    rel = np.linalg.norm(g - g64) / np.linalg.norm(g64)
    assert rel < 1e-3
    big = np.abs(g64) > 1e-2 * np.abs(g64).max()
    assert (np.abs(g - g64)[big] / np.abs(g64)[big]).max() < 1e-3
    assert abs(ga - ga64) / abs(ga64) < 1e-3


def test_adjoint_central_finite_differences(ctx):
    """dF/dtheta_v for 64 random voxels with a non-negligible gradient, and
    dF/dalbedo, against central differences (eps 1e-3) of the FP64 forward with
    the proposal grid held fixed; F = sum(dl * image)."""
    W = H = 32
    spp, depth = 4, 32
    grid, mu, g, ga, g64, ga64, dl = _adjoint_case(ctx, 16, W, H, spp, depth, 11)
    dlp = dl.astype(np.float64).sum(axis=2)
    theta = grid.astype(np.float64)

    def F(th, alb=ALBEDO):
        _, img = orc.diff_render(th, grid, [S, alb, G, mu], LIGHTS, CAM0, W, H, spp, depth)
        return float((dlp * img).sum())

    cand = np.flatnonzero(np.abs(g64.ravel()) > 0.05 * np.abs(g64).max())
    rng = np.random.default_rng(0)
    pick = rng.choice(cand, size=min(64, cand.size), replace=False)
    eps = 1e-3
    for v in pick:
        tp = theta.copy()
        tp.flat[v] += eps
        tm = theta.copy()
        tm.flat[v] -= eps
        fd = (F(tp) - F(tm)) / (2 * eps)
        assert abs(g.flat[v] - fd) <= 1e-3 * abs(fd) + 1e-6, (int(v), float(g.flat[v]), fd)
    # albedo: the estimator is polynomial in alpha; central difference in float32
    # parameters of the FP64 twin (alpha +- 1e-3 represented exactly enough)
    fa = (F(theta, ALBEDO + 1e-3) - F(theta, ALBEDO - 1e-3)) / 2e-3
    assert abs(ga - fa) <= 1e-3 * abs(fa) + 1e-6


@pytest.mark.parametrize("state", ["uniform", "fbm"])
def test_adjoint_grid128_full_size(ctx, state):
    """Full size: 128^3 grid, 256x256, 16 spp; L2 loss against a target from
    an fBm grid (seed 2).  The optimised grid is at the start state (0.5
    everywhere: every ratio-tracking factor is 0, walks end at once) or at a
    non-uniform state (another fBm grid, seed 3: walks of ordinary length).
    GPU gradient vs the FP64 gradient."""
    W = H = 256
    spp, depth = 16, 32
    if state == "uniform":
        theta = np.full((128, 128, 128), 0.5, np.float32)
    else:
        theta = api.gen_fbm_grid(128, seed=3)
    target_grid = api.gen_fbm_grid(128, seed=2)
    mu = S * float(theta.max())
    tv = diff.DiffVolume(ctx, torch.from_numpy(target_grid).cuda(), g=G)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    scene = api.vol_scene(cam)
    target, _ = diff.forward(tv, scene, W, H, spp, depth, S, ALBEDO, S * float(target_grid.max()))
    dv, scene, img, lpath = _setup(ctx, theta, W, H, spp, depth, mu)
    dl = 2.0 * (img - target) / img.numel()
    grad, galpha = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lpath, dl)
    lp64, _ = orc.diff_render(theta.astype(np.float64), theta, [S, ALBEDO, G, mu], LIGHTS, CAM0,
                              W, H, spp, depth)
    g64, ga64 = orc.diff_grad(theta.astype(np.float64), theta, [S, ALBEDO, G, mu], LIGHTS, CAM0,
                              W, H, spp, depth, lp64, dl.cpu().numpy().astype(np.float64).sum(axis=2))
    g = grad.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(g - g64) / np.linalg.norm(g64) < 1e-3
    assert abs(float(galpha.item()) - ga64) / abs(ga64) < 1e-3


def test_autograd_function_and_descent(ctx):
    W = H = 64
    spp, depth = 4, 16
    target_grid = torch.from_numpy(_grid(32, 3)).cuda()
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    scene = api.vol_scene(cam)
    tv = diff.DiffVolume(ctx, target_grid, g=G)
    mu = S
    target, _ = diff.forward(tv, scene, W, H, spp, depth, S, ALBEDO, mu)
    theta = torch.full((32, 32, 32), 0.5, dtype=torch.float32, device="cuda", requires_grad=True)
    albedo = torch.tensor(0.6, dtype=torch.float32, device="cuda", requires_grad=True)
    dv = diff.DiffVolume(ctx, theta.detach(), g=G)
    opt = torch.optim.Adam([theta, albedo], lr=0.02)
    losses = []
    for it in range(8):
        opt.zero_grad()
        img = diff.VolumeRender.apply(theta, albedo, dv, scene, W, H, spp, depth, S, mu)
        loss = ((img - target) ** 2).mean()
        loss.backward()
        assert torch.isfinite(theta.grad).all() and theta.grad.abs().sum() > 0
        losses.append(float(loss))
        opt.step()
        with torch.no_grad():
            theta.clamp_(0.0, 1.0)
            albedo.clamp_(0.05, 1.0)
    assert losses[-1] < losses[0]


@pytest.mark.parametrize("schedule", [1, 2])
@pytest.mark.parametrize("W,H,spp", [(40, 40, 4), (37, 29, 3)])
def test_diff_schedules_agree(ctx, schedule, W, H, spp):
    """One lane per path and persistent path regeneration: identical forward
    (bit for bit) and gradients equal up to the float-atomic summation order,
    with path counts that fill whole 128-path chunks (6,400) and that do not
    (3,219)."""
    grid = _grid(20, 13)
    depth = 32
    mu = S * float(grid.max())
    dv, scene, img1, lp1 = _setup(ctx, grid, W, H, spp, depth, mu)  # default schedule
    cfg_img = torch.empty_like(img1)
    lp = torch.empty_like(lp1)
    img, lp = diff.forward(dv, scene, W, H, spp, depth, S, ALBEDO, mu, out=cfg_img, lpath=lp,
                           schedule=schedule)
    assert torch.equal(img.view(torch.int32), img1.view(torch.int32))
    assert torch.equal(lp.view(torch.int32), lp1.view(torch.int32))
    dl = torch.from_numpy(np.random.default_rng(2).normal(size=(H, W, 3)).astype(np.float32)).cuda()
    g_ref, ga_ref = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lp1, dl, schedule=1)
    g, ga = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lp, dl, schedule=schedule)
    rel = float((g - g_ref).norm() / g_ref.norm())
    assert rel < 1e-5 and abs(float(ga) - float(ga_ref)) <= 1e-5 * abs(float(ga_ref))


@pytest.mark.parametrize("schedule", [1, 2])
def test_walk_tape_matches_replay(ctx, schedule):
    """A forward that records the walk tape renders the same image and path
    values bit for bit, and the adjoint that reads the tape gives the gradient
    of the one that walks every ratio track twice (up to the float-atomic
    summation order).  Covers walks with zero factors: theta at the majorant."""
    W = H = 40
    spp, depth = 4, 32
    for grid in (_grid(20, 17), np.full((20, 20, 20), 0.5, np.float32)):
        mu = S * float(grid.max())
        dv, scene, img0, lp0 = _setup(ctx, grid, W, H, spp, depth, mu)
        tape = diff.new_tape(ctx, W, H, spp, depth)
        assert tape.numel() == W * H * spp * 4 * (depth + 3)
        img, lp = diff.forward(dv, scene, W, H, spp, depth, S, ALBEDO, mu, schedule=schedule, tape=tape)
        assert torch.equal(img.view(torch.int32), img0.view(torch.int32))
        assert torch.equal(lp.view(torch.int32), lp0.view(torch.int32))
        dl = torch.from_numpy(np.random.default_rng(3).normal(size=(H, W, 3)).astype(np.float32)).cuda()
        g_ref, ga_ref = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lp0, dl, schedule=schedule)
        g, ga = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lp, dl, schedule=schedule, tape=tape)
        assert float(g_ref.abs().max()) > 0
        rel = float((g - g_ref).norm() / g_ref.norm())
        assert rel < 1e-5 and abs(float(ga) - float(ga_ref)) <= 1e-5 * abs(float(ga_ref)), rel


LIGHTS_RGB = [1, 2, 2, -2, 50.0, 20.0, 5.0, 1, 0.3, 0.5, 0.7]


def test_chromatic_forward_and_adjoint_match_fp64_twin(ctx):
    """RGB lights: the forward's per-path radiance per channel against the
    twin run on that channel (1e-4 of the path or the mean path), and the
    adjoint, which prices each light in the loss's units in one replay,
    against the sum over channels of the twin's gradients of
    sum(dl_c * image_c) (1e-3)."""
    grid = _grid(20, 19)
    W = H = 40
    spp, depth = 4, 32
    mu = S * float(grid.max())
    dv = diff.DiffVolume(ctx, torch.from_numpy(grid).cuda(), g=G)
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    scene = api.vol_scene(cam, intensity=LIGHTS_RGB[4:7], env=LIGHTS_RGB[8:11])
    img, lpath = diff.forward(dv, scene, W, H, spp, depth, S, ALBEDO, mu)
    lp = lpath.cpu().numpy().astype(np.float64)
    im = img.cpu().numpy().astype(np.float64)
    dl = np.random.default_rng(4).normal(size=(H, W, 3)).astype(np.float32)
    grad, galpha = diff.adjoint(dv, scene, W, H, spp, depth, S, ALBEDO, mu, lpath, torch.from_numpy(dl).cuda())
    g64 = 0.0
    ga64 = 0.0
    th = grid.astype(np.float64)
    for c in range(3):
        lp64, img64 = orc.diff_render(th, grid, [S, ALBEDO, G, mu], LIGHTS_RGB, CAM0, W, H, spp, depth, channel=c)
        scale = max(float(np.abs(lp64).mean()), 1e-6)
        assert (np.abs(lp[:, c] - lp64) / np.maximum(np.abs(lp64), scale)).max() < 1e-4
        assert (np.abs(im[..., c] - img64) / np.maximum(np.abs(img64), 1e-6)).max() < 1e-4
        g, ga = orc.diff_grad(th, grid, [S, ALBEDO, G, mu], LIGHTS_RGB, CAM0, W, H, spp, depth, lp64,
                              dl[..., c].astype(np.float64), channel=c)
        g64 = g64 + g
        ga64 += ga
    got = grad.cpu().numpy().astype(np.float64)
    assert np.abs(g64).max() > 0
    assert np.linalg.norm(got - g64) / np.linalg.norm(g64) < 1e-3
    assert abs(float(galpha.item()) - ga64) / abs(ga64) < 1e-3
