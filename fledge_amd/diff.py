"""Differentiable volume rendering on the GPU through the C ABI.

`render(vol_ctx, density, albedo, ...)` is a torch.autograd.Function whose
forward calls flg_volpath_diff_forward and whose backward calls
flg_volpath_adjoint (path-replay adjoint): gradients with respect to the
density grid and the single-scatter albedo.
"""
import ctypes

import torch

from . import _native as N
from .api import Volume, _ptr, check_buffer, check_device


def _check_forward(ctx, w, h, spp, depth, out, lpath, tape):
    check_buffer(out, "out", torch.float32, None, min_numel=w * h * 3)
    check_buffer(lpath, "lpath", torch.float32, None, min_numel=w * h * spp * 3)
    if tape is not None:
        check_buffer(tape, "tape", torch.uint8, None, min_numel=tape_bytes(ctx, w, h, spp, depth))
    check_device(ctx.device, out=out, lpath=lpath, tape=tape)


def _check_adjoint(ctx, w, h, spp, depth, lpath, dl, grad, grad_numel, galpha, tape):
    check_buffer(lpath, "lpath", torch.float32, None, min_numel=w * h * spp * 3)
    check_buffer(dl, "dl_drgb", torch.float32, None, numel=w * h * 3)
    check_buffer(grad, "grad", torch.float32, None, numel=grad_numel)
    check_buffer(galpha, "galpha", torch.float32, None, min_numel=1)
    if tape is not None:
        check_buffer(tape, "tape", torch.uint8, None, min_numel=tape_bytes(ctx, w, h, spp, depth))
    check_device(ctx.device, lpath=lpath, dl_drgb=dl, grad=grad, galpha=galpha, tape=tape)


class DiffVolume:
    """Holds the native volume (evaluated grid) and, optionally, a separate
    proposal volume for the collision decisions."""

    def __init__(self, ctx, density, bounds=(0, 0, 0, 1, 1, 1), g=0.3):
        self.ctx = ctx
        self.bounds = bounds
        self.g = g
        self.vol = Volume(ctx, density.detach(), bounds, 1.0, 0.0, g)
        self.proposal = None

    @classmethod
    def sparse(cls, ctx, origins, values, shape, bounds=(0, 0, 0, 1, 1, 1), g=0.3):
        """The evaluated grid as a sparse-occupancy volume: 8^3 bricks at
        voxel origins `origins` with values `values` (api.Volume.sparse); a
        gradient over its bricks uses BrickLayout.of_volume(self.vol)."""
        self = cls.__new__(cls)
        self.ctx = ctx
        self.bounds = bounds
        self.g = g
        self.vol = Volume.sparse(ctx, origins, values.detach(), shape, bounds, 1.0, 0.0, g)
        self.proposal = None
        return self

    def set_density(self, density):
        """New densities: a dense grid, or for a sparse volume its brick values."""
        if getattr(self.vol, "n_leaves", None) is not None:
            self.vol.update_sparse(density.detach())
        else:
            self.vol.update(density.detach())

    def set_proposal(self, density):
        if density is None:
            self.proposal = None
        else:
            self.proposal = Volume(self.ctx, density.detach(), self.bounds, 1.0, 0.0, self.g)


def _params(s, albedo, g, majorant, lookups=None):
    p = N.DiffParams()
    p.sigma_t_scale = float(s)
    p.albedo = float(albedo)
    p.g = float(g)
    p.majorant = float(majorant)
    p.d_lookups = lookups.data_ptr() if lookups is not None else None
    return p


def _check_counter(ctx, lookups):
    if lookups is not None:
        check_buffer(lookups, "lookups", torch.int64, ctx.device, min_numel=1)


def _cfg(w, h, spp, depth, schedule=0):
    c = N.RenderCfg()
    c.width, c.height, c.spp, c.max_depth = w, h, spp, depth
    c.schedule = schedule
    return c


def tape_bytes(ctx, w, h, spp, depth):
    """Bytes of the walk tape of one forward (flg_volpath_diff_tape_bytes)."""
    n = int(ctx.lib.flg_volpath_diff_tape_bytes(ctypes.byref(_cfg(w, h, spp, depth))))
    if n < 0:
        raise ValueError("no walk tape for this configuration (max_depth must be <= 63)")
    return n


def new_tape(ctx, w, h, spp, depth):
    """A device buffer for the walk tape of one forward."""
    return torch.empty(tape_bytes(ctx, w, h, spp, depth), dtype=torch.uint8,
                       device=torch.device("cuda", ctx.device))


def forward(dv, scene, w, h, spp, depth, s, albedo, majorant, out=None, lpath=None, schedule=0,
            tape=None, lookups=None):
    """Image and per-path RGB radiance ((w*h*spp, 3)); with `tape` (a buffer from new_tape) also
    the walk tape that the adjoint of this call can read; with `lookups` (an
    int64 CUDA tensor) the call adds its grid lookups to lookups[0]."""
    ctx = dv.ctx
    dev = torch.device("cuda", ctx.device)
    if out is None:
        out = torch.empty((h, w, 3), dtype=torch.float32, device=dev)
    if lpath is None:
        lpath = torch.empty((w * h * spp, 3), dtype=torch.float32, device=dev)
    _check_forward(ctx, w, h, spp, depth, out, lpath, tape)
    _check_counter(ctx, lookups)
    prop = dv.proposal.h if dv.proposal is not None else None
    ctx.check(ctx.lib.flg_volpath_diff_forward(ctx.h, dv.vol.h, prop, ctypes.byref(scene),
                                               ctypes.byref(_cfg(w, h, spp, depth, schedule)),
                                               ctypes.byref(_params(s, albedo, dv.g, majorant, lookups)),
                                               _ptr(out), _ptr(lpath), _ptr(tape)))
    return out, lpath


def adjoint(dv, scene, w, h, spp, depth, s, albedo, majorant, lpath, dl_drgb, grad=None, galpha=None,
            schedule=0, tape=None, lookups=None):
    """Adds dLoss/dDensity and dLoss/dAlbedo into grad and galpha (zeros when
    not given); lpath and tape (or None) from the forward of the same
    arguments."""
    ctx = dv.ctx
    dev = torch.device("cuda", ctx.device)
    nz, ny, nx = dv.vol.shape
    if grad is None:
        grad = torch.zeros((nz, ny, nx), dtype=torch.float32, device=dev)
    if galpha is None:
        galpha = torch.zeros(1, dtype=torch.float32, device=dev)
    dl = dl_drgb.contiguous()
    _check_adjoint(ctx, w, h, spp, depth, lpath, dl, grad, nz * ny * nx, galpha, tape)
    _check_counter(ctx, lookups)
    prop = dv.proposal.h if dv.proposal is not None else None
    ctx.check(ctx.lib.flg_volpath_adjoint(ctx.h, dv.vol.h, prop, ctypes.byref(scene),
                                          ctypes.byref(_cfg(w, h, spp, depth, schedule)),
                                          ctypes.byref(_params(s, albedo, dv.g, majorant, lookups)),
                                          _ptr(lpath), _ptr(tape), _ptr(dl), _ptr(grad), _ptr(galpha)))
    return grad, galpha


class BrickLayout:
    """Active bricks of a dense grid (flg_bricks_build): a brick of `brick`
    voxels per edge is active iff one of its voxels is non-zero.  A gradient
    in this layout is (n_active, brick, brick, brick), brick after brick in
    slot order."""

    def __init__(self, ctx, density, brick=8):
        check_buffer(density, "density", torch.float32, ctx.device, ndim=3)
        nz, ny, nx = density.shape
        self.ctx = ctx
        self.brick = brick
        self.shape = (nz, ny, nx)
        self.nb = tuple((n + brick - 1) // brick for n in (nz, ny, nx))
        self.slot = torch.empty(self.nb, dtype=torch.int32, device=torch.device("cuda", ctx.device))
        n = ctypes.c_int32(0)
        ctx.check(ctx.lib.flg_bricks_build(ctx.h, _ptr(density), nx, ny, nz, brick, _ptr(self.slot),
                                           ctypes.byref(n)))
        self.n_active = int(n.value)
        self.c = N.BrickLayout()
        self.c.brick = brick
        self.c.n_active = self.n_active
        self.c.nx, self.c.ny, self.c.nz = nx, ny, nz
        self.c.d_slot = self.slot.data_ptr()

    @classmethod
    def of_volume(cls, vol):
        """The bricks of a sparse volume (flg_volume_brick_layout): brick 8,
        slot = the brick's index in the volume's leaf list, so a gradient in
        this layout is shaped as the volume's brick values."""
        ctx = vol.ctx
        self = cls.__new__(cls)
        self.ctx = ctx
        self.brick = 8
        self.shape = tuple(vol.shape)
        self.nb = tuple((n + 7) // 8 for n in self.shape)
        self.c = N.BrickLayout()
        ctx.check(ctx.lib.flg_volume_brick_layout(vol.h, ctypes.byref(self.c)))
        self.n_active = int(self.c.n_active)
        # the same map on the torch side (to_dense), from the leaf origins
        o = vol.origins.long()
        self.slot = torch.full(self.nb, -1, dtype=torch.int32, device=torch.device("cuda", ctx.device))
        if self.n_active:
            self.slot[o[:, 2] // 8, o[:, 1] // 8, o[:, 0] // 8] = torch.arange(
                self.n_active, dtype=torch.int32, device=self.slot.device)
        self.vol = vol  # the C layout points into the volume
        return self

    def zeros(self):
        """A gradient buffer in this layout."""
        b = self.brick
        return torch.zeros((self.n_active, b, b, b), dtype=torch.float32, device=self.slot.device)

    def to_dense(self, bricks):
        """Dense (nz, ny, nx) copy of a gradient in this layout (0 on inactive
        bricks)."""
        b = self.brick
        nbz, nby, nbx = self.nb
        full = torch.zeros((nbz, nby, nbx, b, b, b), dtype=bricks.dtype, device=bricks.device)
        act = self.slot >= 0
        full[act] = bricks[self.slot[act].long()]
        dense = full.permute(0, 3, 1, 4, 2, 5).reshape(nbz * b, nby * b, nbx * b)
        nz, ny, nx = self.shape
        return dense[:nz, :ny, :nx].contiguous()


def mesh_forward(dv, bvh, scene, w, h, spp, depth, s, albedo, majorant, out=None, lpath=None, schedule=0,
                 tape=None):
    """The differentiable estimator over `bvh`'s closed mesh as the interface
    around the medium (flg_volpath_mesh_diff_forward): image and per-path
    radiance; with `tape` also the walk tape."""
    ctx = dv.ctx
    dev = torch.device("cuda", ctx.device)
    if out is None:
        out = torch.empty((h, w, 3), dtype=torch.float32, device=dev)
    if lpath is None:
        lpath = torch.empty((w * h * spp, 3), dtype=torch.float32, device=dev)
    _check_forward(ctx, w, h, spp, depth, out, lpath, tape)
    prop = dv.proposal.h if dv.proposal is not None else None
    ctx.check(ctx.lib.flg_volpath_mesh_diff_forward(ctx.h, bvh.h, dv.vol.h, prop, ctypes.byref(scene),
                                                    ctypes.byref(_cfg(w, h, spp, depth, schedule)),
                                                    ctypes.byref(_params(s, albedo, dv.g, majorant)),
                                                    _ptr(out), _ptr(lpath), _ptr(tape)))
    return out, lpath


def mesh_adjoint(dv, bvh, scene, w, h, spp, depth, s, albedo, majorant, lpath, dl_drgb, grad=None, galpha=None,
                 schedule=0, tape=None, bricks=None):
    """Adjoint of mesh_forward (flg_volpath_mesh_adjoint): adds dLoss/dDensity
    into grad, dense or, with `bricks` (a BrickLayout), over its active
    bricks, and dLoss/dAlbedo into galpha."""
    ctx = dv.ctx
    dev = torch.device("cuda", ctx.device)
    if grad is None:
        grad = bricks.zeros() if bricks is not None else torch.zeros(dv.vol.shape, dtype=torch.float32, device=dev)
    if galpha is None:
        galpha = torch.zeros(1, dtype=torch.float32, device=dev)
    dl = dl_drgb.contiguous()
    nz, ny, nx = dv.vol.shape
    gnum = bricks.n_active * bricks.brick ** 3 if bricks is not None else nz * ny * nx
    _check_adjoint(ctx, w, h, spp, depth, lpath, dl, grad, gnum, galpha, tape)
    prop = dv.proposal.h if dv.proposal is not None else None
    ctx.check(ctx.lib.flg_volpath_mesh_adjoint(ctx.h, bvh.h, dv.vol.h, prop, ctypes.byref(scene),
                                               ctypes.byref(_cfg(w, h, spp, depth, schedule)),
                                               ctypes.byref(_params(s, albedo, dv.g, majorant)), _ptr(lpath),
                                               _ptr(tape), _ptr(dl), ctypes.byref(bricks.c) if bricks else None,
                                               _ptr(grad), _ptr(galpha)))
    return grad, galpha


class VolumeRender(torch.autograd.Function):
    """image = VolumeRender.apply(density, albedo, dv, scene, w, h, spp, depth, s, majorant)

    density: (nz, ny, nx) float32 CUDA tensor (the grid evaluated and used as
    the proposal); albedo: 0-d float32 tensor."""

    @staticmethod
    def forward(ctx_, density, albedo, dv, scene, w, h, spp, depth, s, majorant):
        dv.set_density(density)
        a = float(albedo.item())
        tape = new_tape(dv.ctx, w, h, spp, depth) if depth <= 63 else None
        img, lpath = forward(dv, scene, w, h, spp, depth, s, a, majorant, tape=tape)
        ctx_.save_for_backward(lpath)
        ctx_.tape = tape
        ctx_.args = (dv, scene, w, h, spp, depth, s, a, majorant)
        return img

    @staticmethod
    def backward(ctx_, grad_img):
        (lpath,) = ctx_.saved_tensors
        dv, scene, w, h, spp, depth, s, a, majorant = ctx_.args
        grad, galpha = adjoint(dv, scene, w, h, spp, depth, s, a, majorant, lpath,
                               grad_img.float().contiguous(), tape=ctx_.tape)
        return grad, galpha.reshape(()), None, None, None, None, None, None, None, None


def bench_adjoint(ctx, timed, sync, state="uniform", profile=False):
    """128^3 grid, 256x256, 16 spp, depth 32, L2 loss against a target
    rendered from an fBm grid (seed 2); point light and environment as in the
    256^3 volume render, albedo 0.8, g 0.3, sigma_t = 8 density.  The
    optimised grid theta is `state`: "uniform" (0.5 everywhere, the start of
    the optimisation: every ratio-tracking factor 1 - s theta / mu is 0 and
    walks end at once) or "fbm" (a different fBm grid, seed 3: a non-uniform
    state with walks of ordinary length).  Returns timings of the plain
    forward, the differentiable forward and the adjoint at both schedules,
    one fwd+adjoint iteration, the library's kernel times and the grid
    lookups per path of each pass.  With `profile`, runs one plain forward and
    one iteration (forward with the walk tape, then the adjoint) at the
    regenerating schedule and returns None (for hardware-counter runs)."""
    from . import api

    dev = torch.device("cuda", ctx.device)
    W = H = 256
    SPP, DEPTH, S = 16, 32, 8.0
    cam = N.Camera.make((0.5, 0.5, -1.5), (0.5, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0)
    scene = api.vol_scene(cam)
    target_grid = torch.from_numpy(api.gen_fbm_grid(128, seed=2)).to(dev)
    tv = DiffVolume(ctx, target_grid)
    target, _ = forward(tv, scene, W, H, SPP, DEPTH, S, 0.8, S * float(target_grid.max()))
    if state == "uniform":
        theta = torch.full((128, 128, 128), 0.5, dtype=torch.float32, device=dev)
    else:
        theta = torch.from_numpy(api.gen_fbm_grid(128, seed=3)).to(dev)
    dv = DiffVolume(ctx, theta)
    mu = S * float(theta.max())
    plain = Volume(ctx, theta, (0, 0, 0, 1, 1, 1), 0.8 * S, 0.2 * S, 0.3)
    img_plain = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    lpath = torch.empty((W * H * SPP, 3), dtype=torch.float32, device=dev)
    img = torch.empty((H, W, 3), dtype=torch.float32, device=dev)
    grad = torch.zeros((128, 128, 128), dtype=torch.float32, device=dev)
    galpha = torch.zeros(1, dtype=torch.float32, device=dev)
    tape = new_tape(ctx, W, H, SPP, DEPTH)
    if profile:
        plain.render(scene, W, H, SPP, DEPTH, out=img_plain, schedule=2)
        forward(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, out=img, lpath=lpath, schedule=2, tape=tape)
        adjoint(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, lpath, 2.0 * (img - target) / img.numel(), grad, galpha,
                schedule=2, tape=tape)
        sync()
        return None

    # every pass at both schedules (1: one lane per path, 2: persistent lanes
    # with path regeneration); the ratio compares the fastest of each
    t_plain = {sc: timed(lambda: plain.render(scene, W, H, SPP, DEPTH, out=img_plain, schedule=sc), 5, 2, sync)
               for sc in (1, 2)}
    t_fwd = {sc: timed(lambda: forward(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, out=img, lpath=lpath,
                                       schedule=sc, tape=tape), 5, 2, sync) for sc in (1, 2)}
    dl = 2.0 * (img - target) / img.numel()

    def adj(sc, tp=tape, lookups=None):
        grad.zero_()
        galpha.zero_()
        adjoint(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, lpath, dl, grad, galpha, schedule=sc, tape=tp,
                lookups=lookups)

    t_adj = {sc: timed(lambda: adj(sc), 5, 2, sync) for sc in (1, 2)}
    t_adj_notape = {sc: timed(lambda: adj(sc, None), 5, 2, sync) for sc in (1, 2)}
    s_plain = min(t_plain, key=t_plain.get)
    s_fwd = min(t_fwd, key=t_fwd.get)
    s_adj = min(t_adj, key=t_adj.get)

    def iteration():
        forward(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, out=img, lpath=lpath, schedule=s_fwd, tape=tape)
        d = 2.0 * (img - target) / img.numel()
        grad.zero_()
        galpha.zero_()
        adjoint(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, lpath, d, grad, galpha, schedule=s_adj, tape=tape)

    t_iter = timed(iteration, 5, 2, sync)
    # kernel times of one iteration (HIP events on the library's stream)
    ctx.profiling(True)
    ctx.reset_profiling()
    for _ in range(3):
        iteration()
    sync()
    k_fwd, n_fwd = ctx.kernel_time("diff_forward")
    k_adj, n_adj = ctx.kernel_time("diff_adjoint")
    ctx.profiling(False)
    # grid lookups per path of each pass (a separate counting pass)
    cnt = torch.zeros(2, dtype=torch.int64, device=dev)
    forward(dv, scene, W, H, SPP, DEPTH, S, 0.8, mu, out=img, lpath=lpath, schedule=s_fwd, tape=tape,
            lookups=cnt[0:1])
    adj(s_adj, tape, cnt[1:2])
    sync()
    P = W * H * SPP
    return {"state": state, "config": "128^3 grid, 256x256, 16 spp, depth 32",
            "plain_forward_ms": t_plain[s_plain] * 1e3,
            "plain_forward_ms_by_schedule": {str(k): v * 1e3 for k, v in t_plain.items()},
            "diff_forward_ms": t_fwd[s_fwd] * 1e3,
            "diff_forward_ms_by_schedule": {str(k): v * 1e3 for k, v in t_fwd.items()},
            "adjoint_ms": t_adj[s_adj] * 1e3,
            "adjoint_ms_by_schedule": {str(k): v * 1e3 for k, v in t_adj.items()},
            "adjoint_without_tape_ms_by_schedule": {str(k): v * 1e3 for k, v in t_adj_notape.items()},
            "walk_tape_bytes": tape.numel(),
            "fwd_adjoint_ms_per_iter": t_iter * 1e3,
            "fwd_adjoint_vs_plain": t_iter / t_plain[s_plain],
            "kernel_forward_ms": k_fwd / max(n_fwd, 1), "kernel_adjoint_ms": k_adj / max(n_adj, 1),
            "paths": P, "lookups_forward": int(cnt[0]), "lookups_adjoint": int(cnt[1]),
            "majorant": mu}
