"""The Python wrappers refuse buffers the kernels would index out of bounds:
wrong dtype, strided views, short outputs, wrong shapes and tensors off the
context's GPU.  Every check runs before the C ABI is called, so these run on
the CPU with a context stand-in whose library is never reached."""
import types

import pytest
import torch

from fledge_amd import api, diff


def _ctx():
    return types.SimpleNamespace(device=0, lib=None, h=None)


def _obj(cls, **kw):
    o = cls.__new__(cls)
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _raises(fn, match):
    with pytest.raises(ValueError, match=match):
        fn()


def test_check_buffer_order_and_messages():
    t = torch.zeros((4, 8), dtype=torch.float32)
    _raises(lambda: api.check_buffer(t.double(), "x", torch.float32, 0), "dtype")
    _raises(lambda: api.check_buffer(t.t(), "x", torch.float32, 0), "contiguous")
    _raises(lambda: api.check_buffer(t, "x", torch.float32, 0, shape=(4, 7)), "shape")
    _raises(lambda: api.check_buffer(t, "x", torch.float32, 0, numel=31), "elements")
    _raises(lambda: api.check_buffer(t, "x", torch.float32, 0, min_numel=33), "at least")
    _raises(lambda: api.check_buffer(t, "x", torch.float32, 0, ndim=3), "dimensions")
    _raises(lambda: api.check_buffer(t, "x", torch.float32, 0), "expected cuda:0")
    _raises(lambda: api.check_buffer(t.numpy(), "x", torch.float32, 0), "torch tensor")


def test_lbvh_wrappers_refuse_bad_buffers():
    ctx = _ctx()
    b = _obj(api.LBVH, ctx=ctx, h=None, ntris=4)
    rays = torch.zeros((16, 8), dtype=torch.float32)
    _raises(lambda: b.trace(rays.double()), "rays: dtype")
    _raises(lambda: b.trace(torch.zeros((16, 16))[:, ::2]), "rays: must be contiguous")
    _raises(lambda: b.trace(torch.zeros((16, 7))), "rays: shape")
    _raises(lambda: b.trace(torch.zeros(128)), r"rays: shape \(128,\), expected \(.n., 8\)")
    _raises(lambda: b.trace(rays, torch.zeros((15, 2), dtype=torch.int32)), "hits: 30 elements, at least 32")
    _raises(lambda: b.trace(rays, torch.zeros((16, 2), dtype=torch.int64)), "hits: dtype")
    _raises(lambda: b.trace(rays), "expected cuda:0")  # right shape, host memory
    _raises(lambda: b.count(rays.half()), "rays: dtype")
    _raises(lambda: b.autotune_trace(rays, torch.zeros(8, dtype=torch.int32)), "hits")
    out = (torch.zeros(10, dtype=torch.int32), torch.zeros((12, 3), dtype=torch.int32),
           torch.zeros((12, 3), dtype=torch.float32))
    cam = None
    _raises(lambda: b.pathtrace(cam, 2, 2, 3, 2, out=out), "nseg: 10 elements, at least 12")
    out = (torch.zeros(12, dtype=torch.int32), torch.zeros((12, 2), dtype=torch.int32),
           torch.zeros((12, 3), dtype=torch.float32))
    _raises(lambda: b.pathtrace(cam, 2, 2, 3, 2, out=out), "seg_id")
    out = (torch.zeros(12, dtype=torch.int32), torch.zeros((12, 3), dtype=torch.int32),
           torch.zeros((12, 3), dtype=torch.float64))
    _raises(lambda: b.pathtrace(cam, 2, 2, 3, 2, out=out), "seg_t: dtype")


def test_build_refuses_bad_meshes():
    ctx = _ctx()
    v = torch.zeros(9, dtype=torch.float32)
    i = torch.arange(3, dtype=torch.int32)
    _raises(lambda: api.LBVH.build(ctx, v.double(), i), "verts: dtype")
    _raises(lambda: api.LBVH.build(ctx, v, i.long()), "ind: dtype")
    _raises(lambda: api.LBVH.build(ctx, torch.zeros(18)[::2], i), "verts: must be contiguous")
    _raises(lambda: api.LBVH.build(ctx, v[:8], i), "multiples of 3")
    _raises(lambda: api.LBVH.build(ctx, v, i), "verts: on cpu, expected cuda:0")
    _raises(lambda: api.LBVH.build_batched(ctx, []), "no meshes")
    _raises(lambda: api.LBVH.build_batched(ctx, [(v, i), (v, i.double())]), r"ind\[1\]: dtype")


def test_index_range_check():
    v = torch.zeros(9, dtype=torch.float32)
    api.LBVH._check_indices([(v, torch.tensor([0, 1, 2], dtype=torch.int32))])
    _raises(lambda: api.LBVH._check_indices([(v, torch.tensor([0, 1, 3], dtype=torch.int32))]), "outside")
    _raises(lambda: api.LBVH._check_indices([(v, torch.tensor([-1, 1, 2], dtype=torch.int32))]), "outside")


def test_scene_camera_and_volume_wrappers_refuse_bad_buffers():
    ctx = _ctx()
    sa = _obj(api.SceneAccel, ctx=ctx, h=None)
    rays = torch.zeros((8, 8), dtype=torch.float32)
    _raises(lambda: sa.trace(rays, torch.zeros((8, 2), dtype=torch.int32)), "hits: 16 elements, at least 32")
    _raises(lambda: api.camera_rays(ctx, None, 4, 4, out=torch.zeros((15, 8))), "out: 120 elements")
    _raises(lambda: api.Volume(ctx, torch.zeros((4, 4, 4), dtype=torch.float64), [0] * 6, 1, 0, 0), "density: dtype")
    _raises(lambda: api.Volume(ctx, torch.zeros((4, 16)), [0] * 6, 1, 0, 0), "density: 2 dimensions")
    _raises(lambda: api.Volume(ctx, torch.zeros((4, 4, 8))[:, :, ::2], [0] * 6, 1, 0, 0), "contiguous")
    vol = _obj(api.Volume, ctx=ctx, h=None, shape=(4, 4, 4))
    _raises(lambda: vol.update(torch.zeros((4, 4, 5))), "density: shape")
    _raises(lambda: vol.update(torch.zeros((4, 4, 4), dtype=torch.float16)), "density: dtype")
    _raises(lambda: vol.render(None, 8, 8, 1, 4, out=torch.zeros((8, 8, 2))), "out: 128 elements, at least 192")
    _raises(lambda: vol.render_mesh(None, None, 8, 8, 1, 4, out=torch.zeros((8, 8, 3)).double()), "out: dtype")


def test_differentiable_wrappers_refuse_bad_buffers(monkeypatch):
    ctx = _ctx()
    monkeypatch.setattr(diff, "tape_bytes", lambda c, w, h, spp, depth: w * h * spp * 8)
    vol = _obj(api.Volume, ctx=ctx, h=None, shape=(4, 4, 4))
    dv = _obj(diff.DiffVolume, ctx=ctx, vol=vol, proposal=None, g=0.3, bounds=(0, 0, 0, 1, 1, 1))
    W, H, spp = 4, 2, 3
    img = torch.zeros((H, W, 3))
    lp = torch.zeros((W * H * spp, 3))
    _raises(lambda: diff.forward(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, out=img, lpath=lp[:-1]), "lpath")
    _raises(lambda: diff.forward(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, out=img[:, :2], lpath=lp), "out")
    _raises(lambda: diff.forward(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, out=img, lpath=lp,
                                 tape=torch.zeros(10, dtype=torch.uint8)), "tape")
    _raises(lambda: diff.forward(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, out=img, lpath=lp,
                                 tape=torch.zeros(W * H * spp * 8, dtype=torch.int8)), "tape: dtype")
    g = torch.zeros((4, 4, 4))
    ga = torch.zeros(1)
    _raises(lambda: diff.adjoint(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, lp, img, g[:3], ga), "grad")
    _raises(lambda: diff.adjoint(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, lp, img[:1], g, ga), "dl_drgb")
    _raises(lambda: diff.adjoint(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, lp.double(), img, g, ga), "lpath")
    _raises(lambda: diff.adjoint(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, lp, img, g, ga.long()), "galpha")
    _raises(lambda: diff.adjoint(dv, None, W, H, spp, 4, 8.0, 0.8, 4.0, lp, img, g, ga), "expected cuda:0")
    bricks = types.SimpleNamespace(n_active=2, brick=2, c=None)
    _raises(lambda: diff.mesh_adjoint(dv, None, None, W, H, spp, 4, 8.0, 0.8, 4.0, lp, img, g, ga, bricks=bricks),
            "grad: 64 elements, expected 16")
    _raises(lambda: diff.mesh_forward(dv, None, None, W, H, spp, 4, 8.0, 0.8, 4.0, out=img, lpath=lp[:5]), "lpath")
    _raises(lambda: diff.BrickLayout(ctx, torch.zeros((4, 4)), 2), "density: 2 dimensions")
