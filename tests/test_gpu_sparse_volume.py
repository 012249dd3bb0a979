"""Sparse-occupancy volumes (flg_volume_create_sparse: 8^3 bricks in a pool
with a one-voxel apron and a slot map) against the dense volume holding the
same values, on the GPU: the volume-only integrator, the integrator over an
interface mesh, the differentiable forwards and the adjoints give the same
bits (gradients: the same terms, float atomics reorder the sums), at ragged
grid sizes, with no brick, with every brick, after updates; the origins it
refuses.  The reference reads its sparse grid through an OpenVDB accessor
(src/volume.cpp:40-59,76-81); a sparse grid here is defined to equal the dense
grid of the same values, which the dense path's own parity tests pin."""
import numpy as np
import pytest
import torch

from fledge_amd import _native as N
from fledge_amd import api, diff

pytestmark = pytest.mark.gpu

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]
S, ALBEDO, G = 8.0, 0.8, 0.3


def _masked_grid(shape, p, seed):
    """fBm (seeded) on a (nz, ny, nx) grid, zero outside the 8^3 bricks that a
    coin of probability p keeps."""
    nz, ny, nx = shape
    n = max(shape)
    g = api.gen_fbm_grid(n, seed=seed)[:nz, :ny, :nx].copy()
    bz, by, bx = (-(-s // 8) for s in shape)
    keep = np.random.default_rng(seed).random((bz, by, bx)) < p
    m = np.repeat(np.repeat(np.repeat(keep, 8, 0), 8, 1), 8, 2)[:nz, :ny, :nx]
    return (g * m).astype(np.float32)


def _pair(ctx, grid, sigma=(6.4, 1.6)):
    d = torch.from_numpy(grid).cuda()
    dense = api.Volume(ctx, d, [0, 0, 0, 1, 1, 1], sigma[0], sigma[1], G)
    o, v = api.leaves_of(d)
    sparse = api.Volume.sparse(ctx, o, v, grid.shape, [0, 0, 0, 1, 1, 1], sigma[0], sigma[1], G)
    return dense, sparse, o, v


def _scene():
    cam = N.Camera.make(CAM0[0:3], CAM0[3:6], CAM0[6:9], CAM0[9])
    return api.vol_scene(cam)


def _bits(t):
    return t.contiguous().view(torch.int32).cpu().numpy()


def test_leaves_of_round_trip(ctx):
    grid = _masked_grid((44, 52, 60), 0.4, 5)
    o, v = api.leaves_of(torch.from_numpy(grid).cuda())
    back = np.zeros((48, 56, 64), np.float32)
    for (x, y, z), b in zip(o.cpu().numpy(), v.cpu().numpy()):
        back[z:z + 8, y:y + 8, x:x + 8] = b
    assert np.array_equal(back[:44, :52, :60], grid)
    assert 0 < o.shape[0] < 6 * 7 * 8


@pytest.mark.parametrize("shape,p", [((44, 52, 60), 0.4), ((24, 24, 24), 1.0), ((17, 9, 33), 0.0)])
@pytest.mark.parametrize("schedule,film", [(1, 0), (2, 0), (2, 1)])
def test_sparse_render_bit_equal(ctx, shape, p, schedule, film):
    print("sparse render", shape, p, schedule, film, flush=True)
    grid = _masked_grid(shape, p, 11)
    dense, sparse, o, _ = _pair(ctx, grid)
    assert o.shape[0] == 0 if p == 0.0 else o.shape[0] > 0
    sc = _scene()
    a, sa = dense.render(sc, 48, 40, 8, 16, stats=True, schedule=schedule, film=film)
    b, sb = sparse.render(sc, 48, 40, 8, 16, stats=True, schedule=schedule, film=film)
    assert np.array_equal(_bits(a), _bits(b))
    assert sa == sb
    if p > 0:
        assert sa["lookups"] > 0


def test_sparse_update_and_noalloc(ctx):
    grid = _masked_grid((40, 40, 40), 0.5, 2)
    dense, sparse, o, v = _pair(ctx, grid)
    sc = _scene()
    dense.render(sc, 32, 32, 4, 8)
    sparse.render(sc, 32, 32, 4, 8)  # builds the environment table for noalloc
    sparse.update_sparse((v * 0.5).contiguous())
    dense.update(torch.from_numpy(grid * np.float32(0.5)).cuda())
    a = dense.render(sc, 32, 32, 4, 8)
    b = sparse.render(sc, 32, 32, 4, 8)
    assert np.array_equal(_bits(a), _bits(b))
    cfg = N.RenderCfg()
    cfg.width, cfg.height, cfg.spp, cfg.max_depth = 32, 32, 4, 8
    c = torch.empty((32, 32, 3), dtype=torch.float32, device="cuda")
    ctx.check(ctx.lib.flg_volpath_render_noalloc(ctx.h, sparse.h, sc, cfg, api._ptr(c)))
    torch.cuda.synchronize()
    assert np.array_equal(_bits(a), _bits(c))


def test_sparse_refuses_bad_origins(ctx):
    v = torch.ones((2, 8, 8, 8), dtype=torch.float32, device="cuda")
    for bad in ([[0, 0, 0], [0, 0, 0]], [[0, 0, 4], [8, 0, 0]], [[0, 0, 0], [40, 0, 0]], [[-8, 0, 0], [0, 0, 0]]):
        o = torch.tensor(bad, dtype=torch.int32, device="cuda")
        with pytest.raises(RuntimeError):
            api.Volume.sparse(ctx, o, v, (33, 33, 33), [0, 0, 0, 1, 1, 1], 1.0, 0.0, 0.0)
    dense = api.Volume(ctx, torch.zeros((8, 8, 8), device="cuda"), [0, 0, 0, 1, 1, 1], 1.0, 0.0, 0.0)
    with pytest.raises(RuntimeError):
        ctx.check(ctx.lib.flg_volume_update_sparse(ctx.h, dense.h, api._ptr(v)))
    c = N.BrickLayout()
    assert ctx.lib.flg_volume_brick_layout(dense.h, c) != 0


def _mesh(ctx, level=4):
    vv, ind = api.gen_octasphere(level, 0.05, seed=7, base_freq=2.0)
    vv = (vv * 0.5 + 0.5).astype(np.float32)
    return api.LBVH.build(ctx, torch.from_numpy(vv).cuda(), torch.from_numpy(ind).cuda())


@pytest.mark.parametrize("schedule", [1, 2])
def test_sparse_mesh_render_bit_equal(ctx, schedule):
    print("sparse mesh render", schedule, flush=True)
    bvh = _mesh(ctx)
    grid = _masked_grid((40, 48, 56), 0.4, 8)
    dense, sparse, _, _ = _pair(ctx, grid)
    sc = _scene()
    a, sa = dense.render_mesh(bvh, sc, 48, 40, 8, 16, stats=True, schedule=schedule)
    b, sb = sparse.render_mesh(bvh, sc, 48, 40, 8, 16, stats=True, schedule=schedule)
    assert np.array_equal(_bits(a), _bits(b))
    assert sa == sb and sa["rays"] >= 48 * 40 * 8
    assert sa["node_visits"] >= sa["rays"] and sa["tri_tests"] > 0


def test_sparse_differentiable_paths(ctx):
    """Forward bits equal (volume-only and over the mesh); the adjoints'
    gradients equal up to the order of the atomic sums, dense and over the
    bricks (the sparse volume's own layout against flg_bricks_build's)."""
    grid = (_masked_grid((32, 32, 32), 0.4, 3) * 0.9).astype(np.float32)
    gt = torch.from_numpy(grid).cuda()
    o, v = api.leaves_of(gt)
    dd = diff.DiffVolume(ctx, gt, g=G)
    ds = diff.DiffVolume.sparse(ctx, o, v, grid.shape, g=G)
    sc = _scene()
    W, H, spp, D = 40, 32, 4, 16
    mu = S * float(grid.max())
    ia, la = diff.forward(dd, sc, W, H, spp, D, S, ALBEDO, mu)
    ib, lb = diff.forward(ds, sc, W, H, spp, D, S, ALBEDO, mu)
    assert np.array_equal(_bits(la), _bits(lb)) and np.array_equal(_bits(ia), _bits(ib))
    dl = torch.full((H, W, 3), 1.0 / (W * H * 3), dtype=torch.float32, device="cuda")
    ga, gaa = diff.adjoint(dd, sc, W, H, spp, D, S, ALBEDO, mu, la, dl)
    gb, gab = diff.adjoint(ds, sc, W, H, spp, D, S, ALBEDO, mu, lb, dl)
    scale = float(ga.abs().max())
    assert scale > 0 and float((ga - gb).abs().max()) <= 1e-5 * scale
    assert abs(float(gaa) - float(gab)) <= 1e-5 * abs(float(gaa))

    bvh = _mesh(ctx)
    ma, mla = diff.mesh_forward(dd, bvh, sc, W, H, spp, D, S, ALBEDO, mu)
    mb, mlb = diff.mesh_forward(ds, bvh, sc, W, H, spp, D, S, ALBEDO, mu)
    assert np.array_equal(_bits(mla), _bits(mlb)) and np.array_equal(_bits(ma), _bits(mb))
    la_dense = diff.BrickLayout(ctx, gt, 8)
    la_sparse = diff.BrickLayout.of_volume(ds.vol)
    assert la_sparse.n_active == la_dense.n_active == o.shape[0]
    assert torch.equal(la_sparse.slot, la_dense.slot)
    g1, a1 = diff.mesh_adjoint(dd, bvh, sc, W, H, spp, D, S, ALBEDO, mu, mla, dl, bricks=la_dense)
    g2, a2 = diff.mesh_adjoint(ds, bvh, sc, W, H, spp, D, S, ALBEDO, mu, mlb, dl, bricks=la_sparse)
    assert tuple(g2.shape) == tuple(v.shape)
    scale = float(g1.abs().max())
    assert scale > 0 and float((g1 - g2).abs().max()) <= 1e-5 * scale
    assert abs(float(a1) - float(a2)) <= 1e-5 * abs(float(a1))
    # new brick values reach the differentiable path
    ds.set_density((v * 0.5).contiguous())
    dd.set_density(gt * 0.5)
    _, l1 = diff.forward(dd, sc, W, H, spp, D, S, ALBEDO, mu)
    _, l2 = diff.forward(ds, sc, W, H, spp, D, S, ALBEDO, mu)
    assert np.array_equal(_bits(l1), _bits(l2))
