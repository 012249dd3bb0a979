"""The scene accel over several meshes on the GPU against the oracle's
restatement of Scene::intersect (NaiveBVHAccel over the meshes, each mesh's
LBVH underneath): mesh, triangle and t bit for bit, with ties between meshes
decided by the reference's order, up to 64 meshes and beyond."""
import numpy as np
import pytest
import torch

import oracle_lib as orc
from fledge_amd import api
from scene_cases import reference_rank, scene_meshes, scene_rays

pytestmark = pytest.mark.gpu


def _case(ctx, meshes, rays):
    bvhs = api.LBVH.build_batched(ctx, [(torch.from_numpy(v.ravel()).cuda(), torch.from_numpy(i.ravel()).cuda())
                                        for v, i in meshes])
    sa = api.SceneAccel(ctx, bvhs)
    hits = sa.trace(torch.from_numpy(rays).cuda()).cpu().numpy()
    mesh, ids, ts = orc.scene_trace([orc.Lbvh(v, i) for v, i in meshes], rays)
    return sa, hits, mesh, ids, ts


@pytest.mark.parametrize("n,dup,extra", [(1, [], []), (40, [(5, 2), (17, 2), (30, 11)], []),
                                         (150, [], [(7, 3), (90, 40), (140, 41)])])
def test_scene_trace_matches_oracle(ctx, n, dup, extra):
    print("scene accel %d meshes" % n, flush=True)
    meshes = scene_meshes(n, 3, dup, extra)
    rays = np.concatenate([scene_rays(20000, 4), scene_rays(5000, 5, tmax=6.0)])
    sa, hits, mesh, ids, ts = _case(ctx, meshes, rays)
    assert np.array_equal(hits[:, 0], mesh)
    assert np.array_equal(hits[:, 1], ids)
    assert np.array_equal(hits[:, 2], ts.view(np.int32))
    assert (mesh >= 0).mean() > (0.01 if n == 1 else 0.3)
    assert sa.rank() == list(reference_rank(meshes))


def test_scene_accel_rejects_mixed_normal_policies(ctx):
    meshes = scene_meshes(2, 1)
    a = api.LBVH.build(ctx, torch.from_numpy(meshes[0][0].ravel()).cuda(),
                       torch.from_numpy(meshes[0][1].ravel()).cuda())
    b = api.LBVH.build(ctx, torch.from_numpy(meshes[1][0].ravel()).cuda(),
                       torch.from_numpy(meshes[1][1].ravel()).cuda(), compat=True)
    with pytest.raises(Exception):
        api.SceneAccel(ctx, [a, b])


def test_scene_trace_edge_cases(ctx):
    """Single-triangle meshes (including one on the plane of another), rays
    with tmax 0 and rays along the triangles' plane, and an empty batch."""
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    meshes = [(tri + np.float32(k) * np.array([0.3, 0.2, 0.0], np.float32), np.arange(3, dtype=np.int32))
              for k in range(5)]
    meshes.append((tri.copy(), np.arange(3, dtype=np.int32)))  # the first mesh again: ties with it
    rng = np.random.default_rng(8)
    n = 4000
    o = np.concatenate([rng.uniform(-0.5, 2.0, (n, 2)), np.full((n, 1), -1.0)], axis=1).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = np.inf
    rays[:, 4:7] = np.array([0, 0, 1], np.float32)
    rays[: n // 8, 3] = 0.0                              # tmax 0: nothing is closer
    rays[n // 8: n // 4, 2] = 0.0                        # in the plane of the triangles
    rays[n // 8: n // 4, 4:7] = np.array([1, 0, 0], np.float32)
    sa, hits, mesh, ids, ts = _case(ctx, meshes, rays)
    assert np.array_equal(hits[:, 0], mesh) and np.array_equal(hits[:, 1], ids)
    assert np.array_equal(hits[:, 2], ts.view(np.int32))
    assert np.all(mesh[: n // 8] == -1)
    assert not np.any(mesh == 5)                         # the repeated mesh loses every tie
    empty = sa.trace(torch.zeros((0, 8), dtype=torch.float32, device="cuda"))
    assert empty.shape == (0, 4)
