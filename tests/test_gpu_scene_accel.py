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
