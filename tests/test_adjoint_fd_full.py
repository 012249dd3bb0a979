"""The FP64 twin's analytic gradient against central finite differences at the
differentiable-rendering configuration's full size: 128^3 grid at a
non-uniform fBm state, 256x256, 16 spp, depth 32, L2 loss against the image
of another fBm grid.  64 voxels with a non-negligible gradient and the
albedo, eps 1e-3, within 1e-3 relative.  To keep it to seconds on the CPU the
loss sums every 16th pixel (4,096 of 65,536, all 16 spp each); the gradient
is of that same loss.  The GPU gradient is compared with this twin at the
full size in test_gpu_adjoint.py."""
import numpy as np

import oracle_lib as orc
from fledge_amd import api

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]
LIGHTS = [1.0, 2.0, 2.0, -2.0, 50.0, 1.0, 0.5]
S, ALBEDO, G = 8.0, 0.8, 0.3


def test_twin_gradient_central_differences_full_size():
    W = H = 256
    spp, depth, stride = 16, 32, 16
    theta = api.gen_fbm_grid(128, seed=3)
    target = api.gen_fbm_grid(128, seed=2)
    mu = S * float(theta.max())
    _, timg = orc.diff_render_strided(target.astype(np.float64), target, [S, ALBEDO, G, S * float(target.max())],
                                      LIGHTS, CAM0, W, H, spp, depth, stride)
    th = theta.astype(np.float64)
    lp, img = orc.diff_render_strided(th, theta, [S, ALBEDO, G, mu], LIGHTS, CAM0, W, H, spp, depth, stride)
    sel = np.zeros(W * H, bool)
    sel[::stride] = True
    sel = sel.reshape(H, W)
    assert img[sel].min() > 0 and (img[~sel] == 0).all()
    dl = np.where(sel, 2.0 * (img - timg) / sel.sum(), 0.0)
    g, ga = orc.diff_grad_strided(th, theta, [S, ALBEDO, G, mu], LIGHTS, CAM0, W, H, spp, depth, stride, lp, dl)

    def F(t, alb=ALBEDO):
        _, im = orc.diff_render_strided(t, theta, [S, alb, G, mu], LIGHTS, CAM0, W, H, spp, depth, stride)
        return float((dl * im)[sel].sum())

    gmax = np.abs(g).max()
    assert gmax > 0
    cand = np.flatnonzero(np.abs(g.ravel()) > 0.05 * gmax)
    pick = np.random.default_rng(0).choice(cand, size=min(64, cand.size), replace=False)
    assert pick.size == 64
    eps = 1e-3
    for v in pick:
        tp = th.copy()
        tp.flat[v] += eps
        tm = th.copy()
        tm.flat[v] -= eps
        fd = (F(tp) - F(tm)) / (2 * eps)
        assert abs(g.flat[v] - fd) <= 1e-3 * abs(fd) + 1e-6 * gmax, (int(v), float(g.flat[v]), fd)
    fa = (F(th, ALBEDO + 1e-3) - F(th, ALBEDO - 1e-3)) / 2e-3
    assert abs(ga - fa) <= 1e-3 * abs(fa), (ga, fa)
