"""Writes the golden fixtures of tests/golden/ from the CPU oracle (whose
functions the KATs, closed forms and brute-force checks of
tests/test_oracle_cpu.py pin).  Each fixture holds its inputs and expected
outputs, so the GPU parity tests need neither the generators nor the oracle to
reproduce a case.

    python tests/golden/make_golden.py      (needs oracle/build/liboracle.so)

Fixtures:
  lbvh_soup1024.npz   1024-triangle soup (PCG32 42/54): codes, order, topology,
                      node bounds; 64x64 jittered primary rays with hit id and t
  halton_res128.npz   get1D values of 4 samples x 24 dims for 5 pixels
                      (HaltonSampler at 128x128, 16 spp)
  volpath_fbm24.npz   24^3 fBm grid, 16x16, 4 spp, depth 32: per-pixel radiance
  pathtrace_octa3.npz 512-triangle octahedral sphere, 24x24, 2 spp, depth 4:
                      per-path segment ids and t
  adjoint_fbm12.npz   12^3 grid, 12x12, 2 spp: FP64 forward per path and the
                      FP64 gradient of sum(w * image)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as orc  # noqa: E402

CAM0 = [0.5, 0.5, -1.5, 0.5, 0.5, 0.0, 0.0, 1.0, 0.0, 45.0]
LIGHTS = [1.0, 2.0, 2.0, -2.0, 50.0, 1.0, 0.5]


def soup(ntris, initstate=42, initseq=54):
    verts = np.zeros(ntris * 9, np.float32)
    ind = np.zeros(ntris * 3, np.int32)
    orc.lib().orc_gen_soup(initstate, initseq, ntris, verts.ctypes.data, ind.ctypes.data)
    return verts.reshape(-1, 3), ind.reshape(-1, 3)


def lbvh_fixture():
    v, i = soup(1024)
    b = orc.Lbvh(v, i)
    e = b.export()
    rays = orc.camera_rays(CAM0, 64, 64)
    ids, ts = b.trace(rays)
    return dict(verts=v, ind=i, rays=rays, hit_id=ids, hit_t=ts, **e)


def halton_fixture():
    pixels = np.array([[0, 0], [1, 0], [5, 77], [127, 127], [64, 3]], np.int32)
    vals = np.zeros((len(pixels), 4, 24), np.float32)
    index = np.zeros((len(pixels), 4), np.uint32)  # global Halton index of each sample
    stride = 128 * 243  # base scales 2^7 and 3^5 at resolution >= 128
    for k, (px, py) in enumerate(pixels):
        off = orc.lib().orc_halton_offset(128, 128, int(px), int(py))
        for s in range(4):
            vals[k, s] = orc.halton_get(128, 128, 16, int(px), int(py), s, 24)
            index[k, s] = off + s * stride
    return dict(pixels=pixels, values=vals, index=index, res=np.int32(128), spp=np.int32(16))


def fbm(n, seed):
    sys.path.insert(0, ROOT)
    from fledge_amd import api
    return api.gen_fbm_grid(n, seed=seed)


def volpath_fixture():
    grid = fbm(24, 5)
    rgb, lookups, paths = orc.volpath_render(grid, [0, 0, 0, 1, 1, 1], [6.4, 1.6, 0.3], LIGHTS, CAM0,
                                             16, 16, 4, 32)
    return dict(grid=grid, bounds=np.array([0, 0, 0, 1, 1, 1], np.float32),
                medium=np.array([6.4, 1.6, 0.3], np.float32), lights=np.array(LIGHTS, np.float32),
                camera=np.array(CAM0, np.float32), width=np.int32(16), height=np.int32(16),
                spp=np.int32(4), max_depth=np.int32(32), rgb=rgb, lookups=np.uint64(lookups))


def pathtrace_fixture():
    from fledge_amd import api
    v, i = api.gen_octasphere(3, 0.1, seed=7, base_freq=2.0)
    b = orc.Lbvh(v, i)
    cam = [0.0, 0.0, -3.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0, 45.0]
    W, H, spp, depth = 24, 24, 2, 4
    nseg, ids, ts, nrays = b.diffuse_render(cam, W, H, spp, depth, 0, W * H * spp)
    return dict(verts=v.reshape(-1, 3), ind=i.reshape(-1, 3), camera=np.array(cam, np.float32),
                width=np.int32(W), height=np.int32(H), spp=np.int32(spp), max_depth=np.int32(depth),
                nseg=nseg, seg_id=ids, seg_t=ts, segments=np.int64(nrays))


def adjoint_fixture():
    g = (0.1 + 0.8 * fbm(12, 4)).astype(np.float32)
    S, A, G = 8.0, 0.8, 0.3
    mu = S * float(g.max())
    W = H = 12
    spp, depth = 2, 16
    params = [S, A, G, mu]
    lp, img = orc.diff_render(g.astype(np.float64), g, params, LIGHTS, CAM0, W, H, spp, depth)
    w = np.random.default_rng(12).normal(size=(H, W)).astype(np.float64)
    grad, galpha = orc.diff_grad(g.astype(np.float64), g, params, LIGHTS, CAM0, W, H, spp, depth, lp, w)
    return dict(grid=g, params=np.array(params, np.float32), lights=np.array(LIGHTS, np.float32),
                camera=np.array(CAM0, np.float32), width=np.int32(W), height=np.int32(H),
                spp=np.int32(spp), max_depth=np.int32(depth), lpath=lp, image=img, weight=w,
                grad=grad, grad_albedo=np.float64(galpha))


FIXTURES = {
    "lbvh_soup1024.npz": lbvh_fixture,
    "halton_res128.npz": halton_fixture,
    "volpath_fbm24.npz": volpath_fixture,
    "pathtrace_octa3.npz": pathtrace_fixture,
    "adjoint_fbm12.npz": adjoint_fixture,
}


def main():
    manifest = {}
    for name, fn in FIXTURES.items():
        data = fn()
        path = os.path.join(HERE, name)
        np.savez_compressed(path, **data)
        manifest[name] = {k: [str(np.asarray(x).dtype), list(np.asarray(x).shape)] for k, x in data.items()}
        print("wrote", name, os.path.getsize(path), "bytes")
    with open(os.path.join(HERE, "MANIFEST.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
