"""ctypes wrapper of the CPU oracle (oracle/build/liboracle.so) — test
infrastructure only."""
import collections
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "build", "liboracle.so")

P = ctypes.c_void_p
i32 = ctypes.c_int32
i64 = ctypes.c_int64
u32 = ctypes.c_uint32
u64 = ctypes.c_uint64
f32 = ctypes.c_float

_SIGS = {
    "orc_logf": (None, [P, P, i64]),
    "orc_sincosf": (None, [P, P, P, i64]),
    "orc_walk_draw": (None, [P, P, i64]),
    "orc_pcg32": (None, [u64, u64, i64, P]),
    "orc_gen_soup": (None, [u64, u64, i32, P, P]),
    "orc_left_shift3": (u32, [u32]),
    "orc_encode_morton3": (u32, [f32, f32, f32]),
    "orc_lbvh_build": (P, [P, P, i32, i32]),
    "orc_lbvh_free": (None, [P]),
    "orc_lbvh_ntris": (i32, [P]),
    "orc_lbvh_export": (None, [P] + [P] * 12),
    "orc_trace": (None, [P, P, i64, i32, P, P, P, P, P, ctypes.c_int]),
    "orc_brute_trace": (None, [P, P, i32, P, i64, i32, P, P, P, ctypes.c_int]),
    "orc_num_primes": (i32, []),
    "orc_prime": (i32, [i32]),
    "orc_radical_inverse": (f32, [i32, u64]),
    "orc_halton_offset": (u64, [i32, i32, i32, i32]),
    "orc_halton_get": (None, [i32, i32, i32, i32, i32, i64, i32, P]),
    "orc_halton_sequence": (None, [i32, i32, i32, i32, i32, i32, P]),
    "orc_camera_rays": (None, [P, i32, i32, i32, i32, P, ctypes.c_int]),
    "orc_volpath_render": (None, [P, i32, i32, i32, P, P, P, P, i32, i32, i32, i32, P,
                                  ctypes.c_int, P, P]),
    "orc_diffuse_render": (None, [P, i32, P, i32, i32, i32, i32, i64, i64, P, P, P, ctypes.c_int, P]),
    "orc_offset_ray_origin": (None, [P, P, P, P]),
    "orc_camera_generate": (None, [P, f32, f32, i32, i32, P, P]),
    "orc_volpath_render_strided": (None, [P, i32, i32, i32, P, P, P, P, i32, i32, i32, i32, i64, P,
                                          ctypes.c_int, P, P]),
    "orc_diff_render_strided": (None, [P, P, i32, i32, i32, P, P, P, P, i32, i32, i32, i32, i64, P, P, P, P, P,
                                       ctypes.c_int, P, i32]),
    "orc_diffuse_sample": (None, [P, P, P, i64, P, P]),
    "orc_shading_frame": (None, [P, i64, P]),
    "orc_volpath_render_film": (None, [P, i32, i32, i32, P, P, P, P, i32, i32, i32, i32, P, P, P, P, P]),
    "orc_volpath_samples": (None, [P, i32, i32, i32, P, P, P, P, i32, i32, i32, i32, P, P]),
    "orc_mitchell1d": (f32, [f32]),
    "orc_volpath_mesh_render": (None, [P, ctypes.c_int, P, i32, i32, i32, P, P, P, i32, i32, i32, i32, P,
                                       ctypes.c_int, P, i32]),
    "orc_volpath_mesh_render_strided": (None, [P, ctypes.c_int, P, i32, i32, i32, P, P, P, i32, i32, i32, i32,
                                               i64, P, ctypes.c_int, P, P]),
    "orc_halton_stateful": (None, [i32, i32, i32, i32, i32, i32, P]),
    "orc_scene_trace": (None, [P, i32, i32, P, i64, P, P, P, ctypes.c_int]),
    "orc_diff_mesh_render": (None, [P, i32, P, P, i32, i32, i32, P, P, P, i32, i32, i32, i32, P, P, P, P, P,
                                    ctypes.c_int, P, i32]),
}

_lib = None
NTHREADS = max(1, min(8, os.cpu_count() or 1))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle not built (make -C oracle)")
        L = ctypes.CDLL(LIB)
        for n, (r, a) in _SIGS.items():
            f = getattr(L, n, None)
            if f is None:
                continue
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


# Small arrays made inside a call's argument list (bounds, medium, lights,
# camera) would be freed as soon as their address is taken: the last few are
# kept alive here until the call has run.
_alive = collections.deque(maxlen=64)


def _p(a):
    if a is None:
        return None
    if a.nbytes <= 4096:
        _alive.append(a)
    return a.ctypes.data


def lights11(lights):
    """The oracle's light array: point present, position (3), intensity (RGB),
    environment present, radiance (RGB).  A 7-entry grey form (point present,
    position, intensity, environment present, radiance) is widened."""
    a = np.asarray(lights, np.float32).ravel()
    if a.size == 7:
        a = np.array([a[0], a[1], a[2], a[3], a[4], a[4], a[4], a[5], a[6], a[6], a[6]], np.float32)
    if a.size != 11:
        raise ValueError("lights: 7 (grey) or 11 (RGB) entries")
    return a


class Lbvh:
    def __init__(self, verts, ind):
        self.verts = np.ascontiguousarray(verts, np.float32).ravel()
        self.ind = np.ascontiguousarray(ind, np.int32).ravel()
        self.n = self.ind.size // 3
        self.h = lib().orc_lbvh_build(_p(self.verts), _p(self.ind), self.n, self.verts.size // 3)

    def export(self):
        n, ni = self.n, max(self.n - 1, 0)
        out = {
            "bound": np.zeros(6, np.float32), "morton": np.zeros(n, np.uint32),
            "morton_sorted": np.zeros(n, np.uint32), "order": np.zeros(n, np.uint32),
            "split": np.zeros(ni, np.int32), "range_lo": np.zeros(ni, np.int32),
            "range_hi": np.zeros(ni, np.int32), "parent": np.zeros(ni, np.int32),
            "left_leaf": np.zeros(ni, np.uint8), "right_leaf": np.zeros(ni, np.uint8),
            "node_bound": np.zeros((ni, 6), np.float32), "leaf_parent": np.zeros(n, np.int32),
        }
        lib().orc_lbvh_export(self.h, *[_p(out[k]) for k in (
            "bound", "morton", "morton_sorted", "order", "split", "range_lo", "range_hi", "parent",
            "left_leaf", "right_leaf", "node_bound", "leaf_parent")])
        return out

    def trace(self, rays, compat=False, normals=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        ids = np.zeros(n, np.int32)
        ts = np.zeros(n, np.float32)
        ng = np.zeros((n, 3), np.float32) if normals else None
        nodes = np.zeros(1, np.uint64)
        tris = np.zeros(1, np.uint64)
        lib().orc_trace(self.h, _p(rays), n, 1 if compat else 0, _p(ids), _p(ts), _p(ng), _p(nodes),
                        _p(tris), NTHREADS)
        self.last_counts = (int(nodes[0]), int(tris[0]))
        if normals:
            return ids, ts, ng
        return ids, ts

    def diffuse_render(self, cam10, W, H, spp, max_depth, p0, p1, compat=False, record=True,
                       nthreads=None):
        n = p1 - p0
        S = max_depth + 1
        nseg = np.zeros(n, np.int32)
        ids = np.full((n, S), -2, np.int32) if record else None
        ts = np.zeros((n, S), np.float32) if record else None
        rays = np.zeros(1, np.uint64)
        cam = np.asarray(cam10, np.float32)
        lib().orc_diffuse_render(self.h, 1 if compat else 0, _p(cam), W, H, spp, max_depth, p0, p1,
                                 _p(nseg), _p(ids), _p(ts), nthreads or NTHREADS, _p(rays))
        return nseg, ids, ts, int(rays[0])

    def volpath_mesh(self, grid, medium, lights, cam10, W, H, spp, max_depth, compat=False, nthreads=None,
                     film=0):
        """VolPathIntegrator with this mesh as the interface around the grid's
        medium (on the mesh's bounding box): (rgb (H, W, 3), lookups).  film=1:
        through the reference's render loop and film."""
        grid = np.ascontiguousarray(grid, np.float32)
        nz, ny, nx = grid.shape
        rgb = np.zeros((H, W, 3), np.float32)
        lk = np.zeros(1, np.uint64)
        lib().orc_volpath_mesh_render(self.h, 1 if compat else 0, _p(grid), nx, ny, nz,
                                      _p(np.asarray(medium, np.float32)), _p(lights11(lights)),
                                      _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(rgb),
                                      nthreads or NTHREADS, _p(lk), film)
        return rgb, int(lk[0])

    def volpath_mesh_strided(self, grid, medium, lights, cam10, W, H, spp, max_depth, stride, nthreads=None):
        """volpath_mesh on every stride-th pixel only: (rgb, lookups, path
        samples run); the other pixels stay 0."""
        grid = np.ascontiguousarray(grid, np.float32)
        nz, ny, nx = grid.shape
        rgb = np.zeros((H, W, 3), np.float32)
        lk = np.zeros(1, np.uint64)
        pa = np.zeros(1, np.uint64)
        lib().orc_volpath_mesh_render_strided(self.h, 0, _p(grid), nx, ny, nz, _p(np.asarray(medium, np.float32)),
                                              _p(lights11(lights)), _p(np.asarray(cam10, np.float32)),
                                              W, H, spp, max_depth, stride, _p(rgb), nthreads or NTHREADS, _p(lk),
                                              _p(pa))
        return rgb, int(lk[0]), int(pa[0])

    def diff_render(self, theta, theta_p, params, lights, cam10, W, H, spp, max_depth, compat=False,
                    nthreads=None, channel=0):
        """FP64 differentiable forward over this mesh as the interface around
        the medium (on the mesh's bounding box), colour channel `channel` of
        the lights: (per-path radiance, per-pixel mean, lookups)."""
        theta = np.ascontiguousarray(theta, np.float64)
        theta_p = np.ascontiguousarray(theta_p, np.float32)
        nz, ny, nx = theta_p.shape
        lp = np.zeros(W * H * spp, np.float64)
        img = np.zeros(W * H, np.float64)
        lk = np.zeros(1, np.uint64)
        lib().orc_diff_mesh_render(self.h, 1 if compat else 0, _p(theta), _p(theta_p), nx, ny, nz,
                                   _p(np.asarray(params, np.float32)), _p(lights11(lights)),
                                   _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(lp), _p(img),
                                   None, None, None, nthreads or NTHREADS, _p(lk), channel)
        return lp, img.reshape(H, W), int(lk[0])

    def diff_grad(self, theta, theta_p, params, lights, cam10, W, H, spp, max_depth, lpath, dldpix,
                  compat=False, nthreads=None, channel=0):
        """FP64 analytic gradient of sum(dldpix * image) over this mesh, image
        = colour channel `channel`: (dense gradient (nz, ny, nx), d/dalbedo)."""
        theta = np.ascontiguousarray(theta, np.float64)
        theta_p = np.ascontiguousarray(theta_p, np.float32)
        nz, ny, nx = theta_p.shape
        grad = np.zeros(nx * ny * nz, np.float64)
        ga = np.zeros(1, np.float64)
        dl = np.ascontiguousarray(dldpix, np.float64).ravel()
        lp = np.ascontiguousarray(lpath, np.float64)
        lib().orc_diff_mesh_render(self.h, 1 if compat else 0, _p(theta), _p(theta_p), nx, ny, nz,
                                   _p(np.asarray(params, np.float32)), _p(lights11(lights)),
                                   _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(lp), None,
                                   _p(dl), _p(grad), _p(ga), nthreads or NTHREADS, None, channel)
        return grad.reshape(nz, ny, nx), float(ga[0])

    def __del__(self):
        try:
            lib().orc_lbvh_free(self.h)
        except Exception:
            pass


def scene_trace(meshes, rays, compat=False):
    """Scene::intersect over the oracle LBVHs `meshes` (scene order): (mesh,
    id, t) per ray."""
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    hs = (ctypes.c_void_p * len(meshes))(*[m.h for m in meshes])
    mesh = np.zeros(n, np.int32)
    ids = np.zeros(n, np.int32)
    ts = np.zeros(n, np.float32)
    lib().orc_scene_trace(hs, len(meshes), 1 if compat else 0, _p(rays), n, _p(mesh), _p(ids), _p(ts), NTHREADS)
    return mesh, ids, ts


def brute_trace(verts, ind, rays, compat=False):
    verts = np.ascontiguousarray(verts, np.float32).ravel()
    ind = np.ascontiguousarray(ind, np.int32).ravel()
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    n = rays.shape[0]
    ids = np.zeros(n, np.int32)
    ts = np.zeros(n, np.float32)
    ties = np.zeros(n, np.int32)
    lib().orc_brute_trace(_p(verts), _p(ind), ind.size // 3, _p(rays), n, 1 if compat else 0,
                          _p(ids), _p(ts), _p(ties), NTHREADS)
    return ids, ts, ties


def camera_rays(cam10, W, H, spp=1, sample=0):
    out = np.zeros((W * H, 8), np.float32)
    lib().orc_camera_rays(_p(np.asarray(cam10, np.float32)), W, H, spp, sample, _p(out), NTHREADS)
    return out


def volpath_render(grid, bounds, medium, lights, cam10, W, H, spp, max_depth, nthreads=None):
    grid = np.ascontiguousarray(grid, np.float32)
    nz, ny, nx = grid.shape
    rgb = np.zeros((H, W, 3), np.float32)
    lk = np.zeros(1, np.uint64)
    pa = np.zeros(1, np.uint64)
    lib().orc_volpath_render(_p(grid), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                             _p(np.asarray(medium, np.float32)), _p(lights11(lights)),
                             _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(rgb),
                             nthreads or NTHREADS, _p(lk), _p(pa))
    return rgb, int(lk[0]), int(pa[0])


def volpath_render_film(grid, bounds, medium, lights, cam10, W, H, spp, max_depth):
    """The reference's render loop and film: (rgb (H, W, 3), per-task RGB
    radiance (H*W, spp-1, 3), per-task film position (H*W, spp-1, 2),
    lookups)."""
    grid = np.ascontiguousarray(grid, np.float32)
    nz, ny, nx = grid.shape
    rgb = np.zeros((H, W, 3), np.float32)
    T = max(spp - 1, 0)
    tl = np.zeros((H * W, max(T, 1), 3), np.float32)
    tu = np.zeros((H * W, max(T, 1), 2), np.float32)
    lk = np.zeros(1, np.uint64)
    pa = np.zeros(1, np.uint64)
    lib().orc_volpath_render_film(_p(grid), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                                  _p(np.asarray(medium, np.float32)), _p(lights11(lights)),
                                  _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(rgb), _p(tl),
                                  _p(tu), _p(lk), _p(pa))
    return rgb, tl[:, :T], tu[:, :T], int(lk[0])


def volpath_samples(grid, bounds, medium, lights, cam10, W, H, spp, max_depth):
    """Per-sample RGB radiance (H*W, spp, 3) and film position (H*W, spp, 2)
    of the per-sample schedule."""
    grid = np.ascontiguousarray(grid, np.float32)
    nz, ny, nx = grid.shape
    L = np.zeros((H * W, spp, 3), np.float32)
    uv = np.zeros((H * W, spp, 2), np.float32)
    lib().orc_volpath_samples(_p(grid), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                              _p(np.asarray(medium, np.float32)), _p(lights11(lights)),
                              _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(L), _p(uv))
    return L, uv


def halton_stateful(spp, resx, resy, px, py, ndraw):
    out = np.zeros((spp, ndraw), np.float32)
    lib().orc_halton_stateful(spp, resx, resy, px, py, ndraw, _p(out))
    return out


def diff_render(theta, theta_p, params, lights, cam10, W, H, spp, max_depth, bounds=(0, 0, 0, 1, 1, 1),
                nthreads=None, channel=0):
    """FP64 differentiable forward of colour channel `channel` of the lights:
    (per-path radiance, per-pixel mean)."""
    L = lib()
    _diff_sig(L)
    theta = np.ascontiguousarray(theta, np.float64)
    theta_p = np.ascontiguousarray(theta_p, np.float32)
    nz, ny, nx = theta_p.shape
    lp = np.zeros(W * H * spp, np.float64)
    img = np.zeros(W * H, np.float64)
    L.orc_diff_render(_p(theta), _p(theta_p), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                      _p(np.asarray(params, np.float32)), _p(lights11(lights)),
                      _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(lp), _p(img), None,
                      None, None, nthreads or NTHREADS, None, channel)
    return lp, img.reshape(H, W)


def diff_grad(theta, theta_p, params, lights, cam10, W, H, spp, max_depth, lpath, dldpix,
              bounds=(0, 0, 0, 1, 1, 1), nthreads=None, channel=0):
    """FP64 analytic gradient of sum(dldpix * image) w.r.t. theta and albedo,
    image = colour channel `channel` (lpath: that channel's forward)."""
    L = lib()
    _diff_sig(L)
    theta = np.ascontiguousarray(theta, np.float64)
    theta_p = np.ascontiguousarray(theta_p, np.float32)
    nz, ny, nx = theta_p.shape
    grad = np.zeros(nx * ny * nz, np.float64)
    ga = np.zeros(1, np.float64)
    dl = np.ascontiguousarray(dldpix, np.float64).ravel()
    lp = np.ascontiguousarray(lpath, np.float64)
    L.orc_diff_render(_p(theta), _p(theta_p), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                      _p(np.asarray(params, np.float32)), _p(lights11(lights)),
                      _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, _p(lp), None, _p(dl),
                      _p(grad), _p(ga), nthreads or NTHREADS, None, channel)
    return grad.reshape(nz, ny, nx), float(ga[0])


def _diff_sig(L):
    f = L.orc_diff_render
    if f.argtypes is None:
        f.restype = None
        f.argtypes = [P, P, i32, i32, i32, P, P, P, P, i32, i32, i32, i32, P, P, P, P, P,
                      ctypes.c_int, P, i32]


def logf(x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    lib().orc_logf(_p(x), _p(y), x.size)
    return y


def sincosf(x):
    x = np.ascontiguousarray(x, np.float32)
    s = np.zeros_like(x)
    c = np.zeros_like(x)
    lib().orc_sincosf(_p(x), _p(s), _p(c), x.size)
    return s, c


def walk_draw(triples):
    """1 - u of the ratio-tracking walks' stream for (path, walk, n) rows."""
    x = np.ascontiguousarray(triples, np.uint32).reshape(-1, 3)
    y = np.zeros(x.shape[0], np.float32)
    lib().orc_walk_draw(_p(x), _p(y), x.shape[0])
    return y


def halton_get(resx, resy, spp, px, py, sample, n):
    out = np.zeros(n, np.float32)
    lib().orc_halton_get(resx, resy, spp, px, py, sample, n, _p(out))
    return out


def diff_render_strided(theta, theta_p, params, lights, cam10, W, H, spp, max_depth, stride,
                        bounds=(0, 0, 0, 1, 1, 1), nthreads=None, channel=0):
    """diff_render over the pixels 0, stride, 2 stride, ... only: (per-path
    radiance, per-pixel mean); entries of other pixels stay 0."""
    theta = np.ascontiguousarray(theta, np.float64)
    theta_p = np.ascontiguousarray(theta_p, np.float32)
    nz, ny, nx = theta_p.shape
    lp = np.zeros(W * H * spp, np.float64)
    img = np.zeros(W * H, np.float64)
    lib().orc_diff_render_strided(_p(theta), _p(theta_p), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                                  _p(np.asarray(params, np.float32)), _p(lights11(lights)),
                                  _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, stride, _p(lp),
                                  _p(img), None, None, None, nthreads or NTHREADS, None, channel)
    return lp, img.reshape(H, W)


def diff_grad_strided(theta, theta_p, params, lights, cam10, W, H, spp, max_depth, stride, lpath, dldpix,
                      bounds=(0, 0, 0, 1, 1, 1), nthreads=None, channel=0):
    """diff_grad over the pixels 0, stride, 2 stride, ... only."""
    theta = np.ascontiguousarray(theta, np.float64)
    theta_p = np.ascontiguousarray(theta_p, np.float32)
    nz, ny, nx = theta_p.shape
    grad = np.zeros(nx * ny * nz, np.float64)
    ga = np.zeros(1, np.float64)
    dl = np.ascontiguousarray(dldpix, np.float64).ravel()
    lp = np.ascontiguousarray(lpath, np.float64)
    lib().orc_diff_render_strided(_p(theta), _p(theta_p), nx, ny, nz, _p(np.asarray(bounds, np.float32)),
                                  _p(np.asarray(params, np.float32)), _p(lights11(lights)),
                                  _p(np.asarray(cam10, np.float32)), W, H, spp, max_depth, stride, _p(lp), None,
                                  _p(dl), _p(grad), _p(ga), nthreads or NTHREADS, None, channel)
    return grad.reshape(nz, ny, nx), float(ga[0])
