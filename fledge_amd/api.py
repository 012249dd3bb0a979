"""Python front end over the C ABI, with torch tensors as device buffers.

Context        one HIP stream (torch's current stream by default)
LBVH           flg_lbvh_build / flg_lbvh_build_batched / flg_lbvh_export
trace_closest  flg_trace_closest (rays [n, 8] float32, hits [n, 2] int32/float32)
camera_rays    flg_camera_rays
Volume         flg_volume_create / flg_volpath_render / differentiable pair
Scene helpers  numpy generators from libfledge_scene.so
"""
import ctypes
import os

import numpy as np
import torch

from . import _native as N


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def check_buffer(t, name, dtypes, device, numel=None, min_numel=None, ndim=None, shape=None):
    """Raises ValueError unless `t` is a contiguous tensor of one of `dtypes`
    on CUDA device `device` (None: checked later by check_device) holding exactly `numel` (or at least `min_numel`)
    elements, with `ndim` dimensions or the given `shape` (None: any extent): the kernels behind
    the C ABI index these buffers by the sizes passed beside them."""
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s: expected a torch tensor, got %s" % (name, type(t).__name__))
    if isinstance(dtypes, torch.dtype):
        dtypes = (dtypes,)
    if t.dtype not in dtypes:
        raise ValueError("%s: dtype %s, expected %s" % (name, t.dtype, " or ".join(str(d) for d in dtypes)))
    if not t.is_contiguous():
        raise ValueError("%s: must be contiguous (got strides %s for shape %s)"
                         % (name, tuple(t.stride()), tuple(t.shape)))
    if ndim is not None and t.dim() != ndim:
        raise ValueError("%s: %d dimensions, expected %d" % (name, t.dim(), ndim))
    if shape is not None and (t.dim() != len(shape) or
                              any(w is not None and w != g for g, w in zip(t.shape, shape))):
        raise ValueError("%s: shape %s, expected %s" % (name, tuple(t.shape),
                                                        tuple("n" if w is None else w for w in shape)))
    if numel is not None and t.numel() != numel:
        raise ValueError("%s: %d elements, expected %d" % (name, t.numel(), numel))
    if min_numel is not None and t.numel() < min_numel:
        raise ValueError("%s: %d elements, at least %d needed" % (name, t.numel(), min_numel))
    if device is not None:
        check_device(device, **{name: t})
    return t


def check_device(device, **bufs):
    """Raises ValueError unless every tensor given (None skipped) lives on
    CUDA device `device`.  Wrappers call it after the shape checks of all
    their buffers."""
    for name, t in bufs.items():
        if t is not None and (t.device.type != "cuda" or (t.device.index or 0) != device):
            raise ValueError("%s: on %s, expected cuda:%d" % (name, t.device, device))


_HIT_TYPES = (torch.int32, torch.float32)


class Context:
    def __init__(self, device=0, stream=None):
        self.lib = N.hip()
        torch.cuda.set_device(device)
        if stream is None:
            stream = torch.cuda.current_stream(device).cuda_stream
        self.device = device
        self.h = ctypes.c_void_p()
        N.check(self.lib.flg_ctx_create(device, ctypes.c_void_p(stream), ctypes.byref(self.h)))

    def check(self, st):
        N.check(st, self.h)

    def sync(self):
        self.check(self.lib.flg_ctx_synchronize(self.h))

    def profiling(self, on=True):
        self.check(self.lib.flg_ctx_set_profiling(self.h, 1 if on else 0))

    def reset_profiling(self):
        self.check(self.lib.flg_ctx_reset_profiling(self.h))

    def kernel_time(self, name):
        ms = ctypes.c_double()
        n = ctypes.c_int64()
        st = self.lib.flg_ctx_kernel_time(self.h, name.encode(), ctypes.byref(ms), ctypes.byref(n))
        if st != 0:
            return 0.0, 0
        return ms.value, n.value

    def close(self):
        if self.h:
            self.lib.flg_ctx_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class LBVH:
    """A built radix-tree BVH (drop-in for RadixBVHBuilder / NaiveBVHAccel)."""

    def __init__(self, ctx, handle, ntris, keep=()):
        self.ctx = ctx
        self.h = handle
        self.ntris = ntris
        self._keep = keep

    @staticmethod
    def _check_mesh(ctx, verts, ind, k=""):
        check_buffer(verts, "verts" + k, torch.float32, None)
        check_buffer(ind, "ind" + k, torch.int32, None)
        if verts.numel() % 3 or ind.numel() % 3:
            raise ValueError("verts%s / ind%s: sizes must be multiples of 3 (got %d, %d)"
                             % (k, k, verts.numel(), ind.numel()))
        if ind.numel() and verts.numel() == 0:
            raise ValueError("ind%s: indices into an empty vertex array" % k)

    @staticmethod
    def _check_indices(meshes):
        """Every index in [0, nverts): one reduction per mesh, then one sync."""
        bad = []
        for k, (v, ind) in enumerate(meshes):
            if ind.numel():
                lo, hi = torch.aminmax(ind)
                bad.append((k, lo, hi, v.numel() // 3))
        for k, lo, hi, nv in bad:
            if int(lo) < 0 or int(hi) >= nv:
                raise ValueError("mesh %d: vertex indices in [%d, %d] outside [0, %d)" % (k, int(lo), int(hi), nv))

    @classmethod
    def build(cls, ctx, verts, ind, compat=False, check_indices=True):
        """check_indices: reject indices outside the vertex array first (one
        device reduction and a synchronisation)."""
        cls._check_mesh(ctx, verts, ind)
        check_device(ctx.device, verts=verts, ind=ind)
        if check_indices:
            cls._check_indices([(verts, ind)])
        h = ctypes.c_void_p()
        ctx.check(ctx.lib.flg_lbvh_build(ctx.h, _ptr(verts), _ptr(ind), ind.numel() // 3,
                                         verts.numel() // 3, 1 if compat else 0, ctypes.byref(h)))
        return cls(ctx, h, ind.numel() // 3, keep=(verts, ind))

    @classmethod
    def build_batched(cls, ctx, meshes, compat=False, check_indices=True):
        n = len(meshes)
        if n == 0:
            raise ValueError("build_batched: no meshes")
        for i, (v, ind) in enumerate(meshes):
            cls._check_mesh(ctx, v, ind, "[%d]" % i)
        for i, (v, ind) in enumerate(meshes):
            check_device(ctx.device, **{"verts[%d]" % i: v, "ind[%d]" % i: ind})
        if check_indices:
            cls._check_indices(meshes)
        descs = (N.MeshDesc * n)()
        keep = []
        for i, (v, ind) in enumerate(meshes):
            keep += [v, ind]
            descs[i].d_verts = v.data_ptr()
            descs[i].d_ind = ind.data_ptr()
            descs[i].ntris = ind.numel() // 3
            descs[i].nverts = v.numel() // 3
        out = (ctypes.c_void_p * n)()
        ctx.check(ctx.lib.flg_lbvh_build_batched(ctx.h, descs, n, 1 if compat else 0, out))
        keep = tuple(keep)  # one tuple shared by the batch's handles
        return [cls(ctx, ctypes.c_void_p(out[i]), descs[i].ntris, keep=keep) for i in range(n)]

    def info(self):
        inf = N.LbvhInfo()
        self.ctx.check(self.ctx.lib.flg_lbvh_get_info(self.h, ctypes.byref(inf)))
        return inf

    def export(self):
        n = self.ntris
        ni = max(n - 1, 0)
        arrs = {
            "morton": np.zeros(n, np.uint32), "morton_sorted": np.zeros(n, np.uint32),
            "order": np.zeros(n, np.uint32), "split": np.zeros(ni, np.int32),
            "range_lo": np.zeros(ni, np.int32), "range_hi": np.zeros(ni, np.int32),
            "parent": np.zeros(ni, np.int32), "left_leaf": np.zeros(ni, np.uint8),
            "right_leaf": np.zeros(ni, np.uint8), "node_bound": np.zeros((ni, 6), np.float32),
            "leaf_parent": np.zeros(n, np.int32),
        }
        v = N.LbvhView()
        for k, a in arrs.items():
            setattr(v, "h_" + k, a.ctypes.data if a.size else None)
        self.ctx.check(self.ctx.lib.flg_lbvh_export(self.ctx.h, self.h, ctypes.byref(v)))
        arrs["bound"] = np.array(self.info().bound, np.float32)
        return arrs

    def trace(self, rays, hits=None, block_size=0, stack_mode=0, ring_depth=0):
        """rays: (n, 8) float32 (o, tmax, d, pad); hits: (n, 2) int32 or
        float32 (id, t bits)."""
        dev = self.ctx.device
        check_buffer(rays, "rays", torch.float32, None, shape=(None, 8))
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 2), dtype=torch.int32, device=rays.device)
        check_buffer(hits, "hits", _HIT_TYPES, None, min_numel=2 * n)
        check_device(dev, rays=rays, hits=hits)
        cfg = N.TraceCfg()
        cfg.block_size = block_size
        cfg.stack_mode = stack_mode
        cfg.ring_depth = ring_depth
        self.ctx.check(self.ctx.lib.flg_trace_closest(self.ctx.h, self.h, _ptr(rays), _ptr(hits), n,
                                                      ctypes.byref(cfg)))
        return hits

    def count(self, rays):
        check_buffer(rays, "rays", torch.float32, None, shape=(None, 8))
        check_device(self.ctx.device, rays=rays)
        c = torch.zeros(2, dtype=torch.int64, device=rays.device)
        self.ctx.check(self.ctx.lib.flg_trace_count(self.ctx.h, self.h, _ptr(rays), rays.shape[0], _ptr(c)))
        return c

    def pathtrace(self, cam, width, height, spp, max_depth, record=False, trace_block=0,
                  count_visits=False, use_graph=False, blocks_per_cu=0, stack_mode=0, ray_fetch=0,
                  out=None, ring_depth=0):
        """One frame of the diffuse path tracer.  record: also return per-path
        segment counts, ids and t (into `out` = (nseg, seg_id, seg_t) when
        given, so that a replayed graph writes the same buffers)."""
        cfg = N.PtCfg()
        cfg.width, cfg.height, cfg.spp, cfg.max_depth = width, height, spp, max_depth
        cfg.trace_block = trace_block
        cfg.blocks_per_cu = blocks_per_cu
        cfg.stack_mode = stack_mode
        cfg.ray_fetch = ray_fetch
        cfg.ring_depth = ring_depth
        cfg.count_visits = 1 if count_visits else 0
        cfg.use_graph = 1 if use_graph else 0
        npaths = width * height * spp
        dev = torch.device("cuda", self.ctx.device)
        nseg = seg_id = seg_t = None
        if out is not None:
            nseg, seg_id, seg_t = out
            S = max_depth + 1
            if nseg is not None:
                check_buffer(nseg, "nseg", torch.int32, None, min_numel=npaths)
            if seg_id is not None:
                check_buffer(seg_id, "seg_id", torch.int32, None, min_numel=npaths * S)
            if seg_t is not None:
                check_buffer(seg_t, "seg_t", torch.float32, None, min_numel=npaths * S)
            check_device(self.ctx.device, nseg=nseg, seg_id=seg_id, seg_t=seg_t)
        elif record:
            nseg = torch.zeros(npaths, dtype=torch.int32, device=dev)
            seg_id = torch.full((npaths, max_depth + 1), -2, dtype=torch.int32, device=dev)
            seg_t = torch.zeros((npaths, max_depth + 1), dtype=torch.float32, device=dev)
        stats = (ctypes.c_uint64 * 4)()
        self.ctx.check(self.ctx.lib.flg_pathtrace_diffuse(self.ctx.h, self.h, ctypes.byref(cam),
                                                          ctypes.byref(cfg), _ptr(nseg), _ptr(seg_id),
                                                          _ptr(seg_t), stats))
        return {"segments": int(stats[0]), "paths": int(stats[1]), "node_visits": int(stats[2]),
                "tri_tests": int(stats[3]), "nseg": nseg, "seg_id": seg_id, "seg_t": seg_t}

    def autotune(self, cam, width, height, spp, max_depth):
        """Times the path tracer's launch shapes on this frame; later frames of
        this context with trace_block / blocks_per_cu 0 use the fastest."""
        cfg = N.PtCfg()
        cfg.width, cfg.height, cfg.spp, cfg.max_depth = width, height, spp, max_depth
        t = N.Tuning()
        self.ctx.check(self.ctx.lib.flg_autotune_pathtrace(self.ctx.h, self.h, ctypes.byref(cam),
                                                           ctypes.byref(cfg), ctypes.byref(t)))
        return {"block_size": t.block_size, "blocks_per_cu": t.blocks_per_cu,
                "stack_mode": t.stack_mode, "ring_depth": t.ring_depth, "ray_fetch": t.ray_fetch, "ms": t.ms,
                "ms_default": t.ms_default, "candidates": t.candidates}

    def autotune_trace(self, rays, hits):
        check_buffer(rays, "rays", torch.float32, None, shape=(None, 8))
        check_buffer(hits, "hits", _HIT_TYPES, None, min_numel=2 * rays.shape[0])
        check_device(self.ctx.device, rays=rays, hits=hits)
        t = N.Tuning()
        self.ctx.check(self.ctx.lib.flg_autotune_trace(self.ctx.h, self.h, _ptr(rays), _ptr(hits),
                                                       rays.shape[0], ctypes.byref(t)))
        return {"block_size": t.block_size, "stack_mode": t.stack_mode, "ring_depth": t.ring_depth, "ms": t.ms,
                "ms_default": t.ms_default, "candidates": t.candidates}

    def close(self):
        if self.h:
            self.ctx.lib.flg_lbvh_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def load_ply(path):
    """MakeTriangleMesh through flg_ply_load: (verts (n, 3) float32, ind
    (m, 3) int32, normals (n, 3) or None, uv (n, 2) or None)."""
    lib = N.hip()
    m = N.PlyMesh()
    N.check(lib.flg_ply_load(os.fsencode(path), ctypes.byref(m)))
    try:
        def arr(ptr, n, k, dt):
            if not ptr:
                return None
            return np.ctypeslib.as_array(ptr, shape=(n * k,)).astype(dt, copy=True).reshape(n, k)
        return (arr(m.verts, m.nverts, 3, np.float32), arr(m.ind, m.ntris, 3, np.int32),
                arr(m.normals, m.nverts, 3, np.float32), arr(m.uv, m.nverts, 2, np.float32))
    finally:
        lib.flg_ply_free(ctypes.byref(m))


class SceneAccel:
    """The scene's accel over several meshes (flg_scene_accel_build): closest
    hit over the meshes as fledge's Scene::intersect gives it.  Keeps the
    meshes alive."""

    def __init__(self, ctx, meshes):
        self.ctx = ctx
        self.meshes = list(meshes)
        hs = (ctypes.c_void_p * len(self.meshes))(*[m.h.value if isinstance(m.h, ctypes.c_void_p) else m.h
                                                     for m in self.meshes])
        self.h = ctypes.c_void_p()
        ctx.check(ctx.lib.flg_scene_accel_build(ctx.h, hs, len(self.meshes), ctypes.byref(self.h)))

    def rank(self):
        r = (ctypes.c_int32 * len(self.meshes))()
        self.ctx.check(self.ctx.lib.flg_scene_accel_rank(self.h, r))
        return list(r)

    def trace(self, rays, hits=None):
        """rays: (n, 8) float32 CUDA tensor; hits: (n, 4) int32 (mesh, id,
        t bits, 0)."""
        dev = self.ctx.device
        check_buffer(rays, "rays", torch.float32, None, shape=(None, 8))
        n = rays.shape[0]
        if hits is None:
            hits = torch.empty((n, 4), dtype=torch.int32, device=rays.device)
        check_buffer(hits, "hits", _HIT_TYPES, None, min_numel=4 * n)
        check_device(dev, rays=rays, hits=hits)
        self.ctx.check(self.ctx.lib.flg_scene_trace_closest(self.ctx.h, self.h, _ptr(rays), _ptr(hits), n))
        return hits

    def close(self):
        if self.h:
            self.ctx.lib.flg_scene_accel_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def camera_rays(ctx, cam, width, height, spp=1, sample=0, out=None):
    if out is None:
        out = torch.empty((width * height, 8), dtype=torch.float32, device=torch.device("cuda", ctx.device))
    check_buffer(out, "out", torch.float32, ctx.device, min_numel=width * height * 8)
    ctx.check(ctx.lib.flg_camera_rays(ctx.h, ctypes.byref(cam), width, height, spp, sample, _ptr(out)))
    return out


def hits_as_numpy(hits):
    h = hits.cpu().numpy()
    return h[:, 0].copy(), h[:, 1].view(np.float32).copy()


class Volume:
    def __init__(self, ctx, density, bounds, sigma_s, sigma_a, g):
        self.ctx = ctx
        check_buffer(density, "density", torch.float32, ctx.device, ndim=3)
        nz, ny, nx = density.shape
        self.shape = (nz, ny, nx)
        self.n_leaves = None  # dense
        med = N.Medium(sigma_s, sigma_a, g)
        b = (ctypes.c_float * 6)(*[float(x) for x in bounds])
        self.h = ctypes.c_void_p()
        ctx.check(ctx.lib.flg_volume_create(ctx.h, _ptr(density), nx, ny, nz, b, ctypes.byref(med),
                                            ctypes.byref(self.h)))

    @classmethod
    def sparse(cls, ctx, origins, values, shape, bounds, sigma_s, sigma_a, g):
        """Sparse-occupancy grid of `shape` (nz, ny, nx): the 8^3 bricks with
        voxel origins `origins` ((n, 3) int32, x y z, multiples of 8) and
        values `values` ((n, 8, 8, 8) float32, z y x); every other voxel is 0
        (flg_volume_create_sparse).  `leaves_of` makes both from a dense
        grid."""
        nz, ny, nx = (int(x) for x in shape)
        check_buffer(origins, "origins", torch.int32, ctx.device, shape=(None, 3))
        n = origins.shape[0]
        check_buffer(values, "values", torch.float32, ctx.device, shape=(n, 8, 8, 8))
        self = cls.__new__(cls)
        self.ctx = ctx
        self.shape = (nz, ny, nx)
        self.n_leaves = n
        self.origins = origins
        med = N.Medium(sigma_s, sigma_a, g)
        b = (ctypes.c_float * 6)(*[float(x) for x in bounds])
        self.h = ctypes.c_void_p()
        ctx.check(ctx.lib.flg_volume_create_sparse(ctx.h, n, _ptr(origins) if n else None,
                                                   _ptr(values) if n else None, nx, ny, nz, b,
                                                   ctypes.byref(med), ctypes.byref(self.h)))
        return self

    def update(self, density):
        check_buffer(density, "density", torch.float32, self.ctx.device, shape=self.shape)
        self.ctx.check(self.ctx.lib.flg_volume_update(self.ctx.h, self.h, _ptr(density)))

    def update_sparse(self, values):
        """New brick values of a sparse volume (same bricks and order)."""
        check_buffer(values, "values", torch.float32, self.ctx.device, shape=(self.n_leaves, 8, 8, 8))
        self.ctx.check(self.ctx.lib.flg_volume_update_sparse(self.ctx.h, self.h,
                                                             _ptr(values) if self.n_leaves else None))

    def render(self, scene, width, height, spp, max_depth, out=None, stats=False, schedule=0, film=0,
               use_graph=False):
        """RGB image (height, width, 3).  film=1: the reference's render loop
        and Mitchell-filtered film instead of the per-pixel sample mean.
        schedule 3 (the queue-based wavefront) with use_graph replays its
        stage launches from a cached hipGraph."""
        cfg = N.RenderCfg()
        cfg.width, cfg.height, cfg.spp, cfg.max_depth = width, height, spp, max_depth
        cfg.schedule = schedule
        cfg.film = film
        cfg.use_graph = 1 if use_graph else 0
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device=torch.device("cuda", self.ctx.device))
        check_buffer(out, "out", torch.float32, self.ctx.device, min_numel=height * width * 3)
        st = (ctypes.c_uint64 * 2)()
        self.ctx.check(self.ctx.lib.flg_volpath_render(self.ctx.h, self.h, ctypes.byref(scene),
                                                       ctypes.byref(cfg), _ptr(out), st if stats else None))
        if stats:
            return out, {"paths": int(st[0]), "lookups": int(st[1])}
        return out

    def render_mesh(self, bvh, scene, width, height, spp, max_depth, out=None, stats=False, schedule=0, film=0):
        """VolPathIntegrator: `bvh`'s closed mesh as the interface around this
        volume's medium (placed on the mesh's bounding box); RGB image.
        film=1: the reference's render loop and Mitchell-filtered film."""
        cfg = N.RenderCfg()
        cfg.width, cfg.height, cfg.spp, cfg.max_depth = width, height, spp, max_depth
        cfg.schedule = schedule
        cfg.film = film
        if out is None:
            out = torch.empty((height, width, 3), dtype=torch.float32, device=torch.device("cuda", self.ctx.device))
        check_buffer(out, "out", torch.float32, self.ctx.device, min_numel=height * width * 3)
        st = (ctypes.c_uint64 * 5)()
        self.ctx.check(self.ctx.lib.flg_volpath_mesh_render(self.ctx.h, bvh.h, self.h, ctypes.byref(scene),
                                                            ctypes.byref(cfg), _ptr(out), st if stats else None))
        if stats:
            return out, {"paths": int(st[0]), "lookups": int(st[1]), "rays": int(st[2]), "node_visits": int(st[3]),
                         "tri_tests": int(st[4])}
        return out

    def close(self):
        if self.h:
            self.ctx.lib.flg_volume_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def leaves_of(density):
    """The non-zero 8^3 bricks of a dense (nz, ny, nx) float32 grid (any
    device): origins ((n, 3) int32, voxel x y z of each brick's first voxel)
    and values ((n, 8, 8, 8) float32, z y x), bricks in x-fastest brick order.
    A brick is kept iff one of its voxels has bits other than +0, so the
    sparse volume of the result equals the dense one bit for bit."""
    nz, ny, nx = density.shape
    bz, by, bx = (-(-n // 8) for n in (nz, ny, nx))
    full = torch.zeros((bz * 8, by * 8, bx * 8), dtype=torch.float32, device=density.device)
    full[:nz, :ny, :nx] = density
    bricks = full.reshape(bz, 8, by, 8, bx, 8).permute(0, 2, 4, 1, 3, 5).reshape(-1, 8, 8, 8)
    keep = (bricks.view(torch.int32) != 0).reshape(bricks.shape[0], -1).any(dim=1)
    idx = torch.nonzero(keep).flatten()
    x = idx % bx
    y = (idx // bx) % by
    z = idx // (bx * by)
    origins = (torch.stack([x, y, z], dim=1) * 8).to(torch.int32).contiguous()
    return origins, bricks[idx].contiguous()


def _rgb(v, name):
    """A light's colour: one number (grey) or three (RGB)."""
    try:
        c = [float(x) for x in v]
    except TypeError:
        c = [float(v)] * 3
    if len(c) != 3:
        raise ValueError("%s: one value or three (RGB)" % name)
    return c


def vol_scene(cam, point=(2.0, 2.0, -2.0), intensity=50.0, env=0.5, has_point=True, has_env=True):
    """Lights of the volume scenes: a point light of radiant intensity
    `intensity` and a constant environment of radiance `env`, each one value
    (grey) or three (RGB)."""
    s = N.VolScene()
    s.camera = cam
    s.has_point = 1 if has_point else 0
    s.point_pos[:] = [float(x) for x in point]
    s.point_intensity[:] = _rgb(intensity, "intensity")
    s.has_env = 1 if has_env else 0
    s.env_radiance[:] = _rgb(env, "env")
    return s


# ------------------------------------------------------------ generators ---
def gen_soup(ntris, initstate=42, initseq=54):
    L = N.scene()
    v = np.zeros(ntris * 9, np.float32)
    i = np.zeros(ntris * 3, np.int32)
    L.flg_scene_soup(initstate, initseq, ntris, v.ctypes.data, i.ctypes.data)
    return v, i


def gen_octasphere(level, amplitude, seed=7, base_freq=2.0):
    L = N.scene()
    nv = ctypes.c_int64()
    nt = L.flg_scene_octasphere_counts(level, ctypes.byref(nv))
    v = np.zeros(nv.value * 3, np.float32)
    i = np.zeros(nt * 3, np.int32)
    got = L.flg_scene_octasphere(level, amplitude, seed, base_freq, v.ctypes.data, i.ctypes.data)
    assert got == nv.value, (got, nv.value)
    return v, i


def gen_uv_sphere(n_theta, n_phi, radius):
    L = N.scene()
    nt = L.flg_scene_uv_sphere(n_theta, n_phi, radius, None, None)
    v = np.zeros(n_theta * n_phi * 3, np.float32)
    i = np.zeros(nt * 3, np.int32)
    L.flg_scene_uv_sphere(n_theta, n_phi, radius, v.ctypes.data, i.ctypes.data)
    return v, i


def gen_fbm_grid(n, seed=1, base_freq=4.0, octaves=5, bias=0.5, scale=1.0):
    L = N.scene()
    g = np.zeros((n, n, n), np.float32)
    L.flg_scene_fbm_grid(n, n, n, seed, base_freq, octaves, bias, scale, g.ctypes.data)
    return g


def gen_batch_sizes(nmeshes, lo=1024, hi=262144, initstate=42, initseq=56):
    L = N.scene()
    s = np.zeros(nmeshes, np.int32)
    L.flg_scene_batch_sizes(initstate, initseq, nmeshes, lo, hi, s.ctypes.data)
    return s
