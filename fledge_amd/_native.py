"""ctypes bindings of the native libraries (C ABI in include/fledge_hip.h and
include/fledge_scene.h).  Loading never falls back to anything: a missing
library raises immediately."""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libfledge_hip.so")
SCENE_LIB = os.path.join(HERE, "libfledge_scene.so")

c_int = ctypes.c_int
c_i32 = ctypes.c_int32
c_i64 = ctypes.c_int64
c_u32 = ctypes.c_uint32
c_u64 = ctypes.c_uint64
c_f = ctypes.c_float
c_vp = ctypes.c_void_p
c_size = ctypes.c_size_t
P = ctypes.POINTER

FLG_OK = 0
FLG_ERR_INVALID = 1
FLG_ERR_HIP = 2
FLG_ERR_NO_DEVICE = 3
FLG_ERR_OOM = 4
FLG_ERR_RANGE = 5
FLG_NORMAL_FIXED = 0
FLG_NORMAL_COMPAT = 1


class MeshDesc(ctypes.Structure):
    _fields_ = [("d_verts", c_vp), ("d_ind", c_vp), ("ntris", c_i32), ("nverts", c_i32)]


class LbvhInfo(ctypes.Structure):
    _fields_ = [("ntris", c_i32), ("bound", c_f * 6), ("device_bytes", c_u64)]


class LbvhView(ctypes.Structure):
    _fields_ = [(n, c_vp) for n in (
        "h_morton", "h_morton_sorted", "h_order", "h_split", "h_range_lo", "h_range_hi",
        "h_parent", "h_left_leaf", "h_right_leaf", "h_node_bound", "h_leaf_parent")]


class TraceCfg(ctypes.Structure):
    _fields_ = [("block_size", c_i32), ("stack_mode", c_i32), ("ring_depth", c_i32), ("reserved", c_i32 * 5)]


class Camera(ctypes.Structure):
    _fields_ = [("pos", c_f * 3), ("look_at", c_f * 3), ("up", c_f * 3), ("fov_deg", c_f)]

    @classmethod
    def make(cls, pos, look_at, up=(0.0, 1.0, 0.0), fov=45.0):
        c = cls()
        c.pos[:] = [float(x) for x in pos]
        c.look_at[:] = [float(x) for x in look_at]
        c.up[:] = [float(x) for x in up]
        c.fov_deg = float(fov)
        return c

    def as_list(self):
        return list(self.pos) + list(self.look_at) + list(self.up) + [self.fov_deg]


class PtCfg(ctypes.Structure):
    _fields_ = [("width", c_i32), ("height", c_i32), ("spp", c_i32), ("max_depth", c_i32),
                ("trace_block", c_i32), ("use_graph", c_i32), ("count_visits", c_i32),
                ("blocks_per_cu", c_i32), ("stack_mode", c_i32), ("ray_fetch", c_i32),
                ("ring_depth", c_i32), ("reserved", c_i32 * 1)]


class Tuning(ctypes.Structure):
    _fields_ = [("block_size", c_i32), ("blocks_per_cu", c_i32), ("ms", c_f), ("ms_default", c_f),
                ("candidates", c_i32), ("stack_mode", c_i32), ("ray_fetch", c_i32),
                ("ring_depth", c_i32)]


class Medium(ctypes.Structure):
    _fields_ = [("sigma_s", c_f), ("sigma_a", c_f), ("g", c_f)]


class VolScene(ctypes.Structure):
    _fields_ = [("camera", Camera), ("has_point", c_i32), ("point_pos", c_f * 3),
                ("point_intensity", c_f * 3), ("has_env", c_i32), ("env_radiance", c_f * 3)]


class RenderCfg(ctypes.Structure):
    _fields_ = [("width", c_i32), ("height", c_i32), ("spp", c_i32), ("max_depth", c_i32),
                ("schedule", c_i32), ("film", c_i32), ("use_graph", c_i32), ("reserved", c_i32 * 5)]


class DiffParams(ctypes.Structure):
    _fields_ = [("sigma_t_scale", c_f), ("albedo", c_f), ("g", c_f), ("majorant", c_f), ("d_lookups", c_vp)]


class PlyMesh(ctypes.Structure):
    _fields_ = [("nverts", c_i32), ("ntris", c_i32), ("verts", P(c_f)), ("ind", P(c_i32)), ("normals", P(c_f)),
                ("uv", P(c_f))]


class BrickLayout(ctypes.Structure):
    _fields_ = [("brick", c_i32), ("n_active", c_i32), ("nx", c_i32), ("ny", c_i32), ("nz", c_i32),
                ("d_slot", c_vp)]


_SIGS = {
    "flg_ctx_create": (c_int, [c_int, c_vp, P(c_vp)]),
    "flg_ctx_destroy": (c_int, [c_vp]),
    "flg_ctx_synchronize": (c_int, [c_vp]),
    "flg_ctx_stream": (c_vp, [c_vp]),
    "flg_last_error": (ctypes.c_char_p, [c_vp]),
    "flg_status_string": (ctypes.c_char_p, [c_int]),
    "flg_ctx_set_profiling": (c_int, [c_vp, c_int]),
    "flg_ctx_kernel_time": (c_int, [c_vp, ctypes.c_char_p, P(ctypes.c_double), P(c_i64)]),
    "flg_ctx_reset_profiling": (c_int, [c_vp]),
    "flg_malloc": (c_int, [c_vp, c_size, P(c_vp)]),
    "flg_free": (c_int, [c_vp, c_vp]),
    "flg_memcpy_htod": (c_int, [c_vp, c_vp, c_vp, c_size]),
    "flg_memcpy_dtoh": (c_int, [c_vp, c_vp, c_vp, c_size]),
    "flg_lbvh_build": (c_int, [c_vp, c_vp, c_vp, c_i32, c_i32, c_u32, P(c_vp)]),
    "flg_lbvh_build_batched": (c_int, [c_vp, P(MeshDesc), c_i32, c_u32, P(c_vp)]),
    "flg_lbvh_destroy": (c_int, [c_vp]),
    "flg_lbvh_get_info": (c_int, [c_vp, P(LbvhInfo)]),
    "flg_lbvh_export": (c_int, [c_vp, c_vp, P(LbvhView)]),
    "flg_lbvh_download": (c_int, [c_vp, c_vp, c_vp, c_size, c_vp, c_size]),
    "flg_trace_closest": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, P(TraceCfg)]),
    "flg_trace_count": (c_int, [c_vp, c_vp, c_vp, c_i64, c_vp]),
    "flg_autotune_trace": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64, P(Tuning)]),
    "flg_autotune_pathtrace": (c_int, [c_vp, c_vp, P(Camera), P(PtCfg), P(Tuning)]),
    "flg_camera_rays": (c_int, [c_vp, P(Camera), c_i32, c_i32, c_i32, c_i32, c_vp]),
    "flg_pathtrace_diffuse": (c_int, [c_vp, c_vp, P(Camera), P(PtCfg), c_vp, c_vp, c_vp, P(c_u64)]),
    "flg_volume_create": (c_int, [c_vp, c_vp, c_i32, c_i32, c_i32, P(c_f), P(Medium), P(c_vp)]),
    "flg_volume_destroy": (c_int, [c_vp]),
    "flg_volume_update": (c_int, [c_vp, c_vp, c_vp]),
    "flg_volume_create_sparse": (c_int, [c_vp, c_i32, c_vp, c_vp, c_i32, c_i32, c_i32, P(c_f), P(Medium), P(c_vp)]),
    "flg_volume_update_sparse": (c_int, [c_vp, c_vp, c_vp]),
    "flg_volume_brick_layout": (c_int, [c_vp, P(BrickLayout)]),
    "flg_volpath_render": (c_int, [c_vp, c_vp, P(VolScene), P(RenderCfg), c_vp, P(c_u64)]),
    "flg_volpath_render_noalloc": (c_int, [c_vp, c_vp, P(VolScene), P(RenderCfg), c_vp]),
    "flg_volpath_mesh_render": (c_int, [c_vp, c_vp, c_vp, P(VolScene), P(RenderCfg), c_vp, P(c_u64)]),
    "flg_volpath_diff_tape_bytes": (c_i64, [P(RenderCfg)]),
    "flg_volpath_diff_forward": (c_int, [c_vp, c_vp, c_vp, P(VolScene), P(RenderCfg), P(DiffParams),
                                         c_vp, c_vp, c_vp]),
    "flg_volpath_adjoint": (c_int, [c_vp, c_vp, c_vp, P(VolScene), P(RenderCfg), P(DiffParams), c_vp,
                                    c_vp, c_vp, c_vp, c_vp]),
    "flg_ply_load": (c_int, [ctypes.c_char_p, P(PlyMesh)]),
    "flg_ply_free": (None, [P(PlyMesh)]),
    "flg_scene_accel_build": (c_int, [c_vp, P(c_vp), c_i32, P(c_vp)]),
    "flg_scene_accel_destroy": (c_int, [c_vp]),
    "flg_scene_accel_rank": (c_int, [c_vp, P(c_i32)]),
    "flg_scene_trace_closest": (c_int, [c_vp, c_vp, c_vp, c_vp, c_i64]),
    "flg_bricks_build": (c_int, [c_vp, c_vp, c_i32, c_i32, c_i32, c_i32, c_vp, P(c_i32)]),
    "flg_volpath_mesh_diff_forward": (c_int, [c_vp, c_vp, c_vp, c_vp, P(VolScene), P(RenderCfg), P(DiffParams),
                                              c_vp, c_vp, c_vp]),
    "flg_volpath_mesh_adjoint": (c_int, [c_vp, c_vp, c_vp, c_vp, P(VolScene), P(RenderCfg), P(DiffParams), c_vp,
                                         c_vp, c_vp, P(BrickLayout), c_vp, c_vp]),
    "flg_debug_eval": (c_int, [c_vp, c_i32, c_vp, c_vp, c_i64]),
    "flg_film_save_exr": (c_int, [ctypes.c_char_p, c_i32, c_i32, c_vp]),
}

_SCENE_SIGS = {
    "flg_scene_pcg32": (None, [c_u64, c_u64, c_i64, c_vp]),
    "flg_scene_soup": (None, [c_u64, c_u64, c_i32, c_vp, c_vp]),
    "flg_scene_batch_sizes": (c_i32, [c_u64, c_u64, c_i32, c_i32, c_i32, c_vp]),
    "flg_scene_fbm_grid": (None, [c_i32, c_i32, c_i32, c_u32, c_f, c_i32, c_f, c_f, c_vp]),
    "flg_scene_octasphere_counts": (c_i64, [c_i32, P(c_i64)]),
    "flg_scene_octasphere": (c_i32, [c_i32, c_f, c_u32, c_f, c_vp, c_vp]),
    "flg_scene_uv_sphere": (c_i32, [c_i32, c_i32, c_f, c_vp, c_vp]),
}

_hip = None
_scene = None


def declared_symbols():
    return sorted(_SIGS)


def _bind(lib, sigs, strict=True):
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name, None)
        if fn is None:
            if strict:
                raise RuntimeError("%s does not export %s" % (lib._name, name))
            continue
        fn.restype = res
        fn.argtypes = args
    return lib


def hip():
    """The HIP library (raises if it is not built)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError("libfledge_hip.so is not built; run __graft_entry__.build()")
        _hip = _bind(ctypes.CDLL(HIP_LIB), _SIGS, strict=False)
    return _hip


def scene():
    global _scene
    if _scene is None:
        if not os.path.exists(SCENE_LIB):
            raise RuntimeError("libfledge_scene.so is not built; run __graft_entry__.build()")
        _scene = _bind(ctypes.CDLL(SCENE_LIB), _SCENE_SIGS)
    return _scene


class FlgError(RuntimeError):
    pass


def check(status, ctx=None):
    if status != FLG_OK:
        lib = hip()
        msg = lib.flg_last_error(ctx).decode() if ctx else ""
        raise FlgError("fledge: %s (%d): %s" % (lib.flg_status_string(status).decode(), status, msg))
