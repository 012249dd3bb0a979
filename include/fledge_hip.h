/*
 * fledge_hip.h — C ABI of the MI355X (gfx950) ray-tracing hot path:
 * radix-tree LBVH build, closest-hit traversal with the watertight Pluecker
 * test, primary-ray generation with the Halton sampler, the volume-only path
 * integrator (delta + ratio tracking, HG phase, next-event estimation) and
 * its path-replay adjoint; around them the volumetric path integrator over an
 * interface mesh (forward and adjoint over active bricks), the reference's
 * render loop and film, EXR output, PLY reading and the accel over a scene's
 * meshes.
 *
 * Conventions
 *  - Every function returns an int status (FLG_OK == 0); the message of the
 *    last failure on a context is returned by flg_last_error().  No C++
 *    exception crosses this boundary.
 *  - Buffers named d_* are device pointers (hipMalloc / torch CUDA tensors /
 *    flg_malloc); the caller owns them.  Buffers named h_* are host pointers.
 *    Handles (flg_bvh, flg_volume) own the device memory they allocate and
 *    free it in *_destroy.
 *  - All work of a context is issued on its stream, in call order; calls on
 *    one context are not thread-safe, different contexts may run
 *    concurrently.  Functions returning results into host memory synchronise
 *    the stream; the others are asynchronous.
 *  - Floating point follows the reference's operation order with no
 *    contraction; hit ids, Morton codes, sort order and tree topology are
 *    bit-exact with the CPU restatement of fledge's builder and traversal.
 *
 * Reference interfaces replaced (paths relative to the fledge source tree):
 *  - flg_lbvh_build / flg_lbvh_export: RadixBVHBuilder::build
 *    (src/experimental/radix_bvh.cpp:18-24, prestage :32-115,
 *    parallelBuilder :117-200), BVHBuilderBase::build
 *    (src/experimental/base_bvh.hpp:112); the C++ LBVHAccel ctor
 *    (drop-in for NaiveBVHAccel, src/accel.cpp:35-86).
 *  - flg_trace_closest: RadixBVHBuilder::intersect / recursiveIntersect
 *    (radix_bvh.cpp:26-30, 202-245) with PlueckerTriangleIntersect
 *    (src/experimental/intersector.hpp:63-101); Accel::intersect
 *    (src/accel.hpp:12-21, src/primitive.hpp:31-42) batched.
 *  - flg_camera_rays: Camera::generateRay (src/common/camera.h:21-31) with
 *    HaltonSampler::getPixelSample (src/common/sampler.h:183-186).
 *  - flg_volume_create / flg_volpath_render: OpenVDBVolume ctor, tr, sample
 *    (src/volume.cpp:33-148) and SVolIntegrator::Li driven per pixel sample
 *    (src/integrator.cpp:123-235, 328-405).
 *  - flg_pathtrace_diffuse: PathIntegrator::Li's segment loop
 *    (src/integrator.cpp:252-326) over a diffuse mesh.
 *  - flg_volpath_mesh_render: VolPathIntegrator::Li (src/integrator.cpp:
 *    407-510) with VolEstimateTr (:40-64) over one interface mesh.
 *  - flg_film_save_exr: Film::saveBuffer (src/film.cpp:13-49).
 *  - flg_ply_load: MakeTriangleMesh (src/plymesh.cpp:19-86).
 *  - flg_scene_accel_build / flg_scene_trace_closest: Scene::init's accel and
 *    Scene::intersect (src/scene.cpp:46-59) over the scene's MeshPrimitives.
 */
#ifndef FLEDGE_HIP_H_
#define FLEDGE_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(_WIN32)
#define FLG_API __declspec(dllexport)
#else
#define FLG_API __attribute__((visibility("default")))
#endif

/* ------------------------------------------------------------- status --- */
enum {
  FLG_OK = 0,
  FLG_ERR_INVALID = 1,   /* bad argument (null pointer, negative size, ...) */
  FLG_ERR_HIP = 2,       /* a HIP runtime call failed */
  FLG_ERR_NO_DEVICE = 3, /* no usable GPU */
  FLG_ERR_OOM = 4,       /* device allocation failed */
  FLG_ERR_RANGE = 5      /* input beyond a documented limit */
};

typedef struct flg_ctx flg_ctx;
typedef struct flg_bvh flg_bvh;
typedef struct flg_volume flg_volume;

/* ------------------------------------------------------------ context --- */
/* Creates a context on HIP device `device` issuing all work on `stream`
 * (a hipStream_t; NULL selects the device's default stream, which orders the
 * work with other default-stream users such as torch's default stream). */
FLG_API int flg_ctx_create(int device, void* stream, flg_ctx** out);
FLG_API int flg_ctx_destroy(flg_ctx* ctx);
FLG_API int flg_ctx_synchronize(flg_ctx* ctx);
FLG_API void* flg_ctx_stream(flg_ctx* ctx);
FLG_API const char* flg_last_error(flg_ctx* ctx);
FLG_API const char* flg_status_string(int status);

/* Per-kernel timing with HIP events on the context stream (off by default). */
FLG_API int flg_ctx_set_profiling(flg_ctx* ctx, int enable);
/* Total milliseconds and launch count recorded for kernel `name` since the
 * last reset; returns FLG_ERR_INVALID if the name is unknown. */
FLG_API int flg_ctx_kernel_time(flg_ctx* ctx, const char* name, double* ms, int64_t* launches);
FLG_API int flg_ctx_reset_profiling(flg_ctx* ctx);

/* Device memory helpers for callers without HIP headers. */
FLG_API int flg_malloc(flg_ctx* ctx, size_t bytes, void** d_out);
FLG_API int flg_free(flg_ctx* ctx, void* d_ptr);
FLG_API int flg_memcpy_htod(flg_ctx* ctx, void* d_dst, const void* h_src, size_t bytes);
FLG_API int flg_memcpy_dtoh(flg_ctx* ctx, void* h_dst, const void* d_src, size_t bytes);

/* --------------------------------------------------------------- LBVH --- */
/* Normal policy of the watertight test (StableNormalize, vector.h:495-504). */
#define FLG_NORMAL_FIXED 0u  /* x^2 + y^2 + z^2 */
#define FLG_NORMAL_COMPAT 1u /* the reference's x^2 + y^2 + z + z */

/* Builds a Karras radix-tree BVH over an indexed triangle mesh.
 * d_verts: nverts * 3 floats (x,y,z); d_ind: ntris * 3 int32.
 * flags: FLG_NORMAL_* used by later traces of this BVH.
 * Limits: 0 <= ntris < 2^30; ntris == 0 gives an empty BVH (all misses). */
FLG_API int flg_lbvh_build(flg_ctx* ctx, const float* d_verts, const int32_t* d_ind, int32_t ntris,
                           int32_t nverts, uint32_t flags, flg_bvh** out);
FLG_API int flg_lbvh_destroy(flg_bvh* bvh);

typedef struct flg_mesh_desc {
  const float* d_verts;
  const int32_t* d_ind;
  int32_t ntris;
  int32_t nverts;
} flg_mesh_desc;

/* Builds n independent BVHs in one batch (one launch per build stage for all
 * meshes).  Each out[i] equals flg_lbvh_build on mesh i alone, bit for bit. */
FLG_API int flg_lbvh_build_batched(flg_ctx* ctx, const flg_mesh_desc* h_meshes, int32_t n,
                                   uint32_t flags, flg_bvh** out);

typedef struct flg_lbvh_info {
  int32_t ntris;
  float bound[6]; /* global bound lo.xyz, hi.xyz */
  uint64_t device_bytes;
} flg_lbvh_info;
FLG_API int flg_lbvh_get_info(flg_bvh* bvh, flg_lbvh_info* out);

/* Copies the build to host arrays for parity checks; any pointer may be NULL.
 * Node arrays hold ntris-1 entries (Karras node i), leaf arrays ntris
 * (sorted position). */
typedef struct flg_lbvh_view {
  uint32_t* h_morton;      /* codes in original triangle order */
  uint32_t* h_morton_sorted;
  uint32_t* h_order;       /* sorted position -> original triangle */
  int32_t* h_split;        /* gamma */
  int32_t* h_range_lo;     /* min(i, j) */
  int32_t* h_range_hi;     /* max(i, j) */
  int32_t* h_parent;       /* parent of internal node (root: 0) */
  uint8_t* h_left_leaf;
  uint8_t* h_right_leaf;
  float* h_node_bound;     /* 6 per node: lo.xyz, hi.xyz */
  int32_t* h_leaf_parent;
} flg_lbvh_view;
FLG_API int flg_lbvh_export(flg_ctx* ctx, flg_bvh* bvh, const flg_lbvh_view* view);

/* Copies the traversal image of a build to host memory, for host mirrors that
 * run the same closest-hit traversal on the CPU (the C++ Accel drop-in's
 * single-ray intersect; replaces the CPU RadixBVHBuilder tree,
 * src/experimental/radix_bvh.hpp:35-57).
 *   h_nodes: (ntris-1) x 64 bytes, node i = Karras node i:
 *            float L.lo[3], L.hi[3], R.lo[3], R.hi[3]; int32 refL, refR, 0, 0
 *            (ref >= 0: internal node; bit 31 set: leaf at sorted position
 *            ref & 0x7fffffff)
 *   h_tris:  ntris x 48 bytes in sorted order: float v0[3], uint32 original id,
 *            float v1[3], 0, float v2[3], 0
 * The byte counts must equal these sizes; NULL skips an array. */
FLG_API int flg_lbvh_download(flg_ctx* ctx, const flg_bvh* bvh, void* h_nodes, size_t nodes_bytes,
                              void* h_tris, size_t tris_bytes);

/* -------------------------------------------------------------- trace --- */
typedef struct flg_ray {
  float ox, oy, oz, tmax; /* tmax: hits with t <= tmax are accepted */
  float dx, dy, dz, pad;
} flg_ray; /* 32 bytes */

typedef struct flg_hit {
  int32_t id; /* original triangle index, -1 on miss */
  float t;    /* hit distance, tmax on miss */
} flg_hit;    /* 8 bytes */

typedef struct flg_trace_cfg {
  int32_t block_size; /* 64, 128, 256 or 512 threads; 0: the context's tuned
                         value (flg_autotune_trace), else 256 */
  int32_t stack_mode; /* traversal stack: 1 private memory, 2 the ring_depth
                         most recent entries per ray in LDS; 0: tuned value
                         (with its ring depth), else 1 */
  int32_t ring_depth; /* stack_mode 2: 4, 8, 16 or 32 entries in LDS (0: 8);
                         block_size * ring_depth * 8 must not exceed 65536 */
  int32_t reserved[5];
} flg_trace_cfg;

/* Closest hit of n rays; the visit order and tie rule are those of
 * RadixBVHBuilder::recursiveIntersect. cfg may be NULL. */
FLG_API int flg_trace_closest(flg_ctx* ctx, const flg_bvh* bvh, const flg_ray* d_rays,
                              flg_hit* d_hits, int64_t n, const flg_trace_cfg* cfg);

/* Counts node visits and triangle tests of the same traversal (for the
 * roofline's algorithmic bytes); d_counts: 2 uint64 (nodes, triangles). */
FLG_API int flg_trace_count(flg_ctx* ctx, const flg_bvh* bvh, const flg_ray* d_rays, int64_t n,
                            uint64_t* d_counts);

/* ---------------------------------------------------------------- PLY --- */
/* MakeTriangleMesh (src/plymesh.cpp:19-86): a PLY file's "vertex" element
 * (x, y, z; nx, ny, nz and a texture-coordinate pair when present) and its
 * "face" element's "vertex_indices", three per face.  Formats ascii,
 * binary_little_endian and binary_big_endian 1.0, any scalar type.  Returns
 * FLG_ERR_INVALID for a missing or malformed file, a face that is not a
 * triangle or an index outside the vertex range; FLG_ERR_RANGE beyond 2^31
 * vertices or 2^30 faces.  Host memory, released by flg_ply_free; needs no
 * context or GPU. */
typedef struct flg_ply_mesh {
  int32_t nverts, ntris;
  float* verts;   /* nverts * 3 */
  int32_t* ind;   /* ntris * 3 */
  float* normals; /* nverts * 3, or NULL */
  float* uv;      /* nverts * 2, or NULL */
} flg_ply_mesh;
FLG_API int flg_ply_load(const char* path, flg_ply_mesh* out);
FLG_API void flg_ply_free(flg_ply_mesh* mesh);

/* ------------------------------------------------------- scene accel --- */
/* The scene's accel over its meshes: Scene::init's NaiveBVHAccel over the
 * scene's MeshPrimitives (src/scene.cpp:46, src/accel.cpp:35-133), each
 * mesh's own accel (src/primitive.cpp:54-90) being its LBVH.  Built on the
 * host from the meshes' BVHs (flg_lbvh_build or flg_lbvh_build_batched, same
 * context, same FLG_NORMAL_COMPAT), which must outlive it.  Mesh m is
 * h_meshes[m], the scene's order. */
typedef struct flg_scene_accel flg_scene_accel;
FLG_API int flg_scene_accel_build(flg_ctx* ctx, flg_bvh* const* h_meshes, int32_t n, flg_scene_accel** out);
FLG_API int flg_scene_accel_destroy(flg_scene_accel* s);
/* h_rank[m]: mesh m's position in the order in which NaiveBVHAccel asks its
 * leaves (the scene's order up to 64 meshes; beyond, its median splits'
 * order, by std::sort as the reference's). */
FLG_API int flg_scene_accel_rank(const flg_scene_accel* s, int32_t* h_rank);

typedef struct flg_scene_hit {
  int32_t mesh; /* scene index of the mesh hit, -1 on miss */
  int32_t id;   /* original triangle index within that mesh, -1 on miss */
  float t;      /* hit distance, tmax on miss */
  int32_t pad;
} flg_scene_hit; /* 16 bytes */

/* Closest hit of n rays over the scene (Scene::intersect, src/scene.cpp:57):
 * the closest over the meshes of each mesh's flg_trace_closest hit; between
 * meshes at the same distance, the one of lowest rank. */
FLG_API int flg_scene_trace_closest(flg_ctx* ctx, const flg_scene_accel* s, const flg_ray* d_rays,
                                    flg_scene_hit* d_hits, int64_t n);

typedef struct flg_camera {
  float pos[3], look_at[3], up[3];
  float fov_deg;
} flg_camera;

/* Primary rays of sample `sample` (Halton index offset(pixel) + sample *
 * stride, dimensions 0 and 1 for the pixel jitter) for a W x H image,
 * pixel-major: ray[y * W + x].  spp sets the sampler's sample count. */
FLG_API int flg_camera_rays(flg_ctx* ctx, const flg_camera* cam, int32_t width, int32_t height,
                            int32_t spp, int32_t sample, flg_ray* d_rays);

/* Wavefront path tracer over a diffuse mesh with no lights: every path
 * (pixel, sample) traces its camera ray and up to max_depth cosine-weighted
 * bounces.  Optional per-path outputs (path index = pixel * spp + s):
 * d_nseg[paths], d_seg_id[paths * (max_depth+1)], d_seg_t[same].
 * h_stats (may be NULL) receives {total segments, paths}, followed by
 * {node visits, triangle tests} when count_visits is set (4 entries). */
typedef struct flg_pt_cfg {
  int32_t width, height, spp, max_depth;
  int32_t trace_block;   /* 64, 128, 256 or 512 threads per block; 0: the
                            context's tuned value (flg_autotune_pathtrace), else 256 */
  int32_t use_graph;     /* replay the bounce loop from a captured hipGraph
                            (captured on first use, cached per argument set) */
  int32_t count_visits;  /* count traversal work (slower; for the roofline) */
  int32_t blocks_per_cu; /* persistent grid of each bounce = CUs x this (1..64);
                            0: tuned value, else 8 */
  int32_t stack_mode;    /* as flg_trace_cfg.stack_mode */
  int32_t ray_fetch;     /* 1: one fused launch per bounce, each lane traces
                            one queue entry per loop trip; 2: bounces after
                            the first in two launches, a traversal whose lanes
                            take the next entry whenever enough of a wave's
                            lanes are idle (persistent threads), then the
                            shading; 3: every bounce after the first in one
                            launch, each lane following its path to the end
                            and taking a new one when enough of its wave's
                            lanes are idle; 4: the second segment as 2, every
                            later one as 3; 0: tuned value, else 1.  Same
                            results. */
  int32_t ring_depth;    /* as flg_trace_cfg.ring_depth */
  int32_t reserved[1];
} flg_pt_cfg;
FLG_API int flg_pathtrace_diffuse(flg_ctx* ctx, const flg_bvh* bvh, const flg_camera* cam,
                                  const flg_pt_cfg* cfg, int32_t* d_nseg, int32_t* d_seg_id,
                                  float* d_seg_t, uint64_t* h_stats);

/* Autotuner (no counterpart in the reference, which traces one ray per call):
 * times the traversal launch shapes on the caller's own workload and keeps
 * the fastest in the context, where calls whose block_size / trace_block /
 * blocks_per_cu are 0 pick it up.  The tuned calls run the real work: their
 * outputs are written as by a normal call.  Tuned: threads per block, resident
 * blocks per CU and ray fetch (path tracer), where the traversal stack lives
 * and how many of its entries stay in LDS (4, 8, 16 or 32).  The stack's
 * total depth (64, the bound on the tree depth) and node width (binary, which
 * the reference's visiting order needs) are fixed. */
typedef struct flg_tuning {
  int32_t block_size;    /* chosen threads per block */
  int32_t blocks_per_cu; /* chosen persistent blocks per CU (path tracer; 0 for trace) */
  float ms;              /* one call with the chosen shape */
  float ms_default;      /* one call with the untuned shape (256 threads, 8 blocks per CU) */
  int32_t candidates;    /* shapes timed */
  int32_t stack_mode;    /* chosen traversal stack (1 private, 2 LDS ring) */
  int32_t ray_fetch;     /* chosen ray fetch of the path tracer (1..4; 0 for trace) */
  int32_t ring_depth;    /* chosen LDS ring depth (stack_mode 2), else 0 */
} flg_tuning;
FLG_API int flg_autotune_trace(flg_ctx* ctx, const flg_bvh* bvh, const flg_ray* d_rays,
                               flg_hit* d_hits, int64_t n, flg_tuning* out);
FLG_API int flg_autotune_pathtrace(flg_ctx* ctx, const flg_bvh* bvh, const flg_camera* cam,
                                   const flg_pt_cfg* cfg, flg_tuning* out);

/* ------------------------------------------------------------- volume --- */
typedef struct flg_medium {
  float sigma_s, sigma_a; /* sigma_t = sigma_a + sigma_s */
  float g;                /* Henyey-Greenstein asymmetry */
} flg_medium;

/* Dense nx*ny*nz density grid (x fastest) over the box bounds[6]
 * (lo.xyz, hi.xyz).  The majorant is sigma_t * max(density). */
FLG_API int flg_volume_create(flg_ctx* ctx, const float* d_density, int32_t nx, int32_t ny,
                              int32_t nz, const float* h_bounds, const flg_medium* medium,
                              flg_volume** out);
FLG_API int flg_volume_destroy(flg_volume* vol);
/* Replaces the density values (same dimensions); recomputes the majorant. */
FLG_API int flg_volume_update(flg_ctx* ctx, flg_volume* vol, const float* d_density);

/* Sparse-occupancy density grid: the leaf level of the OpenVDB tree that the
 * reference reads through a grid accessor (OpenVDBVolume ctor and
 * OpenVDBVolume::tr/sample, src/volume.cpp:33-59,75-148): n_leaves active
 * 8^3 bricks.  d_leaf_origin (device, 3 * n_leaves int32): voxel (x, y, z) of
 * each brick's first voxel, multiples of 8 inside the grid, no two alike;
 * d_leaf_values (device, 512 * n_leaves floats): each brick's voxels, x
 * fastest (voxels past the grid's edge are ignored).  Every voxel in no
 * brick is 0.  Every render and adjoint entry point takes the result, and
 * gives the bits of the dense volume holding the same values.  Device memory:
 * 4000 B per brick (the brick with a one-voxel apron, so that a lookup reads
 * one brick) plus 32 B per 8^3 cell of the grid (the slot map, one 4-byte
 * read per lookup); n_leaves * 1000 and the cell count * 8 must stay below
 * 2^31 (FLG_ERR_RANGE).  FLG_ERR_INVALID for a bad or repeated origin. */
FLG_API int flg_volume_create_sparse(flg_ctx* ctx, int32_t n_leaves, const int32_t* d_leaf_origin,
                                     const float* d_leaf_values, int32_t nx, int32_t ny, int32_t nz,
                                     const float* h_bounds, const flg_medium* medium, flg_volume** out);
/* New values for a sparse volume's bricks (same bricks, same order);
 * recomputes the majorant.  flg_volume_update refuses a sparse volume, and
 * this call a dense one. */
FLG_API int flg_volume_update_sparse(flg_ctx* ctx, flg_volume* vol, const float* d_leaf_values);

/* Lights of the volume scenes, RGB: every channel runs the reference's
 * Vector3f arithmetic (component-wise), so a grey light gives three equal
 * channels.  The point light's Li is I / d^2 with pdf 1; the environment is
 * InfiniteAreaLight over a ConstTexture of colour env_radiance
 * (light.cpp:93-159), whose sampling tables depend on the colour's norm. */
typedef struct flg_vol_scene {
  flg_camera camera;
  int32_t has_point;
  float point_pos[3];
  float point_intensity[3]; /* RGB radiant intensity */
  int32_t has_env;
  float env_radiance[3]; /* constant environment, RGB */
} flg_vol_scene;

typedef struct flg_render_cfg {
  int32_t width, height, spp, max_depth;
  int32_t schedule; /* 1: one lane per path sample; 2: persistent lanes that
                       take a new path when theirs ends; 3 (flg_volpath_render
                       only): queue-based wavefront, the path state in device
                       memory between stages (primary walk, then collision
                       and shadow walk per bounce), about 80 B per path
                       sample of scratch; 0: 2.  Same results. */
  int32_t film;     /* flg_volpath_render only.  0: pixel = mean of spp path
                       samples, sample s on Halton index offset + s*stride
                       (eval_pixel's per-sample value).  1: the reference's
                       render loop and film (ParallelIntegrator::render,
                       integrator.cpp:136-159; Film::addSample/commitSamples,
                       film.hpp:143-174): a pixel-centre path that is not
                       splatted, then samples 1..spp-1 (sample 1 continues
                       the centre path's Halton index and dimensions, sample
                       s >= 2 uses index s-1) splatted with the Mitchell
                       filter as constructed there (radius 2, B = C = 0) and
                       resolved as sum(w L) / sum(w); race-free and
                       deterministic.  SVolIntegrator's albedo and normal
                       buffers are zero, so none is produced. */
  int32_t use_graph; /* schedule 3: the stage launches are replayed from a
                        hipGraph cached in the context (captured on first
                        use for the same arguments); ignored otherwise */
  int32_t reserved[5];
} flg_render_cfg;

/* Renders width*height RGB float pixels (d_rgb, 3 floats per pixel) with the
 * volume-only integrator (SVolIntegrator::Li, integrator.cpp:328-405).
 * h_stats (may be NULL): {path samples traced, density lookups}. */
FLG_API int flg_volpath_render(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* scene,
                               const flg_render_cfg* cfg, float* d_rgb, uint64_t* h_stats);

/* flg_volpath_render without any device allocation (no counterpart in the
 * reference; for callers that cannot allocate: code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as flg_volpath_render (no statistics).  Its scratch (path counters
 * and per-path RGB radiance) is a fixed array of 192 MiB + 4 KiB per GPU,
 * held for the life of the process: frames of more than 16,777,535 paths
 * (w*h*spp; film = 1: more than 10,066,521 tasks, w*h*(spp-1)) are refused
 * with FLG_ERR_RANGE.
 * The context's Halton tables come from flg_ctx_create; a scene with an
 * environment light needs the volume's environment table built for the
 * norm of the scene's radiance by an earlier flg_volpath_render of that
 * volume (FLG_ERR_INVALID otherwise).  Every call on one GPU shares the array, so
 * calls that overlap on one GPU race: serialise them (one stream). */
FLG_API int flg_volpath_render_noalloc(flg_ctx* ctx, flg_volume* vol, const flg_vol_scene* scene,
                                       const flg_render_cfg* cfg, float* d_rgb);

/* Surfaces and media together: VolPathIntegrator::Li (integrator.cpp:407-510)
 * for the closed mesh `bvh`, whose triangles are index-matched interfaces
 * (no material), with `vol`'s medium inside it, filling the mesh's bounding
 * box as ShapePrimitive::setBound places it (primitive.cpp:17-24); point and
 * environment lights from `scene`.  An interface is crossed in the ray's
 * direction (SInteraction::SpawnRay, interaction.cpp:16-26) without counting
 * a bounce, where the reference dereferences the missing material
 * (integrator.cpp:433,487,501); shadow rays cross interfaces as
 * VolEstimateTr does (integrator.cpp:40-64).  cfg->film 0: per-pixel mean
 * of the path samples on the per-sample Halton schedule; 1: the reference's
 * render loop and film, as flg_volpath_render's film mode.  h_stats (may be
 * NULL): {path samples, density lookups, rays traced on the mesh (camera
 * rays, continuations after a scattering or an interface, shadow-ray
 * segments), their node visits, their triangle tests}, five entries; a call
 * with h_stats runs the counting build of the kernels (slower). */
FLG_API int flg_volpath_mesh_render(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol,
                                    const flg_vol_scene* scene, const flg_render_cfg* cfg, float* d_rgb,
                                    uint64_t* h_stats);

/* Differentiable forward of the same path structure (differential delta
 * tracking on the proposal grid, ratio tracking without roulette on camera
 * and shadow rays, fixed majorant): sigma_t = sigma_t_scale * density,
 * sigma_s = albedo * sigma_t.  Every branch depends only on the random
 * numbers and the proposal grid, so the image is a smooth function of `vol`'s
 * densities.  The main path draws the Halton dimensions; each ratio-tracking
 * walk draws from a counter-based stream of its own (path, walk), so it can
 * stop once its result is fixed.  `proposal` (NULL: `vol` itself) supplies
 * the collision decisions.  d_rgb: width*height RGB (may be NULL); d_lpath:
 * per-path RGB radiance, width*height*spp*3 floats, path = pixel * spp + s
 * (may be NULL; the adjoint needs it); d_tape: walk tape of
 * flg_volpath_diff_tape_bytes(cfg) bytes (may be NULL), the measurement of
 * every ratio-tracking walk, which spares the adjoint one of its two walks
 * per ratio track (no reference counterpart). */
typedef struct flg_diff_params {
  float sigma_t_scale; /* sigma_t per unit density */
  float albedo;        /* single-scatter albedo alpha, > 0 */
  float g;             /* Henyey-Greenstein asymmetry */
  float majorant;      /* fixed majorant per unit length, > 0 */
  unsigned long long* d_lookups; /* NULL, or a device counter to which the call
                                    adds its density-grid lookups (a
                                    measurement pass; slower) */
} flg_diff_params;

/* Bytes of the walk tape of flg_volpath_diff_forward for cfg:
 * width*height*spp * 4 * (max_depth + 3); -1 if cfg is invalid or
 * max_depth > 63 (a tape holds at most 64 walks per path). */
FLG_API int64_t flg_volpath_diff_tape_bytes(const flg_render_cfg* cfg);

FLG_API int flg_volpath_diff_forward(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal,
                                     const flg_vol_scene* scene, const flg_render_cfg* cfg,
                                     const flg_diff_params* params, float* d_rgb, float* d_lpath,
                                     void* d_tape);
/* Path-replay adjoint of flg_volpath_diff_forward (same arguments; d_lpath,
 * and d_tape when not NULL, from that call; with d_tape NULL every ratio
 * track is walked twice): given dLoss/dImage (RGB, d_dl_drgb) adds
 * dLoss/dDensity into d_grad (nx*ny*nz floats, x fastest) and dLoss/dAlbedo
 * into d_grad_albedo (1 float).  The replay prices each light in the loss's
 * units (sum over channels of dLoss/dL_c times the light's channel c), so
 * one replay covers all three channels.  Accumulation uses float atomics, so the
 * summation order (not the terms) varies between runs. */
FLG_API int flg_volpath_adjoint(flg_ctx* ctx, flg_volume* vol, const flg_volume* proposal,
                                const flg_vol_scene* scene, const flg_render_cfg* cfg,
                                const flg_diff_params* params, const float* d_lpath,
                                const void* d_tape, const float* d_dl_drgb, float* d_grad,
                                float* d_grad_albedo);

/* Gradient over the active bricks of a grid (no reference counterpart; a
 * sparse-occupancy grid whose empty bricks are not parameters).  The grid's
 * bricks are cubes of `brick` voxels per edge, ceil(n / brick) per axis,
 * x fastest; d_slot[brick] is the brick's index among the active bricks, -1
 * for an inactive one.  A gradient in this layout is n_active * brick^3
 * floats: active brick after active brick in slot order, each x fastest.
 * Gradient terms that fall on an inactive brick are dropped. */
typedef struct flg_brick_layout {
  int32_t brick;          /* 1, 2, 4, 8, 16 or 32 */
  int32_t n_active;
  int32_t nx, ny, nz;     /* the grid the layout was built for (checked) */
  const int32_t* d_slot;  /* device, one per brick, as flg_bricks_build writes it */
} flg_brick_layout;

/* Active bricks of a dense device grid (nx*ny*nz floats, x fastest): a brick
 * is active iff one of its voxels is non-zero; slots in brick order.  Writes
 * d_slot (ceil(nx/b)*ceil(ny/b)*ceil(nz/b) int32, device) and the number of
 * active bricks (host). */
FLG_API int flg_bricks_build(flg_ctx* ctx, const float* d_density, int32_t nx, int32_t ny, int32_t nz,
                             int32_t brick, int32_t* d_slot, int32_t* h_n_active);

/* The brick layout of a sparse volume's bricks (brick 8, slot = the brick's
 * index in flg_volume_create_sparse's leaf list), for gradients over exactly
 * those bricks (shaped as the leaf values).  d_slot belongs to the volume and
 * stays valid while it lives.  FLG_ERR_INVALID for a dense volume. */
FLG_API int flg_volume_brick_layout(const flg_volume* vol, flg_brick_layout* out);

/* The differentiable estimator of flg_volpath_diff_forward over the scene of
 * flg_volpath_mesh_render: the mesh in `bvh` is the index-matched interface
 * around the medium, which fills the mesh's bounding box
 * (VolPathIntegrator::Li, src/integrator.cpp:407-510, with the reference's
 * draw order: light sample before the phase sample, no roulette).  Each
 * shadow ray's ratio track runs through the interfaces (VolEstimateTr,
 * src/integrator.cpp:40-64) on one walk stream; walk b is the shadow ray of
 * bounce b (b >= 1).  The tape has flg_volpath_diff_tape_bytes(cfg) bytes. */
FLG_API int flg_volpath_mesh_diff_forward(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol,
                                          const flg_volume* proposal, const flg_vol_scene* scene,
                                          const flg_render_cfg* cfg, const flg_diff_params* params,
                                          float* d_rgb, float* d_lpath, void* d_tape);
/* Path-replay adjoint of flg_volpath_mesh_diff_forward (same arguments; its
 * d_lpath, and d_tape when not NULL).  d_grad is dense (nx*ny*nz floats)
 * when bricks is NULL, else in the layout of *bricks.  Adds into d_grad and
 * d_grad_albedo with float atomics. */
FLG_API int flg_volpath_mesh_adjoint(flg_ctx* ctx, const flg_bvh* bvh, flg_volume* vol,
                                     const flg_volume* proposal, const flg_vol_scene* scene,
                                     const flg_render_cfg* cfg, const flg_diff_params* params,
                                     const float* d_lpath, const void* d_tape, const float* d_dl_drgb,
                                     const flg_brick_layout* bricks, float* d_grad, float* d_grad_albedo);

/* --------------------------------------------------------------- film --- */
/* Film::saveBuffer for a ".exr" name (src/film.cpp:13-49): writes h_rgb
 * (host memory, height rows of width RGB floats, row 0 first) as an OpenEXR
 * scanline image of FLOAT channels B, G, R, uncompressed, with the last row of
 * h_rgb as the file's first scanline, as the reference's negative-stride
 * write.  Returns FLG_ERR_INVALID if the file cannot be written.  Needs no
 * context or GPU. */
FLG_API int flg_film_save_exr(const char* path, int32_t width, int32_t height, const float* h_rgb);

/* ------------------------------------------------------------- checks --- */
/* TEST HOOK, not part of the rendering interface and with no counterpart in
 * the reference: evaluates the device copies of the path's scalar helpers so
 * that tests can compare them bit for bit with the CPU restatement.
 * kind 0: flg_logf(x); 1: sin; 2: cos; 3: Halton get1D with x = (index, dim)
 * packed as two uint32 in the input; 4: the ratio-tracking walks' draw
 * (1 - u) with x = (path, walk, n) packed as three uint32.  A build with
 * -DFLG_NO_TEST_HOOKS leaves it out of the library. */
#ifndef FLG_NO_TEST_HOOKS
FLG_API int flg_debug_eval(flg_ctx* ctx, int32_t kind, const void* d_in, void* d_out, int64_t n);
#endif

#ifdef __cplusplus
}
#endif
#endif /* FLEDGE_HIP_H_ */
