/*
 * fledge_scene.h — host-side synthetic workload generators
 * (libfledge_scene.so).  They produce inputs for the benchmarks and tests;
 * no reference interface corresponds (the reference loads PLY/VDB files that
 * do not ship, src/plymesh.cpp:19-86, src/volume.cpp:33-71).
 */
#ifndef FLEDGE_SCENE_H_
#define FLEDGE_SCENE_H_
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* n raw PCG32 outputs of pcg32_srandom(initstate, initseq). */
void flg_scene_pcg32(uint64_t initstate, uint64_t initseq, int64_t n, uint32_t* out);

/* Random triangle soup: per triangle 3 centroid draws in [0,1)^3, then 9
 * vertex offsets u * 0.02 - 0.01; verts[9 * ntris], ind[3 * ntris] = 0..3n-1. */
void flg_scene_soup(uint64_t initstate, uint64_t initseq, int32_t ntris, float* verts,
                    int32_t* ind);

/* Log-uniform mesh sizes in [lo, hi]: round(lo * (hi/lo)^u). */
int32_t flg_scene_batch_sizes(uint64_t initstate, uint64_t initseq, int32_t nmeshes, int32_t lo,
                              int32_t hi, int32_t* sizes);

/* nx*ny*nz grid (x fastest) of clamp(bias + scale * fBm(p), 0, 1) at voxel
 * centres of the unit cube; 5-octave value noise, gain 0.5, lacunarity 2. */
void flg_scene_fbm_grid(int32_t nx, int32_t ny, int32_t nz, uint32_t seed, float base_freq,
                        int32_t octaves, float bias, float scale, float* out);

/* Triangle count 8 * 4^level of the octahedral sphere; *nverts = 4^(level+1)/... */
int64_t flg_scene_octasphere_counts(int32_t level, int64_t* nverts);

/* Octahedral sphere subdivided `level` times, radius 1 + amplitude * fBm(dir).
 * Returns the vertex count. */
int32_t flg_scene_octasphere(int32_t level, float amplitude, uint32_t seed, float base_freq,
                             float* verts, int32_t* ind);

/* UV sphere of the reference's MakeMeshedSphere (src/plymesh.cpp:110-174);
 * returns the triangle count (pass NULL buffers to query it). */
int32_t flg_scene_uv_sphere(int32_t n_theta, int32_t n_phi, float radius, float* verts,
                            int32_t* ind);

#ifdef __cplusplus
}
#endif
#endif
