/*
 * grt_scenes.h -- seeded synthetic scenes for libgrt's tests and benchmarks.
 *
 * The reference loads procedural and file scenes through its rtdb library
 * (src/rtdb/rtdb.cpp:9-287); those scenes are not part of the traced path, so
 * this header defines the generators the benchmarks are specified on instead.
 * Random numbers: splitmix64 seeding -> xoshiro256** (Blackman & Vigna).
 * All functions are host-only and deterministic for a given seed.
 * Geometry arrays hold 9 floats per triangle (v0,v1,v2); normals arrays hold
 * the 3 per-vertex normals (9 floats).  Pass NULL to query the size.
 */
#ifndef GRT_SCENES_H
#define GRT_SCENES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Triangles of one UV-sphere: 16 slices x 9 stacks with pole fans, no
 * degenerate triangles (the reference's loadSphere strips, rtdb.cpp:142-194,
 * emit pole degenerates that PrimitiveBuilder would drop). */
#define GRT_SPHERE_TRIS 256u

/* n_spheres spheres, centres U[0,1]^3, radii U(rmin, rmax).  Returns the
 * triangle count (256 * n_spheres). */
uint32_t grt_gen_spheres(uint64_t seed, uint32_t n_spheres, float rmin, float rmax,
                         float* verts, float* normals);

/* Clustered spheres: n_clusters centres U[0,1]^3, each with
 * spheres_per_cluster spheres at centre + sigma*N(0,1)^3, radii U(rmin,rmax). */
uint32_t grt_gen_sphere_clusters(uint64_t seed, uint32_t n_clusters, uint32_t spheres_per_cluster,
                                 float sigma, float rmin, float rmax, float* verts,
                                 float* normals);

/* "Teapot in a stadium": n_cluster_tris small triangles in Gaussian clusters
 * (cluster c holds triangles [c*tris_per_cluster, ...), centre U[0,1]^3,
 * sigma U(0.01,0.05)); triangle = centroid-free corner v0 ~ centre + sigma*N^3,
 * edges of log-normal length (mu = ln 0.002, s = 0.5) in uniform random
 * directions; then a 2-triangle ground plane at y=-0.25 spanning
 * [-49.5,50.5]^2 in x,z (100x the cluster extent).  centers (3 per cluster)
 * may be NULL.  Returns n_cluster_tris + 2. */
uint32_t grt_gen_stadium(uint64_t seed, uint32_t n_clusters, uint32_t tris_per_cluster,
                         uint32_t n_cluster_tris, float* verts, float* normals, float* centers);

/* Per-frame rigid transforms of the clusters: rotation about the cluster
 * centre by a uniform random angle about a uniform random axis, then a
 * translation U(-max_translation, max_translation)^3.  xf: 12 floats per
 * cluster, row-major 3x4, seeded by (seed, frame). */
void grt_gen_cluster_transforms(uint64_t seed, uint32_t frame, uint32_t n_clusters,
                                const float* centers, float max_translation, float* xf);

/* Uniform double in [0,1) from a seeded xoshiro256** stream position (used by
 * the serving generator for request arrivals). */
double grt_gen_uniform(uint64_t seed, uint64_t index);

#ifdef __cplusplus
}
#endif
#endif
