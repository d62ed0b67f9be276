/*
 * grt_oracle.h -- CPU oracle for the uniform-grid 3D-DDA + Moller-Trumbore path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the reference the parity tests
 * compare the HIP path against, and the CPU baseline that the benchmark times
 * beside it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  The product library (libgrt.so) never links or calls it.
 *
 * It restates, in plain C, the reference CPU tracer of potato3d/grid-rt:
 *   - grid frame:      src/rtplugins/UniformGridAccStructBuilder.cpp:71-102,
 *                      src/rtcore/Aabb.cpp:11-91,
 *                      src/rtplugins/UniformGridAccStruct.cpp:44-63
 *   - grid build:      src/rtplugins/UniformGridAccStructBuilder.cpp:108-145
 *                      (AABB-overlap insertion, refs ascending by triangle id),
 *                      src/rtplugins/UniformGridAccStruct.cpp:92-98 (worldToVoxel)
 *   - ray clip:        src/rtcore/AabbIntersection.cpp:37-63, include/rt/Ray.h:26-37
 *   - 3D-DDA:          src/rtplugins/UniformGridAccStruct.cpp:113-196
 *   - cell loop:       src/rtplugins/UniformGridAccStruct.cpp:362-378
 *   - Moller-Trumbore: src/rtcuda/RayTracing_kernel.cu:56-89,
 *                      shaders/traversalIntersection_FRAG.glsl:110-137
 *   - any-hit:         shaders/shadow_FRAG.glsl:122-245
 *   - camera rays:     src/rtplugins/PinholeCamera.cpp:23-55,131-137
 *   - secondary rays:  src/rtcore/Sample.cpp:7-52,95-114,
 *                      shaders/shadow_FRAG.glsl:259-270,
 *                      shaders/traversalIntersection_FRAG.glsl:285-296
 *   - Phong shading:   shaders/shading_FRAG.glsl:8-36,116-173,
 *                      src/rtplugins/PhongMaterialColor.cpp:19-85,
 *                      src/rtgl/GpuRenderer.cpp:233-297 (reflection blend)
 *
 * Floating-point evaluation order is fixed: every fused multiply-add is an
 * explicit fmaf() call, everything else is a single IEEE-754 binary32
 * operation, and the file is compiled with -ffp-contract=off.  The HIP
 * kernels follow the same order, so results are bit-identical.
 */
#ifndef GRT_ORACLE_H
#define GRT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Grid frame: padded scene box, per-axis resolution, cell size and reciprocal. */
typedef struct {
    float bmin[3];
    float bmax[3];
    float cell[3];
    float inv_cell[3];
    int32_t n[3];
    uint32_t m;          /* n[0]*n[1]*n[2] */
} orc_frame;

/* Per-ray traversal counters: cells visited (C) and triangle tests (T). */
typedef struct {
    uint32_t cells;
    uint32_t tests;
} orc_counters;

/* Pinhole camera in the reference's C-ABI form (cudaSetCameraParameters):
 * position, base direction (ray through pixel corner (0,0)), near-plane U and V. */
typedef struct {
    float pos[3];
    float base[3];
    float nu[3];
    float nv[3];
    uint32_t width;
    uint32_t height;
} orc_camera;

/* Constants of the specification. */
#define ORC_RAY_EPSILON   2e-4f                 /* Context.cpp:505 rayEpsilon: primary tmin */
#define ORC_EXIT_SCALE    (1.0f + 1.0f / 1048576.0f) /* cell-exit bound: exit*(1+2^-20) */

/* ---- a2: grid frame ------------------------------------------------------ */
float orc_cbrtf(float x);
int orc_grid_frame(const float* verts, uint32_t n_tris, float lambda, orc_frame* out);
/* Frame from an explicit box (already padded) and triangle count. */
int orc_grid_frame_from_box(const float bmin_raw[3], const float bmax_raw[3],
                            uint32_t n_tris, float lambda, orc_frame* out);

/* ---- a3-a5: grid build ---------------------------------------------------- */
void orc_tri_cell_range(const orc_frame* f, const float* v9, int32_t lo[3], int32_t hi[3]);
/* Pass 1: per-cell counts (m entries); returns R = total references. */
uint64_t orc_grid_count(const orc_frame* f, const float* verts, uint32_t n_tris, uint32_t* counts);
/* Pass 2: offsets (m+1) and refs (R), refs ascending by triangle id in each cell. */
void orc_grid_fill(const orc_frame* f, const float* verts, uint32_t n_tris,
                   const uint32_t* counts, uint32_t* offsets, uint32_t* refs);

/* ---- a6: primary rays ----------------------------------------------------- */
void orc_camera_lookat(const float eye[3], const float center[3], const float up[3],
                       float fovy_deg, float znear, uint32_t width, uint32_t height,
                       orc_camera* out);
void orc_primary_rays(const orc_camera* cam, float* ox, float* oy, float* oz,
                      float* dx, float* dy, float* dz, float* tmin, float* tmax);

/* ---- a7-a9: closest hit --------------------------------------------------- */
/* verts: 9 floats per triangle (v0,v1,v2).  Hits: id (-1 = miss), t, u, v.
 * cnt may be NULL.  n_threads <= 1 runs serially; otherwise OpenMP over
 * chunks of `chunk` rays with schedule(dynamic). */
void orc_trace_closest(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                       const float* verts, uint32_t n_rays,
                       const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz,
                       const float* tmin, const float* tmax,
                       int32_t* hit_id, float* hit_t, float* hit_u, float* hit_v,
                       orc_counters* cnt, int n_threads, int chunk);

/* Visited cell sequence of the 3D-DDA (test hook). */
uint32_t orc_dda_cells(const orc_frame* f, const float o[3], const float d[3], float tmin,
                       float tmax, uint32_t* out, uint32_t max_cells);

/* ---- a10/a12: any hit ------------------------------------------------------ */
void orc_trace_any(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                   const float* verts, uint32_t n_rays,
                   const float* ox, const float* oy, const float* oz,
                   const float* dx, const float* dy, const float* dz,
                   const float* tmin, const float* tmax,
                   uint8_t* occluded, orc_counters* cnt, int n_threads, int chunk);

/* ---- two-level grid --------------------------------------------------------- */
/* Sub-grids for top cells holding at least min_refs references (resolution
 * from lambda2, at most rmax per axis).  cellsub: 2 words per top cell {first
 * sub-cell, r0 | r1 << 8 | r2 << 16 (0: none)}; returns the sub-cell total. */
uint64_t orc_grid2_plan(const orc_frame* f, const uint32_t* offsets, uint32_t min_refs,
                        float lambda2, uint32_t rmax, uint32_t* cellsub);
uint64_t orc_grid2_count(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                         const float* verts, const uint32_t* cellsub, uint32_t* counts,
                         uint64_t n_sub);
void orc_grid2_fill(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                    const float* verts, const uint32_t* cellsub, const uint32_t* counts,
                    uint64_t n_sub, uint32_t* sub_offsets, uint32_t* sub_refs);
void orc_grid2_trace_closest(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                             const uint32_t* cellsub, const uint32_t* sub_offsets,
                             const uint32_t* sub_refs, const float* verts, uint32_t n_rays,
                             const float* ox, const float* oy, const float* oz,
                             const float* dx, const float* dy, const float* dz,
                             const float* tmin, const float* tmax,
                             int32_t* hit_id, float* hit_t, float* hit_u, float* hit_v,
                             orc_counters* cnt, int n_threads, int chunk);
void orc_grid2_trace_any(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                         const uint32_t* cellsub, const uint32_t* sub_offsets,
                         const uint32_t* sub_refs, const float* verts, uint32_t n_rays,
                         const float* ox, const float* oy, const float* oz,
                         const float* dx, const float* dy, const float* dz,
                         const float* tmin, const float* tmax,
                         uint8_t* occluded, orc_counters* cnt, int n_threads, int chunk);

/* ---- kd-tree traversal (comparison accelerator) ---------------------------- */
/* box: min xyz, max xyz of the tree; nodes: 2 words per node (inner {split
 * bits, (left << 2) | axis}, leaf {first element, 0x80000000 | count});
 * elements: triangle id per leaf element.  Counters: {leaves visited, tests}.
 * Trees deeper than 64 levels are not supported. */
void orc_kd_trace_closest(const float box[6], const uint32_t* nodes, const uint32_t* elements,
                          const float* verts, uint32_t n_rays,
                          const float* ox, const float* oy, const float* oz,
                          const float* dx, const float* dy, const float* dz,
                          const float* tmin, const float* tmax,
                          int32_t* hit_id, float* hit_t, float* hit_u, float* hit_v,
                          orc_counters* cnt, int n_threads, int chunk);
void orc_kd_trace_any(const float box[6], const uint32_t* nodes, const uint32_t* elements,
                      const float* verts, uint32_t n_rays,
                      const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz,
                      const float* tmin, const float* tmax,
                      uint8_t* occluded, orc_counters* cnt, int n_threads, int chunk);

/* Single Moller-Trumbore test (exposed for known-answer tests). Returns 1 on
 * hit inside (tmin, tmax] and writes t,u,v. */
int orc_mt(const float* v9, const float o[3], const float d[3], float tmin, float tmax,
           float* t, float* u, float* v);

/* ---- a10-a12: secondary ray generation from hit records -------------------- */
/* Shading normal (Sample.cpp:100-114) from 9 per-vertex normal floats. */
void orc_shading_normal(const float* n9, float u, float v, float out[3]);

/* Shadow rays toward a point light.  One slot per input ray; inactive slots
 * (miss, or surface facing away from the light) get tmin=1, tmax=-1. */
void orc_gen_shadow(uint32_t n, const float* ox, const float* oy, const float* oz,
                    const float* dx, const float* dy, const float* dz,
                    const int32_t* hit_id, const float* hit_t, const float* hit_u,
                    const float* hit_v, const float* normals, const float light[3],
                    float offset,
                    float* sox, float* soy, float* soz, float* sdx, float* sdy, float* sdz,
                    float* stmin, float* stmax);

/* Mirror reflection rays; inactive slots (miss) get tmin=1, tmax=-1. */
void orc_gen_reflection(uint32_t n, const float* ox, const float* oy, const float* oz,
                        const float* dx, const float* dy, const float* dz,
                        const int32_t* hit_id, const float* hit_t, const float* hit_u,
                        const float* hit_v, const float* normals, float offset,
                        float* rox, float* roy, float* roz, float* rdx, float* rdy, float* rdz,
                        float* rtmin, float* rtmax);

/* Cosine-weighted ambient-occlusion rays, `spp` per input ray, slot i*spp+k. */
void orc_gen_ao(uint32_t n, uint32_t spp, uint64_t seed,
                const float* ox, const float* oy, const float* oz,
                const float* dx, const float* dy, const float* dz,
                const int32_t* hit_id, const float* hit_t, const float* hit_u,
                const float* hit_v, const float* normals, float offset, float ao_tmax,
                float* aox, float* aoy, float* aoz, float* adx, float* ady, float* adz,
                float* atmin, float* atmax);

/* Polynomial sine/cosine of 2*pi*turns for turns in [0,1) (used by orc_gen_ao). */
void orc_sincos_turns(float turns, float* s, float* c);
/* Counter-based uniform in [0,1) with 24 random bits. */
float orc_uniform24(uint64_t seed, uint64_t a, uint64_t b);

/* ---- per-frame rigid cluster transform ------------------------------------- */
/* xf: 12 floats per cluster (row-major 3x4).  Triangle t < n_cluster_tris
 * belongs to cluster t / tris_per_cluster; the rest are left unchanged. */
void orc_transform(uint32_t n_tris, uint32_t tris_per_cluster, uint32_t n_cluster_tris,
                   const float* xf, const float* rest_verts, const float* rest_normals,
                   float* verts, float* normals);

/* ---- FP64 brute force (error statistics and near-tie detection) ------------ */
/* For each ray: nearest and second-nearest candidate t (different triangles)
 * over all triangles in double precision, plus the nearest id and u,v. */
void orc_brute_closest_f64(const float* verts, uint32_t n_tris, uint32_t n_rays,
                           const float* ox, const float* oy, const float* oz,
                           const float* dx, const float* dy, const float* dz,
                           const float* tmin, const float* tmax,
                           int32_t* id, double* t1, double* t2, double* u, double* v,
                           int n_threads);
void orc_brute_any_f64(const float* verts, uint32_t n_tris, uint32_t n_rays,
                       const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz,
                       const float* tmin, const float* tmax,
                       uint8_t* occluded, double* t_near, int n_threads);

/* Headlight shading of a hit (RayTracing_kernel.cu:92-120) to RGBA8. */
uint32_t orc_headlight_rgba8(const float d[3], const float* n9, int32_t id, float u, float v);
/* The same for n rays: pixel i from ray direction (dx,dy,dz)[i] and hit
 * (id,u,v)[i]; normals9 holds 9 floats per triangle. */
void orc_headlight_image(uint32_t n, const float* dx, const float* dy, const float* dz,
                         const float* normals9, const int32_t* id, const float* u,
                         const float* v, uint32_t* rgba);

/* Phong shading of the GPU renderer (shading_FRAG.glsl:8-36,116-173). */
typedef struct {
    float light[3];          /* point light position */
    float light_color[3];
    float ambient[3];
    float specular[3];       /* specular colour */
    uint32_t spec_exp;       /* integer specular exponent; 0: no specular term */
    float shadow_coeff;      /* diffuse factor of a shadowed point */
    float reflect_coeff;     /* weight of each reflection level */
    float background[3];     /* colour of a primary miss */
} orc_phong;

/* x^e by square-and-multiply from the lowest bit of e (x^0 = 1). */
float orc_pow_u32(float x, uint32_t e);
/* Shade the hits of n rays into rgb (3 floats per ray).  depth 0 writes
 * rgb = colour (background for misses); depth d > 0 adds reflect_coeff^d x
 * colour for hits and leaves misses unchanged.  occluded: the flags of the
 * hits' shadow rays (NULL: none).  diffuse_rgb: 3 floats per triangle (NULL:
 * white).  rgba8 (may be NULL) receives the clamped 8-bit colour of the
 * resulting rgb. */
void orc_shade_phong(uint32_t n, const float* ox, const float* oy, const float* oz,
                     const float* dx, const float* dy, const float* dz, const int32_t* hit_id,
                     const float* hit_t, const float* hit_u, const float* hit_v,
                     const uint8_t* occluded, const float* normals, const float* diffuse_rgb,
                     const orc_phong* p, uint32_t depth, float* rgb, uint32_t* rgba8);

/* Wall-clock seconds (for the CPU baseline). */
double orc_now(void);
int orc_max_threads(void);

#ifdef __cplusplus
}
#endif
#endif
