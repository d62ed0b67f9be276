/*
 * grt.h -- C ABI of libgrt, the MI355X (gfx950) uniform-grid ray tracer.
 *
 * This is the drop-in boundary for the reference's CUDA tracer C ABI
 * (src/rtcuda/RayTracing.h:6-44) and the GLSL renderer's pass/buffer contract
 * (src/rtgl/GpuRenderer.cpp:74-319, 598-703).  Signatures use plain pointers
 * and sizes only.  Unless stated otherwise:
 *   - "_host" pointers are host memory and are copied before return;
 *   - ray, hit and image buffers are DEVICE pointers owned by the caller
 *     (structure of arrays, one float/int per ray);
 *   - work is enqueued on the context's stream (grt_set_stream) and the call
 *     returns without synchronising;
 *   - every function returns GRT_OK (0) or an error code, never aborts, and
 *     grt_last_error() describes the last failure of the calling thread.
 * Error codes mirror the reference's RT_OK / RT_NO_MEMORY / RT_INVALID_PARAM
 * (include/rt/common.h:19-21).
 */
#ifndef GRT_GRT_H
#define GRT_GRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GRT_OK             0x0000 /* RT_OK */
#define GRT_NO_MEMORY      0x0001 /* RT_NO_MEMORY */
#define GRT_INVALID_PARAM  0x0003 /* RT_INVALID_PARAM */
#define GRT_DEVICE_ERROR   0x0004
#define GRT_NOT_READY      0x0005 /* grid not built / scene empty */
#define GRT_CAPACITY       0x0006 /* a device-side capacity was exceeded */

typedef struct grt_context_t* grt_context;
typedef struct grt_scene_t* grt_scene;

/* Ray batch, structure of arrays of device pointers (n floats each).
 * Replaces the reference's per-thread Ray (src/rtcuda/KernelDependencies.h:70-78;
 * include/rt/Ray.h:9-24).  A ray is inactive when tmin > tmax. */
typedef struct {
    float* ox; float* oy; float* oz;
    float* dx; float* dy; float* dz;
    float* tmin; float* tmax;
} grt_rays;

/* Closest-hit records (include/rt/Hit.h:9-17; KernelDependencies.h:80-86):
 * id = triangle index or -1 (miss), t = distance (+inf on miss), u,v = the
 * barycentric weights of vertices 1 and 2 (Hit::v1Coord, v2Coord). */
typedef struct {
    int32_t* id; float* t; float* u; float* v;
} grt_hits;

/* Pinhole camera as passed by the reference C ABI cudaSetCameraParameters
 * (RayTracing.h:20-23): position, base direction (ray through pixel corner
 * (0,0)), near-plane U and V vectors; image size. */
typedef struct {
    float pos[3]; float base[3]; float nu[3]; float nv[3];
    uint32_t width; uint32_t height;
} grt_camera;

/* Grid frame and size (UniformGridAccStruct.h:22-33; cudaSetBoxMin/Max and
 * cudaSetGridParameters, RayTracing.h:33-36). */
typedef struct {
    float bmin[3]; float bmax[3]; float cell[3]; float inv_cell[3];
    int32_t n[3];
    uint32_t m;   /* cells */
    uint32_t r;   /* references */
} grt_grid_info;

/* ---- context ---------------------------------------------------------- */
/* rtInitContext (RayTracing.h:8; RayTracing.cu:89-99). */
int grt_init(int device, grt_context* out);
int grt_shutdown(grt_context ctx);
const char* grt_last_error(void);
/* Enqueue on this hipStream_t from now on (NULL = the HIP null stream).  A new
 * context starts on a non-blocking stream of its own. */
int grt_set_stream(grt_context ctx, void* hip_stream);
int grt_synchronize(grt_context ctx);
/* Bytes of device memory currently held by the context's scenes. */
int grt_memory_used(grt_context ctx, uint64_t* bytes);
/* Thread-block size of the traversal kernels (64..256, multiple of 64);
 * cudaSetBlockSize (RayTracing.h:39) was a no-op in the reference. */
int grt_set_block_size(grt_context ctx, uint32_t closest, uint32_t any);

/* Traversal scheduling.  GRT_TRACE_SIMPLE (default): one ray per lane for the
 * whole launch, each lane testing its own cell's triangles.  GRT_TRACE_POOLED:
 * one ray per lane, the triangle tests of all the wave's current cells pooled
 * and spread over the wave's lanes.  GRT_TRACE_PERSISTENT: persistent waves
 * whose idle lanes fetch new rays once `param` (1..64) lanes of a wave are
 * idle, advancing each lane by one triangle test or cell per iteration.
 * GRT_TRACE_SPLIT: `param` (2, 4, 8 or 16) lanes per ray, lane j visiting
 * the ray's cells j, j + param, ... of the serial order; for launches with
 * too few rays to fill the GPU.  For SIMPLE and POOLED `param` is kept as the
 * refill threshold (1..64) of later persistent launches.  The per-lane
 * schedules finish dense cells wave-cooperatively (grt_set_coop).  Results and
 * counters are identical; only the schedule differs.  The serving loop and
 * batches (grt_batch_trace) follow the setting; a batch under GRT_TRACE_SPLIT
 * gives every ray of every item `param` lanes (64 / param rays per wave). */
#define GRT_TRACE_SIMPLE     0
#define GRT_TRACE_PERSISTENT 1
#define GRT_TRACE_POOLED     2
#define GRT_TRACE_SPLIT      3
int grt_set_trace_mode(grt_context ctx, int mode, uint32_t param);
/* Persistent and batched traversal: once a wave finds the ray queue drained,
 * the wave tests the rest of a lane's cell 64 references at a time when at
 * least max(min_refs, 2 x its busy lanes) are left (0 = never; default 16).
 * Results and counters are those of the one-lane loop. */
int grt_set_coop(grt_context ctx, uint32_t min_refs);

/* ---- scene ------------------------------------------------------------ */
int grt_scene_create(grt_context ctx, grt_scene* out);
int grt_scene_destroy(grt_scene scene);
/* cudaTransferTriangleVertices (RayTracing.h:27): 9 floats per triangle
 * (v0,v1,v2), n_tris triangles.  Also packs the 48-byte traversal records. */
int grt_scene_upload_triangles(grt_scene scene, const float* verts_host, uint32_t n_tris);
/* cudaTransferTriangleNormals (RayTracing.h:28): 9 floats per triangle. */
int grt_scene_upload_normals(grt_scene scene, const float* normals_host, uint32_t n_tris);
/* Same from device memory (device-to-device copy). */
int grt_scene_set_triangles_device(grt_scene scene, const float* verts_dev,
                                   const float* normals_dev, uint32_t n_tris);
int grt_scene_triangle_count(grt_scene scene, uint32_t* n_tris);
/* Device pointers of the current vertices / normals (9 floats per triangle). */
int grt_scene_device_pointers(grt_scene scene, float** verts_dev, float** normals_dev);
/* Copy the current vertices / normals to the host (either pointer may be NULL). */
int grt_scene_download_geometry(grt_scene scene, float* verts_host, float* normals_host);
/* Keyframe update on the device (GpuRenderer.cpp:94-130,705-795 uploaded
 * whole keyframes instead): current = M_c * rest for triangle t < n_cluster_tris
 * with c = t / tris_per_cluster, M_c the row-major 3x4 matrix xf_host[12c..];
 * normals get the 3x3 part; other triangles keep the rest pose.  The rest pose
 * is the last uploaded geometry.  The matrices are staged in pinned memory:
 * the call first waits (on an event, not on the whole stream) for the previous
 * transform's host-to-device copy to leave that buffer, so the host can
 * enqueue a frame while the device still works on the previous one. */
int grt_scene_transform_clusters(grt_scene scene, const float* xf_host, uint32_t n_clusters,
                                 uint32_t tris_per_cluster, uint32_t n_cluster_tris);

/* ---- scene ingestion (host) -------------------------------------------- */
/* A host-side triangle soup built the way the reference's rt::Context fed
 * its PrimitiveBuilder (Context.cpp:285-343, PrimitiveBuilder.cpp:71-219):
 * primitives of vertices with a current normal and colour, through a current
 * affine transform.  Assembly: GRT_TRIANGLES takes triples; GRT_TRIANGLE_STRIP
 * takes (i, i+1, i+2) for even i and (i+2, i+1, i) for odd i;
 * GRT_TRIANGLE_FAN and GRT_POLYGON take (0, i, i+1).  A triangle is kept when
 * the nine constants of its Wald accelerator (TriAccel::buildFrom/valid,
 * Triangle.cpp:11-63) are finite, so zero-area and non-finite triangles are
 * dropped and the kept ones get consecutive ids.  The values are those of the
 * reference's RTenum (include/rt/common.h:29-32). */
typedef struct grt_mesh_t* grt_mesh;
#define GRT_TRIANGLES      0x1001
#define GRT_TRIANGLE_STRIP 0x1002
#define GRT_TRIANGLE_FAN   0x1003
#define GRT_POLYGON        0x1004
#define GRT_MESH_DROP_DEGENERATE 1u
int grt_mesh_create(grt_mesh* out);
int grt_mesh_destroy(grt_mesh mesh);
int grt_mesh_clear(grt_mesh mesh);
/* Row-major 3x4 affine matrix applied to the following vertices; normals take
 * the inverse transpose of its 3x3 part and are renormalised (NULL: identity). */
int grt_mesh_set_transform(grt_mesh mesh, const float xf[12]);
/* Diffuse colour of the following triangles (the material's). */
int grt_mesh_set_color(grt_mesh mesh, float r, float g, float b);
int grt_mesh_begin(grt_mesh mesh, int primitive);
int grt_mesh_normal(grt_mesh mesh, float x, float y, float z);
int grt_mesh_vertex(grt_mesh mesh, float x, float y, float z);
int grt_mesh_end(grt_mesh mesh);
/* Wavefront OBJ (GpuObjLoader.cpp:138-151): v, vn, f with 1-based or negative
 * indices; a face of k corners becomes the fan (c0, ci, ci+1); a corner
 * without a normal gets (0,0,0); mtllib/usemtl give the triangles the
 * material's Kd (white without one).  The transform is not applied.  With
 * GRT_MESH_DROP_DEGENERATE the faces go through the degenerate filter too.  A
 * file refused part way (bad vertex or face corner) adds nothing. */
int grt_mesh_load_obj(grt_mesh mesh, const char* path, uint32_t flags);
/* rtdb::loadSphere (rtdb.cpp:142-192) and the unit cube of rtdb::loadCube,
 * through the current transform. */
int grt_mesh_add_sphere(grt_mesh mesh, uint32_t slices, uint32_t stacks);
int grt_mesh_add_cube(grt_mesh mesh);
/* rtdb's scene layouts, through the current transform (the reference's
 * MatrixStack post-multiplies: a layout's translation applies before the
 * current transform).  grt_mesh_add_triangle: rtdb::loadTriangle
 * (rtdb.cpp:9-18), (0,0,0) (1,0,0) (0,1,0), normal +z.  grt_mesh_add_logo:
 * rtdb::loadLogo (rtdb.cpp:20-57), three triangles in z = 0 in the current
 * colour.  grt_mesh_add_mesh: model's triangles as GRT_TRIANGLES primitives
 * (normals through the inverse transpose, the degenerate filter again, each
 * keeping its colour).  grt_mesh_add_even_spaced: rtdb::loadEvenSpaced
 * (rtdb.cpp:215-233), the triangle at x, y, z in {0, 1.5, 3}.
 * grt_mesh_add_mixed: rtdb::loadMixed (rtdb.cpp:235-266) with `model` in
 * place of the teapot (whose data table is not shipped: load it, e.g. from an
 * OBJ).  grt_mesh_add_teapots125: rtdb::loadTeapots125 (rtdb.cpp:268-287),
 * 125 copies of `model` scaled by 5 at offsets 2 x (x, y, z), x, y, z < 5. */
int grt_mesh_add_triangle(grt_mesh mesh);
int grt_mesh_add_logo(grt_mesh mesh);
int grt_mesh_add_mesh(grt_mesh mesh, grt_mesh model);
int grt_mesh_add_even_spaced(grt_mesh mesh);
int grt_mesh_add_mixed(grt_mesh mesh, grt_mesh model);
int grt_mesh_add_teapots125(grt_mesh mesh, grt_mesh model);
/* The degenerate filter on one triangle (9 floats): 1 kept, 0 dropped. */
int grt_triangle_valid(const float v9[9]);
/* Kept triangles, and triangles assembled but dropped as degenerate. */
int grt_mesh_info(grt_mesh mesh, uint32_t* n_tris, uint32_t* n_dropped);
/* Copy out 9 floats of vertices, 9 of normals and 3 of diffuse colour per
 * kept triangle, and each one's index among all assembled triangles (any
 * pointer may be NULL). */
int grt_mesh_copy(grt_mesh mesh, float* verts9, float* normals9, float* diffuse3,
                  uint32_t* source_index);
/* grt_scene_upload_triangles + grt_scene_upload_normals of the mesh. */
int grt_scene_upload_mesh(grt_scene scene, grt_mesh mesh);

/* ---- grid ------------------------------------------------------------- */
/* Device build of the uniform grid (UniformGridAccStructBuilder::buildGeometry,
 * UniformGridAccStructBuilder.cpp:58-149; GpuGrid::rebuildGrid call site
 * GpuRenderer.cpp:122,663-677): lambda replaces the reference's k=6.  Sizes
 * device buffers as needed (synchronises the stream twice). */
int grt_grid_build(grt_scene scene, float lambda);
/* Rebuild without host synchronisation, reusing the capacities of the last
 * grt_grid_build / grt_grid_reserve.  Overflow is flagged on the device and
 * reported by grt_grid_check. */
int grt_grid_rebuild_async(grt_scene scene, float lambda);
/* Capacity for later asynchronous rebuilds: `cells` cells and `refs`
 * references.  A buffer that has to grow is reallocated without its contents,
 * so the grid is then gone (traces report GRT_NOT_READY) until the next
 * grt_grid_build or grt_grid_rebuild_async, and batches holding the scene must
 * be set again. */
int grt_grid_reserve(grt_scene scene, uint32_t cells, uint32_t refs);
/* Synchronise and report the device status of the last (re)build. */
int grt_grid_check(grt_scene scene);
/* Upload a grid built elsewhere (cudaTransferCellPointers /
 * cudaTransferCellTriangleIds / cudaSetBoxMin / cudaSetBoxMax /
 * cudaSetGridParameters, RayTracing.h:31-36): frame from info (bmin, bmax,
 * cell, inv_cell, n), offsets[m+1], refs[offsets[m]] plain triangle ids. */
int grt_grid_upload(grt_scene scene, const grt_grid_info* info, const uint32_t* offsets_host,
                    const uint32_t* refs_host);
int grt_grid_info_get(grt_scene scene, grt_grid_info* out);
/* Second grid level (a two-level grid; not in the reference, whose uniform
 * grid walks every empty cell of a "teapot in a stadium").  From the next
 * grt_grid_build / grt_grid_rebuild_async on, each top cell holding at least
 * min_refs references gets a uniform sub-grid spanning the cell, r0 x r1 x r2
 * cells from the top frame's rule applied to the cell (x = lambda2 n / cell
 * volume, r_a = ceil(cell_a cbrt(x)), at most 16 per axis), rebuilt on the
 * device with the top grid.  grt_trace_closest / grt_trace_any and the _list
 * traces then walk both levels, one ray per lane whatever the trace mode: a
 * subdivided cell is crossed by a DDA of its own, hits bounded by the
 * sub-cell's exit.  The closest hit is the one-level grid's triangle, t, u
 * and v except where two hits lie within the exit bound's 2^-20 of each other;
 * counters count top cells plus sub-cells.  The top grid stays complete, so
 * batches, serving and instance traces (which use it alone) are unchanged.
 * lambda2 = 0 turns the level off.  Its buffers are sized by grt_grid_build
 * (sub-cells + 1/4, references + 1/2, for the rebuilds of later frames) and
 * grt_grid_reserve (the top grid's share); an asynchronous rebuild that
 * outgrows them traces the top grid alone and grt_grid_check reports
 * GRT_CAPACITY. */
int grt_grid_set_subgrid(grt_scene scene, float lambda2, uint32_t min_refs);
typedef struct {
    float lambda2;
    uint32_t min_refs;
    uint32_t built;      /* the last build produced the second level */
    uint32_t cells;      /* sub-cells */
    uint32_t refs;       /* references in sub-cells */
    uint32_t status;     /* 0, or the capacity the last build exceeded */
} grt_subgrid_info;
int grt_subgrid_info_get(grt_scene scene, grt_subgrid_info* out);
/* cellsub_host: 2 words per top cell {first sub-cell, r0 | r1 << 8 | r2 << 16
 * (0: no sub-grid)}; sub-cell (x, y, z) of a cell is first + x + r0 (y + r1 z);
 * an empty top cell reads {0, 0x80000000 | r}, r its skip radius (0..7): no
 * cell within r cells of it on every axis holds a reference, and the walk
 * crosses the r cells ahead of it on every axis in one step;
 * sub_offsets_host: cells + 1 words; sub_refs_host: refs words, ascending
 * triangle ids per sub-cell (any pointer may be NULL). */
int grt_subgrid_download(grt_scene scene, uint32_t* cellsub_host, uint32_t* sub_offsets_host,
                         uint32_t* sub_refs_host);
int grt_grid_download(grt_scene scene, uint32_t* offsets_host, uint32_t* refs_host);

/* ---- tracing ---------------------------------------------------------- */
/* Closest hit for n rays (traverse3ddda + intersectTriangles,
 * UniformGridAccStruct.cpp:30-42,113-196,362-378; kernel rayTrace,
 * RayTracing_kernel.cu:140-275).  image_width != 0 maps each wave to an 8x8
 * tile of a row-major image of that width.  counters_dev: NULL or 2 uint32 per
 * ray {cells visited, triangle tests}.  n = 0 does nothing (the buffers may be
 * NULL); so do the other grt_trace_* calls. */
int grt_trace_closest(grt_scene scene, const grt_rays* rays, const grt_hits* hits, uint32_t n,
                      uint32_t image_width, uint32_t* counters_dev);
/* Any hit in (tmin, tmax] (findAnyHit, shadow_FRAG.glsl:147-245; the
 * reference's IAccStruct::traceAnyInstance stub, IAccStruct.cpp:108-114).
 * occluded_dev: 1 byte per ray. */
int grt_trace_any(grt_scene scene, const grt_rays* rays, uint8_t* occluded_dev, uint32_t n,
                  uint32_t image_width, uint32_t* counters_dev);

/* ---- instances --------------------------------------------------------- */
/* A scene placed in the world by a row-major 3x4 object-to-world matrix
 * (rt::Instance with its Transform, include/rt/Instance.h, Transform.h). */
typedef struct {
    grt_scene scene;
    float xf[12];
} grt_instance;
/* Inverse of a row-major 3x4 affine matrix (binary64 cofactors, rounded to
 * binary32), the matrix the rays are taken through.  GRT_INVALID_PARAM when
 * the 3x3 part is singular. */
int grt_affine_inverse(const float xf[12], float inv[12]);
/* Closest hit over instances (IAccStruct::traceNearestInstance,
 * IAccStruct.cpp:26-73): each world ray goes through every instance's
 * inverse matrix (origin affine, direction linear and not renormalised, so t
 * is the same in both spaces) and is traced in the instance's grid with its
 * interval ending just below the best hit so far; the nearest hit wins, an
 * earlier instance on ties.  hits receive the instance scene's triangle id;
 * instance_dev (int32 per ray) the instance index, -1 for a miss.  All scenes
 * belong to ctx.  Uses 48 bytes of context scratch per ray. */
int grt_trace_instances(grt_context ctx, const grt_instance* instances, uint32_t n_instances,
                        const grt_rays* rays, const grt_hits* hits, int32_t* instance_dev,
                        uint32_t n, uint32_t image_width);

/* ---- kd-tree comparison accelerator ------------------------------------- */
/* The reference's alternative accelerator (KdTreeAccStructBuilder::
 * buildGeometry + TriangleTreeBuilder::buildTree, KdTreeAccStructBuilder.cpp:
 * 21-28, TriangleTreeBuilder.cpp:13-358; KdTreeAccStruct::traceNearestGeometry,
 * KdTreeAccStruct.cpp:122-240): a surface-area-heuristic kd-tree over a
 * scene's current triangles, built on the host (perfect splits, traversal cost
 * 1, intersection cost 1.4, empty-space bias; depth at most 48) and traced on
 * the device front to back.  Hits use the grid traversal's window with the
 * leaf's exit distance in place of the cell's, so a closest hit is the same
 * triangle, t, u and v as the grid's except on exact ties.  The tree keeps its
 * own copy of the triangle records: rebuild it after the scene changes. */
typedef struct grt_kdtree_t* grt_kdtree;
typedef struct {
    float bmin[3]; float bmax[3];   /* the tree's box: tight box of all vertices */
    uint32_t nodes;                 /* inner nodes + leaves */
    uint32_t leaves;
    uint32_t empty_leaves;
    uint32_t elements;              /* triangle references in leaves */
    uint32_t depth;                 /* deepest leaf level (root = 0) */
    double build_ms;                /* host build + upload */
} grt_kdtree_info;
/* Build from the scene's current vertices (synchronises the stream). */
int grt_kdtree_build(grt_scene scene, grt_kdtree* out);
/* The same tree from host vertices (9 floats per triangle), without a device:
 * info and download work, the trace calls return GRT_NOT_READY. */
int grt_kdtree_build_host(const float* verts_host, uint32_t n_tris, grt_kdtree* out);
int grt_kdtree_destroy(grt_kdtree tree);
int grt_kdtree_info_get(grt_kdtree tree, grt_kdtree_info* out);
/* The tree as built: 2 uint32 per node -- inner {split bits, (left << 2) |
 * axis}, the right child at left + 1, nodes in breadth-first order; leaf
 * {first element, 0x80000000 | count} -- and the triangle id of every leaf
 * element (either pointer may be NULL). */
int grt_kdtree_download(grt_kdtree tree, uint32_t* nodes_host, uint32_t* elements_host);
/* Closest hit / any hit, as grt_trace_closest / grt_trace_any; counters_dev:
 * NULL or 2 uint32 per ray {leaves visited, triangle tests}. */
int grt_kdtree_trace_closest(grt_kdtree tree, const grt_rays* rays, const grt_hits* hits,
                             uint32_t n, uint32_t image_width, uint32_t* counters_dev);
int grt_kdtree_trace_any(grt_kdtree tree, const grt_rays* rays, uint8_t* occluded_dev,
                         uint32_t n, uint32_t image_width, uint32_t* counters_dev);

/* ---- ray lists ---------------------------------------------------------- */
/* A secondary stage leaves many slots without a ray (pixels that missed, points
 * facing away from the light).  grt_compact_rays lists the active rays of a
 * batch (tmin <= tmax) in slot order -- 8x8 tiles of an image of width
 * image_width, or ray order when 0 -- into list_dev (n entries) and their
 * number into *count_dev on the device, and writes the next stage's result of
 * every inactive ray as a miss: occluded_dev[i] = 0 when occluded_dev is set,
 * hit id -1, t +inf, u = v = 0 when miss_hits is set.  The _list traces then
 * trace the listed rays only, reading the count on the device (no host
 * synchronisation), so waves are full of active rays in image order.  Results
 * are those of the full traces; counters_dev entries of unlisted rays are not
 * written.  Three kernels, no memset: safe to capture in a graph. */
int grt_compact_rays(grt_context ctx, const grt_rays* rays, uint32_t n, uint32_t image_width,
                     uint32_t* list_dev, uint32_t* count_dev, uint8_t* occluded_dev,
                     const grt_hits* miss_hits);
/* grt_compact_rays without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  The
 * tile sums live in a fixed 64 KiB device array per GPU (set up by grt_init),
 * so a call takes at most 67,100,672 ray slots (n, or the 8x8-tile slots of
 * an image of width image_width) and returns GRT_CAPACITY beyond that.
 * Results equal grt_compact_rays'.  Calls that overlap on one GPU share the
 * array and race: issue them on one stream, or serialise them. */
int grt_compact_rays_noalloc(grt_context ctx, const grt_rays* rays, uint32_t n,
                             uint32_t image_width, uint32_t* list_dev, uint32_t* count_dev,
                             uint8_t* occluded_dev, const grt_hits* miss_hits);
int grt_trace_closest_list(grt_scene scene, const grt_rays* rays, const grt_hits* hits,
                           const uint32_t* list_dev, const uint32_t* count_dev, uint32_t n,
                           uint32_t* counters_dev);
int grt_trace_any_list(grt_scene scene, const grt_rays* rays, uint8_t* occluded_dev,
                       const uint32_t* list_dev, const uint32_t* count_dev, uint32_t n,
                       uint32_t* counters_dev);

/* ---- ray generation --------------------------------------------------- */
/* PinholeCamera::setLookAt/setPerspective/setViewport/newFrame
 * (PinholeCamera.cpp:23-55,91-118), host only. */
int grt_camera_lookat(const float eye[3], const float center[3], const float up[3],
                      float fovy_deg, float znear, uint32_t width, uint32_t height,
                      grt_camera* out);
/* Primary rays, pixel i = y*width + x (PinholeCamera::getRayDirection,
 * PinholeCamera.cpp:131-137): tmin = 2e-4 (Context rayEpsilon), tmax = +inf. */
int grt_gen_primary(grt_context ctx, const grt_camera* cam, const grt_rays* out);
/* Shadow rays toward a point light, one slot per input ray
 * (shadow_FRAG.glsl:259-270; SimplePointLight::illuminate,
 * SimplePointLight.cpp:18-37).  Inactive slots: tmin=1, tmax=-1. */
int grt_gen_shadow(grt_scene scene, const grt_rays* in, const grt_hits* hits, uint32_t n,
                   const float light[3], float offset, const grt_rays* out);
/* Mirror reflection rays (Sample::initReflectionRay, Sample.cpp:44-52). */
int grt_gen_reflection(grt_scene scene, const grt_rays* in, const grt_hits* hits, uint32_t n,
                       float offset, const grt_rays* out);
/* spp cosine-weighted ambient-occlusion rays per input ray, slot i*spp+k. */
int grt_gen_ao(grt_scene scene, const grt_rays* in, const grt_hits* hits, uint32_t n,
               uint32_t spp, uint64_t seed, float offset, float ao_tmax, const grt_rays* out);
/* Headlight shading to RGBA8 (RayTracing_kernel.cu:104-120): the image the
 * reference's cudaRayTrace wrote into its pixel buffer. */
int grt_shade_headlight(grt_scene scene, const grt_rays* rays, const grt_hits* hits, uint32_t n,
                        uint32_t* rgba_dev);

/* Phong shading of the GPU renderer (shading_FRAG.glsl:8-36,116-173;
 * PhongMaterialColor::shade, PhongMaterialColor.cpp:19-85).  A hit is lit
 * when N.L > 0 and its shadow ray is not occluded:
 *   lit:      ambient + max(0, kd*lc*N.L) + specular*lc*max(0, S.L)^spec_exp
 *   shadowed: ambient + max(0, kd*lc*N.L) * shadow_coeff
 * with N the interpolated normal facing the ray, L the unit vector to the
 * light and S = normalise(reflect(d, N)); spec_exp is an integer (0: no
 * specular term), evaluated by square-and-multiply. */
typedef struct {
    float light[3];        /* point light position */
    float light_color[3];
    float ambient[3];
    float specular[3];     /* specular colour */
    uint32_t spec_exp;
    float shadow_coeff;
    float reflect_coeff;   /* weight of each reflection level */
    float background[3];   /* colour of a primary miss */
} grt_phong;
/* The shader's constants: light (1e10,1e10,1e10), white; ambient 0;
 * specular 0.5; exponent 50; shadow 0.2; reflection 0.4; background 1. */
void grt_phong_defaults(grt_phong* p);
/* Shade n hit records into rgb_dev (3 floats per ray, interleaved).  depth 0
 * writes the colour (background for a miss); depth d > 0 adds
 * reflect_coeff^d x colour for each hit of that reflection level and leaves
 * misses unchanged (the additive blend of GpuRenderer.cpp:277-297).
 * occluded_dev: the hits' shadow-ray flags (NULL: unshadowed); diffuse_dev:
 * 3 floats per triangle (NULL: white); rgba8_dev (may be NULL) receives the
 * clamped 8-bit colour of the accumulated rgb. */
int grt_shade_phong(grt_scene scene, const grt_rays* rays, const grt_hits* hits,
                    const uint8_t* occluded_dev, uint32_t n, const float* diffuse_dev,
                    const grt_phong* params, uint32_t depth, float* rgb_dev, uint32_t* rgba8_dev);

/* ---- batched multi-scene tracing ------------------------------------- */
/* The reference traced one scene per launch (cudaRayTrace,
 * RayTracing.cu:227-253).  A batch holds items -- each a ray batch of one
 * scene, closest hit (any = 0, outputs in hits) or any hit (any = 1, output in
 * occluded) -- and traces all of them in one persistent launch whose waves
 * fetch ray slots from a device work queue.  Item pointers are captured at
 * grt_batch_set: call it again after a grt_grid_build or grt_grid_upload of any
 * item's scene (grt_grid_rebuild_async keeps the pointers valid). */
typedef struct grt_batch_t* grt_batch;
typedef struct {
    grt_scene scene;
    grt_rays rays;
    grt_hits hits;
    uint8_t* occluded;
    uint32_t n;
    uint32_t image_width;
    uint32_t any;
} grt_batch_item;
int grt_batch_create(grt_context ctx, uint32_t max_items, grt_batch* out);
int grt_batch_set(grt_batch batch, const grt_batch_item* items, uint32_t n_items);
int grt_batch_trace(grt_batch batch);
int grt_batch_slots(grt_batch batch, uint32_t* slots);
int grt_batch_destroy(grt_batch batch);

/* ---- serving ------------------------------------------------------------ */
/* A render request: a resident scene (with a built grid), a pinhole camera
 * and a depth >= 1 of closest-hit generations (primary + depth-1 mirror
 * bounces), each generation followed by shadow rays toward cfg->light.  The
 * reference rendered one frame at a time with one launch per pass
 * (GpuRenderer.cpp:74-319). */
typedef struct {
    grt_scene scene;
    grt_camera camera;
    uint32_t depth;
    double arrival_s;        /* seconds after the start of the run (sorted) */
    grt_hits out_hits;       /* optional (device, width*height): primary hits */
    uint8_t* out_occluded;   /* optional (device): occlusion of the primary shadow rays */
} grt_request;
typedef struct {
    double start_s;          /* admitted (seconds after the start) */
    double finish_s;         /* results complete on the device */
    uint64_t rays;           /* active rays traced */
} grt_request_result;
typedef struct {
    uint64_t memory_budget;  /* bytes of device memory for request buffers */
    uint32_t max_active;     /* requests in flight (batched mode) */
    uint32_t batched;        /* 1: continuous batching; 0: plain path, one request
                                at a time, one launch per ray batch */
    float light[3];
    float offset;            /* secondary-ray origin offset along the normal */
} grt_serve_config;
/* Runs the requests in real time (arrivals at their arrival_s), blocking until
 * all complete.  Batched: each round traces the next step of every admitted
 * request in one batched launch; up to two rounds are queued on the stream at
 * once (a request's next step waits for its previous one by stream order, not
 * by a host barrier); arrivals are admitted FIFO, while a slot and budget
 * memory are free, into the next round issued.  Plain (batched = 0): one
 * request at a time, one launch per ray batch, the host waiting for every step.
 * Buffers come from one slab of memory_budget bytes, so nothing is allocated
 * inside the loop. */
int grt_serve_run(grt_context ctx, const grt_request* requests, uint32_t n,
                  const grt_serve_config* cfg, grt_request_result* out, double* wall_s);
/* The same loop behind a thread-safe interface, for callers that submit
 * requests as they come (no counterpart in the reference, which rendered one
 * frame per call).  grt_server_create starts a thread that owns the device
 * loop on ctx's stream: the context is reserved to the server until
 * grt_server_destroy, and the requests' scenes and output buffers must stay
 * valid (and their grids unchanged) until their results are taken.
 * grt_server_submit (any thread) checks and queues a request (arrival_s is
 * ignored) and returns its ticket at once; the request joins the next round
 * once a slot and budget memory are free.  grt_server_wait blocks until the
 * ticket's request is complete on the device and returns its result (times in
 * seconds after grt_server_create); grt_server_query returns GRT_NOT_READY
 * while it runs.  A result is taken once.  grt_server_destroy completes every
 * submitted request, then stops the thread.  A device error stops the loop and
 * fails every open ticket with its code. */
typedef struct grt_server_t* grt_server;
int grt_server_create(grt_context ctx, const grt_serve_config* cfg, grt_server* out);
int grt_server_submit(grt_server server, const grt_request* request, uint64_t* ticket);
int grt_server_wait(grt_server server, uint64_t ticket, grt_request_result* out);
int grt_server_query(grt_server server, uint64_t ticket, grt_request_result* out);
int grt_server_destroy(grt_server server);

/* ---- launch graphs ---------------------------------------------------- */
/* Capture every launch enqueued on the context's stream between begin and end
 * into a hipGraph (the stream must not be the null stream; entry points that
 * synchronise or copy from pageable host memory must not be called while
 * capturing: trace, gen, batch_trace and grid_rebuild_async are safe). */
typedef struct grt_graph_t* grt_graph;
int grt_graph_begin(grt_context ctx);
int grt_graph_end(grt_context ctx, grt_graph* out);
int grt_graph_launch(grt_graph graph);
int grt_graph_destroy(grt_graph graph);
/* A non-blocking HIP stream at the default priority (opaque hipStream_t). */
int grt_stream_create(void** out);
int grt_stream_destroy(void* stream);

/* ---- test hooks ------------------------------------------------------- */
/* Exclusive scan of n uint32 (device pointers; out has n+1 entries). */
int grt_debug_scan(grt_context ctx, const uint32_t* in_dev, uint32_t* out_dev, uint32_t n);

#ifdef __cplusplus
}
#endif
#endif
