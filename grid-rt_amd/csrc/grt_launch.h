// grt_launch.h -- host-side launchers for the HIP kernels of libgrt.
// Every launcher enqueues on `stream` and returns the launch status; none
// synchronises.  Buffer sizes are checked by the callers in grt_api.cpp.
#pragma once

#include "grt/grt.h"
#include "grt_common.h"

namespace grt {

// ---- grid build (grt_build.hip) --------------------------------------------
constexpr uint32_t kScanTile = 4096;    // elements per scan block (256 threads x 16)
constexpr uint32_t kBBoxMaxBlocks = 1024;
constexpr uint32_t kSortSmallCells = 16;  // cells up to this size sort in one thread

// Frame from the vertices: block partial boxes, then one block computes the
// frame.  `partial` holds 6*kBBoxMaxBlocks floats.
hipError_t launch_frame_from_verts(const float* verts, uint32_t n_tris, float lambda,
                                   uint32_t cells_cap, float* partial, DevFrame* fr,
                                   hipStream_t s);
// Frame from an explicit (unpadded) box given on the host.
hipError_t launch_frame_from_box(const float bmin[3], const float bmax[3], uint32_t n_tris,
                                 float lambda, uint32_t cells_cap, DevFrame* fr, hipStream_t s);

// Count pass: per-cell counts (atomics), per-scan-tile sums into tile_sums
// (zeroed first; ceil(cells_cap/kScanTile) entries) and the lists of triangles spanning
// more than 16 cells (medium: from the front of big_list, count in
// big_count[0]; huge, over 65536 cells: from the back, count in big_count[3]).
// big_list holds n_tris entries; counts must be zero on entry.
hipError_t launch_grid_count(const DevFrame* fr, const float* verts, uint32_t n_tris,
                             uint32_t* counts, uint32_t* big_list, uint32_t* big_count,
                             uint32_t* tile_sums, uint32_t cells_cap, hipStream_t s);
// Cells the per-cell sort has to visit, collected by the scan: `small` holds
// the cells of 2..kSortSmallCells references, `large` the larger ones;
// counts[0], counts[1] their lengths (zeroed by launch_grid_scan; counts is
// 8-byte aligned, the scan adds to both with one 64-bit atomic).
// `cls` (large_cap entries per class) receives the large list split by size
// for the sort kernels: (16,64], (64,256], (256,1024], (1024,4096], larger;
// counts[2..6] their lengths.
constexpr int kSortClasses = 5;
struct SortLists {
    uint32_t* small;
    uint32_t* large;
    uint32_t* counts;
    uint32_t small_cap;
    uint32_t large_cap;
    uint32_t* cls;
};

// Exclusive scan counts[0..m) -> offsets[0..m], total also to fr->r; flags
// kStatusRefsOver when the total exceeds refs_cap, and then zeroes the counts
// instead of writing offsets (the scatter, which would, does not run).  block_sums holds
// ceil(cells_cap/kScanTile) entries: the per-tile sums the count pass left,
// scanned in place into tile offsets first unless prefixes_ready (a second
// scan of the same counts).  With lists.small set, also fills the sort lists.
hipError_t launch_grid_scan(DevFrame* fr, uint32_t* counts, uint32_t* offsets,
                            uint32_t* block_sums, uint32_t cells_cap, uint32_t refs_cap,
                            const SortLists& lists, bool prefixes_ready, hipStream_t s);
// Scatter pass: refs[offsets[c] + k] = t; returns counts to zero.
hipError_t launch_grid_scatter(const DevFrame* fr, const float* verts, uint32_t n_tris,
                               uint32_t* counts, const uint32_t* offsets, uint32_t* refs,
                               const uint32_t* big_list, const uint32_t* big_count,
                               hipStream_t s);
// Sort the references of every cell on the sort lists ascending.
hipError_t launch_grid_sort(const DevFrame* fr, const uint32_t* offsets, uint32_t* refs,
                            const SortLists& lists, hipStream_t s);

// trir[k] = tri[refs[k]] for k < fr->r (3 float4 per reference).
hipError_t launch_gather_refs(const DevFrame* fr, const uint32_t* refs, const float4* tri,
                              float4* trir, uint32_t refs_cap, hipStream_t s);

// Pack 9-float triangles into the 48-byte traversal record
// {v0.xyz, e1.x} {e1.yz, e2.xy} {e2.z, id, 0, 0}.
hipError_t launch_pack(const float* verts, uint32_t n_tris, float4* tri, hipStream_t s);
// Rigid per-cluster transform of the rest pose, fused with the pack.
hipError_t launch_transform_pack(uint32_t n_tris, uint32_t tris_per_cluster,
                                 uint32_t n_cluster_tris, const float* xf,
                                 const float* rest_v, const float* rest_n, float* verts,
                                 float* normals, float4* tri, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s);

// Exclusive scan of an arbitrary array (exposed for tests of the scan).
hipError_t launch_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* block_sums,
                           DevFrame* tmp_frame, hipStream_t s);

// ---- second grid level (grt_grid2.hip) ----------------------------------------
constexpr uint32_t kSubMax = 16;   // sub-cells per axis of a dense top cell, at most
// Skip radius of an empty top cell, at most: the largest r <= kSkipMax such
// that every cell within r cells of it on every axis (Chebyshev distance)
// holds no reference; cells outside the grid hold none.
constexpr int32_t kSkipMax = 7;
struct G2Build {
    const DevFrame* fr;         // top frame (m cells, status)
    DevFrame* g2;               // level 2: m = sub-cells, r = sub-references, status
    const uint32_t* offsets;    // top offsets[m + 1]
    const uint32_t* refs;       // top references
    const float* verts;         // 9 floats per triangle (current pose)
    float lambda2;
    uint32_t min_refs;
    uint32_t* tile_m;           // per 4096-cell tile: sub-cells, scanned in place
    uint32_t* tile_d;           // per tile: dense cells, scanned in place
    uint2* cellrec;             // per top cell {0x80000000 | first sub-cell, r0|r1<<8|r2<<16}
                                // with a sub-grid, {r, r} (skip radius r) when empty,
                                // else {first ref, end ref}
    uint32_t* dense;            // dense top cells in cell order [dense_cap]
    uint32_t* dense_refs;       // per dense cell: sub-references, then their prefix [dense_cap]
    uint32_t* sub_offsets;      // [sub-cells + 1]
    uint32_t* sub_refs;
    uint32_t* box;              // per top reference: its sub-cell box (dense cells only)
    uint32_t* big;              // sub-cells of more than 1024 references [big_cap]
    uint32_t big_cap;
    uint32_t* longl;            // sub-cells of 17..1024 references [long_cap]
    uint32_t long_cap;
    float* sub_tab;             // [96]: cell_a / r at 16 a + r - 1, 1 / that at 48 + 16 a + r - 1
    uint8_t* skip;              // two planes of skip_stride bytes: per-axis distances
    uint32_t skip_stride;       // >= top cells
};
// Plan, dense list and per-cell counts: g2->m, g2->n[0], g2->r receive the
// sub-cell, dense-cell and sub-reference totals, g2->status the capacities
// exceeded (b.sub_offsets may be NULL).  tile arrays: top_cells_cap /
// kScanTile + 1 entries each.
hipError_t launch_grid2_plan(const G2Build& b, uint32_t top_cells_cap, uint32_t sub_cells_cap,
                             uint32_t sub_refs_cap, uint32_t dense_cap, hipStream_t s);
// The plan in three parts: what needs only the top grid's offsets (cell
// plans, tile sums, the records of the occupied cells, the dense list), the
// skip radii and records of the empty cells (after the plan's first pass;
// nothing of the second level reads them), and what needs the top grid's
// references (per dense cell: boxes, sub-cell counts; the dense scan).
hipError_t launch_grid2_plan_cells(const G2Build& b, uint32_t top_cells_cap,
                                   uint32_t sub_cells_cap, uint32_t dense_cap, hipStream_t s);
hipError_t launch_grid2_skip_records(const G2Build& b, uint32_t top_cells_cap, hipStream_t s);
hipError_t launch_grid2_plan_refs(const G2Build& b, uint32_t sub_refs_cap, uint32_t dense_cap,
                                  hipStream_t s);
// After the plan: sub-cell offsets and references, their sorts and records.
hipError_t launch_grid2_fill(const G2Build& b, uint32_t sub_cells_cap, uint32_t sub_refs_cap,
                             uint32_t dense_cap, const float4* tri, float4* sub_trir,
                             hipStream_t s);
// The whole second level after the top build on the same stream: the plan,
// the skip radii of the empty top cells, then sub-cell offsets, references
// and their 48-byte records.  Over a
// capacity, g2->status is set and every cell record says "no sub-grid".
hipError_t launch_grid2_build(const G2Build& b, uint32_t top_cells_cap, uint32_t sub_cells_cap,
                              uint32_t sub_refs_cap, uint32_t dense_cap, const float4* tri,
                              float4* sub_trir, hipStream_t s);

// ---- traversal and ray generation (grt_traverse.hip, grt_raygen.hip) ---------
struct TraceArgs {
    const DevFrame* fr;
    const uint32_t* offsets;
    const uint32_t* refs;
    const float4* tri;
    const float4* trir;     // per-reference records, 3 float4 per reference
    uint32_t n;
    uint32_t image_width;   // 0: lane i traces ray i; else 8x8 pixel tiles per wave
    uint32_t* counters;     // NULL or 2 per ray {cells, tests}
    uint32_t coop;          // wave-cooperative tests of dense cells: at least this many
                            // references left (0: off)
    uint32_t pool;          // one-ray-per-lane kernels: pool the wave's tests (1) or
                            // test per lane (0)
    const uint32_t* list = nullptr;        // or the rays to trace: slot s traces list[s] ...
    const uint32_t* list_count = nullptr;  // ... for s < *list_count (a device value) and s < n
    // Two-level grid (all set, or cellrec NULL for the top grid alone).  The
    // sub-grids are used only when g2->status is 0.
    const DevFrame* g2 = nullptr;
    const uint2* cellrec = nullptr;        // per top cell, as G2Build::cellrec
    const uint32_t* sub_offsets = nullptr;
    const float4* sub_trir = nullptr;
    const float* sub_tab = nullptr;        // 96 floats: sub-cell sizes, reciprocals (G2Build)
};
hipError_t launch_trace_closest(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                                uint32_t block, hipStream_t s);
// Two-level traversal (a.cellrec set): closest hit into hits, or any hit into
// occluded when it is set; one ray per lane (lanes <= 1) or `lanes` (2..16)
// lanes per ray walking its top cells in turn.
hipError_t launch_trace2(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                         uint8_t* occluded, uint32_t lanes, hipStream_t s);
hipError_t launch_trace_any(const TraceArgs& a, const RaySoA& rays, uint8_t* occluded,
                            uint32_t block, hipStream_t s);
// Persistent traversal (closest hit, or any hit when `any`): lanes refill
// from the slot counter (one uint32, zeroed by the launcher) once `refill`
// lanes of a wave are idle.
hipError_t launch_trace_persistent(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                                   uint8_t* occluded, bool any, uint32_t* slot_counter,
                                   uint32_t block, uint32_t refill, hipStream_t s);
// One item of a batched launch: a ray batch against one scene's grid.  The
// slot range [slot_begin, slot_begin + slots) is padded to whole waves (64).
// Several lanes per ray (2, 4, 8 or 16): lane j of a ray's group visits the
// ray's cells j, j + lanes, ... of the serial order (schedule GRT_TRACE_SPLIT).
hipError_t launch_trace_split(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                              uint8_t* occluded, bool any, uint32_t lanes, uint32_t block,
                              hipStream_t s);

struct alignas(16) BatchTask {
    const DevFrame* fr;
    const uint32_t* offsets;
    const float4* trir;
    RaySoA rays;
    HitSoA hits;            // closest-hit outputs (any == 0)
    uint8_t* occluded;      // any-hit output (any == 1)
    uint32_t n;
    uint32_t image_width;
    uint32_t slot_begin;
    uint32_t slots;
    uint32_t any;
    uint32_t pad[3];
};

// Map each 64-slot block to its task (n_blocks entries).
hipError_t launch_task_map(const BatchTask* tasks, uint32_t n_tasks, uint32_t* map,
                           hipStream_t s);
// Multi-scene traversal over all tasks (mixed closest / any hit) with the
// schedule `mode` (GRT_TRACE_*): one ray per lane (each wave in one task), with
// per-lane or pooled tests; `lanes` lanes per ray (2, 4, 8 or 16); or
// persistent waves refilled from the slot counter.
hipError_t launch_trace_batch(const BatchTask* tasks, const uint32_t* map, uint32_t total_slots,
                              uint32_t* slot_counter, int mode, uint32_t block,
                              uint32_t refill, uint32_t lanes, uint32_t coop, hipStream_t s);
hipError_t launch_gen_primary(const DevCamera& cam, const RaySoAOut& out, hipStream_t s);
hipError_t launch_gen_shadow(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                             const float* normals, float lx, float ly, float lz, float offset,
                             const RaySoAOut& out, hipStream_t s);
hipError_t launch_gen_reflection(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                                 const float* normals, float offset, const RaySoAOut& out,
                                 hipStream_t s);
// Active rays (tmin <= tmax) in slot order (8x8 tiles of an image of width
// img_w, or ray order when 0) -> list[0..*count); the next stage's result of
// every inactive ray written as a miss (occluded = 0 and / or hit id -1,
// t = +inf, u = v = 0, for the outputs that are set).  block_sums holds
// slots / kScanTile + 1 entries.
hipError_t launch_compact_rays(uint32_t n, uint32_t img_w, const float* tmin, const float* tmax,
                               uint32_t* list, uint32_t* count, uint8_t* occluded,
                               const HitSoA& miss_hits, uint32_t* block_sums, hipStream_t s);
hipError_t launch_gen_ao(uint32_t n, uint32_t spp, uint64_t seed, const RaySoA& in,
                         const HitSoAIn& hits, const float* normals, float offset,
                         float ao_tmax, const RaySoAOut& out, hipStream_t s);
hipError_t launch_count_active(const float* tmin, const float* tmax, uint32_t n,
                               unsigned long long* count, hipStream_t s);
// ---- kd-tree comparison accelerator (grt_kdtree_build.cpp, grt_kdtree.hip) ----------
// Deepest level of a kd-tree; the device traversal keeps one stack entry per
// level.
constexpr uint32_t kKdMaxDepth = 48;
struct KdArgs {
    float bmin[3], bmax[3];   // the tree's box (tight box of the vertices)
    const uint2* nodes;       // {split bits | first element, (left << 2) | axis | leaf count}
    const float4* records;    // 3 float4 per leaf element
    uint32_t n;
    uint32_t image_width;     // 0: lane i traces ray i; else 8x8 pixel tiles per wave
    uint32_t* counters;       // NULL or 2 per ray {leaves visited, triangle tests}
};
// rec[k] = tri[elements[k]] (3 float4 each), k < n.
hipError_t launch_kd_gather(const uint32_t* elements, uint32_t n, const float4* tri, float4* rec,
                            hipStream_t s);
// Closest hit into `hits`, or any hit into `occluded` when it is set.
hipError_t launch_kd_trace(const KdArgs& a, const RaySoA& rays, const HitSoA& hits,
                           uint8_t* occluded, uint32_t block, hipStream_t s);

// ---- instance-level tracing (grt_instance.hip) --------------------------------
// Row-major 3x4 affine matrix.
struct Affine {
    float m[12];
};
hipError_t launch_instance_init(uint32_t n, const HitSoA& hits, int32_t* instance, hipStream_t s);
// Object-space rays of one instance, intervals cut below best_t.
hipError_t launch_instance_rays(uint32_t n, const RaySoA& in, const float* best_t,
                                const Affine& inv, const RaySoAOut& out, hipStream_t s);
hipError_t launch_instance_merge(uint32_t n, const HitSoAIn& cand, const HitSoA& hits,
                                 int32_t* instance, int32_t k, hipStream_t s);

// Phong parameters (grt_phong, shading_FRAG.glsl:8-36).
struct DevPhong {
    float light[3];
    float light_color[3];
    float ambient[3];
    float specular[3];
    uint32_t spec_exp;
    float shadow_coeff;
    float reflect_coeff;
    float background[3];
};
// Phong shading into rgb (3 floats per ray, and RGBA8 when rgba is set):
// depth 0 writes, depth d > 0 adds reflect_coeff^d x colour for hits.
hipError_t launch_shade_phong(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                              const uint8_t* occluded, const float* normals,
                              const float* diffuse, const DevPhong& p, uint32_t depth,
                              float* rgb, uint32_t* rgba, hipStream_t s);
hipError_t launch_shade_headlight(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                                  const float* normals, uint32_t* rgba, hipStream_t s);

}  // namespace grt
