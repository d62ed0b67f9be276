// grt_internal.h -- host-side state behind the opaque handles of grt.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <set>
#include <vector>

#include "grt/grt.h"
#include "grt_common.h"

struct grt_context_t;

namespace grt {

int set_error(int code, const char* fmt, ...);
int hip_check(hipError_t e, const char* what);
// Allocation-free compaction (grt_noalloc.hip): take the current device's
// address of the fixed scratch array, once per device (called by grt_init).
int noalloc_setup(int device);
// Size a batch's slot map for `slots` ray slots up front, so later
// grt_batch_set calls of at most that many slots never reallocate it (a
// pipelined caller keeps earlier launches reading it).
int batch_reserve_slots(grt_batch b, uint64_t slots);

// One device allocation, counted against its context.
struct DevMem {
    void* ptr = nullptr;
    size_t bytes = 0;
    int reserve(grt_context_t* ctx, size_t nbytes, bool zero);
    void release(grt_context_t* ctx);
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

// PinholeCamera::setLookAt + newFrame on the host (camera.cpp).  lookat_basis
// gives the camera axes (X right, Y up, Z backwards); camera_from_basis is
// newFrame for a position and axes.
void lookat_basis(const float eye[3], const float center[3], const float up[3], float ax[3],
                  float ay[3], float az[3]);
void camera_from_basis(const float pos[3], const float ax[3], const float ay[3], const float az[3],
                       float fovy_deg, float znear, uint32_t width, uint32_t height,
                       grt_camera* out);
void camera_lookat(const float eye[3], const float center[3], const float up[3], float fovy_deg,
                   float znear, uint32_t width, uint32_t height, grt_camera* out);

}  // namespace grt

struct grt_scene_t;

struct grt_context_t {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t own_stream = nullptr;
    uint64_t bytes_used = 0;
    uint32_t block_closest = 128;
    uint32_t block_any = 128;
    int trace_mode = GRT_TRACE_SIMPLE;
    uint32_t refill = 16;
    uint32_t split_lanes = 4;   // GRT_TRACE_SPLIT: lanes per ray
    uint32_t coop = 16;   // cooperative cell tests in drained persistent waves (0: off)
    std::set<grt_scene_t*> scenes;
    grt::DevMem instance_scratch;   // object-space rays and candidate hits of instance tracing
    grt::DevMem compact_scratch;    // block sums of ray-list compaction
    // A second stream (default priority) and events for the parts of an
    // asynchronous rebuild that run beside the main stream; created on first use.
    hipStream_t aux_stream = nullptr;
    hipEvent_t aux_ev[4] = {nullptr, nullptr, nullptr, nullptr};
    // Per stream the context has traced on: the slot counter of persistent
    // launches (launches on different streams may overlap).
    std::map<hipStream_t, grt::DevMem> stream_counters;
};

struct grt_scene_t {
    grt_context_t* ctx = nullptr;
    uint32_t n_tris = 0;
    bool has_normals = false;
    bool rest_valid = false;
    // Geometry: 9 floats per triangle (current pose), rest pose, 48-byte records.
    grt::DevMem verts, normals, rest_verts, rest_normals, tri, xf;
    void* xf_host = nullptr;
    size_t xf_host_bytes = 0;
    hipEvent_t xf_copied = nullptr;   // recorded after each copy out of xf_host
    bool xf_pending = false;
    // Grid.
    grt::DevMem frame, partial, counts, offsets, refs, block_sums, big_list, sort_list,
        sort_list_small, sort_classes, small_counters, trir;
    uint32_t cells_cap = 0, refs_cap = 0, list_cap = 0;
    // Second grid level: sub-grids of the dense top cells (g2_lambda = 0: off).
    float g2_lambda = 0.0f;
    uint32_t g2_min_refs = 8;
    bool g2_built = false;   // the last build produced the second level's records
    grt::DevMem g2frame, g2tab, cellrec, g2skip, g2tiles, g2dense, g2box, g2big, g2long,
        sub_offsets, sub_refs, sub_trir;
    uint32_t g2_top_cap = 0, g2_box_cap = 0, g2_cells_cap = 0, g2_refs_cap = 0, g2_dense_cap = 0,
             g2_big_cap = 0, g2_long_cap = 0;
    bool grid_valid = false;
    bool info_stale = false;
    float lambda = 0.0f;
    grt::DevFrame info_host{};

    std::vector<grt::DevMem*> all_buffers()
    {
        return {&verts,  &normals,    &rest_verts, &rest_normals, &tri,       &xf,
                &frame,  &partial,    &counts,     &offsets,      &refs,      &block_sums,
                &big_list, &sort_list, &sort_list_small, &sort_classes, &small_counters,
                &trir, &g2frame, &g2tab, &cellrec, &g2skip, &g2tiles, &g2dense, &g2box, &g2big, &g2long,
                &sub_offsets, &sub_refs, &sub_trir};
    }
    ~grt_scene_t()
    {
        if (xf_copied)
            (void)hipEventDestroy(xf_copied);
        if (xf_host)
            (void)hipHostFree(xf_host);
    }
};
