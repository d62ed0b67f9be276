// grt_api.cpp -- C ABI of libgrt (include/grt/grt.h): contexts, scenes, the
// grid build driver and the tracing entry points.  Device memory is owned by
// opaque handles and freed on destroy (the reference's cudaMalloc'd globals
// were never freed, RayTracing.cu:19-22,106-170).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "grt/grt.h"
#include "grt_internal.h"
#include "grt_launch.h"

using namespace grt;

namespace {
thread_local std::string g_err;
}

namespace grt {

int set_error(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

int hip_check(hipError_t e, const char* what)
{
    if (e == hipSuccess)
        return GRT_OK;
    return set_error(e == hipErrorOutOfMemory ? GRT_NO_MEMORY : GRT_DEVICE_ERROR, "%s: %s", what,
                     hipGetErrorString(e));
}

int DevMem::reserve(grt_context_t* ctx, size_t nbytes, bool zero)
{
    if (nbytes <= bytes)
        return GRT_OK;
    release(ctx);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, nbytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GRT_NO_MEMORY, "hipMalloc(%zu bytes) failed: %s", nbytes,
                         hipGetErrorString(e));
    }
    ptr = p;
    bytes = nbytes;
    ctx->bytes_used += nbytes;
    if (zero) {
        e = hipMemsetAsync(ptr, 0, nbytes, ctx->stream);
        if (e != hipSuccess)
            return hip_check(e, "hipMemsetAsync");
    }
    return GRT_OK;
}

void DevMem::release(grt_context_t* ctx)
{
    if (ptr) {
        (void)hipFree(ptr);
        ctx->bytes_used -= bytes;
    }
    ptr = nullptr;
    bytes = 0;
}

int scene_pack(grt_scene_t* sc)
{
    return hip_check(launch_pack(sc->verts.as<float>(), sc->n_tris, sc->tri.as<float4>(),
                                 sc->ctx->stream),
                     "pack");
}

// Ensure the cell-side buffers hold `cells` cells.  counts must stay zero
// outside a build, so a fresh allocation is cleared.
int grid_reserve_cells(grt_scene_t* sc, uint32_t cells)
{
    if (cells <= sc->cells_cap && sc->counts.ptr)
        return GRT_OK;
    const uint64_t cap = ((uint64_t)cells + cells / 8 + kScanTile) / kScanTile * kScanTile;
    if (cap > 0xffffffffull)
        return set_error(GRT_CAPACITY, "grid of %u cells is too large", cells);
    int rc;
    if ((rc = sc->counts.reserve(sc->ctx, cap * 4, true)) != GRT_OK)
        return rc;
    if ((rc = sc->offsets.reserve(sc->ctx, (cap + 1) * 4, false)) != GRT_OK)
        return rc;
    if ((rc = sc->block_sums.reserve(sc->ctx, (cap / kScanTile + 1) * 4, false)) != GRT_OK)
        return rc;
    sc->cells_cap = (uint32_t)cap;
    return GRT_OK;
}

int grid_reserve_refs(grt_scene_t* sc, uint32_t refs)
{
    if (refs <= sc->refs_cap && sc->refs.ptr)
        return GRT_OK;
    uint64_t cap = (uint64_t)refs + refs / 8 + 1024;
    if (cap > 0xfffffff0ull)
        cap = 0xfffffff0ull;
    int rc;
    if ((rc = sc->refs.reserve(sc->ctx, cap * 4, false)) != GRT_OK)
        return rc;
    const uint64_t list_cap = cap / (kSortSmallCells + 1) + 1;
    if ((rc = sc->sort_list.reserve(sc->ctx, list_cap * 4, false)) != GRT_OK)
        return rc;
    if ((rc = sc->sort_list_small.reserve(sc->ctx, (cap / 2 + 1) * 4, false)) != GRT_OK)
        return rc;
    if ((rc = sc->sort_classes.reserve(sc->ctx, list_cap * 4 * kSortClasses, false)) != GRT_OK)
        return rc;
    if ((rc = sc->trir.reserve(sc->ctx, cap * 48, false)) != GRT_OK)
        return rc;
    sc->refs_cap = (uint32_t)cap;
    sc->list_cap = (uint32_t)list_cap;
    return GRT_OK;
}

static int fetch_frame(grt_scene_t* sc, DevFrame* host)
{
    hipError_t e = hipMemcpyAsync(host, sc->frame.ptr, sizeof(DevFrame), hipMemcpyDeviceToHost,
                                  sc->ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(sc->ctx->stream);
    return hip_check(e, "reading the grid frame");
}

static void frame_to_info(const DevFrame& f, grt_grid_info* out)
{
    for (int d = 0; d < 3; ++d) {
        out->bmin[d] = f.bmin[d];
        out->bmax[d] = f.bmax[d];
        out->cell[d] = f.cell[d];
        out->inv_cell[d] = f.inv[d];
        out->n[d] = f.n[d];
    }
    out->m = f.m;
    out->r = f.r;
}

// Count -> scan -> scatter -> sort, all on the stream.
// The per-cell sort lists (empty before the reference buffers exist).
static SortLists sort_lists(grt_scene_t* sc)
{
    if (!sc->sort_list_small.ptr || !sc->sort_list.ptr || !sc->sort_classes.ptr)
        return SortLists{};
    return SortLists{sc->sort_list_small.as<uint32_t>(), sc->sort_list.as<uint32_t>(),
                     sc->small_counters.as<uint32_t>() + 4, sc->refs_cap / 2 + 1, sc->list_cap,
                     sc->sort_classes.as<uint32_t>()};
}

// prefixes_ready: the tile sums were already scanned by an earlier scan of
// the same counts (the synchronous build scans twice).
static int enqueue_scan(grt_scene_t* sc, uint32_t refs_cap, bool with_lists, bool prefixes_ready)
{
    return hip_check(launch_grid_scan(sc->frame.as<DevFrame>(), sc->counts.as<uint32_t>(),
                                      sc->offsets.as<uint32_t>(), sc->block_sums.as<uint32_t>(),
                                      sc->cells_cap, refs_cap,
                                      with_lists ? sort_lists(sc) : SortLists{}, prefixes_ready,
                                      sc->ctx->stream),
                     "grid scan");
}

static int enqueue_build_passes(grt_scene_t* sc, uint32_t refs_cap, bool with_lists)
{
    hipStream_t s = sc->ctx->stream;
    DevFrame* fr = sc->frame.as<DevFrame>();
    const float* verts = sc->verts.as<float>();
    int rc;
    if ((rc = hip_check(launch_grid_count(fr, verts, sc->n_tris, sc->counts.as<uint32_t>(),
                                          sc->big_list.as<uint32_t>(),
                                          sc->small_counters.as<uint32_t>(),
                                          sc->block_sums.as<uint32_t>(), sc->cells_cap, s),
                        "grid count")) != GRT_OK)
        return rc;
    return enqueue_scan(sc, refs_cap, with_lists, false);
}

static int enqueue_scatter_sort_only(grt_scene_t* sc)
{
    hipStream_t s = sc->ctx->stream;
    DevFrame* fr = sc->frame.as<DevFrame>();
    int rc;
    if ((rc = hip_check(launch_grid_scatter(fr, sc->verts.as<float>(), sc->n_tris,
                                            sc->counts.as<uint32_t>(), sc->offsets.as<uint32_t>(),
                                            sc->refs.as<uint32_t>(), sc->big_list.as<uint32_t>(),
                                            sc->small_counters.as<uint32_t>(), s),
                        "grid scatter")) != GRT_OK)
        return rc;
    return hip_check(launch_grid_sort(fr, sc->offsets.as<uint32_t>(), sc->refs.as<uint32_t>(),
                                      sort_lists(sc), s),
                     "grid sort");
}

static int enqueue_scatter_sort(grt_scene_t* sc)
{
    hipStream_t s = sc->ctx->stream;
    DevFrame* fr = sc->frame.as<DevFrame>();
    int rc;
    if ((rc = hip_check(launch_grid_scatter(fr, sc->verts.as<float>(), sc->n_tris,
                                            sc->counts.as<uint32_t>(), sc->offsets.as<uint32_t>(),
                                            sc->refs.as<uint32_t>(), sc->big_list.as<uint32_t>(),
                                            sc->small_counters.as<uint32_t>(), s),
                        "grid scatter")) != GRT_OK)
        return rc;
    if ((rc = hip_check(launch_grid_sort(fr, sc->offsets.as<uint32_t>(), sc->refs.as<uint32_t>(),
                                         sort_lists(sc), s),
                        "grid sort")) != GRT_OK)
        return rc;
    return hip_check(launch_gather_refs(fr, sc->refs.as<uint32_t>(), sc->tri.as<float4>(),
                                        sc->trir.as<float4>(), sc->refs_cap, s),
                     "grid records");
}

// ---- second grid level ------------------------------------------------------
static G2Build g2_args(grt_scene_t* sc, bool with_outputs)
{
    G2Build b;
    b.fr = sc->frame.as<DevFrame>();
    b.g2 = sc->g2frame.as<DevFrame>();
    b.offsets = sc->offsets.as<uint32_t>();
    b.refs = sc->refs.as<uint32_t>();
    b.verts = sc->verts.as<float>();
    b.lambda2 = sc->g2_lambda;
    b.min_refs = sc->g2_min_refs;
    b.tile_m = sc->g2tiles.as<uint32_t>();
    b.tile_d = b.tile_m + (sc->cells_cap / kScanTile + 1);
    b.cellrec = sc->cellrec.as<uint2>();
    b.dense = sc->g2dense.as<uint32_t>();
    b.dense_refs = b.dense + sc->g2_dense_cap;
    b.sub_offsets = with_outputs ? sc->sub_offsets.as<uint32_t>() : nullptr;
    b.sub_refs = with_outputs ? sc->sub_refs.as<uint32_t>() : nullptr;
    b.box = sc->g2box.as<uint32_t>();
    b.big = sc->g2big.as<uint32_t>();
    b.big_cap = sc->g2_big_cap;
    b.longl = sc->g2long.as<uint32_t>();
    b.long_cap = sc->g2_long_cap;
    b.sub_tab = sc->g2tab.as<float>();
    b.skip = sc->g2skip.as<uint8_t>();
    b.skip_stride = sc->g2_top_cap;
    return b;
}

// Cell records, tile sums and reference boxes for the top grid's capacities.
static int g2_reserve_top(grt_scene_t* sc)
{
    int rc;
    if (sc->g2_box_cap < sc->refs_cap || !sc->g2box.ptr) {
        if ((rc = sc->g2box.reserve(sc->ctx, (size_t)sc->refs_cap * 4 + 4, false)) != GRT_OK)
            return rc;
        sc->g2_box_cap = sc->refs_cap;
    }
    if ((rc = sc->g2frame.reserve(sc->ctx, sizeof(DevFrame), true)) != GRT_OK)
        return rc;
    if ((rc = sc->g2tab.reserve(sc->ctx, 96 * sizeof(float), true)) != GRT_OK)
        return rc;
    if ((rc = sc->cellrec.reserve(sc->ctx, (size_t)sc->cells_cap * 8, false)) != GRT_OK)
        return rc;
    if ((rc = sc->g2skip.reserve(sc->ctx, (size_t)sc->cells_cap * 2, false)) != GRT_OK)
        return rc;
    if ((rc = sc->g2tiles.reserve(sc->ctx, 2 * ((size_t)sc->cells_cap / kScanTile + 1) * 4,
                                  false)) != GRT_OK)
        return rc;
    sc->g2_top_cap = sc->cells_cap;
    return GRT_OK;
}

// The dense-cell list (and its reference counts) for `dense` cells.
static int g2_reserve_dense(grt_scene_t* sc, uint32_t dense)
{
    const uint64_t dc = (uint64_t)dense + dense / 4 + 1024;
    if (dc > 0x7ffffff0ull)
        return set_error(GRT_CAPACITY, "%u dense cells are too many", dense);
    if (dc <= sc->g2_dense_cap && sc->g2dense.ptr)
        return GRT_OK;
    int rc;
    if ((rc = sc->g2dense.reserve(sc->ctx, 2 * dc * 4, false)) != GRT_OK)
        return rc;
    sc->g2_dense_cap = (uint32_t)dc;
    return GRT_OK;
}

static int g2_reserve_sub(grt_scene_t* sc, uint32_t cells, uint32_t refs)
{
    // headroom for the rebuilds of later frames (clusters move)
    const uint64_t cc = (uint64_t)cells + cells / 4 + 4096, rc64 = (uint64_t)refs + refs / 2 + 4096;
    if (cc > 0x7ffffff0ull || rc64 > 0xfffffff0ull)
        return set_error(GRT_CAPACITY, "sub-grids of %u cells / %u references are too large",
                         cells, refs);
    int rc;
    if (cc > sc->g2_cells_cap || !sc->sub_offsets.ptr) {
        if ((rc = sc->sub_offsets.reserve(sc->ctx, (cc + 1) * 4, false)) != GRT_OK)
            return rc;
        sc->g2_cells_cap = (uint32_t)cc;
    }
    if (rc64 > sc->g2_refs_cap || !sc->sub_refs.ptr) {
        if ((rc = sc->sub_refs.reserve(sc->ctx, rc64 * 4, false)) != GRT_OK)
            return rc;
        if ((rc = sc->sub_trir.reserve(sc->ctx, rc64 * 48, false)) != GRT_OK)
            return rc;
        // sub-cells of more than 1024 references: at most refs / 1025 of them
        const uint64_t bc = rc64 / 1025 + 1;
        if ((rc = sc->g2big.reserve(sc->ctx, bc * 4, false)) != GRT_OK)
            return rc;
        sc->g2_big_cap = (uint32_t)bc;
        // sub-cells of 17..1024 references: at most refs / 17 of them
        const uint64_t lc = rc64 / 17 + 1;
        if ((rc = sc->g2long.reserve(sc->ctx, lc * 4, false)) != GRT_OK)
            return rc;
        sc->g2_long_cap = (uint32_t)lc;
        sc->g2_refs_cap = (uint32_t)rc64;
    }
    return GRT_OK;
}

// Enqueue the second level after a top build on the same stream (no host
// synchronisation); skipped when its buffers do not cover the top capacity.
// Whether the second level is on and its buffers cover the top capacity.
static bool g2_ready(const grt_scene_t* sc)
{
    return sc->g2_lambda > 0.0f && sc->g2_top_cap >= sc->cells_cap &&
           sc->g2_box_cap >= sc->refs_cap && sc->sub_offsets.ptr && sc->sub_refs.ptr &&
           sc->g2dense.ptr && sc->g2big.ptr && sc->g2long.ptr;
}

static int g2_enqueue(grt_scene_t* sc)
{
    sc->g2_built = false;
    if (!g2_ready(sc))
        return GRT_OK;
    const G2Build b = g2_args(sc, true);
    const int rc = hip_check(launch_grid2_build(b, sc->cells_cap, sc->g2_cells_cap,
                                                sc->g2_refs_cap, sc->g2_dense_cap,
                                                sc->tri.as<float4>(), sc->sub_trir.as<float4>(),
                                                sc->ctx->stream),
                             "sub-grid build");
    sc->g2_built = rc == GRT_OK;
    return rc;
}

// Synchronous second level: plan with unbounded capacity, size the buffers
// with headroom, then build.
static int g2_build_sync(grt_scene_t* sc)
{
    sc->g2_built = false;
    if (!(sc->g2_lambda > 0.0f))
        return GRT_OK;
    int rc;
    if ((rc = g2_reserve_top(sc)) != GRT_OK)
        return rc;
    hipStream_t s = sc->ctx->stream;
    DevFrame g2;
    auto fetch = [&]() {
        hipError_t e = hipMemcpyAsync(&g2, sc->g2frame.ptr, sizeof(g2), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        return hip_check(e, "reading the sub-grid totals");
    };
    // The dense-cell count first (its list sizes the per-cell pass), then
    // the sub-cell and reference totals.
    const uint32_t dense0 = sc->g2dense.ptr ? sc->g2_dense_cap : 0u;
    if ((rc = g2_reserve_dense(sc, dense0)) != GRT_OK)
        return rc;
    for (int pass = 0; pass < 2; ++pass) {
        if ((rc = hip_check(launch_grid2_plan(g2_args(sc, false), sc->cells_cap, 0x7ffffff0u,
                                              0xffffffffu, sc->g2_dense_cap, s),
                            "sub-grid plan")) != GRT_OK)
            return rc;
        if ((rc = fetch()) != GRT_OK)
            return rc;
        if (g2.status == 0)
            break;
        if ((rc = g2_reserve_dense(sc, (uint32_t)g2.n[0])) != GRT_OK)
            return rc;
    }
    if (g2.status != 0)
        return set_error(GRT_CAPACITY, "sub-grid plan status 0x%x", g2.status);
    if ((rc = g2_reserve_dense(sc, (uint32_t)g2.n[0])) != GRT_OK)
        return rc;
    if ((rc = g2_reserve_sub(sc, g2.m, g2.r)) != GRT_OK)
        return rc;
    if ((rc = g2_enqueue(sc)) != GRT_OK)
        return rc;
    return hip_check(hipStreamSynchronize(s), "sub-grid build");
}

int scene_common_buffers(grt_scene_t* sc)
{
    int rc;
    if ((rc = sc->frame.reserve(sc->ctx, sizeof(DevFrame), true)) != GRT_OK)
        return rc;
    if ((rc = sc->partial.reserve(sc->ctx, 6 * kBBoxMaxBlocks * sizeof(float), false)) != GRT_OK)
        return rc;
    if ((rc = sc->small_counters.reserve(sc->ctx, 64, true)) != GRT_OK)
        return rc;
    return GRT_OK;
}

}  // namespace grt

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* grt_last_error(void) { return g_err.c_str(); }

int grt_init(int device, grt_context* out)
{
    if (!out)
        return set_error(GRT_INVALID_PARAM, "grt_init: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return set_error(GRT_DEVICE_ERROR, "grt_init: no HIP device (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return set_error(GRT_INVALID_PARAM, "grt_init: device %d out of range [0,%d)", device,
                         count);
    if ((e = hipSetDevice(device)) != hipSuccess)
        return hip_check(e, "hipSetDevice");
    int rc = noalloc_setup(device);
    if (rc != GRT_OK)
        return rc;
    grt_context_t* ctx = new (std::nothrow) grt_context_t();
    if (!ctx)
        return set_error(GRT_NO_MEMORY, "grt_init: out of host memory");
    ctx->device = device;
    if ((e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking)) != hipSuccess) {
        delete ctx;
        return hip_check(e, "hipStreamCreate");
    }
    ctx->stream = ctx->own_stream;
    if ((rc = ctx->stream_counters[ctx->stream].reserve(ctx, 64, false)) != GRT_OK) {
        (void)hipStreamDestroy(ctx->own_stream);
        delete ctx;
        return rc;
    }
    *out = ctx;
    return GRT_OK;
}

int grt_shutdown(grt_context ctx)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_shutdown: NULL context");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    for (grt_scene_t* sc : std::vector<grt_scene_t*>(ctx->scenes.begin(), ctx->scenes.end()))
        grt_scene_destroy(sc);
    ctx->instance_scratch.release(ctx);
    ctx->compact_scratch.release(ctx);
    for (auto& kv : ctx->stream_counters)
        kv.second.release(ctx);
    ctx->stream_counters.clear();
    if (ctx->aux_stream) {
        (void)hipStreamSynchronize(ctx->aux_stream);
        (void)hipStreamDestroy(ctx->aux_stream);
    }
    for (hipEvent_t ev : ctx->aux_ev)
        if (ev)
            (void)hipEventDestroy(ev);
    if (ctx->own_stream)
        (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
    return GRT_OK;
}

int grt_set_stream(grt_context ctx, void* hip_stream)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_set_stream: NULL context");
    ctx->stream = (hipStream_t)hip_stream;
    // the stream's persistent-launch slot counter, allocated here (not under
    // a capture that a later trace may be part of)
    return ctx->stream_counters[ctx->stream].reserve(ctx, 64, false);
}

int grt_synchronize(grt_context ctx)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_synchronize: NULL context");
    return hip_check(hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
}

int grt_memory_used(grt_context ctx, uint64_t* bytes)
{
    if (!ctx || !bytes)
        return set_error(GRT_INVALID_PARAM, "grt_memory_used: NULL argument");
    *bytes = ctx->bytes_used;
    return GRT_OK;
}

int grt_set_block_size(grt_context ctx, uint32_t closest, uint32_t any)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_set_block_size: NULL context");
    auto ok = [](uint32_t b) { return b >= 64 && b <= 256 && (b % 64) == 0; };
    if (!ok(closest) || !ok(any))
        return set_error(GRT_INVALID_PARAM, "block sizes must be 64..256 in steps of 64");
    ctx->block_closest = closest;
    ctx->block_any = any;
    return GRT_OK;
}

int grt_set_trace_mode(grt_context ctx, int mode, uint32_t param)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_set_trace_mode: NULL context");
    if (mode < GRT_TRACE_SIMPLE || mode > GRT_TRACE_SPLIT)
        return set_error(GRT_INVALID_PARAM, "grt_set_trace_mode: unknown mode %d", mode);
    if (mode == GRT_TRACE_SPLIT) {
        if (param != 2 && param != 4 && param != 8 && param != 16)
            return set_error(GRT_INVALID_PARAM,
                             "grt_set_trace_mode: lanes per ray must be 2, 4, 8 or 16");
        ctx->split_lanes = param;
    } else {
        if (param < 1 || param > 64)
            return set_error(GRT_INVALID_PARAM, "grt_set_trace_mode: refill must be 1..64");
        ctx->refill = param;
    }
    ctx->trace_mode = mode;
    return GRT_OK;
}

int grt_set_coop(grt_context ctx, uint32_t min_refs)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_set_coop: NULL context");
    ctx->coop = min_refs;
    return GRT_OK;
}

int grt_scene_create(grt_context ctx, grt_scene* out)
{
    if (!ctx || !out)
        return set_error(GRT_INVALID_PARAM, "grt_scene_create: NULL argument");
    grt_scene_t* sc = new (std::nothrow) grt_scene_t();
    if (!sc)
        return set_error(GRT_NO_MEMORY, "grt_scene_create: out of host memory");
    sc->ctx = ctx;
    int rc = scene_common_buffers(sc);
    if (rc != GRT_OK) {
        grt_scene_destroy(sc);
        return rc;
    }
    ctx->scenes.insert(sc);
    *out = sc;
    return GRT_OK;
}

int grt_scene_destroy(grt_scene sc)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_scene_destroy: NULL scene");
    grt_context_t* ctx = sc->ctx;
    (void)hipStreamSynchronize(ctx->stream);
    for (DevMem* m : sc->all_buffers())
        m->release(ctx);
    ctx->scenes.erase(sc);
    delete sc;
    return GRT_OK;
}

static int scene_alloc_geometry(grt_scene_t* sc, uint32_t n_tris)
{
    int rc;
    const size_t b9 = (size_t)n_tris * 9 * sizeof(float);
    if ((rc = sc->verts.reserve(sc->ctx, b9, false)) != GRT_OK)
        return rc;
    if ((rc = sc->normals.reserve(sc->ctx, b9, false)) != GRT_OK)
        return rc;
    if ((rc = sc->tri.reserve(sc->ctx, (size_t)n_tris * 48, false)) != GRT_OK)
        return rc;
    if ((rc = sc->big_list.reserve(sc->ctx, (size_t)n_tris * 4, false)) != GRT_OK)
        return rc;
    return GRT_OK;
}

int grt_scene_upload_triangles(grt_scene sc, const float* verts_host, uint32_t n_tris)
{
    if (!sc || (!verts_host && n_tris))
        return set_error(GRT_INVALID_PARAM, "grt_scene_upload_triangles: NULL argument");
    if (n_tris == 0 || n_tris > (1u << 28))
        return set_error(GRT_INVALID_PARAM, "grt_scene_upload_triangles: n_tris=%u out of range",
                         n_tris);
    int rc = scene_alloc_geometry(sc, n_tris);
    if (rc != GRT_OK)
        return rc;
    const size_t b9 = (size_t)n_tris * 9 * sizeof(float);
    hipError_t e = hipMemcpyAsync(sc->verts.ptr, verts_host, b9, hipMemcpyHostToDevice,
                                  sc->ctx->stream);
    if (e != hipSuccess)
        return hip_check(e, "uploading vertices");
    if (sc->n_tris != n_tris || !sc->has_normals) {
        // Until normals are uploaded, use zero normals (secondary rays need them).
        if ((e = hipMemsetAsync(sc->normals.ptr, 0, b9, sc->ctx->stream)) != hipSuccess)
            return hip_check(e, "clearing normals");
        sc->has_normals = false;
    }
    sc->n_tris = n_tris;
    sc->grid_valid = false;
    sc->rest_valid = false;
    if ((rc = scene_pack(sc)) != GRT_OK)
        return rc;
    return hip_check(hipStreamSynchronize(sc->ctx->stream), "upload");
}

int grt_scene_upload_normals(grt_scene sc, const float* normals_host, uint32_t n_tris)
{
    if (!sc || !normals_host)
        return set_error(GRT_INVALID_PARAM, "grt_scene_upload_normals: NULL argument");
    if (n_tris != sc->n_tris)
        return set_error(GRT_INVALID_PARAM,
                         "grt_scene_upload_normals: %u normals for %u triangles", n_tris,
                         sc->n_tris);
    hipError_t e = hipMemcpyAsync(sc->normals.ptr, normals_host, (size_t)n_tris * 36,
                                  hipMemcpyHostToDevice, sc->ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(sc->ctx->stream);
    sc->has_normals = true;
    sc->rest_valid = false;
    return hip_check(e, "uploading normals");
}

int grt_scene_set_triangles_device(grt_scene sc, const float* verts_dev, const float* normals_dev,
                                   uint32_t n_tris)
{
    if (!sc || !verts_dev || n_tris == 0 || n_tris > (1u << 28))
        return set_error(GRT_INVALID_PARAM, "grt_scene_set_triangles_device: bad argument");
    int rc = scene_alloc_geometry(sc, n_tris);
    if (rc != GRT_OK)
        return rc;
    const size_t b9 = (size_t)n_tris * 36;
    hipError_t e = hipMemcpyAsync(sc->verts.ptr, verts_dev, b9, hipMemcpyDeviceToDevice,
                                  sc->ctx->stream);
    if (e == hipSuccess) {
        if (normals_dev)
            e = hipMemcpyAsync(sc->normals.ptr, normals_dev, b9, hipMemcpyDeviceToDevice,
                               sc->ctx->stream);
        else
            e = hipMemsetAsync(sc->normals.ptr, 0, b9, sc->ctx->stream);
    }
    if (e != hipSuccess)
        return hip_check(e, "device triangle copy");
    sc->n_tris = n_tris;
    sc->has_normals = normals_dev != nullptr;
    sc->grid_valid = false;
    sc->rest_valid = false;
    return scene_pack(sc);
}

int grt_scene_triangle_count(grt_scene sc, uint32_t* n)
{
    if (!sc || !n)
        return set_error(GRT_INVALID_PARAM, "grt_scene_triangle_count: NULL argument");
    *n = sc->n_tris;
    return GRT_OK;
}

int grt_scene_device_pointers(grt_scene sc, float** verts_dev, float** normals_dev)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_scene_device_pointers: NULL scene");
    if (verts_dev)
        *verts_dev = sc->verts.as<float>();
    if (normals_dev)
        *normals_dev = sc->normals.as<float>();
    return GRT_OK;
}

int grt_scene_download_geometry(grt_scene sc, float* verts_host, float* normals_host)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_scene_download_geometry: NULL scene");
    const size_t b9 = (size_t)sc->n_tris * 36;
    hipStream_t s = sc->ctx->stream;
    hipError_t e = hipSuccess;
    if (verts_host && b9)
        e = hipMemcpyAsync(verts_host, sc->verts.ptr, b9, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && normals_host && b9)
        e = hipMemcpyAsync(normals_host, sc->normals.ptr, b9, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    return hip_check(e, "geometry download");
}

int grt_scene_transform_clusters(grt_scene sc, const float* xf_host, uint32_t n_clusters,
                                 uint32_t tris_per_cluster, uint32_t n_cluster_tris)
{
    if (!sc || !xf_host || tris_per_cluster == 0)
        return set_error(GRT_INVALID_PARAM, "grt_scene_transform_clusters: bad argument");
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_scene_transform_clusters: empty scene");
    if (n_cluster_tris > sc->n_tris ||
        (uint64_t)n_clusters * tris_per_cluster < (uint64_t)n_cluster_tris)
        return set_error(GRT_INVALID_PARAM,
                         "grt_scene_transform_clusters: %u clusters x %u do not cover %u triangles",
                         n_clusters, tris_per_cluster, n_cluster_tris);
    int rc;
    hipStream_t s = sc->ctx->stream;
    const size_t b9 = (size_t)sc->n_tris * 36;
    if (!sc->rest_valid) {
        if ((rc = sc->rest_verts.reserve(sc->ctx, b9, false)) != GRT_OK)
            return rc;
        if ((rc = sc->rest_normals.reserve(sc->ctx, b9, false)) != GRT_OK)
            return rc;
        hipError_t e = hipMemcpyAsync(sc->rest_verts.ptr, sc->verts.ptr, b9,
                                      hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(sc->rest_normals.ptr, sc->normals.ptr, b9, hipMemcpyDeviceToDevice,
                               s);
        if (e != hipSuccess)
            return hip_check(e, "saving the rest pose");
        sc->rest_valid = true;
    }
    const size_t xb = (size_t)n_clusters * 12 * sizeof(float);
    if ((rc = sc->xf.reserve(sc->ctx, xb, false)) != GRT_OK)
        return rc;
    // Stage through pinned memory so the copy is truly asynchronous.  The
    // previous transform's copy must have left the staging buffer: wait for
    // that copy alone (an event recorded after it), not for the whole stream,
    // so the host can enqueue this frame while the previous one still runs.
    if (!sc->xf_copied) {
        hipError_t e = hipEventCreateWithFlags(&sc->xf_copied, hipEventDisableTiming);
        if (e != hipSuccess) {
            sc->xf_copied = nullptr;
            return hip_check(e, "transform staging event");
        }
    }
    if (sc->xf_pending) {
        hipError_t e = hipEventSynchronize(sc->xf_copied);
        if (e != hipSuccess)
            return hip_check(e, "transform staging");
        sc->xf_pending = false;
    }
    if (sc->xf_host_bytes < xb) {
        if (sc->xf_host)
            (void)hipHostFree(sc->xf_host);
        sc->xf_host = nullptr;
        sc->xf_host_bytes = 0;
        if (hipHostMalloc(&sc->xf_host, xb, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return set_error(GRT_NO_MEMORY, "pinned staging of %zu bytes failed", xb);
        }
        sc->xf_host_bytes = xb;
    }
    std::memcpy(sc->xf_host, xf_host, xb);
    hipError_t e = hipMemcpyAsync(sc->xf.ptr, sc->xf_host, xb, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipEventRecord(sc->xf_copied, s);
    if (e != hipSuccess)
        return hip_check(e, "uploading transforms");
    sc->xf_pending = true;
    sc->grid_valid = false;
    return hip_check(launch_transform_pack(sc->n_tris, tris_per_cluster, n_cluster_tris,
                                           sc->xf.as<float>(), sc->rest_verts.as<float>(),
                                           sc->rest_normals.as<float>(), sc->verts.as<float>(),
                                           sc->normals.as<float>(), sc->tri.as<float4>(), s),
                     "transform");
}

int grt_grid_build(grt_scene sc, float lambda)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_grid_build: NULL scene");
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_grid_build: empty scene");
    if (!(lambda > 0.0f) || !std::isfinite(lambda))
        return set_error(GRT_INVALID_PARAM, "grt_grid_build: lambda must be > 0");
    hipStream_t s = sc->ctx->stream;
    int rc;
    DevFrame host;
    if ((rc = hip_check(launch_frame_from_verts(sc->verts.as<float>(), sc->n_tris, lambda,
                                                0xffffffffu, sc->partial.as<float>(),
                                                sc->frame.as<DevFrame>(), s),
                        "grid frame")) != GRT_OK)
        return rc;
    if ((rc = fetch_frame(sc, &host)) != GRT_OK)
        return rc;
    if (host.status & kStatusFrameInvalid)
        return set_error(GRT_CAPACITY, "grid frame invalid (cells %u x %u x %u)", host.n[0],
                         host.n[1], host.n[2]);
    if ((rc = grid_reserve_cells(sc, host.m)) != GRT_OK)
        return rc;
    // Count and scan to learn the reference total, size the reference
    // buffers, then scan again to collect the sort lists into them.
    if ((rc = enqueue_build_passes(sc, 0xffffffffu, false)) != GRT_OK)
        return rc;
    if ((rc = fetch_frame(sc, &host)) != GRT_OK)
        return rc;
    if (host.status != 0)
        return set_error(GRT_CAPACITY, "grid build status 0x%x", host.status);
    if ((rc = grid_reserve_refs(sc, host.r > 0 ? host.r : 1)) != GRT_OK) {
        // The count pass ran and the scatter will not: return the counts to
        // zero for the next build.
        (void)launch_fill_u32(sc->counts.as<uint32_t>(), 0u, sc->cells_cap, s);
        (void)hipStreamSynchronize(s);
        return rc;
    }
    if ((rc = enqueue_scan(sc, sc->refs_cap, true, true)) != GRT_OK)
        return rc;
    if ((rc = enqueue_scatter_sort(sc)) != GRT_OK)
        return rc;
    if ((rc = fetch_frame(sc, &host)) != GRT_OK)
        return rc;
    if (host.status != 0)
        return set_error(GRT_CAPACITY, "grid build status 0x%x", host.status);
    sc->info_host = host;
    sc->lambda = lambda;
    sc->grid_valid = true;
    return g2_build_sync(sc);
}

int grt_grid_reserve(grt_scene sc, uint32_t cells, uint32_t refs)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_grid_reserve: NULL scene");
    const void* before[4] = {sc->offsets.ptr, sc->refs.ptr, sc->trir.ptr, sc->cellrec.ptr};
    int rc;
    if ((rc = grid_reserve_cells(sc, cells)) == GRT_OK &&
        (rc = grid_reserve_refs(sc, refs)) == GRT_OK && sc->g2_lambda > 0.0f &&
        (sc->g2_top_cap < sc->cells_cap || sc->g2_box_cap < sc->refs_cap))
        rc = g2_reserve_top(sc);
    // A grown buffer is a new allocation without the old contents: the grid
    // is gone until the next build (traces then report GRT_NOT_READY instead
    // of reading uninitialised offsets).
    const void* after[4] = {sc->offsets.ptr, sc->refs.ptr, sc->trir.ptr, sc->cellrec.ptr};
    if (!std::equal(before, before + 4, after)) {
        sc->grid_valid = false;
        sc->g2_built = false;
    }
    return rc;
}

static bool stream_capturing(hipStream_t s)
{
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

static int aux_stream_ready(grt_context_t* ctx)
{
    if (ctx->aux_stream)
        return GRT_OK;
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess)
        return hip_check(e, "hipStreamCreate");
    for (hipEvent_t& ev : ctx->aux_ev) {
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) {
            for (hipEvent_t& done : ctx->aux_ev) {
                if (done)
                    (void)hipEventDestroy(done);
                done = nullptr;
            }
            (void)hipStreamDestroy(st);
            return hip_check(e, "hipEventCreate");
        }
    }
    ctx->aux_stream = st;
    return GRT_OK;
}

// The rest of an asynchronous rebuild of both levels, after the top scan, on
// two streams.  What the second level plans from the top grid's offsets
// alone (cell plans, the occupied cells' records, the dense list) runs on the
// auxiliary stream while the main stream scatters and sorts the top
// references; the skip radii and empty cells' records, then the top records,
// follow on the auxiliary stream while the main stream builds the sub-grids
// from the plan and the sorted references.  The main stream waits for the
// auxiliary one at the end, so callers see one stream-ordered rebuild.
static int rebuild_two_streams_enqueue(grt_scene_t* sc);

static int rebuild_two_streams(grt_scene_t* sc)
{
    const int rc = rebuild_two_streams_enqueue(sc);
    if (rc != GRT_OK) {
        // keep the callers' single-stream view even after a failed enqueue
        grt_context_t* ctx = sc->ctx;
        if (hipEventRecord(ctx->aux_ev[3], ctx->aux_stream) == hipSuccess)
            (void)hipStreamWaitEvent(ctx->stream, ctx->aux_ev[3], 0);
    }
    return rc;
}

static int rebuild_two_streams_enqueue(grt_scene_t* sc)
{
    grt_context_t* ctx = sc->ctx;
    hipStream_t s = ctx->stream, aux = ctx->aux_stream;
    hipEvent_t* ev = ctx->aux_ev;
    sc->g2_built = false;
    const G2Build b = g2_args(sc, true);
    int rc;
    // offsets ready -> the plan on aux
    if ((rc = hip_check(hipEventRecord(ev[0], s), "hipEventRecord")) != GRT_OK ||
        (rc = hip_check(hipStreamWaitEvent(aux, ev[0], 0), "hipStreamWaitEvent")) != GRT_OK ||
        (rc = hip_check(launch_grid2_plan_cells(b, sc->cells_cap, sc->g2_cells_cap,
                                                sc->g2_dense_cap, aux),
                        "sub-grid plan")) != GRT_OK ||
        (rc = hip_check(hipEventRecord(ev[1], aux), "hipEventRecord")) != GRT_OK ||
        (rc = hip_check(launch_grid2_skip_records(b, sc->cells_cap, aux), "skip radii")) !=
            GRT_OK)
        return rc;
    // top references scattered and sorted on the main stream
    if ((rc = enqueue_scatter_sort_only(sc)) != GRT_OK)
        return rc;
    // sorted references -> the top records on aux
    if ((rc = hip_check(hipEventRecord(ev[2], s), "hipEventRecord")) != GRT_OK ||
        (rc = hip_check(hipStreamWaitEvent(aux, ev[2], 0), "hipStreamWaitEvent")) != GRT_OK ||
        (rc = hip_check(launch_gather_refs(sc->frame.as<DevFrame>(), sc->refs.as<uint32_t>(),
                                           sc->tri.as<float4>(), sc->trir.as<float4>(),
                                           sc->refs_cap, aux),
                        "grid records")) != GRT_OK ||
        (rc = hip_check(hipEventRecord(ev[3], aux), "hipEventRecord")) != GRT_OK)
        return rc;
    // the sub-grids from the plan and the sorted references on the main stream
    if ((rc = hip_check(hipStreamWaitEvent(s, ev[1], 0), "hipStreamWaitEvent")) != GRT_OK ||
        (rc = hip_check(launch_grid2_plan_refs(b, sc->g2_refs_cap, sc->g2_dense_cap, s),
                        "sub-grid plan")) != GRT_OK ||
        (rc = hip_check(launch_grid2_fill(b, sc->g2_cells_cap, sc->g2_refs_cap, sc->g2_dense_cap,
                                          sc->tri.as<float4>(), sc->sub_trir.as<float4>(), s),
                        "sub-grid build")) != GRT_OK)
        return rc;
    if ((rc = hip_check(hipStreamWaitEvent(s, ev[3], 0), "hipStreamWaitEvent")) != GRT_OK)
        return rc;
    sc->g2_built = true;
    return GRT_OK;
}

int grt_grid_rebuild_async(grt_scene sc, float lambda)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_grid_rebuild_async: NULL scene");
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_grid_rebuild_async: empty scene");
    if (!sc->counts.ptr || !sc->refs.ptr)
        return set_error(GRT_NOT_READY,
                         "grt_grid_rebuild_async: call grt_grid_build or grt_grid_reserve first");
    if (!(lambda > 0.0f) || !std::isfinite(lambda))
        return set_error(GRT_INVALID_PARAM, "grt_grid_rebuild_async: lambda must be > 0");
    hipStream_t s = sc->ctx->stream;
    int rc;
    if ((rc = hip_check(launch_frame_from_verts(sc->verts.as<float>(), sc->n_tris, lambda,
                                                sc->cells_cap, sc->partial.as<float>(),
                                                sc->frame.as<DevFrame>(), s),
                        "grid frame")) != GRT_OK)
        return rc;
    if ((rc = enqueue_build_passes(sc, sc->refs_cap, true)) != GRT_OK)
        return rc;
    sc->lambda = lambda;
    sc->grid_valid = true;
    sc->info_stale = true;
    if (g2_ready(sc) && !stream_capturing(s) && aux_stream_ready(sc->ctx) == GRT_OK)
        return rebuild_two_streams(sc);
    if ((rc = enqueue_scatter_sort(sc)) != GRT_OK)
        return rc;
    return g2_enqueue(sc);
}

int grt_grid_check(grt_scene sc)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_grid_check: NULL scene");
    if (!sc->grid_valid)
        return set_error(GRT_NOT_READY, "grt_grid_check: no grid");
    DevFrame host;
    int rc = fetch_frame(sc, &host);
    if (rc != GRT_OK)
        return rc;
    sc->info_host = host;
    sc->info_stale = false;
    // A rebuild over the reference capacity zeroed the counts on the device
    // (the scan's apply pass), so the next build needs no clean-up here.
    if (host.status == 0 && sc->g2_built) {
        DevFrame g2;
        hipError_t e = hipMemcpyAsync(&g2, sc->g2frame.ptr, sizeof(g2), hipMemcpyDeviceToHost,
                                      sc->ctx->stream);
        if (e == hipSuccess)
            e = hipStreamSynchronize(sc->ctx->stream);
        if (e != hipSuccess)
            return hip_check(e, "reading the sub-grid status");
        // Over its capacity the second level is left out (every cell record
        // says "no sub-grid"): traces stay correct on the top grid alone.
        if (g2.status != 0)
            return set_error(GRT_CAPACITY,
                             "sub-grids exceeded their capacity (status 0x%x: %u cells, %u "
                             "references); this frame traced the top grid alone",
                             g2.status, g2.m, g2.r);
    }
    if (host.status != 0) {
        sc->grid_valid = false;
        return set_error(GRT_CAPACITY, "grid rebuild exceeded capacity (status 0x%x)",
                         host.status);
    }
    return GRT_OK;
}

int grt_grid_upload(grt_scene sc, const grt_grid_info* info, const uint32_t* offsets_host,
                    const uint32_t* refs_host)
{
    if (!sc || !info || !offsets_host)
        return set_error(GRT_INVALID_PARAM, "grt_grid_upload: NULL argument");
    uint64_t m = 1;
    for (int d = 0; d < 3; ++d) {
        if (info->n[d] < 1)
            return set_error(GRT_INVALID_PARAM, "grt_grid_upload: n[%d]=%d", d, info->n[d]);
        m *= (uint64_t)info->n[d];
    }
    if (m > (1ull << 30))
        return set_error(GRT_INVALID_PARAM, "grt_grid_upload: too many cells");
    const uint32_t r = offsets_host[m];
    if (r && !refs_host)
        return set_error(GRT_INVALID_PARAM, "grt_grid_upload: refs is NULL");
    if (offsets_host[0] != 0)
        return set_error(GRT_INVALID_PARAM, "grt_grid_upload: offsets[0] must be 0");
    for (uint64_t c = 0; c < m; ++c)
        if (offsets_host[c + 1] < offsets_host[c])
            return set_error(GRT_INVALID_PARAM, "grt_grid_upload: offsets not monotone at %llu",
                             (unsigned long long)c);
    for (uint32_t k = 0; k < r; ++k)
        if (refs_host[k] >= sc->n_tris)
            return set_error(GRT_INVALID_PARAM, "grt_grid_upload: ref %u = %u >= %u triangles", k,
                             refs_host[k], sc->n_tris);
    int rc;
    if ((rc = grid_reserve_cells(sc, (uint32_t)m)) != GRT_OK)
        return rc;
    if ((rc = grid_reserve_refs(sc, r > 0 ? r : 1)) != GRT_OK)
        return rc;
    DevFrame f;
    std::memset(&f, 0, sizeof(f));
    for (int d = 0; d < 3; ++d) {
        f.bmin[d] = info->bmin[d];
        f.bmax[d] = info->bmax[d];
        f.cell[d] = info->cell[d];
        f.inv[d] = info->inv_cell[d];
        f.n[d] = info->n[d];
    }
    f.m = (uint32_t)m;
    f.r = r;
    hipStream_t s = sc->ctx->stream;
    hipError_t e = hipMemcpyAsync(sc->frame.ptr, &f, sizeof(f), hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = hipMemcpyAsync(sc->offsets.ptr, offsets_host, (m + 1) * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess && r)
        e = hipMemcpyAsync(sc->refs.ptr, refs_host, (size_t)r * 4, hipMemcpyHostToDevice, s);
    if (e == hipSuccess)
        e = launch_gather_refs(sc->frame.as<DevFrame>(), sc->refs.as<uint32_t>(),
                               sc->tri.as<float4>(), sc->trir.as<float4>(), sc->refs_cap, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return hip_check(e, "grid upload");
    sc->info_host = f;
    sc->grid_valid = true;
    sc->g2_built = false;   // an uploaded grid has no second level
    sc->info_stale = false;
    return GRT_OK;
}

int grt_grid_set_subgrid(grt_scene sc, float lambda2, uint32_t min_refs)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_grid_set_subgrid: NULL scene");
    if (!(lambda2 >= 0.0f) || !std::isfinite(lambda2))
        return set_error(GRT_INVALID_PARAM, "grt_grid_set_subgrid: lambda2 must be >= 0");
    if (lambda2 > 0.0f && min_refs < 2)
        return set_error(GRT_INVALID_PARAM, "grt_grid_set_subgrid: min_refs must be >= 2");
    sc->g2_lambda = lambda2;
    sc->g2_min_refs = min_refs;
    sc->g2_built = false;   // takes effect with the next build
    return GRT_OK;
}

int grt_subgrid_info_get(grt_scene sc, grt_subgrid_info* out)
{
    if (!sc || !out)
        return set_error(GRT_INVALID_PARAM, "grt_subgrid_info_get: NULL argument");
    std::memset(out, 0, sizeof(*out));
    out->lambda2 = sc->g2_lambda;
    out->min_refs = sc->g2_min_refs;
    out->built = sc->g2_built ? 1u : 0u;
    if (!sc->g2_built)
        return GRT_OK;
    DevFrame g2;
    hipError_t e = hipMemcpyAsync(&g2, sc->g2frame.ptr, sizeof(g2), hipMemcpyDeviceToHost,
                                  sc->ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(sc->ctx->stream);
    if (e != hipSuccess)
        return hip_check(e, "reading the sub-grid totals");
    out->cells = g2.m;
    out->refs = g2.r;
    out->status = g2.status;
    return GRT_OK;
}

int grt_subgrid_download(grt_scene sc, uint32_t* cellsub_host, uint32_t* sub_offsets_host,
                         uint32_t* sub_refs_host)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "grt_subgrid_download: NULL scene");
    grt_subgrid_info si;
    int rc = grt_subgrid_info_get(sc, &si);
    if (rc != GRT_OK)
        return rc;
    if (!si.built || si.status != 0)
        return set_error(GRT_NOT_READY, "grt_subgrid_download: no second level built");
    grt_grid_info gi;
    if ((rc = grt_grid_info_get(sc, &gi)) != GRT_OK)
        return rc;
    hipStream_t s = sc->ctx->stream;
    hipError_t e = hipSuccess;
    if (cellsub_host) {
        std::vector<uint32_t> rec((size_t)gi.m * 2);
        e = hipMemcpyAsync(rec.data(), sc->cellrec.ptr, rec.size() * 4, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess)
            e = hipStreamSynchronize(s);
        for (size_t c = 0; e == hipSuccess && c < gi.m; ++c) {
            const bool sub = (rec[2 * c] >> 31) != 0;
            const bool empty = rec[2 * c] == rec[2 * c + 1];
            cellsub_host[2 * c] = sub ? rec[2 * c] & 0x7fffffffu : 0u;
            cellsub_host[2 * c + 1] = sub ? rec[2 * c + 1] : empty ? 0x80000000u | rec[2 * c] : 0u;
        }
    }
    if (e == hipSuccess && sub_offsets_host)
        e = hipMemcpyAsync(sub_offsets_host, sc->sub_offsets.ptr, ((size_t)si.cells + 1) * 4,
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && sub_refs_host && si.refs)
        e = hipMemcpyAsync(sub_refs_host, sc->sub_refs.ptr, (size_t)si.refs * 4,
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    return hip_check(e, "sub-grid download");
}

int grt_grid_info_get(grt_scene sc, grt_grid_info* out)
{
    if (!sc || !out)
        return set_error(GRT_INVALID_PARAM, "grt_grid_info_get: NULL argument");
    if (!sc->grid_valid)
        return set_error(GRT_NOT_READY, "grt_grid_info_get: no grid");
    if (sc->info_stale) {
        int rc = grt_grid_check(sc);
        if (rc != GRT_OK)
            return rc;
    }
    frame_to_info(sc->info_host, out);
    return GRT_OK;
}

int grt_grid_download(grt_scene sc, uint32_t* offsets_host, uint32_t* refs_host)
{
    grt_grid_info info;
    int rc = grt_grid_info_get(sc, &info);
    if (rc != GRT_OK)
        return rc;
    hipStream_t s = sc->ctx->stream;
    hipError_t e = hipSuccess;
    if (offsets_host)
        e = hipMemcpyAsync(offsets_host, sc->offsets.ptr, ((size_t)info.m + 1) * 4,
                           hipMemcpyDeviceToHost, s);
    if (e == hipSuccess && refs_host && info.r)
        e = hipMemcpyAsync(refs_host, sc->refs.ptr, (size_t)info.r * 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    return hip_check(e, "grid download");
}

static bool rays_ok(const grt_rays* r)
{
    return r && r->ox && r->oy && r->oz && r->dx && r->dy && r->dz && r->tmin && r->tmax;
}
static bool hits_ok(const grt_hits* h) { return h && h->id && h->t && h->u && h->v; }

static RaySoA to_soa(const grt_rays* r)
{
    return RaySoA{r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
}
static RaySoAOut to_soa_out(const grt_rays* r)
{
    return RaySoAOut{r->ox, r->oy, r->oz, r->dx, r->dy, r->dz, r->tmin, r->tmax};
}
static HitSoAIn to_hits_in(const grt_hits* h, uint32_t n_tris)
{
    return HitSoAIn{h->id, h->t, h->u, h->v, n_tris};
}

static int trace_ready(grt_scene sc, const char* who)
{
    if (!sc)
        return set_error(GRT_INVALID_PARAM, "%s: NULL scene", who);
    if (sc->n_tris == 0 || !sc->grid_valid)
        return set_error(GRT_NOT_READY, "%s: scene has no grid", who);
    return GRT_OK;
}

static TraceArgs trace_args(grt_scene sc, uint32_t n, uint32_t image_width, uint32_t* counters)
{
    TraceArgs a;
    a.fr = sc->frame.as<DevFrame>();
    a.offsets = sc->offsets.as<uint32_t>();
    a.refs = sc->refs.as<uint32_t>();
    a.tri = sc->tri.as<float4>();
    a.trir = sc->trir.as<float4>();
    a.n = n;
    a.image_width = image_width;
    a.counters = counters;
    a.coop = sc->ctx->coop;
    a.pool = sc->ctx->trace_mode == GRT_TRACE_POOLED ? 1u : 0u;
    if (sc->g2_built) {
        a.g2 = sc->g2frame.as<DevFrame>();
        a.cellrec = sc->cellrec.as<uint2>();
        a.sub_offsets = sc->sub_offsets.as<uint32_t>();
        a.sub_trir = sc->sub_trir.as<float4>();
        a.sub_tab = sc->g2tab.as<float>();
    }
    return a;
}

// Lanes per ray of a two-level trace: the split schedule's, else one.
static uint32_t trace2_lanes(const grt_context_t* ctx)
{
    return ctx->trace_mode == GRT_TRACE_SPLIT ? ctx->split_lanes : 1u;
}

// The slot counter of a persistent launch on the context's current stream.
static uint32_t* stream_counter(grt_context_t* ctx, int& rc)
{
    grt::DevMem& m = ctx->stream_counters[ctx->stream];   // allocated by grt_set_stream
    rc = m.reserve(ctx, 64, false);
    return rc == GRT_OK ? m.as<uint32_t>() : nullptr;
}

// One closest-hit launch of the context's schedule.
static int dispatch_closest(grt_scene sc, const TraceArgs& a, const grt_rays* rays,
                            const grt_hits* hits)
{
    const HitSoA h{hits->id, hits->t, hits->u, hits->v};
    grt_context_t* ctx = sc->ctx;
    if (a.cellrec)
        return hip_check(launch_trace2(a, to_soa(rays), h, nullptr, trace2_lanes(ctx),
                                       ctx->stream),
                         "trace_closest");
    if (ctx->trace_mode == GRT_TRACE_PERSISTENT) {
        int rc = GRT_OK;
        uint32_t* counter = stream_counter(ctx, rc);
        if (rc != GRT_OK)
            return rc;
        return hip_check(launch_trace_persistent(a, to_soa(rays), h, nullptr, false, counter,
                                                 ctx->block_closest, ctx->refill, ctx->stream),
                         "trace_closest");
    }
    if (ctx->trace_mode == GRT_TRACE_SPLIT)
        return hip_check(launch_trace_split(a, to_soa(rays), h, nullptr, false, ctx->split_lanes,
                                            ctx->block_closest, ctx->stream),
                         "trace_closest");
    return hip_check(launch_trace_closest(a, to_soa(rays), h, ctx->block_closest, ctx->stream),
                     "trace_closest");
}

// One any-hit launch of the context's schedule.
static int dispatch_any(grt_scene sc, const TraceArgs& a, const grt_rays* rays,
                        uint8_t* occluded_dev)
{
    grt_context_t* ctx = sc->ctx;
    if (a.cellrec)
        return hip_check(launch_trace2(a, to_soa(rays), HitSoA{}, occluded_dev, trace2_lanes(ctx),
                                       ctx->stream),
                         "trace_any");
    if (ctx->trace_mode == GRT_TRACE_PERSISTENT) {
        int rc = GRT_OK;
        uint32_t* counter = stream_counter(ctx, rc);
        if (rc != GRT_OK)
            return rc;
        return hip_check(launch_trace_persistent(a, to_soa(rays), HitSoA{}, occluded_dev, true,
                                                 counter, ctx->block_any, ctx->refill,
                                                 ctx->stream),
                         "trace_any");
    }
    if (ctx->trace_mode == GRT_TRACE_SPLIT)
        return hip_check(launch_trace_split(a, to_soa(rays), HitSoA{}, occluded_dev, true,
                                            ctx->split_lanes, ctx->block_any, ctx->stream),
                         "trace_any");
    return hip_check(launch_trace_any(a, to_soa(rays), occluded_dev, ctx->block_any, ctx->stream),
                     "trace_any");
}

int grt_trace_closest(grt_scene sc, const grt_rays* rays, const grt_hits* hits, uint32_t n,
                      uint32_t image_width, uint32_t* counters_dev)
{
    int rc = trace_ready(sc, "grt_trace_closest");
    if (rc != GRT_OK)
        return rc;
    if (n == 0)
        return GRT_OK;   // nothing to trace; the buffers may be NULL
    if (!rays_ok(rays) || !hits_ok(hits))
        return set_error(GRT_INVALID_PARAM, "grt_trace_closest: NULL ray or hit buffer");
    return dispatch_closest(sc, trace_args(sc, n, image_width, counters_dev), rays, hits);
}

int grt_trace_any(grt_scene sc, const grt_rays* rays, uint8_t* occluded_dev, uint32_t n,
                  uint32_t image_width, uint32_t* counters_dev)
{
    int rc = trace_ready(sc, "grt_trace_any");
    if (rc != GRT_OK)
        return rc;
    if (n == 0)
        return GRT_OK;   // nothing to trace; the buffers may be NULL
    if (!rays_ok(rays) || !occluded_dev)
        return set_error(GRT_INVALID_PARAM, "grt_trace_any: NULL ray or output buffer");
    return dispatch_any(sc, trace_args(sc, n, image_width, counters_dev), rays, occluded_dev);
}

int grt_compact_rays(grt_context ctx, const grt_rays* rays, uint32_t n, uint32_t image_width,
                     uint32_t* list_dev, uint32_t* count_dev, uint8_t* occluded_dev,
                     const grt_hits* miss_hits)
{
    if (!ctx || !rays_ok(rays) || !list_dev || !count_dev || (miss_hits && !hits_ok(miss_hits)))
        return set_error(GRT_INVALID_PARAM, "grt_compact_rays: bad argument");
    const uint64_t slots = slot_count(n, image_width);
    if (slots > 0xffffffffull - kScanTile)
        return set_error(GRT_INVALID_PARAM, "grt_compact_rays: too many rays");
    int rc = ctx->compact_scratch.reserve(ctx, (slots / kScanTile + 2) * 4, false);
    if (rc != GRT_OK)
        return rc;
    const HitSoA mh = miss_hits ? HitSoA{miss_hits->id, miss_hits->t, miss_hits->u, miss_hits->v}
                                : HitSoA{nullptr, nullptr, nullptr, nullptr};
    return hip_check(launch_compact_rays(n, image_width, rays->tmin, rays->tmax, list_dev,
                                         count_dev, occluded_dev, mh,
                                         ctx->compact_scratch.as<uint32_t>(), ctx->stream),
                     "compact_rays");
}

int grt_trace_closest_list(grt_scene sc, const grt_rays* rays, const grt_hits* hits,
                           const uint32_t* list_dev, const uint32_t* count_dev, uint32_t n,
                           uint32_t* counters_dev)
{
    int rc = trace_ready(sc, "grt_trace_closest_list");
    if (rc != GRT_OK)
        return rc;
    if (n == 0)
        return GRT_OK;   // nothing to trace; the buffers may be NULL
    if (!rays_ok(rays) || !hits_ok(hits) || !list_dev || !count_dev)
        return set_error(GRT_INVALID_PARAM, "grt_trace_closest_list: bad argument");
    TraceArgs a = trace_args(sc, n, 0, counters_dev);
    a.list = list_dev;
    a.list_count = count_dev;
    return dispatch_closest(sc, a, rays, hits);
}

int grt_trace_any_list(grt_scene sc, const grt_rays* rays, uint8_t* occluded_dev,
                       const uint32_t* list_dev, const uint32_t* count_dev, uint32_t n,
                       uint32_t* counters_dev)
{
    int rc = trace_ready(sc, "grt_trace_any_list");
    if (rc != GRT_OK)
        return rc;
    if (n == 0)
        return GRT_OK;   // nothing to trace; the buffers may be NULL
    if (!rays_ok(rays) || !occluded_dev || !list_dev || !count_dev)
        return set_error(GRT_INVALID_PARAM, "grt_trace_any_list: bad argument");
    TraceArgs a = trace_args(sc, n, 0, counters_dev);
    a.list = list_dev;
    a.list_count = count_dev;
    return dispatch_any(sc, a, rays, occluded_dev);
}

int grt_camera_lookat(const float eye[3], const float center[3], const float up[3],
                      float fovy_deg, float znear, uint32_t width, uint32_t height,
                      grt_camera* out)
{
    if (!eye || !center || !up || !out || width == 0 || height == 0)
        return set_error(GRT_INVALID_PARAM, "grt_camera_lookat: bad argument");
    camera_lookat(eye, center, up, fovy_deg, znear, width, height, out);
    return GRT_OK;
}

int grt_gen_primary(grt_context ctx, const grt_camera* cam, const grt_rays* out)
{
    if (!ctx || !cam || !rays_ok(out))
        return set_error(GRT_INVALID_PARAM, "grt_gen_primary: bad argument");
    if ((uint64_t)cam->width * cam->height > 0xffffffffull)
        return set_error(GRT_INVALID_PARAM, "grt_gen_primary: image too large");
    DevCamera c;
    std::memcpy(c.pos, cam->pos, sizeof(c.pos));
    std::memcpy(c.base, cam->base, sizeof(c.base));
    std::memcpy(c.nu, cam->nu, sizeof(c.nu));
    std::memcpy(c.nv, cam->nv, sizeof(c.nv));
    c.width = cam->width;
    c.height = cam->height;
    return hip_check(launch_gen_primary(c, to_soa_out(out), ctx->stream), "gen_primary");
}

int grt_gen_shadow(grt_scene sc, const grt_rays* in, const grt_hits* hits, uint32_t n,
                   const float light[3], float offset, const grt_rays* out)
{
    if (!sc || !rays_ok(in) || !hits_ok(hits) || !rays_ok(out) || !light)
        return set_error(GRT_INVALID_PARAM, "grt_gen_shadow: bad argument");
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_gen_shadow: empty scene");
    return hip_check(launch_gen_shadow(n, to_soa(in), to_hits_in(hits, sc->n_tris), sc->normals.as<float>(),
                                       light[0], light[1], light[2], offset, to_soa_out(out),
                                       sc->ctx->stream),
                     "gen_shadow");
}

int grt_gen_reflection(grt_scene sc, const grt_rays* in, const grt_hits* hits, uint32_t n,
                       float offset, const grt_rays* out)
{
    if (!sc || !rays_ok(in) || !hits_ok(hits) || !rays_ok(out))
        return set_error(GRT_INVALID_PARAM, "grt_gen_reflection: bad argument");
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_gen_reflection: empty scene");
    return hip_check(launch_gen_reflection(n, to_soa(in), to_hits_in(hits, sc->n_tris),
                                           sc->normals.as<float>(), offset, to_soa_out(out),
                                           sc->ctx->stream),
                     "gen_reflection");
}

int grt_gen_ao(grt_scene sc, const grt_rays* in, const grt_hits* hits, uint32_t n, uint32_t spp,
               uint64_t seed, float offset, float ao_tmax, const grt_rays* out)
{
    if (!sc || !rays_ok(in) || !hits_ok(hits) || !rays_ok(out) || spp == 0)
        return set_error(GRT_INVALID_PARAM, "grt_gen_ao: bad argument");
    if ((uint64_t)n * spp > 0xffffffffull)
        return set_error(GRT_INVALID_PARAM, "grt_gen_ao: too many rays");
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_gen_ao: empty scene");
    return hip_check(launch_gen_ao(n, spp, seed, to_soa(in), to_hits_in(hits, sc->n_tris),
                                   sc->normals.as<float>(), offset, ao_tmax, to_soa_out(out),
                                   sc->ctx->stream),
                     "gen_ao");
}

int grt_shade_headlight(grt_scene sc, const grt_rays* rays, const grt_hits* hits, uint32_t n,
                        uint32_t* rgba_dev)
{
    if (!sc || !rays_ok(rays) || !hits_ok(hits) || !rgba_dev)
        return set_error(GRT_INVALID_PARAM, "grt_shade_headlight: bad argument");
    return hip_check(launch_shade_headlight(n, to_soa(rays), to_hits_in(hits, sc->n_tris),
                                            sc->normals.as<float>(), rgba_dev, sc->ctx->stream),
                     "shade_headlight");
}

int grt_affine_inverse(const float xf[12], float inv[12])
{
    if (!xf || !inv)
        return set_error(GRT_INVALID_PARAM, "grt_affine_inverse: NULL argument");
    double m[3][3], c[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
            m[r][k] = xf[4 * r + k];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            c[r][k] = m[r1][k1] * m[r2][k2] - m[r1][k2] * m[r2][k1];
        }
    const double det = m[0][0] * c[0][0] + m[0][1] * c[0][1] + m[0][2] * c[0][2];
    if (!(det != 0.0) || !std::isfinite(det))
        return set_error(GRT_INVALID_PARAM, "grt_affine_inverse: singular matrix");
    // inverse = transpose(cofactors) / det; translation = -inverse * t
    double a[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k)
            a[r][k] = c[k][r] / det;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k)
            inv[4 * r + k] = (float)a[r][k];
        inv[4 * r + 3] = (float)(-(a[r][0] * xf[3] + a[r][1] * xf[7] + a[r][2] * xf[11]));
    }
    return GRT_OK;
}

int grt_trace_instances(grt_context ctx, const grt_instance* instances, uint32_t n_instances,
                        const grt_rays* rays, const grt_hits* hits, int32_t* instance_dev,
                        uint32_t n, uint32_t image_width)
{
    if (!ctx || (!instances && n_instances) || !rays_ok(rays) || !hits_ok(hits) || !instance_dev)
        return set_error(GRT_INVALID_PARAM, "grt_trace_instances: bad argument");
    std::vector<Affine> inv(n_instances);
    for (uint32_t k = 0; k < n_instances; ++k) {
        grt_scene sc = instances[k].scene;
        int rc = trace_ready(sc, "grt_trace_instances");
        if (rc != GRT_OK)
            return rc;
        if (sc->ctx != ctx)
            return set_error(GRT_INVALID_PARAM, "grt_trace_instances: instance %u of another context", k);
        if ((rc = grt_affine_inverse(instances[k].xf, inv[k].m)) != GRT_OK)
            return rc;
    }
    hipStream_t s = ctx->stream;
    const HitSoA h{hits->id, hits->t, hits->u, hits->v};
    int rc = hip_check(launch_instance_init(n, h, instance_dev, s), "instance init");
    if (rc != GRT_OK || n == 0 || n_instances == 0)
        return rc;
    if ((rc = ctx->instance_scratch.reserve(ctx, (size_t)n * 48, false)) != GRT_OK)
        return rc;
    float* f = ctx->instance_scratch.as<float>();
    const size_t N = n;
    const grt_rays local = {f, f + N, f + 2 * N, f + 3 * N, f + 4 * N, f + 5 * N, f + 6 * N, f + 7 * N};
    const grt_hits cand = {reinterpret_cast<int32_t*>(f + 8 * N), f + 9 * N, f + 10 * N, f + 11 * N};
    for (uint32_t k = 0; k < n_instances; ++k) {
        grt_scene sc = instances[k].scene;
        if ((rc = hip_check(launch_instance_rays(n, to_soa(rays), hits->t, inv[k], to_soa_out(&local),
                                                 s),
                            "instance rays")) != GRT_OK)
            return rc;
        if ((rc = grt_trace_closest(sc, &local, &cand, n, image_width, nullptr)) != GRT_OK)
            return rc;
        if ((rc = hip_check(launch_instance_merge(n, to_hits_in(&cand, sc->n_tris), h, instance_dev,
                                                  (int32_t)k, s),
                            "instance merge")) != GRT_OK)
            return rc;
    }
    return GRT_OK;
}

void grt_phong_defaults(grt_phong* p)
{
    if (!p)
        return;
    for (int c = 0; c < 3; ++c) {
        p->light[c] = 1e10f;
        p->light_color[c] = 1.0f;
        p->ambient[c] = 0.0f;
        p->specular[c] = 0.5f;
        p->background[c] = 1.0f;
    }
    p->spec_exp = 50;
    p->shadow_coeff = 0.2f;
    p->reflect_coeff = 0.4f;
}

int grt_shade_phong(grt_scene sc, const grt_rays* rays, const grt_hits* hits,
                    const uint8_t* occluded_dev, uint32_t n, const float* diffuse_dev,
                    const grt_phong* params, uint32_t depth, float* rgb_dev, uint32_t* rgba8_dev)
{
    if (!sc || !rays_ok(rays) || !hits_ok(hits) || !params || !rgb_dev)
        return set_error(GRT_INVALID_PARAM, "grt_shade_phong: bad argument");
    if (!sc->has_normals)
        return set_error(GRT_NOT_READY, "grt_shade_phong: scene has no normals");
    DevPhong p;
    for (int c = 0; c < 3; ++c) {
        p.light[c] = params->light[c];
        p.light_color[c] = params->light_color[c];
        p.ambient[c] = params->ambient[c];
        p.specular[c] = params->specular[c];
        p.background[c] = params->background[c];
    }
    p.spec_exp = params->spec_exp;
    p.shadow_coeff = params->shadow_coeff;
    p.reflect_coeff = params->reflect_coeff;
    return hip_check(launch_shade_phong(n, to_soa(rays), to_hits_in(hits, sc->n_tris),
                                        occluded_dev, sc->normals.as<float>(), diffuse_dev, p,
                                        depth, rgb_dev, rgba8_dev, sc->ctx->stream),
                     "shade_phong");
}

int grt_debug_scan(grt_context ctx, const uint32_t* in_dev, uint32_t* out_dev, uint32_t n)
{
    if (!ctx || !in_dev || !out_dev || n == 0)
        return set_error(GRT_INVALID_PARAM, "grt_debug_scan: bad argument");
    DevMem bs;
    int rc = bs.reserve(ctx, ((size_t)n / kScanTile + 2) * 4, false);
    if (rc != GRT_OK)
        return rc;
    hipError_t e = launch_scan_u32(in_dev, out_dev, n, bs.as<uint32_t>(), nullptr, ctx->stream);
    if (e == hipSuccess)
        e = hipStreamSynchronize(ctx->stream);
    bs.release(ctx);
    return hip_check(e, "debug scan");
}

}  // extern "C"
