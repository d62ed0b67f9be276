// grt_batch.cpp -- batched multi-scene tracing (device work queue) and hipGraph
// capture of launch sequences.  The reference traced one scene per frame with
// one kernel launch (cudaRayTrace, RayTracing.cu:227-253); a batch puts many
// scenes and ray batches of uneven size into one persistent launch.
#include <hip/hip_runtime.h>

#include <cstring>
#include <new>
#include <vector>

#include "grt/grt.h"
#include "grt_internal.h"
#include "grt_launch.h"

using namespace grt;

struct grt_batch_t {
    grt_context_t* ctx = nullptr;
    uint32_t max_items = 0;
    uint32_t n_items = 0;
    uint32_t total_slots = 0;
    DevMem tasks, map, counter;
    BatchTask* staging = nullptr;   // pinned
    hipEvent_t staged = nullptr;    // staging may be reused once this completes
    bool staged_pending = false;
};

struct grt_graph_t {
    grt_context_t* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

int grt::batch_reserve_slots(grt_batch b, uint64_t slots)
{
    if (!b)
        return set_error(GRT_INVALID_PARAM, "batch_reserve_slots: NULL batch");
    if (slots > 0xffffffc0ull)
        slots = 0xffffffc0ull;
    return b->map.reserve(b->ctx, (slots / 64 + 1) * 4, false);
}

extern "C" {

int grt_batch_create(grt_context ctx, uint32_t max_items, grt_batch* out)
{
    if (!ctx || !out || max_items == 0 || max_items > (1u << 20))
        return set_error(GRT_INVALID_PARAM, "grt_batch_create: bad argument");
    grt_batch_t* b = new (std::nothrow) grt_batch_t();
    if (!b)
        return set_error(GRT_NO_MEMORY, "grt_batch_create: out of host memory");
    b->ctx = ctx;
    b->max_items = max_items;
    int rc = b->tasks.reserve(ctx, sizeof(BatchTask) * (size_t)max_items, false);
    if (rc == GRT_OK)
        rc = b->counter.reserve(ctx, 64, true);
    if (rc == GRT_OK && hipHostMalloc((void**)&b->staging, sizeof(BatchTask) * max_items,
                                      hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        rc = set_error(GRT_NO_MEMORY, "grt_batch_create: pinned staging failed");
    }
    if (rc == GRT_OK && hipEventCreateWithFlags(&b->staged, hipEventDisableTiming) != hipSuccess)
        rc = set_error(GRT_DEVICE_ERROR, "grt_batch_create: event creation failed");
    if (rc != GRT_OK) {
        grt_batch_destroy(b);
        return rc;
    }
    *out = b;
    return GRT_OK;
}

int grt_batch_destroy(grt_batch b)
{
    if (!b)
        return set_error(GRT_INVALID_PARAM, "grt_batch_destroy: NULL batch");
    (void)hipStreamSynchronize(b->ctx->stream);
    b->tasks.release(b->ctx);
    b->map.release(b->ctx);
    b->counter.release(b->ctx);
    if (b->staging)
        (void)hipHostFree(b->staging);
    if (b->staged)
        (void)hipEventDestroy(b->staged);
    delete b;
    return GRT_OK;
}

int grt_batch_set(grt_batch b, const grt_batch_item* items, uint32_t n_items)
{
    if (!b || (!items && n_items))
        return set_error(GRT_INVALID_PARAM, "grt_batch_set: NULL argument");
    if (n_items > b->max_items)
        return set_error(GRT_CAPACITY, "grt_batch_set: %u items > capacity %u", n_items,
                         b->max_items);
    // The previous set's upload must have left the staging buffer.
    if (b->staged_pending) {
        if (hipEventSynchronize(b->staged) != hipSuccess)
            return set_error(GRT_DEVICE_ERROR, "grt_batch_set: waiting for the staging copy");
        b->staged_pending = false;
    }
    uint64_t slots = 0;
    for (uint32_t i = 0; i < n_items; ++i) {
        const grt_batch_item& it = items[i];
        grt_scene_t* sc = it.scene;
        if (!sc || sc->ctx != b->ctx)
            return set_error(GRT_INVALID_PARAM, "grt_batch_set: item %u has no scene of this "
                             "context", i);
        if (!sc->grid_valid || sc->n_tris == 0)
            return set_error(GRT_NOT_READY, "grt_batch_set: item %u scene has no grid", i);
        const grt_rays& r = it.rays;
        if (!r.ox || !r.oy || !r.oz || !r.dx || !r.dy || !r.dz || !r.tmin || !r.tmax)
            return set_error(GRT_INVALID_PARAM, "grt_batch_set: item %u ray buffer is NULL", i);
        if (it.any ? !it.occluded : (!it.hits.id || !it.hits.t || !it.hits.u || !it.hits.v))
            return set_error(GRT_INVALID_PARAM, "grt_batch_set: item %u output is NULL", i);
        uint64_t s = it.n;
        if (it.image_width) {
            const uint64_t h = (it.n + (uint64_t)it.image_width - 1) / it.image_width;
            s = (uint64_t)((it.image_width + 7u) >> 3) * ((h + 7u) >> 3) * 64u;
        }
        s = (s + 63) / 64 * 64;
        BatchTask& t = b->staging[i];
        std::memset(&t, 0, sizeof(t));
        t.fr = sc->frame.as<DevFrame>();
        t.offsets = sc->offsets.as<uint32_t>();
        t.trir = sc->trir.as<float4>();
        t.rays = RaySoA{r.ox, r.oy, r.oz, r.dx, r.dy, r.dz, r.tmin, r.tmax};
        t.hits = HitSoA{it.hits.id, it.hits.t, it.hits.u, it.hits.v};
        t.occluded = it.occluded;
        t.n = it.n;
        t.image_width = it.image_width;
        t.slot_begin = (uint32_t)slots;
        t.slots = (uint32_t)s;
        t.any = it.any ? 1u : 0u;
        slots += s;
        if (slots > 0xffffffc0ull)
            return set_error(GRT_CAPACITY, "grt_batch_set: more than 2^32 ray slots");
    }
    int rc = b->map.reserve(b->ctx, (slots / 64 + 1) * 4, false);
    if (rc != GRT_OK)
        return rc;
    hipStream_t s = b->ctx->stream;
    hipError_t e = hipSuccess;
    if (n_items) {
        e = hipMemcpyAsync(b->tasks.ptr, b->staging, sizeof(BatchTask) * n_items,
                           hipMemcpyHostToDevice, s);
        if (e == hipSuccess)
            e = hipEventRecord(b->staged, s);
        if (e == hipSuccess)
            b->staged_pending = true;
        if (e == hipSuccess)
            e = launch_task_map(b->tasks.as<BatchTask>(), n_items, b->map.as<uint32_t>(), s);
    }
    if (e != hipSuccess)
        return hip_check(e, "grt_batch_set");
    b->n_items = n_items;
    b->total_slots = (uint32_t)slots;
    return GRT_OK;
}

int grt_batch_trace(grt_batch b)
{
    if (!b)
        return set_error(GRT_INVALID_PARAM, "grt_batch_trace: NULL batch");
    grt_context_t* ctx = b->ctx;
    return hip_check(launch_trace_batch(b->tasks.as<BatchTask>(), b->map.as<uint32_t>(),
                                        b->total_slots, b->counter.as<uint32_t>(),
                                        ctx->trace_mode, ctx->block_closest, ctx->refill,
                                        ctx->split_lanes, ctx->coop, ctx->stream),
                     "grt_batch_trace");
}

int grt_batch_slots(grt_batch b, uint32_t* slots)
{
    if (!b || !slots)
        return set_error(GRT_INVALID_PARAM, "grt_batch_slots: NULL argument");
    *slots = b->total_slots;
    return GRT_OK;
}

int grt_graph_begin(grt_context ctx)
{
    if (!ctx)
        return set_error(GRT_INVALID_PARAM, "grt_graph_begin: NULL context");
    if (ctx->stream == nullptr)
        return set_error(GRT_INVALID_PARAM,
                         "grt_graph_begin: the null stream cannot be captured; set a stream");
    return hip_check(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                     "hipStreamBeginCapture");
}

int grt_graph_end(grt_context ctx, grt_graph* out)
{
    if (!ctx || !out)
        return set_error(GRT_INVALID_PARAM, "grt_graph_end: NULL argument");
    grt_graph_t* g = new (std::nothrow) grt_graph_t();
    if (!g)
        return set_error(GRT_NO_MEMORY, "grt_graph_end: out of host memory");
    g->ctx = ctx;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g->graph);
    if (e == hipSuccess)
        e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g->graph)
            (void)hipGraphDestroy(g->graph);
        delete g;
        return hip_check(e, "graph capture");
    }
    *out = g;
    return GRT_OK;
}

int grt_graph_launch(grt_graph g)
{
    if (!g)
        return set_error(GRT_INVALID_PARAM, "grt_graph_launch: NULL graph");
    return hip_check(hipGraphLaunch(g->exec, g->ctx->stream), "hipGraphLaunch");
}

int grt_graph_destroy(grt_graph g)
{
    if (!g)
        return set_error(GRT_INVALID_PARAM, "grt_graph_destroy: NULL graph");
    if (g->exec)
        (void)hipGraphExecDestroy(g->exec);
    if (g->graph)
        (void)hipGraphDestroy(g->graph);
    delete g;
    return GRT_OK;
}

int grt_stream_create(void** out)
{
    if (!out)
        return set_error(GRT_INVALID_PARAM, "grt_stream_create: NULL argument");
    hipStream_t s = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess)
        return hip_check(e, "hipStreamCreate");
    *out = (void*)s;
    return GRT_OK;
}

int grt_stream_destroy(void* stream)
{
    if (!stream)
        return set_error(GRT_INVALID_PARAM, "grt_stream_destroy: NULL stream");
    return hip_check(hipStreamDestroy((hipStream_t)stream), "hipStreamDestroy");
}

}  // extern "C"
