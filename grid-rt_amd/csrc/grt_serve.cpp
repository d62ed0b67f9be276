// grt_serve.cpp -- device-resident render-request loop with continuous
// batching, and the plain one-request-at-a-time path it is measured against.
//
// A render request is a scene (already resident with its grid), a pinhole
// camera and a depth d >= 1 (closest-hit generations: primary + d-1 mirror
// bounces, each followed by a shadow ray toward the light).  It runs as d+1
// steps; step j depends only on step j-1's closest hits:
//   step 0      primary rays, closest hit
//   step j>=1   shadow rays from generation j-1 (any hit), and, if j < d, the
//               bounce rays of generation j (closest hit)
// The reference rendered one frame at a time, one launch per pass
// (GpuRenderer.cpp:74-319; cudaRayTrace, RayTracing.cu:227-253).
//
// Continuous batching: each round gathers the current step of every admitted
// request into one batched persistent launch (grt_batch); requests that arrive
// meanwhile are admitted at the next round boundary as device memory (the
// budget) and request slots free up, so their work overlaps on the device with
// requests already in flight.  Buffers come from one slab of `memory_budget`
// bytes allocated up front, sub-allocated first-fit, so no hipMalloc/hipFree
// runs inside the loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

#include "grt/grt.h"
#include "grt_internal.h"
#include "grt_launch.h"

using namespace grt;

namespace {

using Clock = std::chrono::steady_clock;

// First-fit sub-allocator over one device slab (offsets in bytes, 256-aligned).
class Slab {
public:
    bool init(grt_context_t* ctx, size_t bytes)
    {
        if (mem_.reserve(ctx, bytes, false) != GRT_OK)
            return false;
        ctx_ = ctx;
        free_.clear();
        free_[0] = bytes;
        return true;
    }
    void release() { mem_.release(ctx_); }
    bool alloc(size_t bytes, size_t* off)
    {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t roff = it->first + bytes;
                free_.erase(it);
                if (rest)
                    free_[roff] = rest;
                return true;
            }
        }
        return false;
    }
    void free(size_t off, size_t bytes)
    {
        bytes = (bytes + 255) & ~size_t(255);
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    char* base() const { return static_cast<char*>(mem_.ptr); }

private:
    grt_context_t* ctx_ = nullptr;
    DevMem mem_;
    std::map<size_t, size_t> free_;
};

struct RaySet {
    float* f[8];
    grt_rays api() const { return grt_rays{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]}; }
    RaySoA in() const { return RaySoA{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]}; }
    RaySoAOut out() const { return RaySoAOut{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]}; }
};
struct HitSet {
    int32_t* id;
    float *t, *u, *v;
    grt_hits api() const { return grt_hits{id, t, u, v}; }
    HitSoA out() const { return HitSoA{id, t, u, v}; }
    HitSoAIn in(uint32_t n_tris) const { return HitSoAIn{id, t, u, v, n_tris}; }
};

struct Job {
    uint32_t idx = 0;
    grt_scene_t* sc = nullptr;
    DevCamera cam{};
    uint32_t n = 0, width = 0, depth = 1, step = 0;
    size_t off = 0, bytes = 0;
    RaySet ray[2], sh;
    HitSet hit[2];
    uint8_t* occ = nullptr;
};

// Buffers of a request with n pixels, as carved below: two ray sets and two
// hit sets (closest-hit generations ping-pong: 2 x (8 + 4) arrays), one shadow
// ray set (8 arrays) -- 32 arrays of n 4-byte values -- and n occlusion bytes,
// each array 256-byte aligned.
size_t job_bytes(uint32_t n)
{
    const size_t a = ((size_t)n * 4 + 255) & ~size_t(255);
    const size_t b = ((size_t)n + 255) & ~size_t(255);
    return 32 * a + b;
}

void carve(Job& j, char* base)
{
    const size_t a = ((size_t)j.n * 4 + 255) & ~size_t(255);
    char* p = base + j.off;
    for (int s = 0; s < 2; ++s) {
        for (int k = 0; k < 8; ++k, p += a)
            j.ray[s].f[k] = reinterpret_cast<float*>(p);
        j.hit[s].id = reinterpret_cast<int32_t*>(p);
        p += a;
        j.hit[s].t = reinterpret_cast<float*>(p);
        p += a;
        j.hit[s].u = reinterpret_cast<float*>(p);
        p += a;
        j.hit[s].v = reinterpret_cast<float*>(p);
        p += a;
    }
    for (int k = 0; k < 8; ++k, p += a)
        j.sh.f[k] = reinterpret_cast<float*>(p);
    j.occ = reinterpret_cast<uint8_t*>(p);
}

struct Item {
    grt_scene_t* sc;
    RaySet rays;
    HitSet hits;
    uint8_t* occ;
    uint32_t n, width;
    bool any;
};

// Enqueue the ray generation of a job's current step and append its traces.
int issue_step(Job& j, const grt_serve_config& cfg, hipStream_t s, unsigned long long* count,
               std::vector<Item>& items)
{
    hipError_t e = hipSuccess;
    if (j.step == 0) {
        e = launch_gen_primary(j.cam, j.ray[0].out(), s);
        items.push_back(Item{j.sc, j.ray[0], j.hit[0], nullptr, j.n, j.width, false});
    } else {
        const int c = (j.step - 1) & 1;
        const HitSoAIn hin = j.hit[c].in(j.sc->n_tris);
        const float* nrm = j.sc->normals.as<float>();
        e = launch_gen_shadow(j.n, j.ray[c].in(), hin, nrm, cfg.light[0], cfg.light[1],
                              cfg.light[2], cfg.offset, j.sh.out(), s);
        if (e == hipSuccess && count)
            e = launch_count_active(j.sh.f[6], j.sh.f[7], j.n, count, s);
        items.push_back(Item{j.sc, j.sh, HitSet{}, j.occ, j.n, j.width, true});
        if (e == hipSuccess && j.step < j.depth) {
            e = launch_gen_reflection(j.n, j.ray[c].in(), hin, nrm, cfg.offset,
                                      j.ray[1 - c].out(), s);
            if (e == hipSuccess && count)
                e = launch_count_active(j.ray[1 - c].f[6], j.ray[1 - c].f[7], j.n, count, s);
            items.push_back(Item{j.sc, j.ray[1 - c], j.hit[1 - c], nullptr, j.n, j.width, false});
        }
    }
    return hip_check(e, "serve: ray generation");
}

double seconds_since(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

}  // namespace

extern "C" {

int grt_serve_run(grt_context ctx, const grt_request* reqs, uint32_t n_reqs,
                  const grt_serve_config* cfg, grt_request_result* out, double* wall_s)
{
    if (!ctx || (!reqs && n_reqs) || !cfg || (!out && n_reqs))
        return set_error(GRT_INVALID_PARAM, "grt_serve_run: NULL argument");
    for (uint32_t i = 0; i < n_reqs; ++i) {
        const grt_request& r = reqs[i];
        if (!r.scene || r.scene->ctx != ctx || !r.scene->grid_valid)
            return set_error(GRT_NOT_READY, "grt_serve_run: request %u scene has no grid", i);
        if (r.depth < 1 || r.depth > 16 || r.camera.width == 0 || r.camera.height == 0 ||
            (uint64_t)r.camera.width * r.camera.height > (1u << 26))
            return set_error(GRT_INVALID_PARAM, "grt_serve_run: request %u out of range", i);
        if (i && r.arrival_s < reqs[i - 1].arrival_s)
            return set_error(GRT_INVALID_PARAM, "grt_serve_run: arrivals must be sorted");
        if (job_bytes(r.camera.width * r.camera.height) > cfg->memory_budget)
            return set_error(GRT_CAPACITY, "grt_serve_run: request %u exceeds the memory budget",
                             i);
    }
    hipStream_t s = ctx->stream;
    Slab slab;
    if (!slab.init(ctx, (size_t)cfg->memory_budget))
        return GRT_NO_MEMORY;
    DevMem counts;
    int rc = counts.reserve(ctx, sizeof(unsigned long long) * (n_reqs + 1), true);
    grt_batch batch = nullptr;
    if (rc == GRT_OK && cfg->batched)
        rc = grt_batch_create(ctx, 2 * std::max<uint32_t>(1, cfg->max_active), &batch);
    hipEvent_t done = nullptr;
    if (rc == GRT_OK && hipEventCreateWithFlags(&done, hipEventDisableTiming) != hipSuccess)
        rc = set_error(GRT_DEVICE_ERROR, "grt_serve_run: event creation failed");
    if (rc == GRT_OK)
        rc = hip_check(hipStreamSynchronize(s), "serve: start");
    const uint32_t max_active = std::max<uint32_t>(1, cfg->batched ? cfg->max_active : 1);

    std::deque<uint32_t> waiting;
    std::vector<Job> active;
    std::vector<Item> items;
    std::vector<grt_batch_item> bitems;
    uint32_t next = 0, finished = 0;
    unsigned long long* cnt = counts.as<unsigned long long>();
    const Clock::time_point t0 = Clock::now();

    while (rc == GRT_OK && finished < n_reqs) {
        const double now = seconds_since(t0);
        while (next < n_reqs && reqs[next].arrival_s <= now)
            waiting.push_back(next++);
        // Admission: FIFO, while a request slot and budget memory are free.
        while (!waiting.empty() && active.size() < max_active) {
            const uint32_t i = waiting.front();
            const grt_request& r = reqs[i];
            Job j;
            j.idx = i;
            j.sc = r.scene;
            std::memcpy(j.cam.pos, r.camera.pos, sizeof(j.cam.pos));
            std::memcpy(j.cam.base, r.camera.base, sizeof(j.cam.base));
            std::memcpy(j.cam.nu, r.camera.nu, sizeof(j.cam.nu));
            std::memcpy(j.cam.nv, r.camera.nv, sizeof(j.cam.nv));
            j.cam.width = r.camera.width;
            j.cam.height = r.camera.height;
            j.width = r.camera.width;
            j.n = r.camera.width * r.camera.height;
            j.depth = r.depth;
            j.bytes = job_bytes(j.n);
            if (!slab.alloc(j.bytes, &j.off))
                break;
            carve(j, slab.base());
            out[i].start_s = seconds_since(t0);
            out[i].rays = 0;
            waiting.pop_front();
            active.push_back(j);
        }
        if (active.empty()) {
            if (next < n_reqs) {
                const double wait = reqs[next].arrival_s - seconds_since(t0);
                if (wait > 2e-4)
                    std::this_thread::sleep_for(std::chrono::duration<double>(wait - 1e-4));
            }
            continue;
        }
        // One round: the current step of every active request.
        items.clear();
        for (Job& j : active) {
            if ((rc = issue_step(j, *cfg, s, cnt + j.idx, items)) != GRT_OK)
                break;
        }
        if (rc != GRT_OK)
            break;
        if (cfg->batched) {
            bitems.resize(items.size());
            for (size_t k = 0; k < items.size(); ++k) {
                const Item& it = items[k];
                grt_batch_item& b = bitems[k];
                std::memset(&b, 0, sizeof(b));
                b.scene = it.sc;
                b.rays = it.rays.api();
                b.hits = it.hits.api();
                b.occluded = it.occ;
                b.n = it.n;
                b.image_width = it.width;
                b.any = it.any ? 1u : 0u;
            }
            if ((rc = grt_batch_set(batch, bitems.data(), (uint32_t)bitems.size())) != GRT_OK)
                break;
            if ((rc = grt_batch_trace(batch)) != GRT_OK)
                break;
        } else {
            // Plain path: one launch per ray batch, with the context's
            // stand-alone schedule.
            for (const Item& it : items) {
                TraceArgs a;
                a.fr = it.sc->frame.as<DevFrame>();
                a.offsets = it.sc->offsets.as<uint32_t>();
                a.refs = it.sc->refs.as<uint32_t>();
                a.tri = it.sc->tri.as<float4>();
                a.trir = it.sc->trir.as<float4>();
                a.n = it.n;
                a.image_width = it.width;
                a.counters = nullptr;
                a.coop = ctx->coop;
                a.pool = ctx->trace_mode == GRT_TRACE_POOLED ? 1u : 0u;
                hipError_t e;
                if (ctx->trace_mode == GRT_TRACE_PERSISTENT)
                    e = launch_trace_persistent(a, it.rays.in(), it.hits.out(), it.occ, it.any,
                                                it.sc->small_counters.as<uint32_t>() + 2,
                                                it.any ? ctx->block_any : ctx->block_closest,
                                                ctx->refill, s);
                else if (ctx->trace_mode == GRT_TRACE_SPLIT)
                    e = launch_trace_split(a, it.rays.in(), it.hits.out(), it.occ, it.any,
                                           ctx->split_lanes,
                                           it.any ? ctx->block_any : ctx->block_closest, s);
                else if (it.any)
                    e = launch_trace_any(a, it.rays.in(), it.occ, ctx->block_any, s);
                else
                    e = launch_trace_closest(a, it.rays.in(), it.hits.out(), ctx->block_closest,
                                             s);
                if ((rc = hip_check(e, "serve: trace")) != GRT_OK)
                    break;
            }
            if (rc != GRT_OK)
                break;
        }
        if ((rc = hip_check(hipEventRecord(done, s), "serve: event")) != GRT_OK)
            break;
        if ((rc = hip_check(hipEventSynchronize(done), "serve: round")) != GRT_OK)
            break;
        const double t_done = seconds_since(t0);
        // Advance; retire finished requests and free their buffers.
        size_t w = 0;
        for (size_t k = 0; k < active.size(); ++k) {
            Job& j = active[k];
            const grt_request& r = reqs[j.idx];
            hipError_t e = hipSuccess;
            if (j.step == 0) {
                out[j.idx].rays += j.n;
                // Primary hits are complete; the next rounds only read them
                // until step 2 reuses the buffer, and copies are stream-ordered.
                if (r.out_hits.id && r.out_hits.t && r.out_hits.u && r.out_hits.v) {
                    const size_t b = (size_t)j.n * 4;
                    e = hipMemcpyAsync(r.out_hits.id, j.hit[0].id, b, hipMemcpyDeviceToDevice, s);
                    if (e == hipSuccess)
                        e = hipMemcpyAsync(r.out_hits.t, j.hit[0].t, b, hipMemcpyDeviceToDevice, s);
                    if (e == hipSuccess)
                        e = hipMemcpyAsync(r.out_hits.u, j.hit[0].u, b, hipMemcpyDeviceToDevice, s);
                    if (e == hipSuccess)
                        e = hipMemcpyAsync(r.out_hits.v, j.hit[0].v, b, hipMemcpyDeviceToDevice, s);
                }
            } else if (j.step == 1 && r.out_occluded) {
                e = hipMemcpyAsync(r.out_occluded, j.occ, j.n, hipMemcpyDeviceToDevice, s);
            }
            if (e != hipSuccess) {
                rc = hip_check(e, "serve: result copy");
                break;
            }
            ++j.step;
            if (j.step > j.depth) {
                out[j.idx].finish_s = t_done;
                slab.free(j.off, j.bytes);
                ++finished;
            } else {
                active[w++] = j;
            }
        }
        active.resize(w);
    }
    const double wall = seconds_since(t0);
    if (rc == GRT_OK && n_reqs) {
        std::vector<unsigned long long> host(n_reqs);
        hipError_t e = hipMemcpy(host.data(), cnt, sizeof(unsigned long long) * n_reqs,
                                 hipMemcpyDeviceToHost);
        if (e == hipSuccess)
            for (uint32_t i = 0; i < n_reqs; ++i)
                out[i].rays += host[i];
        else
            rc = hip_check(e, "serve: ray counts");
    }
    if (wall_s)
        *wall_s = wall;
    (void)hipStreamSynchronize(s);
    if (batch)
        grt_batch_destroy(batch);
    if (done)
        (void)hipEventDestroy(done);
    counts.release(ctx);
    slab.release();
    return rc;
}

}  // extern "C"
