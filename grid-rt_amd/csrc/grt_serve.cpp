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
// Continuous batching: each round gathers the next step of every admitted
// request into one batched persistent launch (grt_batch).  Rounds are
// pipelined on the stream -- up to two in flight, step j+1 of a request queued
// behind step j by stream order alone -- so the host never waits for a round
// before issuing the next, and only retires requests (their last round's event)
// and admits new ones, FIFO, as device memory (the budget) and request slots
// free up; their work overlaps on the device with requests already in flight.
// Buffers come from one slab of `memory_budget` bytes allocated up front,
// sub-allocated first-fit, so no hipMalloc/hipFree runs inside the loop.
//
// Two drivers share the engine: grt_serve_run replays a request list with
// arrival times (the measurement harness), and grt_server runs the loop on a
// thread of its own behind a thread-safe submit / wait interface.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
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

int check_request(grt_context_t* ctx, const grt_request& r, const grt_serve_config& cfg,
                  const char* who, uint32_t i)
{
    if (!r.scene || r.scene->ctx != ctx || !r.scene->grid_valid)
        return set_error(GRT_NOT_READY, "%s: request %u scene has no grid", who, i);
    if (r.depth < 1 || r.depth > 16 || r.camera.width == 0 || r.camera.height == 0 ||
        (uint64_t)r.camera.width * r.camera.height > (1u << 26))
        return set_error(GRT_INVALID_PARAM, "%s: request %u out of range", who, i);
    if (job_bytes(r.camera.width * r.camera.height) > cfg.memory_budget)
        return set_error(GRT_CAPACITY, "%s: request %u exceeds the memory budget", who, i);
    return GRT_OK;
}

// The device loop shared by both drivers.  Not thread-safe: one thread drives
// it.  Keys identify requests to the caller.
class Engine {
public:
    struct Done {
        uint64_t key;
        grt_request_result res;
    };

    int init(grt_context_t* ctx, const grt_serve_config& cfg)
    {
        ctx_ = ctx;
        cfg_ = cfg;
        batched_ = cfg.batched != 0;
        max_active_ = std::max<uint32_t>(1, batched_ ? cfg.max_active : 1);
        depth_ = batched_ ? 2u : 1u;
        if (!slab_.init(ctx, (size_t)cfg.memory_budget))
            return GRT_NO_MEMORY;
        int rc = counts_.reserve(ctx, sizeof(unsigned long long) * max_active_, true);
        if (rc == GRT_OK &&
            hipHostMalloc((void**)&pinned_, sizeof(unsigned long long) * max_active_) !=
                hipSuccess) {
            pinned_ = nullptr;
            rc = set_error(GRT_NO_MEMORY, "serve: pinned result counters");
        }
        if (rc == GRT_OK && batched_) {
            rc = grt_batch_create(ctx, 2 * max_active_, &batch_);
            // the slot map of the largest round the budget admits: two ray
            // batches per request of at most budget / 128 pixels in all, each
            // padded to whole waves
            if (rc == GRT_OK)
                rc = batch_reserve_slots(batch_, 2 * (cfg.memory_budget / 128) +
                                                      (uint64_t)2 * max_active_ * 64);
        }
        for (uint32_t k = 0; rc == GRT_OK && k < depth_ + 1; ++k) {
            hipEvent_t e = nullptr;
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                rc = set_error(GRT_DEVICE_ERROR, "serve: event creation failed");
            else
                events_.push_back(e);
        }
        for (uint32_t k = 0; k < max_active_; ++k)
            free_slots_.push_back(max_active_ - 1 - k);
        if (rc == GRT_OK)
            rc = hip_check(hipStreamSynchronize(ctx->stream), "serve: start");
        return rc;
    }

    // Times are reported in seconds after t0 (set once the buffers exist).
    void start(Clock::time_point t0) { t0_ = t0; }

    void release()
    {
        if (ctx_)
            (void)hipStreamSynchronize(ctx_->stream);
        if (batch_)
            grt_batch_destroy(batch_);
        batch_ = nullptr;
        for (hipEvent_t e : events_)
            (void)hipEventDestroy(e);
        events_.clear();
        for (Round& r : rounds_)
            (void)hipEventDestroy(r.ev);
        rounds_.clear();
        if (pinned_)
            (void)hipHostFree(pinned_);
        pinned_ = nullptr;
        if (ctx_)
            counts_.release(ctx_);
        slab_.release();
    }

    bool idle() const { return active_.empty() && rounds_.empty(); }
    bool pipeline_full() const { return rounds_.size() >= depth_; }
    bool in_flight() const { return !rounds_.empty(); }

    // Admit a request now if a slot and budget memory are free.
    int admit(const grt_request& r, uint64_t key, bool* admitted)
    {
        *admitted = false;
        if (free_slots_.empty())
            return GRT_OK;
        Live j;
        j.key = key;
        j.req = r;
        j.idx = free_slots_.back();
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
        if (!slab_.alloc(j.bytes, &j.off))
            return GRT_OK;
        carve(j, slab_.base());
        free_slots_.pop_back();
        const hipError_t e = hipMemsetAsync(counts_.as<unsigned long long>() + j.idx, 0,
                                            sizeof(unsigned long long), ctx_->stream);
        if (e != hipSuccess)
            return hip_check(e, "serve: counter reset");
        j.start_s = seconds_since(t0_);
        active_.push_back(j);
        *admitted = true;
        return GRT_OK;
    }

    // One round: the next step of every admitted request that has one left,
    // queued behind the rounds in flight.  Does nothing when the pipeline is
    // full or no request has a step left to issue.
    int issue()
    {
        if (pipeline_full())
            return GRT_OK;
        hipStream_t s = ctx_->stream;
        items_.clear();
        std::vector<uint64_t> fin;
        std::vector<Live*> stepped;
        int rc = GRT_OK;
        for (Live& j : active_) {
            if (j.issued > j.depth)
                continue;
            j.step = j.issued;
            if ((rc = issue_step(j, cfg_, s, counts_.as<unsigned long long>() + j.idx, items_)) !=
                GRT_OK)
                return rc;
            stepped.push_back(&j);
        }
        if (items_.empty())
            return GRT_OK;
        if ((rc = trace(items_)) != GRT_OK)
            return rc;
        for (Live* jp : stepped) {
            Live& j = *jp;
            const grt_request& r = j.req;
            hipError_t e = hipSuccess;
            if (j.issued == 0 && r.out_hits.id && r.out_hits.t && r.out_hits.u && r.out_hits.v) {
                // primary hits: copied before step 2 reuses the buffer (stream order)
                const size_t b = (size_t)j.n * 4;
                e = hipMemcpyAsync(r.out_hits.id, j.hit[0].id, b, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(r.out_hits.t, j.hit[0].t, b, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(r.out_hits.u, j.hit[0].u, b, hipMemcpyDeviceToDevice, s);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(r.out_hits.v, j.hit[0].v, b, hipMemcpyDeviceToDevice, s);
            } else if (j.issued == 1 && r.out_occluded) {
                e = hipMemcpyAsync(r.out_occluded, j.occ, j.n, hipMemcpyDeviceToDevice, s);
            }
            ++j.issued;
            if (e == hipSuccess && j.issued > j.depth) {
                // last step: its ray count comes back with the round
                e = hipMemcpyAsync(pinned_ + j.idx, counts_.as<unsigned long long>() + j.idx,
                                   sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
                fin.push_back(j.key);
            }
            if (e != hipSuccess)
                return hip_check(e, "serve: result copy");
        }
        hipEvent_t ev = events_.back();
        const hipError_t e = hipEventRecord(ev, s);
        if (e != hipSuccess)
            return hip_check(e, "serve: event");
        events_.pop_back();
        rounds_.push_back(Round{ev, std::move(fin)});
        return GRT_OK;
    }

    // Retire completed rounds, oldest first (waiting for the oldest when
    // `block`); finished requests go to `done` and free their slot and memory.
    int retire(bool block, std::vector<Done>& done)
    {
        while (!rounds_.empty()) {
            Round& r = rounds_.front();
            hipError_t e = block ? hipEventSynchronize(r.ev) : hipEventQuery(r.ev);
            if (e == hipErrorNotReady)
                return GRT_OK;
            if (e != hipSuccess)
                return hip_check(e, "serve: round");
            block = false;
            const double t = seconds_since(t0_);
            for (uint64_t key : r.fin) {
                auto it = std::find_if(active_.begin(), active_.end(),
                                       [&](const Live& j) { return j.key == key; });
                if (it == active_.end())
                    continue;
                grt_request_result res;
                res.start_s = it->start_s;
                res.finish_s = t;
                res.rays = (uint64_t)it->n + pinned_[it->idx];
                done.push_back(Done{key, res});
                slab_.free(it->off, it->bytes);
                free_slots_.push_back(it->idx);
                active_.erase(it);
            }
            events_.push_back(r.ev);
            rounds_.pop_front();
        }
        return GRT_OK;
    }

private:
    struct Live : Job {
        uint64_t key = 0;
        grt_request req{};
        uint32_t issued = 0;   // steps enqueued so far
        double start_s = 0.0;
    };
    struct Round {
        hipEvent_t ev;
        std::vector<uint64_t> fin;   // requests whose last step this round carries
    };

    int trace(const std::vector<Item>& items)
    {
        hipStream_t s = ctx_->stream;
        if (batched_) {
            bitems_.resize(items.size());
            for (size_t k = 0; k < items.size(); ++k) {
                const Item& it = items[k];
                grt_batch_item& b = bitems_[k];
                std::memset(&b, 0, sizeof(b));
                b.scene = it.sc;
                b.rays = it.rays.api();
                b.hits = it.hits.api();
                b.occluded = it.occ;
                b.n = it.n;
                b.image_width = it.width;
                b.any = it.any ? 1u : 0u;
            }
            int rc = grt_batch_set(batch_, bitems_.data(), (uint32_t)bitems_.size());
            return rc == GRT_OK ? grt_batch_trace(batch_) : rc;
        }
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
            a.coop = ctx_->coop;
            a.pool = ctx_->trace_mode == GRT_TRACE_POOLED ? 1u : 0u;
            hipError_t e;
            if (ctx_->trace_mode == GRT_TRACE_PERSISTENT)
                e = launch_trace_persistent(a, it.rays.in(), it.hits.out(), it.occ, it.any,
                                            it.sc->small_counters.as<uint32_t>() + 2,
                                            it.any ? ctx_->block_any : ctx_->block_closest,
                                            ctx_->refill, s);
            else if (ctx_->trace_mode == GRT_TRACE_SPLIT)
                e = launch_trace_split(a, it.rays.in(), it.hits.out(), it.occ, it.any,
                                       ctx_->split_lanes,
                                       it.any ? ctx_->block_any : ctx_->block_closest, s);
            else if (it.any)
                e = launch_trace_any(a, it.rays.in(), it.occ, ctx_->block_any, s);
            else
                e = launch_trace_closest(a, it.rays.in(), it.hits.out(), ctx_->block_closest, s);
            if (e != hipSuccess)
                return hip_check(e, "serve: trace");
        }
        return GRT_OK;
    }

    grt_context_t* ctx_ = nullptr;
    grt_serve_config cfg_{};
    Clock::time_point t0_;
    bool batched_ = true;
    uint32_t max_active_ = 1, depth_ = 1;
    Slab slab_;
    DevMem counts_;
    unsigned long long* pinned_ = nullptr;
    grt_batch batch_ = nullptr;
    std::vector<hipEvent_t> events_;
    std::deque<Round> rounds_;
    std::vector<uint32_t> free_slots_;
    std::deque<Live> active_;
    std::vector<Item> items_;
    std::vector<grt_batch_item> bitems_;
};

// Sleep about `s` seconds (short waits spin on yield).
void nap(double s)
{
    if (s > 2e-4)
        std::this_thread::sleep_for(std::chrono::duration<double>(s - 1e-4));
    else
        std::this_thread::yield();
}

}  // namespace

// Thread-safe front of the engine: a thread of its own drives the device loop;
// submit / wait / query may be called from any thread.
struct grt_server_t {
    grt_context_t* ctx = nullptr;
    grt_serve_config cfg{};
    Clock::time_point t0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::pair<uint64_t, grt_request>> submitted;
    struct Ticket {
        bool done = false;
        int rc = GRT_OK;
        std::string err;
        grt_request_result res{};
    };
    std::unordered_map<uint64_t, Ticket> tickets;
    uint64_t next_ticket = 1;
    bool stopping = false;
    int fatal = GRT_OK;
    std::string fatal_err;

    void run(Engine& eng)
    {
        std::deque<std::pair<uint64_t, grt_request>> waiting;
        std::vector<Engine::Done> done;
        for (;;) {
            {
                std::unique_lock<std::mutex> lk(mu);
                if (eng.idle() && waiting.empty() && submitted.empty()) {
                    if (stopping)
                        return;
                    cv_work.wait(lk, [&] { return stopping || !submitted.empty(); });
                }
                while (!submitted.empty()) {
                    waiting.push_back(submitted.front());
                    submitted.pop_front();
                }
            }
            int rc = GRT_OK;
            while (rc == GRT_OK && !waiting.empty()) {
                bool admitted = false;
                rc = eng.admit(waiting.front().second, waiting.front().first, &admitted);
                if (!admitted)
                    break;
                waiting.pop_front();
            }
            if (rc == GRT_OK)
                rc = eng.issue();
            done.clear();
            if (rc == GRT_OK)
                rc = eng.retire(eng.pipeline_full(), done);
            {
                std::lock_guard<std::mutex> lk(mu);
                for (const Engine::Done& d : done) {
                    Ticket& t = tickets[d.key];
                    t.done = true;
                    t.res = d.res;
                }
                if (rc != GRT_OK) {
                    // a device error ends the loop: every open request fails
                    fatal = rc;
                    fatal_err = grt_last_error();
                    for (auto& kv : tickets)
                        if (!kv.second.done) {
                            kv.second.done = true;
                            kv.second.rc = rc;
                            kv.second.err = fatal_err;
                        }
                    submitted.clear();
                }
            }
            if (!done.empty() || rc != GRT_OK)
                cv_done.notify_all();
            if (rc != GRT_OK)
                return;
            if (!eng.pipeline_full() && eng.in_flight())
                nap(2e-5);   // new submissions may still join the next round
        }
    }
};

extern "C" {

int grt_serve_run(grt_context ctx, const grt_request* reqs, uint32_t n_reqs,
                  const grt_serve_config* cfg, grt_request_result* out, double* wall_s)
{
    if (!ctx || (!reqs && n_reqs) || !cfg || (!out && n_reqs))
        return set_error(GRT_INVALID_PARAM, "grt_serve_run: NULL argument");
    for (uint32_t i = 0; i < n_reqs; ++i) {
        int rc = check_request(ctx, reqs[i], *cfg, "grt_serve_run", i);
        if (rc != GRT_OK)
            return rc;
        if (i && reqs[i].arrival_s < reqs[i - 1].arrival_s)
            return set_error(GRT_INVALID_PARAM, "grt_serve_run: arrivals must be sorted");
    }
    Engine eng;
    int rc = eng.init(ctx, *cfg);
    const Clock::time_point t0 = Clock::now();
    eng.start(t0);
    std::deque<uint32_t> waiting;
    std::vector<Engine::Done> done;
    uint32_t next = 0, finished = 0;
    while (rc == GRT_OK && finished < n_reqs) {
        const double now = seconds_since(t0);
        while (next < n_reqs && reqs[next].arrival_s <= now)
            waiting.push_back(next++);
        // Admission: FIFO, while a request slot and budget memory are free.
        while (rc == GRT_OK && !waiting.empty()) {
            bool admitted = false;
            rc = eng.admit(reqs[waiting.front()], waiting.front(), &admitted);
            if (!admitted)
                break;
            waiting.pop_front();
        }
        if (rc == GRT_OK)
            rc = eng.issue();
        done.clear();
        if (rc == GRT_OK)
            rc = eng.retire(eng.pipeline_full(), done);
        for (const Engine::Done& d : done) {
            out[d.key] = d.res;
            ++finished;
        }
        if (rc != GRT_OK || finished == n_reqs)
            break;
        if (eng.idle()) {
            // nothing on the device: admit what has arrived, else sleep until
            // the next arrival
            if (waiting.empty() && next < n_reqs)
                nap(reqs[next].arrival_s - seconds_since(t0));
        } else if (!eng.pipeline_full()) {
            nap(2e-5);
        }
    }
    if (wall_s)
        *wall_s = seconds_since(t0);
    eng.release();
    return rc;
}

int grt_server_create(grt_context ctx, const grt_serve_config* cfg, grt_server* out)
{
    if (!ctx || !cfg || !out)
        return set_error(GRT_INVALID_PARAM, "grt_server_create: NULL argument");
    *out = nullptr;
    grt_server_t* srv = new (std::nothrow) grt_server_t();
    if (!srv)
        return set_error(GRT_NO_MEMORY, "grt_server_create: out of host memory");
    srv->ctx = ctx;
    srv->cfg = *cfg;
    // The engine is created on the calling thread (allocation errors come back
    // here) and handed to the worker.
    int rc = GRT_OK;
    Engine* eng = new (std::nothrow) Engine();
    if (!eng)
        rc = set_error(GRT_NO_MEMORY, "grt_server_create: out of host memory");
    if (rc == GRT_OK)
        rc = eng->init(ctx, *cfg);
    srv->t0 = Clock::now();
    if (eng)
        eng->start(srv->t0);
    if (rc != GRT_OK) {
        if (eng) {
            eng->release();
            delete eng;
        }
        delete srv;
        return rc;
    }
    try {
        srv->worker = std::thread([srv, eng] {
            (void)hipSetDevice(srv->ctx->device);
            srv->run(*eng);
            eng->release();
            delete eng;
        });
    } catch (...) {
        eng->release();
        delete eng;
        delete srv;
        return set_error(GRT_NO_MEMORY, "grt_server_create: thread creation failed");
    }
    *out = srv;
    return GRT_OK;
}

int grt_server_submit(grt_server srv, const grt_request* req, uint64_t* ticket)
{
    if (!srv || !req || !ticket)
        return set_error(GRT_INVALID_PARAM, "grt_server_submit: NULL argument");
    int rc = check_request(srv->ctx, *req, srv->cfg, "grt_server_submit", 0);
    if (rc != GRT_OK)
        return rc;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->stopping)
            return set_error(GRT_INVALID_PARAM, "grt_server_submit: server is shutting down");
        if (srv->fatal != GRT_OK)
            return set_error(srv->fatal, "grt_server_submit: server stopped: %s",
                             srv->fatal_err.c_str());
        *ticket = srv->next_ticket++;
        srv->tickets[*ticket] = grt_server_t::Ticket{};
        srv->submitted.emplace_back(*ticket, *req);
    }
    srv->cv_work.notify_one();
    return GRT_OK;
}

static int take_result(grt_server_t* srv, std::unordered_map<uint64_t, grt_server_t::Ticket>::iterator it,
                       grt_request_result* out)
{
    const grt_server_t::Ticket t = it->second;
    srv->tickets.erase(it);
    if (t.rc != GRT_OK)
        return set_error(t.rc, "%s", t.err.c_str());
    if (out)
        *out = t.res;
    return GRT_OK;
}

int grt_server_query(grt_server srv, uint64_t ticket, grt_request_result* out)
{
    if (!srv)
        return set_error(GRT_INVALID_PARAM, "grt_server_query: NULL server");
    std::lock_guard<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end())
        return set_error(GRT_INVALID_PARAM, "grt_server_query: unknown ticket %llu",
                         (unsigned long long)ticket);
    if (!it->second.done)
        return set_error(GRT_NOT_READY, "grt_server_query: request %llu still running",
                         (unsigned long long)ticket);
    return take_result(srv, it, out);
}

int grt_server_wait(grt_server srv, uint64_t ticket, grt_request_result* out)
{
    if (!srv)
        return set_error(GRT_INVALID_PARAM, "grt_server_wait: NULL server");
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->tickets.find(ticket);
    if (it == srv->tickets.end())
        return set_error(GRT_INVALID_PARAM, "grt_server_wait: unknown ticket %llu",
                         (unsigned long long)ticket);
    srv->cv_done.wait(lk, [&] { return it->second.done; });
    return take_result(srv, it, out);
}

int grt_server_destroy(grt_server srv)
{
    if (!srv)
        return GRT_OK;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stopping = true;
    }
    srv->cv_work.notify_all();
    if (srv->worker.joinable())
        srv->worker.join();
    const int rc = srv->fatal;
    const std::string err = srv->fatal_err;
    delete srv;
    return rc == GRT_OK ? GRT_OK : set_error(rc, "grt_server_destroy: %s", err.c_str());
}

}  // extern "C"
