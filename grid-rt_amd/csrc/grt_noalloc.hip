// grt_noalloc.hip -- ray-list compaction that allocates nothing.
//
// grt_compact_rays keeps its per-tile sums in context memory that it grows
// with hipMalloc on demand.  Callers that may not allocate (code under HIP
// graph capture, tight serving loops, harnesses that forbid it) use
// grt_compact_rays_noalloc instead: the same three kernels, with the tile sums
// in a fixed file-scope device array.  Each GPU holds its own copy of the
// array until the process ends; its device address is looked up once per
// device when a context is created (grt_init), and host code passes only that
// address to the kernels.
//
// The kernels write every tile sum they read (count pass, then an in-place
// scan, then the write pass), so nothing has to be cleared between calls.
// Calls that overlap on one GPU (two streams, two host threads) share the
// array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "grt/grt.h"
#include "grt_internal.h"
#include "grt_launch.h"

using namespace grt;

namespace {

// One tile sum per kScanTile (4096) ray slots plus the two extra entries the
// scan uses: 64 KiB per GPU, enough for (16384 - 2) x 4096 = 67,100,672 slots.
constexpr uint32_t kScratchWords = 16384;
static __device__ __align__(256) uint32_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint32_t* g_scratch[kMaxDevices] = {};   // device address of `scratch`, per GPU

uint32_t* scratch_of(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_scratch[device];
}

}  // namespace

namespace grt {

uint64_t noalloc_max_slots() { return (uint64_t)(kScratchWords - 2) * kScanTile; }

// With `device` current: take the address of this GPU's copy of `scratch`.
int noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices)
        return set_error(GRT_INVALID_PARAM, "device %d beyond the %d supported", device,
                         kMaxDevices);
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_scratch[device])
        return GRT_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return set_error(GRT_DEVICE_ERROR, "hipGetSymbolAddress(scratch) on device %d: %s",
                         device, hipGetErrorString(e));
    }
    g_scratch[device] = static_cast<uint32_t*>(p);
    return GRT_OK;
}

}  // namespace grt

extern "C" int grt_compact_rays_noalloc(grt_context ctx, const grt_rays* rays, uint32_t n,
                                        uint32_t image_width, uint32_t* list_dev,
                                        uint32_t* count_dev, uint8_t* occluded_dev,
                                        const grt_hits* miss_hits)
{
    const bool rays_ok = rays && rays->ox && rays->oy && rays->oz && rays->dx && rays->dy &&
                         rays->dz && rays->tmin && rays->tmax;
    if (!ctx || !rays_ok || !list_dev || !count_dev ||
        (miss_hits && (!miss_hits->id || !miss_hits->t || !miss_hits->u || !miss_hits->v)))
        return set_error(GRT_INVALID_PARAM, "grt_compact_rays_noalloc: bad argument");
    const uint64_t slots = slot_count(n, image_width);
    if (slots > noalloc_max_slots())
        return set_error(GRT_CAPACITY,
                         "grt_compact_rays_noalloc: %llu ray slots exceed the fixed scratch "
                         "(at most %llu slots); use grt_compact_rays",
                         (unsigned long long)slots, (unsigned long long)noalloc_max_slots());
    uint32_t* ptr = scratch_of(ctx->device);
    if (!ptr)
        return set_error(GRT_NOT_READY,
                         "grt_compact_rays_noalloc: no scratch address for device %d",
                         ctx->device);
    const HitSoA mh = miss_hits ? HitSoA{miss_hits->id, miss_hits->t, miss_hits->u, miss_hits->v}
                                : HitSoA{nullptr, nullptr, nullptr, nullptr};
    return hip_check(launch_compact_rays(n, image_width, rays->tmin, rays->tmax, list_dev,
                                         count_dev, occluded_dev, mh, ptr, ctx->stream),
                     "compact_rays_noalloc");
}
