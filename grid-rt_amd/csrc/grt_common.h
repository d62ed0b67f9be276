// grt_common.h -- device-side records and launch helpers shared by the HIP
// translation units of libgrt (grid build, traversal, ray generation).
//
// Floating-point rules (identical to the CPU oracle): every fused
// multiply-add is an explicit __builtin_fmaf, everything else is one IEEE-754
// binary32 operation in the written order (all HIP code is compiled with
// -ffp-contract=off), division and sqrtf are correctly rounded, and min/max are
// the ternaries a<b?a:b / a>b?a:b.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace grt {

// Grid frame in HBM: padded scene box, per-axis resolution, cell size and
// reciprocal (UniformGridAccStruct.cpp:44-63), plus the build's outputs.
struct alignas(64) DevFrame {
    float bmin[3];
    float bmax[3];
    float cell[3];
    float inv[3];
    int32_t n[3];
    uint32_t m;        // cells = n0*n1*n2
    uint32_t r;        // references = offsets[m]
    uint32_t status;   // 0 ok; bit0 frame invalid; bit1 cells over capacity; bit2 refs over capacity
};
static_assert(sizeof(DevFrame) == 128, "DevFrame is two 64-byte lines");

enum : uint32_t {
    kStatusFrameInvalid = 1u,
    kStatusCellsOver = 2u,
    kStatusRefsOver = 4u,
};

// Ray batch, structure of arrays (device pointers).
struct RaySoA {
    const float* ox;
    const float* oy;
    const float* oz;
    const float* dx;
    const float* dy;
    const float* dz;
    const float* tmin;
    const float* tmax;
};
struct RaySoAOut {
    float* ox;
    float* oy;
    float* oz;
    float* dx;
    float* dy;
    float* dz;
    float* tmin;
    float* tmax;
};

// Closest-hit records, structure of arrays (device pointers).
struct HitSoA {
    int32_t* id;
    float* t;
    float* u;
    float* v;
};
struct HitSoAIn {
    const int32_t* id;
    const float* t;
    const float* u;
    const float* v;
    uint32_t n_tris;   // ids outside [0, n_tris) are treated as misses
};

// Pinhole camera in the reference C-ABI form (RayTracing.h:20-23).
struct DevCamera {
    float pos[3];
    float base[3];
    float nu[3];
    float nv[3];
    uint32_t width;
    uint32_t height;
};

constexpr float kRayEpsilon = 2e-4f;                          // Context.cpp:505
constexpr float kExitScale = 1.0f + 1.0f / 1048576.0f;        // cell-exit bound factor

__device__ __forceinline__ float tmin_f(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float tmax_f(float a, float b) { return a > b ? a : b; }

__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz)
{
    return __builtin_fmaf(az, bz, __builtin_fmaf(ay, by, ax * bx));
}

// worldToVoxel (UniformGridAccStruct.cpp:92-98) clamped to [0, n-1] in float.
__device__ __forceinline__ int32_t vox(float val, float bmin, float inv, int32_t n)
{
    float x = (val - bmin) * inv;
    x = x > 0.0f ? x : 0.0f;
    const float hi = (float)(n - 1);
    x = x < hi ? x : hi;
    return (int32_t)x;
}

// Inclusive prefix sum over a full wave64 with DPP: row_shr 1,2,4,8 inside each
// row of 16 lanes, then row_bcast15 (rows 1,3) and row_bcast31 (rows 2,3).
// All 64 lanes must be active.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, true);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
    return v;
}

// Number of set bits of a wave mask below the calling lane.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                     __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Slot -> ray index.  With img_w != 0 each group of 64 consecutive slots is an
// 8x8 pixel tile of a row-major image of that width.
__device__ __forceinline__ uint32_t slot_ray(uint32_t slot, uint32_t img_w, uint32_t n,
                                             bool& valid)
{
    if (img_w == 0) {
        valid = slot < n;
        return slot;
    }
    const uint32_t wave = slot >> 6, lane = slot & 63u;
    const uint32_t tiles_x = (img_w + 7u) >> 3;
    const uint32_t tx = wave % tiles_x, ty = wave / tiles_x;
    const uint32_t px = tx * 8u + (lane & 7u), py = ty * 8u + (lane >> 3);
    const uint32_t i = py * img_w + px;
    valid = (px < img_w) && (i < n);
    return i;
}

// Lanes of one wave exchange data through LDS: order the accesses around this
// point (a wave's LDS operations execute in program order).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Slots covering n rays: n, or whole 8x8 tiles of an image of width img_w.
inline uint64_t slot_count(uint32_t n, uint32_t img_w)
{
    if (img_w == 0)
        return n;
    const uint64_t h = (n + (uint64_t)img_w - 1) / img_w;
    return (uint64_t)((img_w + 7u) >> 3) * ((h + 7u) >> 3) * 64u;
}

}  // namespace grt
