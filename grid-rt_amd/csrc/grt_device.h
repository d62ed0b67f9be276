// grt_device.h -- device primitives shared by the grid-build translation units
// (grt_build.hip, grt_grid2.hip): block scan, streaming stores and the
// specification's cube root.
#pragma once

#include "grt_common.h"

namespace grt {

// ---------------------------------------------------------------------------
// Wave64 / block primitives
// ---------------------------------------------------------------------------

// Exclusive block scan (BLOCK threads, BLOCK/64 waves); lds needs BLOCK/64+1
// words.  Returns the exclusive prefix; `total` receives the block sum.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds, uint32_t& total)
{
    constexpr int NW = BLOCK / 64;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63)
        lds[wid] = inc;
    __syncthreads();
    if (wid == 0) {
        const uint32_t w = lane < NW ? lds[lane] : 0u;
        const uint32_t wi = wave_incl_scan(w);
        if (lane < NW)
            lds[lane] = wi - w;
        if (lane == NW - 1)
            lds[NW] = wi;
    }
    __syncthreads();
    const uint32_t res = inc - v + lds[wid];
    total = lds[NW];
    __syncthreads();
    return res;
}

// Streaming 16-byte store that is not kept in the caches (the grid's dense
// arrays are written once per build and read later by traversal).
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void stream_store(uint4* p, uint4 v)
{
    __builtin_nontemporal_store(u32x4{v.x, v.y, v.z, v.w}, reinterpret_cast<u32x4*>(p));
}
__device__ __forceinline__ void stream_store(float4* p, float4 v)
{
    stream_store(reinterpret_cast<uint4*>(p),
                 make_uint4(__float_as_uint(v.x), __float_as_uint(v.y), __float_as_uint(v.z),
                            __float_as_uint(v.w)));
}

// Cube root of the specification: binary64 Newton iteration with basic IEEE
// operations only, so host and device agree bit for bit (the reference's
// powf(x, 1.0f/3.0f), UniformGridAccStructBuilder.cpp:93, is not reproducible
// across math libraries).
__device__ __forceinline__ float spec_cbrtf(float xf)
{
    double x = (double)xf;
    if (!(x > 0.0) || __builtin_isinf(x))
        return xf;
    int e;
    double m = frexp(x, &e);
    const int q = (e >= 0) ? (e / 3) : -((-e + 2) / 3);
    const int r = e - 3 * q;
    m = ldexp(m, r);
    double y = 1.0;
#pragma unroll 1
    for (int i = 0; i < 8; ++i)
        y = (2.0 * y + m / (y * y)) / 3.0;
    return (float)ldexp(y, q);
}

}  // namespace grt
