// grt_device.h -- device primitives shared by the grid-build translation units
// (grt_build.hip, grt_grid2.hip): block scan, streaming stores, the
// specification's cube root and the ascending sorts of cell lists.
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

// ---------------------------------------------------------------------------
// Ascending sorts of short id lists (grid cells and sub-cells)
// ---------------------------------------------------------------------------
constexpr uint32_t kSortTile = 16384;     // LDS tile of a block sort (64 KB)
constexpr uint32_t kSortWaveMax = 4096;   // longest list one wave sorts in registers

__device__ __forceinline__ void cswap(uint32_t& a, uint32_t& b)
{
    const uint32_t lo = a < b ? a : b;
    const uint32_t hi = a < b ? b : a;
    a = lo;
    b = hi;
}

// Bitonic layer (k = 2^LK, j = 2^LJ) over N registers of one thread, then the
// next smaller j: element i is compared with i^j, ascending when (i & k) == 0.
template <int N, int LK, int LJ>
__device__ __forceinline__ void thread_bitonic_layer(uint32_t (&a)[N])
{
    constexpr int k = 1 << LK, j = 1 << LJ;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        if ((i & j) == 0) {
            if ((i & k) == 0)
                cswap(a[i], a[i | j]);
            else
                cswap(a[i | j], a[i]);
        }
    }
    if constexpr (LJ > 0)
        thread_bitonic_layer<N, LK, LJ - 1>(a);
}

template <int N, int LK>
__device__ __forceinline__ void thread_bitonic(uint32_t (&a)[N])
{
    thread_bitonic_layer<N, LK, LK - 1>(a);
    if constexpr ((1 << LK) < N)
        thread_bitonic<N, LK + 1>(a);
}

// A cell of 2..16 references sorted by one thread in registers (slots past
// the count hold 0xffffffff, which sorts last).
template <int N>
__device__ __forceinline__ void sort_cell_thread(uint32_t* __restrict__ seg, uint32_t cnt)
{
    uint32_t a[N];
#pragma unroll
    for (int k = 0; k < N; ++k)
        a[k] = (uint32_t)k < cnt ? seg[k] : 0xffffffffu;
    bool sorted = true;
#pragma unroll
    for (int k = 1; k < N; ++k)
        sorted = sorted && (a[k - 1] <= a[k]);
    if (sorted)
        return;
    thread_bitonic<N, 1>(a);
#pragma unroll
    for (int k = 0; k < N; ++k)
        if ((uint32_t)k < cnt)
            seg[k] = a[k];
}

__device__ __forceinline__ void sort_small_cell(const uint32_t* __restrict__ offsets,
                                                uint32_t* __restrict__ refs, uint32_t c)
{
    const uint32_t b = offsets[c];
    const uint32_t cnt = offsets[c + 1] - b;
    if (cnt <= 8)
        sort_cell_thread<8>(refs + b, cnt);
    else
        sort_cell_thread<16>(refs + b, cnt);
}

// Bitonic sort written with ascending comparators only (the first comparator
// of each merge mirrors the block: i <-> i^(k-1)), so the virtual padding up to
// a power of two is +inf at the top and never needs storage: a comparator whose
// upper index is >= n is a no-op.
__device__ __forceinline__ void cex(uint32_t* a, uint32_t i, uint32_t j)
{
    const uint32_t x = a[i], y = a[j];
    if (y < x) {
        a[i] = y;
        a[j] = x;
    }
}

// Full sort of lds[0..n), n <= kSortTile, by the calling block.
__device__ inline void lds_sort(uint32_t* lds, uint32_t n)
{
    uint32_t p = 2;
    while (p < n)
        p <<= 1;
    for (uint32_t k = 2; k <= p; k <<= 1) {
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t i = threadIdx.x; i < (p >> 1); i += blockDim.x) {
                const uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
                const uint32_t hi = (j == (k >> 1)) ? (lo ^ (k - 1)) : (lo | j);
                const uint32_t a = lo < hi ? lo : hi, b = lo < hi ? hi : lo;
                if (b < n)
                    cex(lds, a, b);
            }
            __syncthreads();
        }
    }
}

// Half-cleaners j = jmax .. 1 on lds[0..len) where the tile starts at global
// index `g0` of a cell of n elements (comparators mirror nothing here).
__device__ inline void lds_half_cleaners(uint32_t* lds, uint32_t len, uint32_t jmax)
{
    for (uint32_t j = jmax; j > 0; j >>= 1) {
        for (uint32_t i = threadIdx.x; i < (kSortTile >> 1); i += blockDim.x) {
            const uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
            const uint32_t hi = lo | j;
            if (hi < len)
                cex(lds, lo, hi);
        }
        __syncthreads();
    }
}

// Global comparator pass over refs[b..b+n) by one block.
__device__ inline void global_pass(uint32_t* __restrict__ a, uint32_t n, uint32_t p, uint32_t k,
                                   uint32_t j)
{
    for (uint32_t i = threadIdx.x; i < (p >> 1); i += blockDim.x) {
        const uint32_t lo = ((i & ~(j - 1)) << 1) | (i & (j - 1));
        const uint32_t hi = (j == (k >> 1)) ? (lo ^ (k - 1)) : (lo | j);
        const uint32_t x = lo < hi ? lo : hi, y = lo < hi ? hi : lo;
        if (y < n)
            cex(a, x, y);
    }
    __syncthreads();
}

// One comparator layer (k = 2^LK, j = 2^LJ) of the wave sort below, then the
// next smaller j.
template <int E, int LK, int LJ>
__device__ __forceinline__ void bitonic_layer(uint32_t (&x)[E], uint32_t lane)
{
    constexpr uint32_t k = 1u << LK, j = 1u << LJ;
    if constexpr (j >= 64u) {
        constexpr uint32_t je = j >> 6;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            if (((uint32_t)e & je) == 0) {
                const int f = e ^ (int)je;
                const bool up = (((uint32_t)e * 64u) & k) == 0;
                const uint32_t lo = x[e] < x[f] ? x[e] : x[f];
                const uint32_t hi = x[e] < x[f] ? x[f] : x[e];
                x[e] = up ? lo : hi;
                x[f] = up ? hi : lo;
            }
        }
    } else {
        const bool lower = (lane & j) == 0;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint32_t o = (uint32_t)__shfl_xor((int)x[e], (int)j, 64);
            const bool up = ((((uint32_t)e * 64u) | lane) & k) == 0;
            const uint32_t lo = x[e] < o ? x[e] : o;
            const uint32_t hi = x[e] < o ? o : x[e];
            x[e] = (lower == up) ? lo : hi;
        }
    }
    if constexpr (LJ > 0)
        bitonic_layer<E, LK, LJ - 1>(x, lane);
}

// Merge stages k = 2^LK .. 64*E.
template <int E, int LK>
__device__ __forceinline__ void bitonic_stages(uint32_t (&x)[E], uint32_t lane)
{
    bitonic_layer<E, LK, LK - 1>(x, lane);
    if constexpr ((1u << LK) < 64u * E)
        bitonic_stages<E, LK + 1>(x, lane);
}

// One wave sorts one cell of up to 64*E references in registers: element
// i = e*64 + lane sits in lane `lane`, register e; the bitonic network compares
// i with i^j ascending when (i & k) == 0 -- across lanes by shuffles (j < 64),
// within a lane between registers (j >= 64).  Slots past n hold 0xffffffff,
// which sorts last.
template <int E>
__device__ __forceinline__ void wave_sort(uint32_t* __restrict__ seg, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t x[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = (uint32_t)e * 64u + lane;
        x[e] = i < n ? seg[i] : 0xffffffffu;
    }
    bitonic_stages<E, 1>(x, lane);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const uint32_t i = (uint32_t)e * 64u + lane;
        if (i < n)
            seg[i] = x[e];
    }
}


// Sort of seg[0..n), n <= 1024 * (BLOCK / 64), by a block of BLOCK threads:
// every wave sorts one run of 1024 ids in registers (the wave sort's
// network), then runs are merged pairwise through LDS -- src and dst
// ping-pong, n words each -- every thread writing 16 consecutive outputs from
// its merge-path split (a binary search on the diagonal; ties take the left
// run first).  Run boundaries are multiples of 2048 and the outputs of a
// thread start at a multiple of 16, so no thread's outputs straddle two
// merges.
template <int BLOCK>
__device__ inline void block_merge_sort(uint32_t* __restrict__ seg, uint32_t n, uint32_t* src,
                                        uint32_t* dst)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t run = (threadIdx.x >> 6) * 1024u;
    if (run < n) {
        uint32_t x[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t i = run + (uint32_t)e * 64u + lane;
            x[e] = i < n ? seg[i] : 0xffffffffu;
        }
        bitonic_stages<16, 1>(x, lane);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const uint32_t i = run + (uint32_t)e * 64u + lane;
            if (i < n)
                src[i] = x[e];
        }
    }
    __syncthreads();
    for (uint32_t w = 1024u; w < n; w <<= 1) {
        const uint32_t p0 = threadIdx.x * 16u;
        if (p0 < n) {
            const uint32_t base = p0 & ~(2u * w - 1u);
            const uint32_t rest = n - base;
            const uint32_t la = rest < w ? rest : w;
            const uint32_t lb = rest > w ? (rest - w < w ? rest - w : w) : 0u;
            const uint32_t* ra = src + base;
            const uint32_t* rb = ra + la;
            const uint32_t q = p0 - base;
            uint32_t lo = q > lb ? q - lb : 0u, hi = q < la ? q : la;
            while (lo < hi) {
                const uint32_t mid = (lo + hi) >> 1;
                if (ra[mid] <= rb[q - 1u - mid])
                    lo = mid + 1u;
                else
                    hi = mid;
            }
            uint32_t ia = lo, ib = q - lo;
            const uint32_t end = p0 + 16u < n ? p0 + 16u : n;
            for (uint32_t p = p0; p < end; ++p) {
                const bool take_a = ib >= lb || (ia < la && ra[ia] <= rb[ib]);
                dst[p] = take_a ? ra[ia++] : rb[ib++];
            }
        }
        __syncthreads();
        uint32_t* t = src;
        src = dst;
        dst = t;
    }
    for (uint32_t i = threadIdx.x; i < n; i += BLOCK)
        seg[i] = src[i];
    __syncthreads();
}

// Sort of one cell segment seg[0..n), n > kSortWaveMax, by the whole block:
// LDS bitonic sort when it fits a tile; otherwise each tile is sorted in LDS
// and the merge stages with k > kSortTile run their j >= kSortTile
// comparators in global memory and the rest tile by tile in LDS.
__device__ inline void block_sort_cell(uint32_t* __restrict__ seg, uint32_t n, uint32_t* lds)
{
    if (n <= kSortTile) {
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            lds[i] = seg[i];
        __syncthreads();
        lds_sort(lds, n);
        for (uint32_t i = threadIdx.x; i < n; i += blockDim.x)
            seg[i] = lds[i];
        __syncthreads();
        return;
    }
    uint32_t p = kSortTile;
    while (p < n)
        p <<= 1;
    for (uint32_t t0 = 0; t0 < n; t0 += kSortTile) {
        const uint32_t len = (n - t0) < kSortTile ? (n - t0) : kSortTile;
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
            lds[i] = seg[t0 + i];
        __syncthreads();
        lds_sort(lds, len);
        for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
            seg[t0 + i] = lds[i];
        __syncthreads();
    }
    for (uint32_t k = 2 * kSortTile; k <= p; k <<= 1) {
        for (uint32_t j = k >> 1; j >= kSortTile; j >>= 1)
            global_pass(seg, n, p, k, j);
        for (uint32_t t0 = 0; t0 < n; t0 += kSortTile) {
            const uint32_t len = (n - t0) < kSortTile ? (n - t0) : kSortTile;
            for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
                lds[i] = seg[t0 + i];
            __syncthreads();
            lds_half_cleaners(lds, len, kSortTile >> 1);
            for (uint32_t i = threadIdx.x; i < len; i += blockDim.x)
                seg[t0 + i] = lds[i];
            __syncthreads();
        }
    }
}

}  // namespace grt
