// grt_build.hip -- per-frame uniform-grid construction on gfx950.
//
// Pipeline (one stream, no host round trip needed once buffers are sized):
//   1. k_bbox_partial + k_frame : scene box -> padded grid frame
//        (Aabb.cpp:11-91, UniformGridAccStructBuilder.cpp:73-102,
//         UniformGridAccStruct.cpp:44-63)
//   2. k_grid_pass<false>       : per-cell reference counts
//        (UniformGridAccStructBuilder.cpp:108-145; countCellVP.glsl:31-58)
//   3. k_scan_*                 : exclusive scan of the counts (the count pass
//        also sums each 4096-cell tile; DPP wave scan, LDS block scan)
//   4. k_grid_pass<true>        : scatter triangle ids into their cells
//        (gridVP_sem_GS.glsl:56-106)
//   5. k_sort_small/large       : every cell's ids ascending, the order the
//        reference's push_back loop produces (UniformGridAccStructBuilder.cpp:141)
//
// Triangles spanning more than kSmallCells cells (ground planes) are put on a
// list and expanded by whole blocks, so one thread never loops over 10^5 cells.
#include "grt_device.h"
#include "grt_launch.h"

namespace grt {

constexpr uint32_t kSmallCells = 64;
constexpr uint32_t kHugeCells = 65536;
constexpr uint32_t kHugeMerge = 4;   // huge triangles applied together in one cell walk

// ---------------------------------------------------------------------------
// 1. Frame
// ---------------------------------------------------------------------------

// Aabb::scaleBy(0.01f) + resolution heuristic + setResolution.
__device__ void frame_from_box(const float mn[3], const float mx[3], uint32_t n_tris,
                               float lambda, uint32_t cells_cap, DevFrame* fr)
{
    uint32_t status = 0;
    float diag[3];
    for (int d = 0; d < 3; ++d) {
        const float s = (mn[d] == mx[d]) ? 0.01f : fabsf(mx[d] - mn[d]) * 0.01f;
        fr->bmin[d] = mn[d] - s;
        fr->bmax[d] = mx[d] + s;
        diag[d] = fr->bmax[d] - fr->bmin[d];
    }
    const float vol = (diag[0] * diag[1]) * diag[2];
    const float x = (lambda * (float)n_tris) / vol;
    const float fac = spec_cbrtf(x);
    uint64_t m = 1;
    for (int d = 0; d < 3; ++d) {
        const float c = ceilf(diag[d] * fac);
        int32_t n = 1;
        if (!(c < 1048576.0f))
            status |= kStatusFrameInvalid;
        else
            n = (int32_t)c;
        if (n < 1)
            n = 1;
        fr->n[d] = n;
        m *= (uint64_t)n;
    }
    if (n_tris == 0 || !(lambda > 0.0f) || m > (1ull << 30))
        status |= kStatusFrameInvalid;
    if (m > (uint64_t)cells_cap)
        status |= kStatusCellsOver;
    fr->m = (uint32_t)(m > 0xffffffffull ? 0xffffffffull : m);
    for (int d = 0; d < 3; ++d) {
        fr->cell[d] = diag[d] / (float)fr->n[d];
        fr->inv[d] = 1.0f / fr->cell[d];
    }
    fr->r = 0;
    fr->status = status;
}

__global__ __launch_bounds__(256) void k_bbox_partial(const float* __restrict__ verts,
                                                      uint32_t n_tris, float* __restrict__ partial)
{
    __shared__ float lds[6][4];
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n_tris;
         t += gridDim.x * blockDim.x) {
        const float* v = verts + 9ull * t;
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const float x = v[3 * k + d];
                mn[d] = x < mn[d] ? x : mn[d];
                mx[d] = x > mx[d] ? x : mx[d];
            }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float a = __shfl_xor(mn[d], off);
            const float b = __shfl_xor(mx[d], off);
            mn[d] = a < mn[d] ? a : mn[d];
            mx[d] = b > mx[d] ? b : mx[d];
        }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int d = 0; d < 3; ++d) {
            lds[d][wid] = mn[d];
            lds[3 + d][wid] = mx[d];
        }
    __syncthreads();
    if (threadIdx.x < 6) {
        const int d = threadIdx.x;
        float r = lds[d][0];
        for (int w = 1; w < 4; ++w) {
            const float x = lds[d][w];
            r = d < 3 ? (x < r ? x : r) : (x > r ? x : r);
        }
        partial[6 * blockIdx.x + d] = r;
    }
}

__global__ __launch_bounds__(64) void k_frame(const float* __restrict__ partial, uint32_t nparts,
                                              uint32_t n_tris, float lambda, uint32_t cells_cap,
                                              DevFrame* __restrict__ fr)
{
    // One wave reduces the block partials (min/max are exact, so the order
    // does not matter), then lane 0 computes the frame.
    float mn[3] = {INFINITY, INFINITY, INFINITY};
    float mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t p = threadIdx.x; p < nparts; p += 64)
        for (int d = 0; d < 3; ++d) {
            const float a = partial[6 * p + d], b = partial[6 * p + 3 + d];
            mn[d] = a < mn[d] ? a : mn[d];
            mx[d] = b > mx[d] ? b : mx[d];
        }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const float a = __shfl_xor(mn[d], off);
            const float b = __shfl_xor(mx[d], off);
            mn[d] = a < mn[d] ? a : mn[d];
            mx[d] = b > mx[d] ? b : mx[d];
        }
    if (threadIdx.x == 0)
        frame_from_box(mn, mx, n_tris, lambda, cells_cap, fr);
}

__global__ void k_frame_box(float x0, float y0, float z0, float x1, float y1, float z1,
                            uint32_t n_tris, float lambda, uint32_t cells_cap, DevFrame* fr)
{
    if (threadIdx.x != 0)
        return;
    const float mn[3] = {x0, y0, z0}, mx[3] = {x1, y1, z1};
    frame_from_box(mn, mx, n_tris, lambda, cells_cap, fr);
}

// ---------------------------------------------------------------------------
// 2/4. Count and scatter passes
// ---------------------------------------------------------------------------
struct CellBox {
    int32_t lo[3];
    int32_t w[3];
};

__device__ __forceinline__ CellBox tri_cells(const DevFrame* fr, const float* v)
{
    CellBox b;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const float a = v[d], c1 = v[3 + d], c2 = v[6 + d];
        float mn = a, mx = a;
        mn = c1 < mn ? c1 : mn;
        mx = c1 > mx ? c1 : mx;
        mn = c2 < mn ? c2 : mn;
        mx = c2 > mx ? c2 : mx;
        const int32_t lo = vox(mn, fr->bmin[d], fr->inv[d], fr->n[d]);
        const int32_t hi = vox(mx, fr->bmin[d], fr->inv[d], fr->n[d]);
        b.lo[d] = lo;
        b.w[d] = hi - lo + 1;
    }
    return b;
}

// The count pass also sums the counts of every scan tile (kScanTile cells),
// so the scan needs no reduction pass over the count array.
constexpr uint32_t kScanTileShift = 12;
static_assert((1u << kScanTileShift) == kScanTile, "scan tile");

// Adds 1 to the tile sum of every active lane's cell: lanes in the tile of
// the first active lane add together, the others one by one.  Called by the
// active lanes of a wave together.
__device__ __forceinline__ void tile_add_wave(uint32_t* tile_sums, bool active, uint32_t c)
{
    const uint64_t am = __ballot(active);
    if (am == 0)
        return;
    const uint32_t f = (uint32_t)__builtin_ctzll(am);
    const uint32_t tile = c >> kScanTileShift;
    const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)tile, (int)f);
    const bool same = active && tile == t0;
    const uint64_t sm = __ballot(same);
    if ((threadIdx.x & 63u) == f)
        atomicAdd(&tile_sums[t0], (uint32_t)__popcll(sm));
    if (active && !same)
        atomicAdd(&tile_sums[tile], 1u);
}

// As tile_add_wave, adding m (per lane) instead of 1.
__device__ __forceinline__ void tile_add_wave_n(uint32_t* tile_sums, bool active, uint32_t c,
                                                uint32_t m)
{
    const uint64_t am = __ballot(active);
    if (am == 0)
        return;
    const uint32_t f = (uint32_t)__builtin_ctzll(am);
    const uint32_t tile = c >> kScanTileShift;
    const uint32_t t0 = (uint32_t)__builtin_amdgcn_readlane((int)tile, (int)f);
    const bool same = active && tile == t0;
    const uint32_t sum =
        (uint32_t)__builtin_amdgcn_readlane((int)wave_incl_scan(same ? m : 0u), 63);
    if ((threadIdx.x & 63u) == f)
        atomicAdd(&tile_sums[t0], sum);
    if (active && !same)
        atomicAdd(&tile_sums[tile], m);
}

template <bool SCATTER>
__device__ __forceinline__ void cell_op(uint32_t c, uint32_t t, uint32_t* counts,
                                        const uint32_t* offsets, uint32_t* refs,
                                        uint32_t* tile_sums)
{
    if (SCATTER) {
        const uint32_t slot = atomicSub(&counts[c], 1u) - 1u;
        refs[offsets[c] + slot] = t;
    } else {
        atomicAdd(&counts[c], 1u);
        tile_add_wave(tile_sums, true, c);
    }
}

// Per-block LDS hash of the cells touched by the block's triangles (key: cell
// index, or kHashEmpty).  Keys are written once and never removed, so a cell
// that could not be inserted within kHashProbes probes is never found later
// either, and such (triangle, cell) pairs fall back to global atomics.
constexpr uint32_t kHashBits = 11;
constexpr uint32_t kHashSlots = 1u << kHashBits;
constexpr uint32_t kHashEmpty = 0xffffffffu;
constexpr uint32_t kHashProbes = 16;

__device__ __forceinline__ uint32_t hash_slot(uint32_t c)
{
    return (c * 2654435761u) >> (32 - kHashBits);
}

__device__ __forceinline__ uint32_t hash_insert(uint32_t* keys, uint32_t c)
{
    uint32_t h = hash_slot(c);
    for (uint32_t p = 0; p < kHashProbes; ++p) {
        const uint32_t prev = atomicCAS(&keys[h], kHashEmpty, c);
        if (prev == kHashEmpty || prev == c)
            return h;
        h = (h + 1) & (kHashSlots - 1);
    }
    return kHashSlots;
}

__device__ __forceinline__ uint32_t hash_find(const uint32_t* keys, uint32_t c)
{
    uint32_t h = hash_slot(c);
    for (uint32_t p = 0; p < kHashProbes; ++p) {
        const uint32_t k = keys[h];
        if (k == c)
            return h;
        if (k == kHashEmpty)
            return kHashSlots;
        h = (h + 1) & (kHashSlots - 1);
    }
    return kHashSlots;
}

// Count (SCATTER = false) or scatter pass for triangles of at most kSmallCells
// cells, one thread per triangle.  Consecutive triangles are neighbours in
// space, so a block's (triangle, cell) pairs are first combined per cell in
// LDS and each distinct cell costs one global atomic per block: the count
// adds the block's total; the scatter reserves a range of the cell's slots
// (counting down, which returns counts[] to zero) and hands positions out of
// it with LDS atomics.  The order within a cell is fixed by the sort pass.
constexpr uint32_t kTileSlots = 32;   // per-block tile sums of the count pass

template <bool SCATTER>
__global__ __launch_bounds__(256) void k_grid_pass(const DevFrame* __restrict__ fr,
                                                   const float* __restrict__ verts, uint32_t n_tris,
                                                   uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ offsets,
                                                   uint32_t* __restrict__ refs,
                                                   uint32_t* __restrict__ big_list,
                                                   uint32_t* __restrict__ big_count,
                                                   uint32_t* __restrict__ tile_sums)
{
    __shared__ uint32_t keys[kHashSlots];
    __shared__ uint32_t slot_n[kHashSlots];   // count, then the next absolute position
    __shared__ uint32_t tile_key[kTileSlots];
    __shared__ uint32_t tile_n[kTileSlots];
    const uint32_t bad = SCATTER ? (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)
                                 : (kStatusFrameInvalid | kStatusCellsOver);
    if (fr->status & bad)
        return;   // uniform over the block
    for (uint32_t i = threadIdx.x; i < kHashSlots; i += blockDim.x) {
        keys[i] = kHashEmpty;
        slot_n[i] = 0;
    }
    if (!SCATTER && threadIdx.x < kTileSlots) {
        tile_key[threadIdx.x] = kHashEmpty;
        tile_n[threadIdx.x] = 0;
    }
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    CellBox b = {};
    bool small = false;
    if (t < n_tris) {
        b = tri_cells(fr, verts + 9ull * t);
        const uint32_t k = (uint32_t)b.w[0] * (uint32_t)b.w[1] * (uint32_t)b.w[2];
        if (k > kSmallCells) {
            // Medium triangles fill the list from the front, huge ones from
            // the back (each triangle is on at most one list).
            if (!SCATTER) {
                if (k > kHugeCells)
                    big_list[n_tris - 1u - atomicAdd(big_count + 3, 1u)] = t;
                else
                    big_list[atomicAdd(big_count, 1u)] = t;
            }
        } else {
            small = true;
        }
    }
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    if (small) {
        for (int32_t z = 0; z < b.w[2]; ++z)
            for (int32_t y = 0; y < b.w[1]; ++y) {
                const uint32_t row =
                    (uint32_t)b.lo[0] + nx * ((uint32_t)(b.lo[1] + y) + ny * (uint32_t)(b.lo[2] + z));
                for (int32_t x = 0; x < b.w[0]; ++x) {
                    const uint32_t c = row + (uint32_t)x;
                    const uint32_t h = hash_insert(keys, c);
                    if (h < kHashSlots) {
                        atomicAdd(&slot_n[h], 1u);
                    } else if (SCATTER) {
                        const uint32_t slot = atomicSub(&counts[c], 1u) - 1u;
                        refs[offsets[c] + slot] = t;
                    } else {
                        atomicAdd(&counts[c], 1u);
                        atomicAdd(&tile_sums[c >> kScanTileShift], 1u);
                    }
                }
            }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < kHashSlots; i += blockDim.x) {
        const uint32_t c = keys[i];
        if (c == kHashEmpty)
            continue;
        const uint32_t n = slot_n[i];
        if (SCATTER) {
            slot_n[i] = offsets[c] + atomicSub(&counts[c], n) - n;
        } else {
            atomicAdd(&counts[c], n);
            // the block's tile sums in LDS, a global atomic when they are full
            const uint32_t tile = c >> kScanTileShift;
            uint32_t h = tile & (kTileSlots - 1);
            bool done = false;
            for (uint32_t p = 0; p < kTileSlots && !done; ++p) {
                const uint32_t prev = atomicCAS(&tile_key[h], kHashEmpty, tile);
                if (prev == kHashEmpty || prev == tile) {
                    atomicAdd(&tile_n[h], n);
                    done = true;
                }
                h = (h + 1) & (kTileSlots - 1);
            }
            if (!done)
                atomicAdd(&tile_sums[tile], n);
        }
    }
    if (!SCATTER) {
        __syncthreads();
        if (threadIdx.x < kTileSlots && tile_key[threadIdx.x] != kHashEmpty)
            atomicAdd(&tile_sums[tile_key[threadIdx.x]], tile_n[threadIdx.x]);
        return;
    }
    __syncthreads();
    if (small) {
        for (int32_t z = 0; z < b.w[2]; ++z)
            for (int32_t y = 0; y < b.w[1]; ++y) {
                const uint32_t row =
                    (uint32_t)b.lo[0] + nx * ((uint32_t)(b.lo[1] + y) + ny * (uint32_t)(b.lo[2] + z));
                for (int32_t x = 0; x < b.w[0]; ++x) {
                    const uint32_t h = hash_find(keys, row + (uint32_t)x);
                    if (h < kHashSlots)
                        refs[atomicAdd(&slot_n[h], 1u)] = t;
                }
            }
    }
}

template <bool SCATTER>
__global__ __launch_bounds__(256) void k_grid_pass_huge(const DevFrame* __restrict__ fr,
                                                        const float* __restrict__ verts,
                                                        uint32_t n_tris,
                                                        uint32_t* __restrict__ counts,
                                                        const uint32_t* __restrict__ offsets,
                                                        uint32_t* __restrict__ refs,
                                                        const uint32_t* __restrict__ big_list,
                                                        const uint32_t* __restrict__ big_count,
                                                        uint32_t* __restrict__ tile_sums)
{
    const uint32_t bad = SCATTER ? (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)
                                 : (kStatusFrameInvalid | kStatusCellsOver);
    if (fr->status & bad)
        return;
    const uint32_t nh = big_count[3];
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    if (nh > 0 && nh <= kHugeMerge) {
        // Few huge triangles (ground planes) whose boxes overlap: walk the
        // union of their boxes once, one thread per cell and blocks over
        // whole rows, and apply every huge triangle of the cell with plain
        // loads and stores -- no other thread touches the cell in this kernel.
        CellBox bx[kHugeMerge];
        uint32_t tid[kHugeMerge];
        int32_t ulo[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff}, uhi[3] = {-1, -1, -1};
        uint64_t vol = 0;
#pragma unroll
        for (uint32_t i = 0; i < kHugeMerge; ++i) {
            tid[i] = 0;
            bx[i] = CellBox{{0, 0, 0}, {0, 0, 0}};
            if (i < nh) {
                tid[i] = big_list[n_tris - 1u - i];
                bx[i] = tri_cells(fr, verts + 9ull * tid[i]);
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    ulo[d] = bx[i].lo[d] < ulo[d] ? bx[i].lo[d] : ulo[d];
                    const int32_t h = bx[i].lo[d] + bx[i].w[d] - 1;
                    uhi[d] = h > uhi[d] ? h : uhi[d];
                }
                vol += (uint64_t)bx[i].w[0] * (uint64_t)bx[i].w[1] * (uint64_t)bx[i].w[2];
            }
        }
        const uint32_t wx = (uint32_t)(uhi[0] - ulo[0] + 1), wy = (uint32_t)(uhi[1] - ulo[1] + 1),
                       wz = (uint32_t)(uhi[2] - ulo[2] + 1);
        if ((uint64_t)wx * wy * wz <= 2 * vol) {
            const uint32_t rows = wy * wz;
            for (uint32_t row = blockIdx.x; row < rows; row += gridDim.x) {
                const uint32_t y = (uint32_t)ulo[1] + row % wy, z = (uint32_t)ulo[2] + row / wy;
                const uint32_t row_c = nx * (y + ny * z);
                for (uint32_t x0 = 0; x0 < wx; x0 += blockDim.x) {
                    const uint32_t x = (uint32_t)ulo[0] + x0 + threadIdx.x;
                    const bool in_row = x0 + threadIdx.x < wx;
                    bool hit[kHugeMerge];
                    uint32_t m = 0;
#pragma unroll
                    for (uint32_t i = 0; i < kHugeMerge; ++i) {
                        hit[i] = in_row && i < nh &&
                                 (uint32_t)((int32_t)x - bx[i].lo[0]) < (uint32_t)bx[i].w[0] &&
                                 (uint32_t)((int32_t)y - bx[i].lo[1]) < (uint32_t)bx[i].w[1] &&
                                 (uint32_t)((int32_t)z - bx[i].lo[2]) < (uint32_t)bx[i].w[2];
                        m += hit[i] ? 1u : 0u;
                    }
                    const uint32_t c = x + row_c;
                    if (SCATTER) {
                        if (m != 0) {
                            const uint32_t left = counts[c], off = offsets[c];
                            uint32_t q = 0;
#pragma unroll
                            for (uint32_t i = 0; i < kHugeMerge; ++i) {
                                if (hit[i]) {
                                    refs[off + left - 1u - q] = tid[i];
                                    ++q;
                                }
                            }
                            counts[c] = left - m;
                        }
                    } else {
                        if (m != 0)
                            counts[c] += m;
                        tile_add_wave_n(tile_sums, m != 0, c, m);
                    }
                }
            }
            return;
        }
    }
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i = 0; i < nh; ++i) {
        const uint32_t t = big_list[n_tris - 1u - i];
        const CellBox b = tri_cells(fr, verts + 9ull * t);
        const uint32_t wx = (uint32_t)b.w[0], wy = (uint32_t)b.w[1];
        const uint32_t k = wx * wy * (uint32_t)b.w[2];
        for (uint32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < k; j += stride) {
            const uint32_t x = j % wx;
            const uint32_t yz = j / wx;
            const uint32_t y = yz % wy;
            const uint32_t z = yz / wy;
            const uint32_t c = ((uint32_t)b.lo[0] + x) +
                               nx * (((uint32_t)b.lo[1] + y) + ny * ((uint32_t)b.lo[2] + z));
            cell_op<SCATTER>(c, t, counts, offsets, refs, tile_sums);
        }
    }
}

__global__ void k_zero_build_counters(uint32_t* __restrict__ c)
{
    if (threadIdx.x == 0) {
        c[0] = 0u;
        c[3] = 0u;
    }
}

// Triangles on the medium list: one block per triangle, threads stride its cells.
template <bool SCATTER>
__global__ __launch_bounds__(256) void k_grid_pass_big(const DevFrame* __restrict__ fr,
                                                       const float* __restrict__ verts,
                                                       uint32_t* __restrict__ counts,
                                                       const uint32_t* __restrict__ offsets,
                                                       uint32_t* __restrict__ refs,
                                                       const uint32_t* __restrict__ big_list,
                                                       const uint32_t* __restrict__ big_count,
                                                       uint32_t* __restrict__ tile_sums)
{
    const uint32_t bad = SCATTER ? (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)
                                 : (kStatusFrameInvalid | kStatusCellsOver);
    if (fr->status & bad)
        return;
    const uint32_t nb = *big_count;
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    for (uint32_t i = blockIdx.x; i < nb; i += gridDim.x) {
        const uint32_t t = big_list[i];
        const CellBox b = tri_cells(fr, verts + 9ull * t);
        const uint32_t wx = (uint32_t)b.w[0], wy = (uint32_t)b.w[1];
        const uint32_t k = wx * wy * (uint32_t)b.w[2];
        for (uint32_t j = threadIdx.x; j < k; j += blockDim.x) {
            const uint32_t x = j % wx;
            const uint32_t yz = j / wx;
            const uint32_t y = yz % wy;
            const uint32_t z = yz / wy;
            const uint32_t c = ((uint32_t)b.lo[0] + x) +
                               nx * (((uint32_t)b.lo[1] + y) + ny * ((uint32_t)b.lo[2] + z));
            cell_op<SCATTER>(c, t, counts, offsets, refs, tile_sums);
        }
    }
}

// ---------------------------------------------------------------------------
// 3. Exclusive scan (reduce -> scan block sums -> apply)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool scan_len(const DevFrame* fr, uint32_t n_host, uint32_t& n)
{
    if (fr) {
        if (fr->status & (kStatusFrameInvalid | kStatusCellsOver))
            return false;
        n = fr->m;
    } else {
        n = n_host;
    }
    return true;
}

__device__ __forceinline__ void load16(const uint32_t* __restrict__ in, uint32_t i0, uint32_t n,
                                       uint32_t v[16])
{
    if (i0 + 16 <= n) {
        const uint4* p = reinterpret_cast<const uint4*>(in + i0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint4 w = p[q];
            v[4 * q] = w.x;
            v[4 * q + 1] = w.y;
            v[4 * q + 2] = w.z;
            v[4 * q + 3] = w.w;
        }
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            v[k] = (i0 + k < n) ? in[i0 + k] : 0u;
    }
}

__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* __restrict__ in,
                                                     uint32_t n_host, const DevFrame* fr,
                                                     uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t lds[8];
    uint32_t n;
    if (!scan_len(fr, n_host, n))
        return;
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n)
        return;
    uint32_t v[16];
    load16(in, base + threadIdx.x * 16u, n, v);
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        s += v[k];
    uint32_t total;
    block_excl_scan<256>(s, lds, total);
    if (threadIdx.x == 0)
        bsum[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void k_scan_blocks(uint32_t* __restrict__ bsum, uint32_t n_host,
                                                      DevFrame* fr, uint32_t* __restrict__ out,
                                                      uint32_t refs_cap)
{
    __shared__ uint32_t lds[17];
    uint32_t n;
    if (!scan_len(fr, n_host, n))
        return;
    const uint32_t nb = (n + kScanTile - 1) / kScanTile;
    uint64_t carry = 0;
    // 4096 sums per round, four consecutive ones per thread.
    for (uint32_t base = 0; base < nb; base += 4096) {
        const uint32_t i = base + 4u * threadIdx.x;
        uint32_t v[4] = {0u, 0u, 0u, 0u};
        const bool vec = i + 4 <= nb;
        if (vec) {
            const uint4 x = *reinterpret_cast<const uint4*>(bsum + i);
            v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w;
        } else {
            for (uint32_t k = 0; k < 4; ++k)
                if (i + k < nb)
                    v[k] = bsum[i + k];
        }
        uint32_t tot;
        const uint32_t ex = (uint32_t)carry + block_excl_scan<1024>(v[0] + v[1] + v[2] + v[3], lds, tot);
        const uint32_t o[4] = {ex, ex + v[0], ex + v[0] + v[1], ex + v[0] + v[1] + v[2]};
        if (vec) {
            *reinterpret_cast<uint4*>(bsum + i) = make_uint4(o[0], o[1], o[2], o[3]);
        } else {
            for (uint32_t k = 0; k < 4; ++k)
                if (i + k < nb)
                    bsum[i + k] = o[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) {
        out[n] = (uint32_t)carry;
        if (fr) {
            fr->r = (uint32_t)carry;
            if (carry > (uint64_t)refs_cap)
                fr->status |= kStatusRefsOver;
        }
    }
}

// With lists.small set, also appends every cell with 2..kSortSmallCells
// references to lists.small and every larger cell to lists.large (the cells
// the sort pass has to visit), one atomic per block and list.
// When the total exceeds the reference capacity (kStatusRefsOver) the scatter
// will not run, so nothing would return the counts to zero: with `clear` set
// (the grid's counts, the same array as `in`), each block then zeroes the
// counts of its tile instead of writing offsets, and the next build starts
// from clean counts whether or not the host checks the status first.
__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* __restrict__ in, uint32_t n_host,
                                                    const DevFrame* fr,
                                                    const uint32_t* __restrict__ bsum,
                                                    uint32_t* __restrict__ out, SortLists lists,
                                                    uint32_t* clear)
{
    __shared__ uint32_t lds[8];
    __shared__ uint32_t lds2[8];
    __shared__ uint32_t list_base[2];
    uint32_t n;
    if (!scan_len(fr, n_host, n))
        return;
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= n)
        return;
    const uint32_t i0 = base + threadIdx.x * 16u;
    // A tile whose references sum to zero (bsum holds exclusive tile prefixes,
    // out[n] the total) has only empty cells: every offset is the tile's
    // prefix, and its counts need not be read.
    const uint32_t pre = bsum[blockIdx.x];
    const uint32_t next = base + kScanTile < n ? bsum[blockIdx.x + 1] : out[n];
    if (next == pre) {
        if (base + kScanTile <= n) {
            // Lane-contiguous 16-byte stores: each store of a wave covers 1 KB.
            uint4* q = reinterpret_cast<uint4*>(out + base) + threadIdx.x;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                stream_store(q + 256 * k, make_uint4(pre, pre, pre, pre));
        } else {
            for (uint32_t k = 0; k < 16; ++k)
                if (i0 + k < n)
                    out[i0 + k] = pre;
        }
        return;
    }
    const bool whole = base + kScanTile <= n;
    if (clear && (fr->status & kStatusRefsOver)) {
        // Tiles with a zero sum (above) hold zero counts already.
        if (whole) {
            uint4* q = reinterpret_cast<uint4*>(clear + base) + threadIdx.x;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                q[256 * k] = make_uint4(0u, 0u, 0u, 0u);
        } else {
            for (uint32_t k = 0; k < 16; ++k)
                if (i0 + k < n)
                    clear[i0 + k] = 0u;
        }
        return;
    }
    // Whole tiles move through LDS so that global loads and stores are
    // lane-contiguous (each 16-byte access of a wave covers 1 KB); thread t
    // then scans cells [16t, 16t + 16) of the tile from LDS.
    __shared__ uint4 tile[kScanTile / 4];
    uint32_t v[16];
    if (whole) {
        const uint4* src = reinterpret_cast<const uint4*>(in + base);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            tile[256 * k + threadIdx.x] = src[256 * k + threadIdx.x];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const uint4 x = tile[4 * threadIdx.x + k];
            v[4 * k] = x.x;
            v[4 * k + 1] = x.y;
            v[4 * k + 2] = x.z;
            v[4 * k + 3] = x.w;
        }
    } else {
        load16(in, i0, n, v);
    }
    uint32_t s = 0;
#pragma unroll
    for (int k = 0; k < 16; ++k)
        s += v[k];
    if (lists.small && !(fr->status & kStatusRefsOver)) {
        uint32_t ns = 0, nl = 0;
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            ns += (v[k] >= 2u && v[k] <= kSortSmallCells) ? 1u : 0u;
            nl += (v[k] > kSortSmallCells) ? 1u : 0u;
        }
        uint32_t ts, tl;
        uint32_t ps = block_excl_scan<256>(ns, lds, ts);
        uint32_t pl = block_excl_scan<256>(nl, lds2, tl);
        if (threadIdx.x == 0) {
            // both list lengths in one 64-bit atomic: every block adds to
            // the same counters, and such atomics serialise (each length
            // stays below 2^32, so the low half never carries)
            const unsigned long long old =
                (ts | tl) ? atomicAdd(reinterpret_cast<unsigned long long*>(lists.counts),
                                      (unsigned long long)ts | ((unsigned long long)tl << 32))
                          : 0ull;
            list_base[0] = (uint32_t)old;
            list_base[1] = (uint32_t)(old >> 32);
        }
        __syncthreads();
        ps += list_base[0];
        pl += list_base[1];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (v[k] >= 2u && v[k] <= kSortSmallCells) {
                if (ps < lists.small_cap)
                    lists.small[ps] = i0 + (uint32_t)k;
                ++ps;
            } else if (v[k] > kSortSmallCells) {
                if (pl < lists.large_cap)
                    lists.large[pl] = i0 + (uint32_t)k;
                ++pl;
            }
        }
    }
    uint32_t total;
    uint32_t p = block_excl_scan<256>(s, lds, total) + bsum[blockIdx.x];
    uint32_t o[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
        o[k] = p;
        p += v[k];
    }
    if (whole) {
#pragma unroll
        for (int k = 0; k < 4; ++k)
            tile[4 * threadIdx.x + k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
        __syncthreads();
        uint4* dst = reinterpret_cast<uint4*>(out + base);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            stream_store(dst + 256 * k + threadIdx.x, tile[256 * k + threadIdx.x]);
    } else if (i0 + 16 <= n) {
        uint4* q = reinterpret_cast<uint4*>(out + i0);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            q[k] = make_uint4(o[4 * k], o[4 * k + 1], o[4 * k + 2], o[4 * k + 3]);
    } else {
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (i0 + k < n)
                out[i0 + k] = o[k];
    }
}

// ---------------------------------------------------------------------------
// Ray-list compaction (the scan applied to ray batches)
// ---------------------------------------------------------------------------
// A block takes kScanTile consecutive slots in 16 rounds of 256, so that the
// lanes of a wave read consecutive slots; the list keeps slot order.

__device__ __forceinline__ bool compact_active(uint32_t slot, uint32_t n, uint32_t img_w,
                                               const float* __restrict__ tmin,
                                               const float* __restrict__ tmax, uint32_t& i,
                                               bool& valid)
{
    i = slot_ray(slot, img_w, n, valid);
    return valid && tmin[i] <= tmax[i];
}

__global__ __launch_bounds__(256) void k_compact_count(uint32_t n, uint32_t img_w,
                                                       uint32_t n_slots,
                                                       const float* __restrict__ tmin,
                                                       const float* __restrict__ tmax,
                                                       uint32_t* __restrict__ bsum)
{
    __shared__ uint32_t lds[8];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t c = 0;
    for (uint32_t r = 0; r < kScanTile / 256; ++r) {
        const uint32_t slot = base + r * 256 + threadIdx.x;
        uint32_t i;
        bool valid;
        c += (slot < n_slots && compact_active(slot, n, img_w, tmin, tmax, i, valid)) ? 1u : 0u;
    }
    uint32_t total;
    block_excl_scan<256>(c, lds, total);
    if (threadIdx.x == 0)
        bsum[blockIdx.x] = total;
}

// Exclusive scan of the nb block counts in place; the total to *count.
__global__ __launch_bounds__(1024) void k_compact_offsets(uint32_t* __restrict__ bsum, uint32_t nb,
                                                          uint32_t* __restrict__ count)
{
    __shared__ uint32_t lds[17];
    uint32_t carry = 0;
    for (uint32_t b0 = 0; b0 < nb; b0 += 1024) {
        const uint32_t b = b0 + threadIdx.x;
        const uint32_t v = b < nb ? bsum[b] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<1024>(v, lds, tot);
        if (b < nb)
            bsum[b] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0)
        count[0] = carry;
}

__global__ __launch_bounds__(256) void k_compact_write(uint32_t n, uint32_t img_w, uint32_t n_slots,
                                                       const float* __restrict__ tmin,
                                                       const float* __restrict__ tmax,
                                                       const uint32_t* __restrict__ bsum,
                                                       uint32_t* __restrict__ list,
                                                       uint8_t* __restrict__ occluded,
                                                       HitSoA miss)
{
    __shared__ uint32_t lds[8];
    const uint32_t base = blockIdx.x * kScanTile;
    uint32_t pos = bsum[blockIdx.x];
    for (uint32_t r = 0; r < kScanTile / 256; ++r) {
        const uint32_t slot = base + r * 256 + threadIdx.x;
        uint32_t i = 0;
        bool valid = false;
        const bool act = slot < n_slots && compact_active(slot, n, img_w, tmin, tmax, i, valid);
        uint32_t total;
        const uint32_t ex = block_excl_scan<256>(act ? 1u : 0u, lds, total);
        if (act) {
            list[pos + ex] = i;
        } else if (valid) {
            if (occluded)
                occluded[i] = 0;
            if (miss.id) {
                miss.id[i] = -1;
                miss.t[i] = INFINITY;
                miss.u[i] = 0.0f;
                miss.v[i] = 0.0f;
            }
        }
        pos += total;
    }
}

// ---------------------------------------------------------------------------
// 5. Per-cell sort
// ---------------------------------------------------------------------------
// Cells of 2..kSortSmallCells references (lists.small), one thread per cell.
__global__ __launch_bounds__(256) void k_sort_small(const DevFrame* __restrict__ fr,
                                                    const uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ refs, SortLists lists)
{
    if (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver))
        return;
    uint32_t nl = lists.counts[0];
    nl = nl < lists.small_cap ? nl : lists.small_cap;
    for (uint32_t li = blockIdx.x * blockDim.x + threadIdx.x; li < nl;
         li += gridDim.x * blockDim.x)
        sort_small_cell(offsets, refs, lists.small[li]);
}



// Split lists.large by size class (see SortLists).  The waves of a block
// reserve their runs in an LDS count per class, and the block then takes one
// global atomic per class it holds: many blocks adding to the same few
// counters serialise, so the fewer such atomics the better.
__global__ __launch_bounds__(256) void k_sort_classify(const DevFrame* __restrict__ fr,
                                                       const uint32_t* __restrict__ offsets,
                                                       SortLists lists)
{
    __shared__ uint32_t block_cnt[kSortClasses];
    __shared__ uint32_t block_base[kSortClasses];
    if (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver))
        return;
    uint32_t cnt = lists.counts[1];
    cnt = cnt < lists.large_cap ? cnt : lists.large_cap;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * blockDim.x;
    // the loop runs over whole blocks (li0 is the block's first entry), so
    // every thread takes part in the ballots and barriers
    for (uint32_t li0 = blockIdx.x * blockDim.x; li0 < cnt; li0 += stride) {
        const uint32_t li = li0 + threadIdx.x;
        int cls = -1;
        uint32_t c = 0;
        if (li < cnt) {
            c = lists.large[li];
            const uint32_t n = offsets[c + 1] - offsets[c];
            cls = n <= 64u ? 0 : n <= 256u ? 1 : n <= 1024u ? 2 : n <= kSortWaveMax ? 3 : 4;
        }
        if (threadIdx.x < (uint32_t)kSortClasses)
            block_cnt[threadIdx.x] = 0u;
        __syncthreads();
        uint32_t at = 0;
        uint64_t mine = 0;
#pragma unroll
        for (int q = 0; q < kSortClasses; ++q) {
            const uint64_t m = __ballot(cls == q);
            if (m == 0)
                continue;
            const uint32_t leader = (uint32_t)__builtin_ctzll(m);
            uint32_t w = 0;
            if (lane == leader)
                w = atomicAdd(block_cnt + q, (uint32_t)__popcll(m));
            w = (uint32_t)__builtin_amdgcn_readlane((int)w, (int)leader);
            if (cls == q) {
                at = w;
                mine = m;
            }
        }
        __syncthreads();
        if (threadIdx.x < (uint32_t)kSortClasses) {
            const uint32_t k = block_cnt[threadIdx.x];
            block_base[threadIdx.x] = k ? atomicAdd(lists.counts + 2 + threadIdx.x, k) : 0u;
        }
        __syncthreads();
        if (cls >= 0)
            lists.cls[(size_t)cls * lists.large_cap + block_base[cls] + at + lanes_below(mine)] = c;
    }
}

// Cells of size class Q (E = 4^Q: 16E < n <= 64E references), one wave each.
template <int E, int Q>
__global__ __launch_bounds__(256) void k_sort_wave(const DevFrame* __restrict__ fr,
                                                   const uint32_t* __restrict__ offsets,
                                                   uint32_t* __restrict__ refs, SortLists lists)
{
    if (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver))
        return;
    uint32_t cnt = lists.counts[2 + Q];
    cnt = cnt < lists.large_cap ? cnt : lists.large_cap;
    const uint32_t* __restrict__ list = lists.cls + (size_t)Q * lists.large_cap;
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t li = wave; li < cnt; li += n_waves) {
        const uint32_t c = list[li];
        const uint32_t b = offsets[c];
        wave_sort<E>(refs + b, offsets[c + 1] - b);
    }
}

// Cells of 1025..kSortWaveMax references (size class 3), one 256-thread
// block each: four register-sorted runs merged through LDS.
__global__ __launch_bounds__(256) void k_sort_block(const DevFrame* __restrict__ fr,
                                                    const uint32_t* __restrict__ offsets,
                                                    uint32_t* __restrict__ refs, SortLists lists)
{
    __shared__ uint32_t lds[2 * kSortWaveMax];
    if (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver))
        return;
    uint32_t cnt = lists.counts[2 + 3];
    cnt = cnt < lists.large_cap ? cnt : lists.large_cap;
    const uint32_t* __restrict__ list = lists.cls + (size_t)3 * lists.large_cap;
    for (uint32_t li = blockIdx.x; li < cnt; li += gridDim.x) {
        const uint32_t c = list[li];
        const uint32_t b = offsets[c];
        block_merge_sort<256>(refs + b, offsets[c + 1] - b, lds, lds + kSortWaveMax);
    }
}

// Cells of more than kSortWaveMax references (size class 4), one 1024-thread
// block each: up to kSortTile by sixteen register-sorted runs merged through
// LDS, longer ones by the tiled bitonic sort.
__global__ __launch_bounds__(1024) void k_sort_large(const DevFrame* __restrict__ fr,
                                                     const uint32_t* __restrict__ offsets,
                                                     uint32_t* __restrict__ refs, SortLists lists)
{
    __shared__ uint32_t lds[2 * kSortTile];
    if (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver))
        return;
    uint32_t cnt = lists.counts[2 + 4];
    cnt = cnt < lists.large_cap ? cnt : lists.large_cap;
    const uint32_t* __restrict__ list = lists.cls + (size_t)4 * lists.large_cap;
    for (uint32_t li = blockIdx.x; li < cnt; li += gridDim.x) {
        const uint32_t c = list[li];
        const uint32_t b = offsets[c];
        const uint32_t n = offsets[c + 1] - b;
        if (n <= kSortTile)
            block_merge_sort<1024>(refs + b, n, lds, lds + kSortTile);
        else
            block_sort_cell(refs + b, n, lds);
    }
}

// ---------------------------------------------------------------------------
// Triangle packing and per-frame rigid transform
// ---------------------------------------------------------------------------
__device__ __forceinline__ void pack_one(uint32_t t, const float v[9], float4* __restrict__ tri)
{
    const float e1x = v[3] - v[0], e1y = v[4] - v[1], e1z = v[5] - v[2];
    const float e2x = v[6] - v[0], e2y = v[7] - v[1], e2z = v[8] - v[2];
    tri[3ull * t + 0] = make_float4(v[0], v[1], v[2], e1x);
    tri[3ull * t + 1] = make_float4(e1y, e1z, e2x, e2y);
    tri[3ull * t + 2] = make_float4(e2z, __uint_as_float(t), 0.0f, 0.0f);
}

__global__ __launch_bounds__(256) void k_pack(const float* __restrict__ verts, uint32_t n_tris,
                                              float4* __restrict__ tri)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tris)
        return;
    float v[9];
#pragma unroll
    for (int k = 0; k < 9; ++k)
        v[k] = verts[9ull * t + k];
    pack_one(t, v, tri);
}

__global__ __launch_bounds__(256) void k_transform_pack(uint32_t n_tris, uint32_t tpc, uint32_t nct,
                                                        const float* __restrict__ xf,
                                                        const float* __restrict__ rest_v,
                                                        const float* __restrict__ rest_n,
                                                        float* __restrict__ verts,
                                                        float* __restrict__ normals,
                                                        float4* __restrict__ tri)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_tris)
        return;
    float v[9], nn[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        v[k] = rest_v[9ull * t + k];
        nn[k] = rest_n[9ull * t + k];
    }
    if (t < nct) {
        const float* m = xf + 12ull * (t / tpc);
        float mm[12];
#pragma unroll
        for (int k = 0; k < 12; ++k)
            mm[k] = m[k];
        float ov[9], on[9];
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                ov[3 * k + r] = __builtin_fmaf(mm[4 * r], v[3 * k],
                                __builtin_fmaf(mm[4 * r + 1], v[3 * k + 1],
                                __builtin_fmaf(mm[4 * r + 2], v[3 * k + 2], mm[4 * r + 3])));
                on[3 * k + r] = __builtin_fmaf(mm[4 * r], nn[3 * k],
                                __builtin_fmaf(mm[4 * r + 1], nn[3 * k + 1],
                                mm[4 * r + 2] * nn[3 * k + 2]));
            }
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            v[k] = ov[k];
            nn[k] = on[k];
        }
    }
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        verts[9ull * t + k] = v[k];
        normals[9ull * t + k] = nn[k];
    }
    pack_one(t, v, tri);
}

// Per-reference triangle records: trir[k] = tri[refs[k]], so a cell's
// triangles are contiguous and traversal needs no index indirection.
__global__ __launch_bounds__(256) void k_gather_refs(const DevFrame* __restrict__ fr,
                                                     const uint32_t* __restrict__ refs,
                                                     const float4* __restrict__ tri,
                                                     float4* __restrict__ trir)
{
    // Each wave takes 64 consecutive references at a time: lane l gathers the
    // record of reference k0 + l into the wave's LDS slice, then the wave
    // writes the 64 records (3 KB, contiguous in trir) with lane-contiguous
    // 16-byte stores.
    __shared__ float4 stage[4][192];
    if (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver))
        return;
    const uint64_t r_total = fr->r;
    const uint32_t lane = threadIdx.x & 63u;
    float4* st = stage[threadIdx.x >> 6];
    const uint64_t stride = (uint64_t)gridDim.x * (blockDim.x >> 6) * 64u;
    for (uint64_t k0 = (blockIdx.x * (uint64_t)(blockDim.x >> 6) + (threadIdx.x >> 6)) * 64u;
         k0 < r_total; k0 += stride) {
        const uint64_t k = k0 + lane;
        if (k < r_total) {
            const uint64_t t = refs[k];
            st[3 * lane] = tri[3 * t];
            st[3 * lane + 1] = tri[3 * t + 1];
            st[3 * lane + 2] = tri[3 * t + 2];
        }
        wave_lds_sync();
        const uint32_t m = 3u * (uint32_t)(r_total - k0 < 64u ? r_total - k0 : 64u);
        for (uint32_t j = lane; j < m; j += 64u)
            stream_store(trir + 3 * k0 + j, st[j]);
        wave_lds_sync();
    }
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t value, uint64_t n)
{
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static uint32_t grid_for(uint64_t n, uint32_t block, uint32_t cap = 0x7fffffffu)
{
    uint64_t g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    return (uint32_t)(g < cap ? g : cap);
}

hipError_t launch_frame_from_verts(const float* verts, uint32_t n_tris, float lambda,
                                   uint32_t cells_cap, float* partial, DevFrame* fr,
                                   hipStream_t s)
{
    const uint32_t nb = grid_for(n_tris, 256, kBBoxMaxBlocks);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, s, verts, n_tris, partial);
    hipLaunchKernelGGL(k_frame, dim3(1), dim3(64), 0, s, (const float*)partial, nb, n_tris, lambda,
                       cells_cap, fr);
    return hipGetLastError();
}

hipError_t launch_frame_from_box(const float bmin[3], const float bmax[3], uint32_t n_tris,
                                 float lambda, uint32_t cells_cap, DevFrame* fr, hipStream_t s)
{
    hipLaunchKernelGGL(k_frame_box, dim3(1), dim3(64), 0, s, bmin[0], bmin[1], bmin[2], bmax[0],
                       bmax[1], bmax[2], n_tris, lambda, cells_cap, fr);
    return hipGetLastError();
}

hipError_t launch_grid_count(const DevFrame* fr, const float* verts, uint32_t n_tris,
                             uint32_t* counts, uint32_t* big_list, uint32_t* big_count,
                             uint32_t* tile_sums, uint32_t cells_cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_zero_build_counters, dim3(1), dim3(64), 0, s, big_count);
    const uint32_t nb = grid_for(cells_cap, kScanTile);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(nb, 256, 1024)), dim3(256), 0, s, tile_sums, 0u,
                       (uint64_t)nb);
    hipLaunchKernelGGL(k_grid_pass<false>, dim3(grid_for(n_tris, 256)), dim3(256), 0, s, fr, verts,
                       n_tris, counts, (const uint32_t*)nullptr, (uint32_t*)nullptr, big_list,
                       big_count, tile_sums);
    hipLaunchKernelGGL(k_grid_pass_big<false>, dim3(2048), dim3(256), 0, s, fr, verts, counts,
                       (const uint32_t*)nullptr, (uint32_t*)nullptr, (const uint32_t*)big_list,
                       (const uint32_t*)big_count, tile_sums);
    hipLaunchKernelGGL(k_grid_pass_huge<false>, dim3(2048), dim3(256), 0, s, fr, verts, n_tris,
                       counts, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                       (const uint32_t*)big_list, (const uint32_t*)big_count, tile_sums);
    return hipGetLastError();
}

hipError_t launch_grid_scan(DevFrame* fr, uint32_t* counts, uint32_t* offsets,
                            uint32_t* block_sums, uint32_t cells_cap, uint32_t refs_cap,
                            const SortLists& lists, bool prefixes_ready, hipStream_t s)
{
    const uint32_t nb = grid_for(cells_cap, kScanTile);
    if (lists.small)
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, lists.counts, 0u,
                           (uint64_t)(2 + kSortClasses));
    // block_sums holds the count pass's tile sums: scan them into tile
    // offsets (in place), unless an earlier scan of the same counts did.
    if (!prefixes_ready)
        hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, block_sums, 0u, fr, offsets,
                           refs_cap);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, (const uint32_t*)counts, 0u,
                       (const DevFrame*)fr, (const uint32_t*)block_sums, offsets, lists, counts);
    return hipGetLastError();
}

hipError_t launch_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* block_sums,
                           DevFrame* tmp_frame, hipStream_t s)
{
    (void)tmp_frame;
    const uint32_t nb = grid_for(n, kScanTile);
    hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(256), 0, s, in, n, (const DevFrame*)nullptr,
                       block_sums);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, block_sums, n, (DevFrame*)nullptr,
                       out, 0xffffffffu);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(256), 0, s, in, n, (const DevFrame*)nullptr,
                       (const uint32_t*)block_sums, out, SortLists{}, (uint32_t*)nullptr);
    return hipGetLastError();
}

hipError_t launch_compact_rays(uint32_t n, uint32_t img_w, const float* tmin, const float* tmax,
                               uint32_t* list, uint32_t* count, uint8_t* occluded,
                               const HitSoA& miss_hits, uint32_t* block_sums, hipStream_t s)
{
    const uint64_t slots = slot_count(n, img_w);
    const uint32_t nb = (uint32_t)((slots + kScanTile - 1) / kScanTile);
    if (nb > 0) {
        hipLaunchKernelGGL(k_compact_count, dim3(nb), dim3(256), 0, s, n, img_w, (uint32_t)slots,
                           tmin, tmax, block_sums);
    }
    hipLaunchKernelGGL(k_compact_offsets, dim3(1), dim3(1024), 0, s, block_sums, nb, count);
    if (nb > 0) {
        hipLaunchKernelGGL(k_compact_write, dim3(nb), dim3(256), 0, s, n, img_w, (uint32_t)slots,
                           tmin, tmax, (const uint32_t*)block_sums, list, occluded, miss_hits);
    }
    return hipGetLastError();
}

hipError_t launch_grid_scatter(const DevFrame* fr, const float* verts, uint32_t n_tris,
                               uint32_t* counts, const uint32_t* offsets, uint32_t* refs,
                               const uint32_t* big_list, const uint32_t* big_count,
                               hipStream_t s)
{
    hipLaunchKernelGGL(k_grid_pass<true>, dim3(grid_for(n_tris, 256)), dim3(256), 0, s, fr, verts,
                       n_tris, counts, offsets, refs, (uint32_t*)nullptr, (uint32_t*)nullptr,
                       (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_grid_pass_big<true>, dim3(2048), dim3(256), 0, s, fr, verts, counts, offsets,
                       refs, big_list, big_count, (uint32_t*)nullptr);
    hipLaunchKernelGGL(k_grid_pass_huge<true>, dim3(2048), dim3(256), 0, s, fr, verts, n_tris,
                       counts, offsets, refs, big_list, big_count, (uint32_t*)nullptr);
    return hipGetLastError();
}

hipError_t launch_grid_sort(const DevFrame* fr, const uint32_t* offsets, uint32_t* refs,
                            const SortLists& lists, hipStream_t s)
{
    hipLaunchKernelGGL(k_sort_small, dim3(2048), dim3(256), 0, s, fr, offsets, refs, lists);
    hipLaunchKernelGGL(k_sort_classify, dim3(512), dim3(256), 0, s, fr, offsets, lists);
    hipLaunchKernelGGL((k_sort_wave<1, 0>), dim3(2048), dim3(256), 0, s, fr, offsets, refs, lists);
    hipLaunchKernelGGL((k_sort_wave<4, 1>), dim3(2048), dim3(256), 0, s, fr, offsets, refs, lists);
    hipLaunchKernelGGL((k_sort_wave<16, 2>), dim3(1024), dim3(256), 0, s, fr, offsets, refs, lists);
    hipLaunchKernelGGL(k_sort_block, dim3(1024), dim3(256), 0, s, fr, offsets, refs, lists);
    hipLaunchKernelGGL(k_sort_large, dim3(256), dim3(1024), 0, s, fr, offsets, refs, lists);
    return hipGetLastError();
}

hipError_t launch_gather_refs(const DevFrame* fr, const uint32_t* refs, const float4* tri,
                              float4* trir, uint32_t refs_cap, hipStream_t s)
{
    hipLaunchKernelGGL(k_gather_refs, dim3(grid_for(refs_cap, 256, 16384)), dim3(256), 0, s, fr,
                       refs, tri, trir);
    return hipGetLastError();
}

hipError_t launch_pack(const float* verts, uint32_t n_tris, float4* tri, hipStream_t s)
{
    hipLaunchKernelGGL(k_pack, dim3(grid_for(n_tris, 256)), dim3(256), 0, s, verts, n_tris, tri);
    return hipGetLastError();
}

hipError_t launch_transform_pack(uint32_t n_tris, uint32_t tris_per_cluster,
                                 uint32_t n_cluster_tris, const float* xf,
                                 const float* rest_v, const float* rest_n, float* verts,
                                 float* normals, float4* tri, hipStream_t s)
{
    hipLaunchKernelGGL(k_transform_pack, dim3(grid_for(n_tris, 256)), dim3(256), 0, s, n_tris,
                       tris_per_cluster, n_cluster_tris, xf, rest_v, rest_n, verts, normals, tri);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t* p, uint32_t value, uint64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, p, value, n);
    return hipGetLastError();
}

}  // namespace grt
