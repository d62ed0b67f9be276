// grt_grid2.hip -- the second level of the two-level grid on gfx950: a uniform
// sub-grid inside every dense top cell, rebuilt on the device after the top
// grid of the same frame.
//
// A top cell holding n >= min_refs references gets r0 x r1 x r2 sub-cells
// spanning exactly the cell, from the top frame's own rule applied to the cell
// (x = lambda2 n / cell volume, f = cbrt(x), r_a = ceil(cell_a f) in
// [1, kSubMax]); no sub-grid when r = (1,1,1).  Its frame is bmin_a =
// (float)c_a cell_a + bmin_a (the DDA's boundary formula), cell'_a = cell_a /
// r_a, inv'_a = 1 / cell'_a.  A reference goes into every sub-cell of its
// triangle's AABB box in that frame (clamp, truncate), each sub-cell listing
// ascending triangle ids.  The top grid stays complete, so every path that
// ignores the second level still traces correctly.
//
// Work layout: the plan of every top cell comes from 4096-cell tiles (16
// consecutive cells per thread); the dense cells -- few, and clustered where
// the scene is -- then get one 256-thread block each.  In such a block thread
// j owns sub-cell row j (y = j % r1, z = j / r1; r1 r2 <= 256 rows) and keeps
// the row's counters or write cursors in registers, while the cell's
// references stream through LDS in ascending order, 1024 at a time, as packed
// 4-bit sub-cell ranges.  A thread appends a reference to its own rows only,
// in stream order, so every sub-cell's list comes out ascending without
// atomics or a sort.
//   k_g2_plan<false>  per tile: sub-cells and dense cells (tile sums)
//   k_g2_scan_tiles   exclusive scan of both; totals, capacity
//   k_g2_plan<true>   per tile: cell records, the ordered dense-cell list
//   k_g2_cells<false> per dense cell: its sub-references
//   k_g2_scan_dense   exclusive scan of those; total, capacity
//   k_g2_cells<true>  per dense cell: sub-cell offsets and references
//   k_gather_refs     48-byte record of every sub-reference
#include "grt_device.h"
#include "grt_launch.h"

namespace grt {

constexpr uint32_t kG2Chunk = 1024;   // references staged in LDS per round

// Packed resolution r0 | r1 << 8 | r2 << 16 of a top cell with n references
// (0: no sub-grid) and its sub-cell count m.
__device__ __forceinline__ uint32_t g2_plan(const DevFrame* fr, uint32_t n, float lambda2,
                                            uint32_t min_refs, uint32_t& m)
{
    m = 0;
    if (n == 0 || n < min_refs)
        return 0u;
    const float vol = (fr->cell[0] * fr->cell[1]) * fr->cell[2];
    const float fac = spec_cbrtf((lambda2 * (float)n) / vol);
    uint32_t r[3], mm = 1;
    for (int a = 0; a < 3; ++a) {
        float x = ceilf(fr->cell[a] * fac);
        x = x < 1.0f ? 1.0f : x;
        x = x > (float)kSubMax ? (float)kSubMax : x;
        r[a] = (uint32_t)x;
        mm *= r[a];
    }
    if (mm <= 1)
        return 0u;
    m = mm;
    return r[0] | (r[1] << 8) | (r[2] << 16);
}

__device__ __forceinline__ uint32_t g2_vox(float val, float bmin, float inv, uint32_t r)
{
    float x = (val - bmin) * inv;
    x = x > 0.0f ? x : 0.0f;
    const float hi = (float)(r - 1);
    x = x < hi ? x : hi;
    return (uint32_t)x;
}

// One dense cell's references, streamed through LDS in order: COUNT adds to the
// owned row's counters, else appends to the owned row's cursors.  Ranges are
// read 4 references per 16-byte LDS load (the chunk's last nk % 4 one by one).
// A range has lo <= hi on every axis, so the unsigned tests x - lo <= hi - lo
// are exact.
template <bool COUNT>
__device__ __forceinline__ void g2_ref(uint32_t p, uint32_t t, uint32_t y, uint32_t z,
                                       uint32_t (&cnt)[kSubMax], uint32_t* sub_refs)
{
    const uint32_t ylo = (p >> 8) & 15u, yhi = (p >> 12) & 15u;
    const uint32_t zlo = (p >> 16) & 15u, zhi = (p >> 20) & 15u;
    if (y - ylo <= yhi - ylo && z - zlo <= zhi - zlo) {
        const uint32_t xlo = p & 15u, xw = ((p >> 4) & 15u) - xlo;
#pragma unroll
        for (uint32_t x = 0; x < kSubMax; ++x) {
            const bool in = x - xlo <= xw;
            if (COUNT) {
                cnt[x] += in ? 1u : 0u;
            } else if (in) {
                sub_refs[cnt[x]] = t;
                ++cnt[x];
            }
        }
    }
}

template <bool COUNT>
__device__ __forceinline__ void g2_stream(const G2Build& b, uint32_t kb, uint32_t ke,
                                         const float sb[3], const float si[3], const uint32_t r[3],
                                         bool owner, uint32_t y, uint32_t z,
                                         uint32_t (&cnt)[kSubMax], uint32_t* s_rng,
                                         uint32_t* s_id)
{
    for (uint32_t k0 = kb; k0 < ke; k0 += kG2Chunk) {
        const uint32_t nk = ke - k0 < kG2Chunk ? ke - k0 : kG2Chunk;
        for (uint32_t i = threadIdx.x; i < nk; i += blockDim.x) {
            const uint32_t t = b.refs[k0 + i];
            const float* v = b.verts + 9ull * t;
            uint32_t p = 0;
            for (int a = 0; a < 3; ++a) {
                const float x = v[a], yv = v[3 + a], zv = v[6 + a];
                float mn = x, mx = x;
                mn = yv < mn ? yv : mn;
                mx = yv > mx ? yv : mx;
                mn = zv < mn ? zv : mn;
                mx = zv > mx ? zv : mx;
                const uint32_t lo = g2_vox(mn, sb[a], si[a], r[a]);
                const uint32_t hi = g2_vox(mx, sb[a], si[a], r[a]);
                p |= (lo | (hi << 4)) << (8 * a);
            }
            s_rng[i] = p;
            s_id[i] = t;
        }
        __syncthreads();
        if (owner) {
            const uint4* r4 = reinterpret_cast<const uint4*>(s_rng);
            const uint4* t4 = reinterpret_cast<const uint4*>(s_id);
            const uint32_t n4 = nk / 4u;
            for (uint32_t i = 0; i < n4; ++i) {
                const uint4 p = r4[i];
                const uint4 t = t4[i];
                g2_ref<COUNT>(p.x, t.x, y, z, cnt, b.sub_refs);
                g2_ref<COUNT>(p.y, t.y, y, z, cnt, b.sub_refs);
                g2_ref<COUNT>(p.z, t.z, y, z, cnt, b.sub_refs);
                g2_ref<COUNT>(p.w, t.w, y, z, cnt, b.sub_refs);
            }
            for (uint32_t i = 4u * n4; i < nk; ++i)
                g2_ref<COUNT>(s_rng[i], s_id[i], y, z, cnt, b.sub_refs);
        }
        __syncthreads();
    }
}

// Status bits of the second level in g2->status.
__device__ __forceinline__ bool top_ok(const DevFrame* fr)
{
    return (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)) == 0;
}

// 1 (LIST = false): per 4096-cell tile, the sub-cells and dense cells it holds.
// 3 (LIST = true): per tile, every cell's record {refs begin, end, first
// sub-cell, plan} and the tile's dense cells into the global dense list, in
// cell order.  16 consecutive cells per thread.
template <bool LIST>
__global__ __launch_bounds__(256) void k_g2_plan(G2Build b)
{
    __shared__ uint32_t s_scan[8];
    const DevFrame* fr = b.fr;
    if (!top_ok(fr))
        return;
    const uint32_t m_top = fr->m;
    const uint32_t base = blockIdx.x * kScanTile;
    if (base >= m_top) {
        if (!LIST && threadIdx.x == 0) {
            b.tile_m[blockIdx.x] = 0u;
            b.tile_d[blockIdx.x] = 0u;
        }
        return;
    }
    const bool sub_ok = !LIST || b.g2->status == 0;
    const uint32_t c0 = base + 16u * threadIdx.x;
    uint32_t off[17];
#pragma unroll
    for (uint32_t k = 0; k < 17; ++k)
        off[k] = c0 + k <= m_top ? b.offsets[c0 + k] : 0u;
    uint32_t packed[16], ms[16], nd = 0, msum = 0;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        ms[k] = 0;
        packed[k] = (c0 + k < m_top && sub_ok)
                        ? g2_plan(fr, off[k + 1] - off[k], b.lambda2, b.min_refs, ms[k])
                        : 0u;
        nd += packed[k] ? 1u : 0u;
        msum += ms[k];
    }
    uint32_t nd_total, m_total;
    uint32_t dpos = block_excl_scan<256>(nd, s_scan, nd_total);
    uint32_t mpos = block_excl_scan<256>(msum, s_scan, m_total);
    if (!LIST) {
        if (threadIdx.x == 0) {
            b.tile_m[blockIdx.x] = m_total;
            b.tile_d[blockIdx.x] = nd_total;
        }
        return;
    }
    mpos += b.tile_m[blockIdx.x];
    dpos += b.tile_d[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        if (c0 + k < m_top) {
            b.cellrec[c0 + k] = make_uint4(off[k], off[k + 1], packed[k] ? mpos : 0u, packed[k]);
            if (packed[k])
                b.dense[dpos++] = c0 + k;
            mpos += ms[k];
        }
    }
}

// 2: exclusive scan of the tiles' sub-cell and dense-cell sums (in place);
// totals into g2 (m sub-cells, n[0] dense cells) and their capacity bits.
__global__ __launch_bounds__(1024) void k_g2_scan_tiles(G2Build b, uint32_t n_tiles,
                                                        uint32_t sub_cells_cap, uint32_t dense_cap)
{
    __shared__ uint32_t lds[17];
    const DevFrame* fr = b.fr;
    DevFrame* g2 = b.g2;
    if (!top_ok(fr)) {
        if (threadIdx.x == 0) {
            g2->m = 0;
            g2->r = 0;
            g2->n[0] = 0;
            g2->status = kStatusFrameInvalid;
        }
        return;
    }
    const uint32_t used = (fr->m + kScanTile - 1) / kScanTile;
    const uint32_t nt = used < n_tiles ? used : n_tiles;
    uint64_t cm = 0, cd = 0;
    for (uint32_t t0 = 0; t0 < nt; t0 += 1024) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t vm = t < nt ? b.tile_m[t] : 0u;
        const uint32_t vd = t < nt ? b.tile_d[t] : 0u;
        uint32_t tm, td;
        const uint32_t em = block_excl_scan<1024>(vm, lds, tm);
        const uint32_t ed = block_excl_scan<1024>(vd, lds, td);
        if (t < nt) {
            b.tile_m[t] = (uint32_t)cm + em;
            b.tile_d[t] = (uint32_t)cd + ed;
        }
        cm += tm;
        cd += td;
    }
    if (threadIdx.x == 0) {
        uint32_t status = 0;
        if (cm > (uint64_t)sub_cells_cap || cd > (uint64_t)dense_cap)
            status |= kStatusCellsOver;
        g2->m = (uint32_t)(cm > 0xffffffffull ? 0xffffffffull : cm);
        g2->n[0] = (int32_t)(cd > 0x7fffffffull ? 0x7fffffffull : cd);
        g2->r = 0;
        g2->status = status;
    }
}

// Frame of dense cell c's sub-grid and this thread's row.
struct G2Cell {
    uint32_t kb, ke, pk, m;
    uint32_t r[3];
    float sb[3], si[3];
    bool owner;
    uint32_t y, z;
};

__device__ __forceinline__ void g2_cell(const G2Build& b, uint32_t c, G2Cell& g)
{
    const DevFrame* fr = b.fr;
    g.kb = b.offsets[c];
    g.ke = b.offsets[c + 1];
    g.pk = g2_plan(fr, g.ke - g.kb, b.lambda2, b.min_refs, g.m);
    g.r[0] = g.pk & 255u;
    g.r[1] = (g.pk >> 8) & 255u;
    g.r[2] = (g.pk >> 16) & 255u;
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    const uint32_t cc[3] = {c % nx, (c / nx) % ny, c / (nx * ny)};
    for (int a = 0; a < 3; ++a) {
        g.sb[a] = (float)cc[a] * fr->cell[a] + fr->bmin[a];
        const float sc = fr->cell[a] / (float)g.r[a];
        g.si[a] = 1.0f / sc;
    }
    const uint32_t row = threadIdx.x;
    g.owner = row < g.r[1] * g.r[2];
    g.y = g.r[1] ? row % g.r[1] : 0u;
    g.z = g.r[1] ? row / g.r[1] : 0u;
}

// 4 (FILL = false): per dense cell (one block each, grid-strided), its
// sub-references into dense_refs.  6 (FILL = true): per dense cell, the
// sub-cell offsets and the references, ascending per sub-cell.
template <bool FILL>
__global__ __launch_bounds__(256) void k_g2_cells(G2Build b)
{
    __shared__ __align__(16) uint32_t s_rng[kG2Chunk];
    __shared__ __align__(16) uint32_t s_id[kG2Chunk];
    __shared__ uint32_t s_scan[8];
    if (!top_ok(b.fr) || (b.g2->status & ~(FILL ? 0u : kStatusRefsOver)) != 0)
        return;
    const uint32_t nd = (uint32_t)b.g2->n[0];
    for (uint32_t d = blockIdx.x; d < nd; d += gridDim.x) {
        const uint32_t c = b.dense[d];
        G2Cell g;
        g2_cell(b, c, g);
        const uint32_t sub_base = b.cellrec[c].z;
        uint32_t* row_off = b.sub_offsets ? b.sub_offsets + sub_base + g.r[0] * threadIdx.x
                                          : nullptr;
        uint32_t cnt[kSubMax];
        if (!FILL) {
            // The counts go into the row's offset slots for the fill pass.
#pragma unroll
            for (uint32_t x = 0; x < kSubMax; ++x)
                cnt[x] = 0;
            g2_stream<true>(b, g.kb, g.ke, g.sb, g.si, g.r, g.owner, g.y, g.z, cnt, s_rng, s_id);
        } else {
#pragma unroll
            for (uint32_t x = 0; x < kSubMax; ++x)
                cnt[x] = (g.owner && x < g.r[0]) ? row_off[x] : 0u;
        }
        uint32_t tot = 0;
#pragma unroll
        for (uint32_t x = 0; x < kSubMax; ++x)
            tot += cnt[x];
        uint32_t cell_total;
        const uint32_t row_pre = block_excl_scan<256>(tot, s_scan, cell_total);
        if (!FILL) {
            if (threadIdx.x == 0)
                b.dense_refs[d] = cell_total;
            if (row_off && g.owner) {
#pragma unroll
                for (uint32_t x = 0; x < kSubMax; ++x)
                    if (x < g.r[0])
                        row_off[x] = cnt[x];
            }
            continue;
        }
        uint32_t p = b.dense_refs[d] + row_pre;
#pragma unroll
        for (uint32_t x = 0; x < kSubMax; ++x) {
            const uint32_t n = cnt[x];
            cnt[x] = p;
            p += n;
            if (g.owner && x < g.r[0])
                row_off[x] = cnt[x];
        }
        g2_stream<false>(b, g.kb, g.ke, g.sb, g.si, g.r, g.owner, g.y, g.z, cnt, s_rng, s_id);
    }
}

// 5: exclusive scan of the dense cells' sub-references (in place); the total
// into g2->r, its capacity bit, and the end of the last sub-cell's range.
__global__ __launch_bounds__(1024) void k_g2_scan_dense(G2Build b, uint32_t sub_refs_cap)
{
    __shared__ uint32_t lds[17];
    DevFrame* g2 = b.g2;
    if (!top_ok(b.fr) || g2->status != 0)
        return;
    const uint32_t nd = (uint32_t)g2->n[0];
    uint64_t cr = 0;
    for (uint32_t t0 = 0; t0 < nd; t0 += 1024) {
        const uint32_t t = t0 + threadIdx.x;
        const uint32_t v = t < nd ? b.dense_refs[t] : 0u;
        uint32_t tot;
        const uint32_t e = block_excl_scan<1024>(v, lds, tot);
        if (t < nd)
            b.dense_refs[t] = (uint32_t)cr + e;
        cr += tot;
    }
    if (threadIdx.x == 0) {
        g2->r = (uint32_t)(cr > 0xffffffffull ? 0xffffffffull : cr);
        if (cr > (uint64_t)sub_refs_cap)
            g2->status |= kStatusRefsOver;
        else if (b.sub_offsets)
            b.sub_offsets[g2->m] = (uint32_t)cr;
    }
}

hipError_t launch_grid2_plan(const G2Build& b, uint32_t top_cells_cap, uint32_t sub_cells_cap,
                             uint32_t sub_refs_cap, uint32_t dense_cap, hipStream_t s)
{
    const uint32_t nt = top_cells_cap / kScanTile + 1;
    hipLaunchKernelGGL(k_g2_plan<false>, dim3(nt), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_g2_scan_tiles, dim3(1), dim3(1024), 0, s, b, nt, sub_cells_cap,
                       dense_cap);
    hipLaunchKernelGGL(k_g2_plan<true>, dim3(nt), dim3(256), 0, s, b);
    const uint32_t nb = dense_cap < 4096u ? (dense_cap ? dense_cap : 1u) : 4096u;
    hipLaunchKernelGGL(k_g2_cells<false>, dim3(nb), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_g2_scan_dense, dim3(1), dim3(1024), 0, s, b, sub_refs_cap);
    return hipGetLastError();
}

hipError_t launch_grid2_build(const G2Build& b, uint32_t top_cells_cap, uint32_t sub_cells_cap,
                              uint32_t sub_refs_cap, uint32_t dense_cap, const float4* tri,
                              float4* sub_trir, hipStream_t s)
{
    hipError_t e0 = launch_grid2_plan(b, top_cells_cap, sub_cells_cap, sub_refs_cap, dense_cap, s);
    if (e0 != hipSuccess)
        return e0;
    const uint32_t nb = dense_cap < 4096u ? (dense_cap ? dense_cap : 1u) : 4096u;
    hipLaunchKernelGGL(k_g2_cells<true>, dim3(nb), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    return launch_gather_refs(b.g2, b.sub_refs, tri, sub_trir, sub_refs_cap, s);
}

}  // namespace grt
