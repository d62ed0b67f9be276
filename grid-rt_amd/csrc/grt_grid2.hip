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
// the scene is -- then get one 1024-thread block each, one reference per
// thread: the count pass adds to LDS counters over the reference's sub-cell
// box (and keeps the box, 4 bits per bound, at the reference's position), the
// fill pass scatters through LDS cursors, and every sub-cell's ids are then
// sorted ascending over the whole level -- by a thread, a wave or, past 1024
// ids, a block.
//   k_g2_plan<false>  per tile: sub-cells and dense cells (tile sums); the
//                     x pass of the empty cells' skip radii
//   k_g2_scan_tiles   exclusive scan of both; totals, capacity
//   k_g2_plan<true>   per tile: records of the occupied cells, the ordered
//                     dense-cell list
//   k_g2_skip_line    y and z passes of the skip radii; the z pass writes the
//                     empty cells' records
//   k_g2_count        per dense cell: boxes, sub-cell counts, sub-references
//   k_g2_scan_dense   exclusive scan of those; total, capacity
//   k_g2_fill         per dense cell: sub-cell offsets, references
//   k_g2_sort_thread  per sub-cell: ids ascending (2..16), longer ones listed
//   k_g2_sort_wave    sub-cells of 17..1024 references, a wave each
//   k_g2_sort_big     sub-cells of more than 1024 references
//   k_gather_refs     48-byte record of every sub-reference
#include "grt_device.h"
#include "grt_launch.h"

namespace grt {

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

// Status bits of the second level in g2->status.
__device__ __forceinline__ bool top_ok(const DevFrame* fr)
{
    return (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)) == 0;
}

// Skip radii: the Chebyshev distance D from each top cell to the nearest cell
// holding a reference (cells outside the grid hold none), capped at C =
// kSkipMax + 1, axis by axis -- D_x along the cell's x row (in the first plan
// pass, from the occupancy bits of its tile), then D_xy = min_k max(|k|,
// D_x(y + k)), then D = min_k max(|k|, D_xy(z + k)) -- and an empty cell's
// record is {D - 1, D - 1} (the second plan pass).  The y and z passes stage
// their tile plus a halo of kSkipMax rows per side in LDS; a cell's scan stops
// once |k| reaches the best distance found.  The frame's dimensions are device
// values, so those launches are fixed grids that loop over the tiles.
constexpr uint32_t kSkipHalo = (uint32_t)kSkipMax;
constexpr uint32_t kSkipRows = 64;      // rows per line-pass tile (64 lines wide)

// D_x of the cell at bit p of an occupancy bit array (bit q: the cell p - q
// positions before it; the array holds 7 bits either side), x its column in a
// row of n0 cells: the nearest occupied cell of its row within kSkipMax.
__device__ __forceinline__ uint32_t row_distance(const uint64_t* occ, uint32_t p, uint32_t x,
                                                 uint32_t n0)
{
    constexpr uint32_t C = (uint32_t)kSkipMax + 1u;
    const uint32_t lo = p - kSkipHalo, sh = lo & 63u;
    const uint64_t w0 = occ[lo >> 6], w1 = occ[(lo >> 6) + 1u];
    const uint32_t win = (uint32_t)((sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0) & 0x7fffu);
    if (win & (1u << kSkipHalo))
        return 0u;
    const uint32_t right = n0 - 1u - x;
    uint32_t d = C;
    // below: bit j is the cell 7 - j to the left, within the row
    const uint32_t below = win & (x >= kSkipHalo ? 0x7fu : (0x7fu << (kSkipHalo - x)) & 0x7fu);
    if (below)
        d = kSkipHalo - (31u - (uint32_t)__clz(below));
    // above: bit i is the cell i + 1 to the right, within the row
    const uint32_t above =
        (win >> (kSkipHalo + 1u)) & (right >= kSkipHalo ? 0x7fu : (1u << right) - 1u);
    if (above) {
        const uint32_t ka = (uint32_t)__ffs(above);
        d = ka < d ? ka : d;
    }
    return d;
}

// y (AXIS 1: plane 0 -> plane 1 of skip) or z (AXIS 2: plane 1 -> the records
// of the empty cells) pass over tiles of 64 lines x kSkipRows rows: the lines
// of a tile start at 64 consecutive cells of coordinate 0 on the axis (x of one
// z plane for y; x + n0 y for z).  Distances are stored as thermometer bytes, bit t set when
// D <= t (D = 8: no bit), so that D_out <= t, i.e. some cell within t rows has
// D_in <= t, is the OR over k <= t of bit t of the neighbours at +-k: per row
// out = in | OR_k ((0xff << k) & (in(-k) | in(+k))), four lines per 32-bit
// operation.  Rows outside the grid read 0 (no reference there).  Bit 0 of the
// final byte is the cell's own occupancy (D = 0), so the z pass writes {r, r},
// r = D - 1, for exactly the cells without references, and the records of the
// other cells come from k_g2_plan<true> alone.
template <int AXIS>
__global__ __launch_bounds__(256) void k_g2_skip_line(G2Build b)
{
    __shared__ uint32_t s_in[kSkipRows + 2 * kSkipHalo][16];   // 64 lines of bytes per row
    const DevFrame* fr = b.fr;
    if (!top_ok(fr))
        return;
    const uint32_t n0 = (uint32_t)fr->n[0], n1 = (uint32_t)fr->n[1], n2 = (uint32_t)fr->n[2];
    const uint32_t step = AXIS == 1 ? n0 : n0 * n1;        // cell stride along the axis
    const uint32_t len = AXIS == 1 ? n1 : n2;
    const uint32_t span = AXIS == 1 ? n0 : n0 * n1;        // line starts per plane
    const uint32_t planes = AXIS == 1 ? n2 : 1u;
    // rows per tile: the axis when it is short, else kSkipRows
    const uint32_t rows = len < kSkipRows ? len : kSkipRows;
    const uint32_t gx = (span + 63u) / 64u, segs = (len + rows - 1u) / rows;
    const uint64_t work = (uint64_t)gx * planes * segs;
    const uint8_t* in = AXIS == 1 ? b.skip : b.skip + b.skip_stride;
    uint8_t* out = b.skip + b.skip_stride;
    const uint32_t lane = threadIdx.x & 63u, row0 = threadIdx.x >> 6;
    const uint32_t col = threadIdx.x & 15u, rsub = threadIdx.x >> 4;
    uint8_t* s_bytes = reinterpret_cast<uint8_t*>(&s_in[0][0]);
    for (uint64_t w = blockIdx.x; w < work; w += gridDim.x) {
        const uint32_t seg = (uint32_t)(w % segs);
        const uint64_t g = w / segs;
        const uint32_t a0 = (uint32_t)(g % gx) * 64u;
        const uint64_t plane = (g / gx) * (uint64_t)n0 * n1;
        const int64_t q0 = (int64_t)seg * rows - (int64_t)kSkipHalo;
        {
            const uint32_t a = a0 + lane;
            for (uint32_t r = row0; r < rows + 2u * kSkipHalo; r += 4u) {
                const int64_t q = q0 + (int64_t)r;
                s_bytes[64u * r + lane] = a < span && q >= 0 && q < (int64_t)len
                                              ? in[(uint64_t)a + plane + (uint64_t)q * step]
                                              : (uint8_t)0;
            }
        }
        __syncthreads();
        for (uint32_t r = rsub; r < rows; r += 16u) {
            const uint32_t q = seg * rows + r;
            if (q >= len)
                break;
            const uint32_t j = r + kSkipHalo;
            uint32_t t = s_in[j][col];
#pragma unroll
            for (uint32_t k = 1; k <= kSkipHalo; ++k)
                t |= (((0xffu << k) & 0xffu) * 0x01010101u) & (s_in[j - k][col] | s_in[j + k][col]);
            const uint64_t c = plane + (uint64_t)q * step;
#pragma unroll
            for (uint32_t i = 0; i < 4u; ++i) {
                const uint32_t a = a0 + 4u * col + i;
                const uint32_t v = (t >> (8u * i)) & 0xffu;
                if (a >= span)
                    continue;
                if (AXIS == 1) {
                    out[c + a] = (uint8_t)v;
                } else if (!(v & 1u)) {
                    // skip radius D - 1, D the lowest bit of the byte (8: none)
                    const uint32_t r = (v ? (uint32_t)__ffs(v) - 1u : (uint32_t)kSkipMax + 1u) - 1u;
                    b.cellrec[c + a] = make_uint2(r, r);
                }
            }
        }
        __syncthreads();
    }
}

// 1 (LIST = false): per 4096-cell tile, the sub-cells and dense cells it holds.
// 3 (LIST = true): per tile, the 8-byte record of every cell with references
// -- {0x80000000 | first sub-cell, plan} for a cell with a sub-grid, else
// {refs begin, end} -- and the tile's dense cells into the global dense list,
// in cell order.  Slot k of thread j is cell base + 256 k + j, so loads and
// record stores are coalesced; in-tile positions come from a wave scan per slot
// (sub-cells in bits 0..23, the dense flag in bits 24..31: a wave's sums stay
// below 2^18 and 64) and a scan of the 64 (slot, wave) totals in slot-major
// order.
template <bool LIST>
__global__ __launch_bounds__(256) void k_g2_plan(G2Build b)
{
    __shared__ uint32_t s_tot[2][64];
    __shared__ uint32_t s_scan[8];
    __shared__ uint64_t s_occ[kScanTile / 64u + 2u];   // occupancy, 64 cells either side
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
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    uint32_t packed[16], ms[16], nd = 0, msum = 0, empty = 0;
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        const uint32_t c = base + 256u * k + threadIdx.x;
        const uint32_t n = c < m_top ? b.offsets[c + 1] - b.offsets[c] : 0u;
        ms[k] = 0;
        packed[k] = sub_ok ? g2_plan(fr, n, b.lambda2, b.min_refs, ms[k]) : 0u;
        empty |= (c < m_top && n == 0u ? 1u : 0u) << k;
        nd += packed[k] ? 1u : 0u;
        msum += ms[k];
        if (!LIST) {
            const uint64_t w = __ballot(n != 0u);   // cells base + 64 (4 k + wave) ...
            if (lane == 0u)
                s_occ[1u + 4u * k + wave] = w;
        }
    }
    if (!LIST) {
        // D_x of the tile's cells: the occupancy of 64 cells either side too
        if (wave < 2u) {
            const int64_t c = wave == 0u ? (int64_t)base - 64 + (int64_t)lane
                                         : (int64_t)base + kScanTile + lane;
            const bool occ = c >= 0 && c < (int64_t)m_top && b.offsets[c + 1] != b.offsets[c];
            const uint64_t w = __ballot(occ);
            if (lane == 0u)
                s_occ[wave == 0u ? 0u : kScanTile / 64u + 1u] = w;
        }
        __syncthreads();
        const uint32_t n0 = (uint32_t)fr->n[0], inc = 256u % n0;
        uint32_t x = (base + threadIdx.x) % n0;
        for (uint32_t k = 0; k < 16; ++k) {
            const uint32_t c = base + 256u * k + threadIdx.x;
            if (c < m_top)   // thermometer byte: bit t set when D_x <= t
                b.skip[c] =
                    (uint8_t)(0xffu << row_distance(s_occ, 256u * k + threadIdx.x + 64u, x, n0));
            x += inc;
            x -= x >= n0 ? n0 : 0u;
        }
        uint32_t nd_total, m_total;
        block_excl_scan<256>(nd, s_scan, nd_total);
        block_excl_scan<256>(msum, s_scan, m_total);
        if (threadIdx.x == 0) {
            b.tile_m[blockIdx.x] = m_total;
            b.tile_d[blockIdx.x] = nd_total;
        }
        return;
    }
    // ms[k] becomes the exclusive prefix within (slot k, this wave)
    uint32_t ed[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        const uint32_t v = ms[k] | ((packed[k] ? 1u : 0u) << 24);
        const uint32_t inc = wave_incl_scan(v);
        const uint32_t ex = inc - v;
        ms[k] = ex & 0xffffffu;
        ed[k] = ex >> 24;
        if (lane == 63u) {
            s_tot[0][4u * k + wave] = inc & 0xffffffu;
            s_tot[1][4u * k + wave] = inc >> 24;
        }
    }
    __syncthreads();
    if (wave == 0) {
        const uint32_t x0 = s_tot[0][lane], x1 = s_tot[1][lane];
        s_tot[0][lane] = wave_incl_scan(x0) - x0;
        s_tot[1][lane] = wave_incl_scan(x1) - x1;
    }
    __syncthreads();
    const uint32_t tm = b.tile_m[blockIdx.x], td = b.tile_d[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k) {
        const uint32_t c = base + 256u * k + threadIdx.x;
        if (c < m_top) {
            if (packed[k]) {
                const uint32_t mpos = tm + s_tot[0][4u * k + wave] + ms[k];
                b.cellrec[c] = make_uint2(0x80000000u | mpos, packed[k]);
                b.dense[td + s_tot[1][4u * k + wave] + ed[k]] = c;
            } else if (!((empty >> k) & 1u)) {
                b.cellrec[c] = make_uint2(b.offsets[c], b.offsets[c + 1]);
            }
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
            g2->n[1] = 0;
            g2->n[2] = 0;
            g2->status = kStatusFrameInvalid;
        }
        return;
    }
    if (threadIdx.x < 48u) {
        // sub-cell size and reciprocal per axis and resolution (traversal table)
        const uint32_t ax = threadIdx.x >> 4, r = (threadIdx.x & 15u) + 1u;
        const float sc = fr->cell[ax] / (float)r;
        b.sub_tab[threadIdx.x] = sc;
        b.sub_tab[48u + threadIdx.x] = 1.0f / sc;
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
        g2->n[1] = 0;
        g2->n[2] = 0;
        g2->r = 0;
        g2->status = status;
    }
}

// Frame of dense cell c's sub-grid.
struct G2Cell {
    uint32_t kb, ke, pk, m;
    uint32_t r[3];
    float sb[3], si[3];
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
}

// Sub-cell box of triangle t in cell g's frame: lo | hi << 4 per axis, axis a
// in bits 8a..8a+7.
__device__ __forceinline__ uint32_t g2_box(const float* __restrict__ verts, uint32_t t,
                                           const G2Cell& g)
{
    const float* v = verts + 9ull * t;
    uint32_t p = 0;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float x = v[a], yv = v[3 + a], zv = v[6 + a];
        float mn = x, mx = x;
        mn = yv < mn ? yv : mn;
        mx = yv > mx ? yv : mx;
        mn = zv < mn ? zv : mn;
        mx = zv > mx ? zv : mx;
        const uint32_t lo = g2_vox(mn, g.sb[a], g.si[a], g.r[a]);
        const uint32_t hi = g2_vox(mx, g.sb[a], g.si[a], g.r[a]);
        p |= (lo | (hi << 4)) << (8 * a);
    }
    return p;
}

// 4: per dense cell (one block each, grid-strided), every reference's box
// into b.box (at its top-reference position), the sub-cell counts into the
// cell's sub_offsets slots (when present) and the cell's sub-reference total
// into dense_refs.  References go one per thread with LDS atomic counters.
__global__ __launch_bounds__(1024) void k_g2_count(G2Build b)
{
    __shared__ uint32_t s_cnt[kSubMax * kSubMax * kSubMax];
    __shared__ uint32_t s_scan[17];
    if (!top_ok(b.fr) || (b.g2->status & ~kStatusRefsOver) != 0)
        return;
    const uint32_t nd = (uint32_t)b.g2->n[0];
    for (uint32_t d = blockIdx.x; d < nd; d += gridDim.x) {
        const uint32_t c = b.dense[d];
        G2Cell g;
        g2_cell(b, c, g);
        for (uint32_t s = threadIdx.x; s < g.m; s += blockDim.x)
            s_cnt[s] = 0u;
        __syncthreads();
        const uint32_t r0 = g.r[0], r01 = g.r[0] * g.r[1];
        uint32_t tot = 0;
        for (uint32_t k = g.kb + threadIdx.x; k < g.ke; k += blockDim.x) {
            const uint32_t p = g2_box(b.verts, b.refs[k], g);
            b.box[k] = p;
            const uint32_t xlo = p & 15u, xhi = (p >> 4) & 15u;
            const uint32_t ylo = (p >> 8) & 15u, yhi = (p >> 12) & 15u;
            const uint32_t zlo = (p >> 16) & 15u, zhi = (p >> 20) & 15u;
            tot += (xhi - xlo + 1u) * (yhi - ylo + 1u) * (zhi - zlo + 1u);
            for (uint32_t z = zlo; z <= zhi; ++z)
                for (uint32_t y = ylo; y <= yhi; ++y)
                    for (uint32_t x = xlo; x <= xhi; ++x)
                        atomicAdd(s_cnt + z * r01 + y * r0 + x, 1u);
        }
        uint32_t cell_total;
        block_excl_scan<1024>(tot, s_scan, cell_total);   // barrier: counts complete
        if (threadIdx.x == 0)
            b.dense_refs[d] = cell_total;
        if (b.sub_offsets) {
            uint32_t* out = b.sub_offsets + (b.cellrec[c].x & 0x7fffffffu);
            for (uint32_t s = threadIdx.x; s < g.m; s += blockDim.x)
                out[s] = s_cnt[s];
        }
        __syncthreads();
    }
}

// 6: per dense cell, the sub-cell offsets (scan of the counts left in
// sub_offsets) and the references scattered through LDS cursors.
constexpr uint32_t kG2WaveSortMax = 1024;

__global__ __launch_bounds__(1024) void k_g2_fill(G2Build b)
{
    constexpr uint32_t kM = kSubMax * kSubMax * kSubMax;
    __shared__ uint32_t s_cur[kM];
    __shared__ uint32_t s_scan[17];
    if (!top_ok(b.fr) || b.g2->status != 0)
        return;
    const uint32_t nd = (uint32_t)b.g2->n[0];
    for (uint32_t d = blockIdx.x; d < nd; d += gridDim.x) {
        const uint32_t c = b.dense[d];
        G2Cell g;
        g2_cell(b, c, g);
        const uint32_t sub_base = b.cellrec[c].x & 0x7fffffffu;
        uint32_t* out = b.sub_offsets + sub_base;
        // offsets: 4 consecutive sub-cells per thread (m <= 4096)
        const uint32_t s0 = 4u * threadIdx.x;
        uint32_t v[4], sum = 0;
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            v[i] = s0 + i < g.m ? out[s0 + i] : 0u;
            sum += v[i];
        }
        uint32_t cell_total;
        uint32_t o = b.dense_refs[d] + block_excl_scan<1024>(sum, s_scan, cell_total);
#pragma unroll
        for (uint32_t i = 0; i < 4; ++i) {
            if (s0 + i < g.m) {
                s_cur[s0 + i] = o;
                out[s0 + i] = o;
            }
            o += v[i];
        }
        (void)cell_total;
        __syncthreads();
        const uint32_t r0 = g.r[0], r01 = g.r[0] * g.r[1];
        for (uint32_t k = g.kb + threadIdx.x; k < g.ke; k += blockDim.x) {
            const uint32_t p = b.box[k];
            const uint32_t t = b.refs[k];
            const uint32_t xlo = p & 15u, xhi = (p >> 4) & 15u;
            const uint32_t ylo = (p >> 8) & 15u, yhi = (p >> 12) & 15u;
            const uint32_t zlo = (p >> 16) & 15u, zhi = (p >> 20) & 15u;
            for (uint32_t z = zlo; z <= zhi; ++z)
                for (uint32_t y = ylo; y <= yhi; ++y)
                    for (uint32_t x = xlo; x <= xhi; ++x)
                        b.sub_refs[atomicAdd(s_cur + z * r01 + y * r0 + x, 1u)] = t;
        }
        __syncthreads();
    }
}

// 7: every sub-cell's ids ascending (the scatter's atomics leave them in any
// order), one thread per sub-cell over the whole second level: 2..16 sorted in
// registers; 17..kG2WaveSortMax queued on b.longl for k_g2_sort_wave, longer
// ones on b.big for k_g2_sort_big.  The loop runs over whole blocks so every
// lane takes part in the ballots.
__global__ __launch_bounds__(256) void k_g2_sort_thread(G2Build b)
{
    if (!top_ok(b.fr) || b.g2->status != 0)
        return;
    const uint32_t m = b.g2->m;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t s0 = blockIdx.x * blockDim.x; s0 < m; s0 += stride) {
        const uint32_t s = s0 + threadIdx.x;
        uint32_t n = 0, o = 0;
        if (s < m) {
            o = b.sub_offsets[s];
            n = b.sub_offsets[s + 1] - o;
        }
        if (n >= 2u && n <= 8u)
            sort_cell_thread<8>(b.sub_refs + o, n);
        else if (n > 8u && n <= 16u)
            sort_cell_thread<16>(b.sub_refs + o, n);
        const bool is_long = n > 16u && n <= kG2WaveSortMax;
        const bool is_big = n > kG2WaveSortMax;
        const uint64_t ml = __ballot(is_long), mb = __ballot(is_big);
        if (ml) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(ml);
            uint32_t base = 0;
            if (lane == leader)
                base = atomicAdd(reinterpret_cast<uint32_t*>(&b.g2->n[2]), (uint32_t)__popcll(ml));
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader);
            if (is_long)
                b.longl[base + lanes_below(ml)] = s;
        }
        if (mb) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(mb);
            uint32_t base = 0;
            if (lane == leader)
                base = atomicAdd(reinterpret_cast<uint32_t*>(&b.g2->n[1]), (uint32_t)__popcll(mb));
            base = (uint32_t)__builtin_amdgcn_readlane((int)base, (int)leader);
            if (is_big)
                b.big[base + lanes_below(mb)] = s;
        }
    }
}

// 8: sub-cells of 17..kG2WaveSortMax references (b.longl), one wave each, in
// registers.
__global__ __launch_bounds__(256) void k_g2_sort_wave(G2Build b)
{
    if (!top_ok(b.fr) || b.g2->status != 0)
        return;
    const uint32_t nl = (uint32_t)b.g2->n[2];
    const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint32_t n_waves = (gridDim.x * blockDim.x) >> 6;
    for (uint32_t li = wave; li < nl; li += n_waves) {
        const uint32_t sc = b.longl[li];
        const uint32_t o = b.sub_offsets[sc];
        const uint32_t n = b.sub_offsets[sc + 1] - o;
        if (n <= 64u)
            wave_sort<1>(b.sub_refs + o, n);
        else if (n <= 256u)
            wave_sort<4>(b.sub_refs + o, n);
        else
            wave_sort<16>(b.sub_refs + o, n);
    }
}

// 9: sub-cells of more than kG2WaveSortMax references (b.big), one block each.
__global__ __launch_bounds__(256) void k_g2_sort_big(G2Build b)
{
    __shared__ uint32_t lds[kSortTile];
    if (!top_ok(b.fr) || b.g2->status != 0)
        return;
    const uint32_t nb = (uint32_t)b.g2->n[1];
    for (uint32_t li = blockIdx.x; li < nb; li += gridDim.x) {
        const uint32_t s = b.big[li];
        const uint32_t o = b.sub_offsets[s];
        block_sort_cell(b.sub_refs + o, b.sub_offsets[s + 1] - o, lds);
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

hipError_t launch_grid2_plan_cells(const G2Build& b, uint32_t top_cells_cap,
                                   uint32_t sub_cells_cap, uint32_t dense_cap, hipStream_t s)
{
    if (b.skip_stride < top_cells_cap)
        return hipErrorInvalidValue;
    const uint32_t nt = top_cells_cap / kScanTile + 1;
    hipLaunchKernelGGL(k_g2_plan<false>, dim3(nt), dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_g2_scan_tiles, dim3(1), dim3(1024), 0, s, b, nt, sub_cells_cap,
                       dense_cap);
    hipLaunchKernelGGL(k_g2_plan<true>, dim3(nt), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_grid2_skip_records(const G2Build& b, uint32_t top_cells_cap, hipStream_t s)
{
    if (b.skip_stride < top_cells_cap)
        return hipErrorInvalidValue;
    const uint32_t nsk = top_cells_cap / 1024u + 1u;
    const dim3 gsk(nsk < 4096u ? nsk : 4096u);
    hipLaunchKernelGGL(k_g2_skip_line<1>, gsk, dim3(256), 0, s, b);
    hipLaunchKernelGGL(k_g2_skip_line<2>, gsk, dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_grid2_plan_refs(const G2Build& b, uint32_t sub_refs_cap, uint32_t dense_cap,
                                  hipStream_t s)
{
    const uint32_t nb = dense_cap < 2048u ? (dense_cap ? dense_cap : 1u) : 2048u;
    hipLaunchKernelGGL(k_g2_count, dim3(nb), dim3(1024), 0, s, b);
    hipLaunchKernelGGL(k_g2_scan_dense, dim3(1), dim3(1024), 0, s, b, sub_refs_cap);
    return hipGetLastError();
}

hipError_t launch_grid2_plan(const G2Build& b, uint32_t top_cells_cap, uint32_t sub_cells_cap,
                             uint32_t sub_refs_cap, uint32_t dense_cap, hipStream_t s)
{
    hipError_t e = launch_grid2_plan_cells(b, top_cells_cap, sub_cells_cap, dense_cap, s);
    if (e == hipSuccess)
        e = launch_grid2_skip_records(b, top_cells_cap, s);
    return e != hipSuccess ? e : launch_grid2_plan_refs(b, sub_refs_cap, dense_cap, s);
}

hipError_t launch_grid2_fill(const G2Build& b, uint32_t sub_cells_cap, uint32_t sub_refs_cap,
                             uint32_t dense_cap, const float4* tri, float4* sub_trir,
                             hipStream_t s)
{
    const uint32_t nb = dense_cap < 2048u ? (dense_cap ? dense_cap : 1u) : 2048u;
    hipLaunchKernelGGL(k_g2_fill, dim3(nb), dim3(1024), 0, s, b);
    const uint32_t nt = sub_cells_cap / 256u + 1u;
    hipLaunchKernelGGL(k_g2_sort_thread, dim3(nt < 4096u ? nt : 4096u), dim3(256), 0, s, b);
    const uint32_t nl = b.long_cap / 4u + 1u;
    hipLaunchKernelGGL(k_g2_sort_wave, dim3(nl < 2048u ? nl : 2048u), dim3(256), 0, s, b);
    const uint32_t nbig = b.big_cap < 256u ? (b.big_cap ? b.big_cap : 1u) : 256u;
    hipLaunchKernelGGL(k_g2_sort_big, dim3(nbig), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return e;
    return launch_gather_refs(b.g2, b.sub_refs, tri, sub_trir, sub_refs_cap, s);
}

hipError_t launch_grid2_build(const G2Build& b, uint32_t top_cells_cap, uint32_t sub_cells_cap,
                              uint32_t sub_refs_cap, uint32_t dense_cap, const float4* tri,
                              float4* sub_trir, hipStream_t s)
{
    const hipError_t e = launch_grid2_plan(b, top_cells_cap, sub_cells_cap, sub_refs_cap,
                                           dense_cap, s);
    return e != hipSuccess ? e
                           : launch_grid2_fill(b, sub_cells_cap, sub_refs_cap, dense_cap, tri,
                                               sub_trir, s);
}

}  // namespace grt
