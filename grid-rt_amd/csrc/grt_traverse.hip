// grt_traverse.hip -- one-ray-per-lane 3D-DDA traversal of the uniform grid
// with Moller-Trumbore intersection, on gfx950 (wave64).
//
// Reference semantics:
//   clip + DDA set-up   AabbIntersection.cpp:37-63, UniformGridAccStruct.cpp:118-149
//   DDA loop            UniformGridAccStruct.cpp:160-195 (single-axis step,
//                       ties to y then z)
//   cell loop           UniformGridAccStruct.cpp:362-378 (hits bounded by the
//                       current cell's exit)
//   Moller-Trumbore     RayTracing_kernel.cu:56-89 (no determinant epsilon)
//   any hit             shadow_FRAG.glsl:122-245
//
// Data in HBM: rays and hits are structures of arrays (coalesced 4-byte
// lanes); the grid is offsets[m+1] (uint32) plus one 48-byte record per
// reference, trir[k] = {v0.xyz, e1.x}{e1.yz, e2.xy}{e2.z, id, -, -}, stored
// contiguously per cell, so a cell's triangles stream without an index
// indirection.  Each traversal loop loads the next record while the current
// one is tested.
//
// Two schedules with identical per-ray results:
//   k_trace_closest / k_trace_any   one ray per lane for the whole launch;
//       when every active lane of a wave is in the same cell the records are
//       read with wave-uniform (scalar) loads;
//   k_trace_persistent              persistent waves that refill idle lanes
//       from a shared slot counter and advance each lane by one work item
//       (one triangle test or one cell) per iteration.
#include "grt_launch.h"
#include "grt_mt.h"

namespace grt {

// ---------------------------------------------------------------------------
// Ray indexing
// ---------------------------------------------------------------------------

// Slot -> ray: through the ray list when the launch has one (slots beyond
// the list's device-side count are idle), else the image-tile mapping.
__device__ __forceinline__ uint32_t trace_ray(const TraceArgs& a, uint32_t slot, bool& valid)
{
    if (a.list) {
        const uint32_t c = *a.list_count;
        valid = slot < c && slot < a.n;
        return valid ? a.list[slot] : 0u;
    }
    return slot_ray(slot, a.image_width, a.n, valid);
}

// ---------------------------------------------------------------------------
// DDA
// ---------------------------------------------------------------------------
struct Dda {
    int32_t cx, cy, cz;
    int32_t sx, sy, sz;
    int32_t ox, oy, oz;   // out-of-grid limits
    float tx, ty, tz;     // t at the next boundary per axis
    float dtx, dty, dtz;
};

__device__ __forceinline__ bool grid_ok(const DevFrame* __restrict__ fr)
{
    return (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)) == 0;
}

// Ray::update + clipRay + DDA set-up.  False when the ray misses the grid box
// (or its interval is empty).
__device__ __forceinline__ bool dda_setup(const DevFrame* __restrict__ fr, float o0, float o1,
                                          float o2, float d0, float d1, float d2, float tmin,
                                          float tmax, Dda& s)
{
    const float i0 = 1.0f / d0, i1 = 1.0f / d1, i2 = 1.0f / d2;
    float tn = tmin, tf = tmax;
    {
        const float a = (fr->bmin[0] - o0) * i0, b = (fr->bmax[0] - o0) * i0;
        tn = tmax_f(tn, tmin_f(a, b));
        tf = tmin_f(tf, tmax_f(a, b));
    }
    {
        const float a = (fr->bmin[1] - o1) * i1, b = (fr->bmax[1] - o1) * i1;
        tn = tmax_f(tn, tmin_f(a, b));
        tf = tmin_f(tf, tmax_f(a, b));
    }
    {
        const float a = (fr->bmin[2] - o2) * i2, b = (fr->bmax[2] - o2) * i2;
        tn = tmax_f(tn, tmin_f(a, b));
        tf = tmin_f(tf, tmax_f(a, b));
    }
    if (tn > tf)
        return false;
    const float p0 = o0 + d0 * tn, p1 = o1 + d1 * tn, p2 = o2 + d2 * tn;
    s.cx = vox(p0, fr->bmin[0], fr->inv[0], fr->n[0]);
    s.cy = vox(p1, fr->bmin[1], fr->inv[1], fr->n[1]);
    s.cz = vox(p2, fr->bmin[2], fr->inv[2], fr->n[2]);
    const bool n0 = __builtin_signbit(d0), n1 = __builtin_signbit(d1), n2 = __builtin_signbit(d2);
    s.sx = n0 ? -1 : 1;
    s.sy = n1 ? -1 : 1;
    s.sz = n2 ? -1 : 1;
    s.ox = n0 ? -1 : fr->n[0];
    s.oy = n1 ? -1 : fr->n[1];
    s.oz = n2 ? -1 : fr->n[2];
    const float w0 = (float)(s.cx + (n0 ? 0 : 1)) * fr->cell[0] + fr->bmin[0];
    const float w1 = (float)(s.cy + (n1 ? 0 : 1)) * fr->cell[1] + fr->bmin[1];
    const float w2 = (float)(s.cz + (n2 ? 0 : 1)) * fr->cell[2] + fr->bmin[2];
    s.tx = (w0 - o0) * i0;
    s.ty = (w1 - o1) * i1;
    s.tz = (w2 - o2) * i2;
    s.dtx = fabsf(fr->cell[0] * i0);
    s.dty = fabsf(fr->cell[1] * i1);
    s.dtz = fabsf(fr->cell[2] * i2);
    return true;
}

// Single-axis advance with the reference tie order.  False when leaving.
__device__ __forceinline__ bool dda_step(Dda& s)
{
    if (s.tx < s.ty && s.tx < s.tz) {
        s.cx += s.sx;
        if (s.cx == s.ox)
            return false;
        s.tx += s.dtx;
    } else if (s.ty < s.tz) {
        s.cy += s.sy;
        if (s.cy == s.oy)
            return false;
        s.ty += s.dty;
    } else {
        s.cz += s.sz;
        if (s.cz == s.oz)
            return false;
        s.tz += s.dtz;
    }
    return true;
}


// ---------------------------------------------------------------------------
// Wave-cooperative cell tests
// ---------------------------------------------------------------------------
//
// Once a persistent wave finds the ray queue drained, its last rays would run
// their remaining triangle tests one per iteration on single lanes while the
// other lanes idle; a ray through a dense cell (thousands of references) then
// sets the length of the whole launch.  In that phase the wave instead tests
// the rest of such a cell 64 references at a time, one per lane.  The result
// is the serial one: closest hit keeps the smallest t and, among equal t, the
// earliest reference (a lexicographic (t, reference) minimum, since the serial
// loop only replaces on a strictly smaller t); any hit stops at the first
// occluding reference in order, and the test counter counts the references the
// serial loop would have tested.

__device__ __forceinline__ float bcast_f(float v, uint32_t l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)l));
}

__device__ __forceinline__ uint32_t bcast_u(uint32_t v, uint32_t l)
{
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)l);
}

template <class T>
__device__ __forceinline__ T* bcast_p(T* p, uint32_t l)
{
    const uint64_t x = (uint64_t)p;
    const uint64_t lo = bcast_u((uint32_t)x, l), hi = bcast_u((uint32_t)(x >> 32), l);
    return (T*)((hi << 32) | lo);
}

__device__ __forceinline__ float wave_min_f(float v)
{
    for (int o = 32; o > 0; o >>= 1)
        v = tmin_f(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ uint32_t wave_min_u(uint32_t v)
{
    for (int o = 32; o > 0; o >>= 1) {
        const uint32_t w = (uint32_t)__shfl_xor((int)v, o, 64);
        v = w < v ? w : v;
    }
    return v;
}

// State of one lane's ray that the cooperative pass reads and updates.
struct LaneRay {
    float o0, o1, o2, d0, d1, d2, tmin, tmax, bound;
    uint32_t k, e;        // references [k, e) of the current cell are left
    int32_t id;
    float best, bu, bv;
    uint32_t tests;
};

// The whole wave (every lane, uniform control flow) finishes the current cell
// of lane l: records [k, e) of `trir` (lane l's grid).  Updates lane l's k,
// hit and test count.
__device__ __forceinline__ void coop_cell(uint32_t l, bool any, const float4* trir, LaneRay& r)
{
    const uint32_t lane = threadIdx.x & 63u;
    const float o0 = bcast_f(r.o0, l), o1 = bcast_f(r.o1, l), o2 = bcast_f(r.o2, l);
    const float d0 = bcast_f(r.d0, l), d1 = bcast_f(r.d1, l), d2 = bcast_f(r.d2, l);
    const float tmin = bcast_f(r.tmin, l), tmax = bcast_f(r.tmax, l);
    const uint32_t kb = bcast_u(r.k, l), ke = bcast_u(r.e, l);
    const float4* __restrict__ tr = bcast_p(trir, l);
    if (any) {
        uint32_t first = 0xffffffffu;
        for (uint32_t base = kb; base < ke; base += 64u) {
            const uint32_t kk = base + lane;
            bool h = false;
            if (kk < ke) {
                const float4 A = tr[3ull * kk], B = tr[3ull * kk + 1], C = tr[3ull * kk + 2];
                h = any_test(A, B, C, o0, o1, o2, d0, d1, d2, tmin, tmax);
            }
            const uint64_t hm = __ballot(h);
            if (hm != 0) {
                first = base + (uint32_t)__builtin_ctzll(hm);
                break;
            }
        }
        if (lane == l) {
            if (first != 0xffffffffu) {
                r.tests += first - kb + 1u;
                r.id = 1;
            } else {
                r.tests += ke - kb;
            }
            r.k = ke;
        }
        return;
    }
    const float bound = bcast_f(r.bound, l), best0 = bcast_f(r.best, l);
    float lb = best0, lu = 0.0f, lv = 0.0f;
    int32_t lid = -1;
    uint32_t lk = 0xffffffffu;
    for (uint32_t base = kb; base < ke; base += 64u) {
        const uint32_t kk = base + lane;
        if (kk < ke) {
            const float4 A = tr[3ull * kk], B = tr[3ull * kk + 1], C = tr[3ull * kk + 2];
            const MtPre m = mt_pre(A, B, C, o0, o1, o2, d0, d1, d2);
            if (mt_maybe(m, tmin, tmin_f(tmin_f(tmax, bound), lb))) {
                float t, u, v;
                if (mt_exact(m, t, u, v) && t > tmin && t <= tmax && t <= bound && t < lb) {
                    lb = t;
                    lu = u;
                    lv = v;
                    lid = (int32_t)__float_as_uint(C.y);
                    lk = kk;
                }
            }
        }
    }
    const float mt = wave_min_f(lb);
    const uint32_t mk = wave_min_u((lk != 0xffffffffu && lb == mt) ? lk : 0xffffffffu);
    if (mk != 0xffffffffu) {
        const uint32_t w = (uint32_t)__builtin_ctzll(__ballot(lk == mk));
        const int32_t wid = (int32_t)bcast_u((uint32_t)lid, w);
        const float wu = bcast_f(lu, w), wv = bcast_f(lv, w);
        if (lane == l) {
            r.best = mt;
            r.id = wid;
            r.bu = wu;
            r.bv = wv;
        }
    }
    if (lane == l) {
        r.tests += ke - kb;
        r.k = ke;
    }
}

// Lanes whose current cell is worth a cooperative pass now: the queue is
// drained for this wave and at least max(coop, 2 x busy lanes) references
// are left.
__device__ __forceinline__ uint64_t coop_lanes(bool busy, bool drained_lane, uint32_t k,
                                               uint32_t e, uint32_t coop)
{
    if (coop == 0 || __ballot(drained_lane) == 0)
        return 0;
    const uint32_t nb = (uint32_t)__popcll(__ballot(busy));
    const uint32_t thr = coop > 2u * nb ? coop : 2u * nb;
    return __ballot(busy && k < e && e - k >= thr);
}

// ---------------------------------------------------------------------------
// Pooled tests (one ray per lane)
// ---------------------------------------------------------------------------
//
// With rays in different cells, a per-lane loop over each lane's references
// leaves most lanes idle (the wave runs as long as its fullest cell).  Pooled,
// the (lane, reference) pairs of all the wave's current cells are numbered
// through an exclusive scan of the lanes' counts and tested 64 at a time, one
// pair per lane, the owner's ray read from LDS.  Closest hit keeps, per owner,
// the minimum of (t, reference) in an LDS 64-bit atomic -- t mapped to an
// order-preserving integer with -0 folded into +0 -- which is the reference
// the serial loop would keep (it replaces only on a strictly smaller t); the
// winner's u, v are then recomputed by the owner with the same arithmetic.
// Any hit keeps the first occluding offset; the test counter counts what the
// serial loop would have tested.

// LDS staging of one wave.
struct WavePool {
    float o0[64], o1[64], o2[64], d0[64], d1[64], d2[64], tmin[64], tmax[64], bound[64];
    uint32_t pre[64];                // exclusive prefix of the lanes' reference counts
    uint32_t kb[64];                 // first reference of each lane's cell
    unsigned long long win[64];      // per owner: (t, reference) minimum / first occluder
};

__device__ __forceinline__ uint32_t ordered_bits(float t)
{
    const uint32_t b = __float_as_uint(t == 0.0f ? 0.0f : t);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ void pool_store_ray(WavePool& P, const LaneRay& r)
{
    const uint32_t lane = threadIdx.x & 63u;
    P.o0[lane] = r.o0;
    P.o1[lane] = r.o1;
    P.o2[lane] = r.o2;
    P.d0[lane] = r.d0;
    P.d1[lane] = r.d1;
    P.d2[lane] = r.d2;
    P.tmin[lane] = r.tmin;
    P.tmax[lane] = r.tmax;
}

// Owner of pair p: the last lane whose prefix is <= p.
__device__ __forceinline__ uint32_t pool_owner(const WavePool& P, uint32_t p)
{
    uint32_t o = 0;
#pragma unroll
    for (uint32_t step = 32; step > 0; step >>= 1)
        if (P.pre[o + step] <= p)
            o += step;
    return o;
}

// The whole wave tests every lane's references [k, k + c) (c = 0 for lanes
// without work); each lane's hit and test count are updated as by the serial
// loop and its cell is consumed.
template <bool ANY>
__device__ __forceinline__ void pool_cells(WavePool& P, const float4* __restrict__ trir,
                                           LaneRay& r, uint32_t c)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t total = bcast_u(incl, 63);
    P.pre[lane] = incl - c;
    P.kb[lane] = r.k;
    P.bound[lane] = r.bound;
    P.win[lane] = ~0ull;
    wave_lds_sync();
    for (uint32_t base = 0; base < total; base += 64u) {
        const uint32_t p = base + lane;
        if (p < total) {
            const uint32_t o = pool_owner(P, p);
            const uint32_t j = p - P.pre[o];
            const uint32_t ref = P.kb[o] + j;
            if (ANY && P.win[o] < j)
                continue;   // the owner has an earlier occluder already
            const float4 A = trir[3ull * ref], B = trir[3ull * ref + 1], C = trir[3ull * ref + 2];
            const float o0 = P.o0[o], o1 = P.o1[o], o2 = P.o2[o];
            const float d0 = P.d0[o], d1 = P.d1[o], d2 = P.d2[o];
            const float tmin = P.tmin[o], tmax = P.tmax[o];
            if (ANY) {
                if (any_test(A, B, C, o0, o1, o2, d0, d1, d2, tmin, tmax))
                    atomicMin(&P.win[o], (unsigned long long)j);
            } else {
                const float bound = P.bound[o];
                const MtPre m = mt_pre(A, B, C, o0, o1, o2, d0, d1, d2);
                if (mt_maybe(m, tmin, tmin_f(tmax, bound))) {
                    float t, u, v;
                    if (mt_exact(m, t, u, v) && t > tmin && t <= tmax && t <= bound)
                        atomicMin(&P.win[o], ((unsigned long long)ordered_bits(t) << 32) | ref);
                }
            }
        }
    }
    wave_lds_sync();
    if (c == 0)
        return;
    const unsigned long long w = P.win[lane];
    if (ANY) {
        if (w != ~0ull) {
            r.id = 1;
            r.tests += (uint32_t)w + 1u;
        } else {
            r.tests += c;
        }
    } else {
        if (w != ~0ull) {
            const uint32_t ref = (uint32_t)w;
            const float4 A = trir[3ull * ref], B = trir[3ull * ref + 1], C = trir[3ull * ref + 2];
            const MtPre m = mt_pre(A, B, C, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2);
            float t, u, v;
            mt_exact(m, t, u, v);
            r.best = t;
            r.bu = u;
            r.bv = v;
            r.id = (int32_t)__float_as_uint(C.y);
        }
        r.tests += c;
    }
    r.k = r.e;
}

// ---------------------------------------------------------------------------
// One ray per lane
// ---------------------------------------------------------------------------
//
// Each iteration of the wave-uniform loop handles one cell per busy lane:
//   1. the cooperative pass (cells with many references, when the busy lanes
//      of the wave are not all in one cell),
//   2. the remaining references of each lane's cell -- with wave-uniform
//      (scalar) record loads when every busy lane is in the same cell,
//   3. finish (hit in this cell / ray leaves its interval or the grid) or step
//      to the next cell.

// Lanes for a cooperative pass in the one-ray-per-lane kernels: none when all
// busy lanes share one cell (the scalar path serves them better), else those
// with at least max(coop, 2 x busy lanes) references left.
__device__ __forceinline__ uint64_t coop_lanes_simple(bool busy, uint32_t k, uint32_t e,
                                                      uint32_t coop)
{
    if (coop == 0)
        return 0;
    const uint64_t bm = __ballot(busy);
    const uint32_t f = (uint32_t)__builtin_ctzll(bm);
    const uint32_t k0 = bcast_u(k, f), e0 = bcast_u(e, f);
    if (__ballot(busy && (k != k0 || e != e0)) == 0)
        return 0;
    const uint32_t nb = (uint32_t)__popcll(bm);
    const uint32_t thr = coop > 2u * nb ? coop : 2u * nb;
    return __ballot(busy && k < e && e - k >= thr);
}

// Current cell of a lane: exit distance, reference range, hit bound.
__device__ __forceinline__ void load_cell(const Dda& s, const uint32_t* __restrict__ offsets,
                                          uint32_t nx, uint32_t ny, float& exit_t, float& bound,
                                          uint32_t& k, uint32_t& e)
{
    exit_t = tmin_f(tmin_f(s.tx, s.ty), s.tz);
    const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
    k = offsets[ci];
    e = offsets[ci + 1];
    bound = exit_t * kExitScale;
}

// Closest hit of ray i (valid lanes) against one grid; every lane of the wave
// calls it together.
template <bool COUNT, bool POOL>
__device__ __forceinline__ void trace_closest_lane(const DevFrame* __restrict__ fr,
                                                   const uint32_t* __restrict__ offsets,
                                                   const float4* __restrict__ trir,
                                                   const RaySoA& rays, const HitSoA& hits,
                                                   uint32_t* __restrict__ counters, uint32_t coop,
                                                   uint32_t i, bool valid, WavePool* pool)
{
    const KF4 ktr = (KF4)trir;
    LaneRay r = {};
    r.id = -1;
    r.best = INFINITY;
    uint32_t cells = 0;
    Dda s = {};
    bool busy = false;
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    if (valid) {
        r.o0 = rays.ox[i];
        r.o1 = rays.oy[i];
        r.o2 = rays.oz[i];
        r.d0 = rays.dx[i];
        r.d1 = rays.dy[i];
        r.d2 = rays.dz[i];
        r.tmin = rays.tmin[i];
        r.tmax = rays.tmax[i];
        busy = grid_ok(fr) && dda_setup(fr, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin, r.tmax, s);
    }
    if (POOL)
        pool_store_ray(*pool, r);
    float exit_t = 0.0f;
    if (busy) {
        load_cell(s, offsets, nx, ny, exit_t, r.bound, r.k, r.e);
        cells = 1;
    }
    while (__ballot(busy) != 0) {
        bool pooled = false;
        if (POOL) {
            const uint32_t c = (busy && r.k < r.e) ? r.e - r.k : 0u;
            const uint64_t cm = __ballot(c != 0);
            if (cm != 0) {
                const uint32_t f = (uint32_t)__builtin_ctzll(cm);
                const uint32_t k0 = bcast_u(r.k, f), e0 = bcast_u(r.e, f);
                if (__ballot(c != 0 && (r.k != k0 || r.e != e0)) != 0) {
                    pool_cells<false>(*pool, trir, r, c);
                    pooled = true;
                }
            }
        } else if (const uint64_t hv0 = coop_lanes_simple(busy, r.k, r.e, coop)) {
            for (uint64_t hv = hv0; hv != 0; hv &= hv - 1)
                coop_cell((uint32_t)__builtin_ctzll(hv), false, trir, r);
        }
        if (!pooled && busy && r.k < r.e) {
            const uint32_t b = r.k, e = r.e;
            const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
            const uint32_t e0 = __builtin_amdgcn_readfirstlane(e);
            if (__ballot(b != b0 || e != e0) == 0) {
                // Every lane here is in this cell: wave-uniform records.
                test_records<false>(ktr, b0, e0, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin,
                                    r.tmax, r.bound, r.id, r.best, r.bu, r.bv, r.tests);
            } else {
                test_records<false>(trir, b, e, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin,
                                    r.tmax, r.bound, r.id, r.best, r.bu, r.bv, r.tests);
            }
            r.k = e;
        }
        if (busy) {
            if (r.id >= 0 || exit_t > r.tmax || !dda_step(s)) {
                busy = false;
            } else {
                load_cell(s, offsets, nx, ny, exit_t, r.bound, r.k, r.e);
                ++cells;
            }
        }
    }
    if (valid) {
        hits.id[i] = r.id;
        hits.t[i] = r.best;
        hits.u[i] = r.bu;
        hits.v[i] = r.bv;
        if (COUNT) {
            counters[2ull * i] = cells;
            counters[2ull * i + 1] = r.tests;
        }
    }
}

// Any hit of ray i (valid lanes) against one grid; every lane of the wave calls
// it together.
template <bool COUNT, bool POOL>
__device__ __forceinline__ void trace_any_lane(const DevFrame* __restrict__ fr,
                                               const uint32_t* __restrict__ offsets,
                                               const float4* __restrict__ trir,
                                               const RaySoA& rays, uint8_t* __restrict__ occluded,
                                               uint32_t* __restrict__ counters, uint32_t coop,
                                               uint32_t i, bool valid, WavePool* pool)
{
    const KF4 ktr = (KF4)trir;
    LaneRay r = {};
    r.id = -1;
    uint32_t cells = 0;
    Dda s = {};
    bool busy = false;
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    if (valid) {
        r.o0 = rays.ox[i];
        r.o1 = rays.oy[i];
        r.o2 = rays.oz[i];
        r.d0 = rays.dx[i];
        r.d1 = rays.dy[i];
        r.d2 = rays.dz[i];
        r.tmin = rays.tmin[i];
        r.tmax = rays.tmax[i];
        busy = grid_ok(fr) && dda_setup(fr, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin, r.tmax, s);
    }
    if (POOL)
        pool_store_ray(*pool, r);
    float exit_t = 0.0f;
    if (busy) {
        load_cell(s, offsets, nx, ny, exit_t, r.bound, r.k, r.e);
        cells = 1;
    }
    while (__ballot(busy) != 0) {
        bool pooled = false;
        if (POOL) {
            const uint32_t c = (busy && r.k < r.e) ? r.e - r.k : 0u;
            const uint64_t cm = __ballot(c != 0);
            if (cm != 0) {
                const uint32_t f = (uint32_t)__builtin_ctzll(cm);
                const uint32_t k0 = bcast_u(r.k, f), e0 = bcast_u(r.e, f);
                if (__ballot(c != 0 && (r.k != k0 || r.e != e0)) != 0) {
                    pool_cells<true>(*pool, trir, r, c);
                    pooled = true;
                }
            }
        } else if (const uint64_t hv0 = coop_lanes_simple(busy, r.k, r.e, coop)) {
            for (uint64_t hv = hv0; hv != 0; hv &= hv - 1)
                coop_cell((uint32_t)__builtin_ctzll(hv), true, trir, r);
        }
        if (!pooled && busy && r.k < r.e) {
            const uint32_t b = r.k, e = r.e;
            const uint32_t b0 = __builtin_amdgcn_readfirstlane(b);
            const uint32_t e0 = __builtin_amdgcn_readfirstlane(e);
            int32_t nid = -1;
            float nt = 0.0f, nu = 0.0f, nv = 0.0f;
            // references tested: up to and including an occluder
            const bool occ =
                __ballot(b != b0 || e != e0) == 0
                    ? test_records<true>(ktr, b0, e0, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin,
                                         r.tmax, 0.0f, nid, nt, nu, nv, r.tests)
                    : test_records<true>(trir, b, e, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin,
                                         r.tmax, 0.0f, nid, nt, nu, nv, r.tests);
            if (occ)
                r.id = 1;
            r.k = e;
        }
        if (busy) {
            if (r.id >= 0 || exit_t > r.tmax || !dda_step(s)) {
                busy = false;
            } else {
                load_cell(s, offsets, nx, ny, exit_t, r.bound, r.k, r.e);
                ++cells;
            }
        }
    }
    if (valid) {
        occluded[i] = r.id >= 0 ? 1 : 0;
        if (COUNT) {
            counters[2ull * i] = cells;
            counters[2ull * i + 1] = r.tests;
        }
    }
}

template <bool COUNT, bool POOL>
__global__ __launch_bounds__(256) void k_trace_closest(TraceArgs a, RaySoA rays, HitSoA hits)
{
    __shared__ WavePool pools[POOL ? 4 : 1];
    bool valid;
    const uint32_t i = trace_ray(a, blockIdx.x * blockDim.x + threadIdx.x, valid);
    trace_closest_lane<COUNT, POOL>(a.fr, a.offsets, a.trir, rays, hits, a.counters, a.coop, i,
                                    valid, POOL ? &pools[threadIdx.x >> 6] : nullptr);
}

template <bool COUNT, bool POOL>
__global__ __launch_bounds__(256) void k_trace_any(TraceArgs a, RaySoA rays,
                                                   uint8_t* __restrict__ occluded)
{
    __shared__ WavePool pools[POOL ? 4 : 1];
    bool valid;
    const uint32_t i = trace_ray(a, blockIdx.x * blockDim.x + threadIdx.x, valid);
    trace_any_lane<COUNT, POOL>(a.fr, a.offsets, a.trir, rays, occluded, a.counters, a.coop, i,
                                valid, POOL ? &pools[threadIdx.x >> 6] : nullptr);
}

// Batched multi-scene traversal, one ray per lane: every 64-slot group belongs
// to one task (map[slot / 64]), so each wave runs the single-scene loop of its
// task with wave-uniform scene data.
template <bool POOL>
__global__ __launch_bounds__(256) void k_trace_batch_simple(const BatchTask* __restrict__ tasks,
                                                            const uint32_t* __restrict__ map,
                                                            uint32_t total_slots, uint32_t coop)
{
    __shared__ WavePool pools[POOL ? 4 : 1];
    WavePool* pool = POOL ? &pools[threadIdx.x >> 6] : nullptr;
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t group = __builtin_amdgcn_readfirstlane(slot >> 6);
    if ((group << 6) >= total_slots)
        return;   // whole waves only: total_slots is a multiple of 64
    const uint32_t task = __builtin_amdgcn_readfirstlane(map[group]);
    const BatchTask& T = tasks[task];
    bool valid;
    const uint32_t i = slot_ray(slot - T.slot_begin, T.image_width, T.n, valid);
    if (T.any)
        trace_any_lane<false, POOL>(T.fr, T.offsets, T.trir, T.rays, T.occluded, nullptr, coop, i,
                                    valid, pool);
    else
        trace_closest_lane<false, POOL>(T.fr, T.offsets, T.trir, T.rays, T.hits, nullptr, coop, i,
                                        valid, pool);
}

// ---------------------------------------------------------------------------
// Several lanes per ray
// ---------------------------------------------------------------------------
//
// A launch with few rays (secondary rays of a few pixels, a small request)
// leaves most SIMDs with one wave or none, and each lane then pays the memory
// latency of every cell and record of its ray in series.  Here a group of
// `lanes` (2..16, a power of two) adjacent lanes shares one ray: lane j of the
// group visits the ray's cells j, j + lanes, j + 2 lanes, ... of the serial
// DDA order, stepping the DDA on registers in between.  After each round the
// group keeps the earliest cell (lowest ordinal) that holds a hit -- the cell
// where the serial loop would have stopped -- and discards later cells, so
// hits, occlusion flags and the {cells, tests} counters equal the serial
// ones.  Dense cells get the wave-cooperative pass as in the one-ray-per-lane
// kernels.

// Advance by `count` cells of the serial order: false when the ray's interval
// ends in the current cell (its exit beyond tmax) or the ray leaves the grid.
__device__ __forceinline__ bool dda_skip(Dda& s, uint32_t count, float tmax)
{
    for (uint32_t q = 0; q < count; ++q)
        if (tmin_f(tmin_f(s.tx, s.ty), s.tz) > tmax || !dda_step(s))
            return false;
    return true;
}

// Serial tests of the rest of a lane's current cell, per-lane (vector) loads.
template <bool ANY>
__device__ __forceinline__ void lane_cell(const float4* __restrict__ trir, LaneRay& r)
{
    if (ANY) {
        int32_t nid = -1;
        float nt = 0.0f, nu = 0.0f, nv = 0.0f;
        if (test_records<true>(trir, r.k, r.e, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin,
                               r.tmax, 0.0f, nid, nt, nu, nv, r.tests))
            r.id = 1;
    } else {
        test_records<false>(trir, r.k, r.e, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin, r.tmax,
                            r.bound, r.id, r.best, r.bu, r.bv, r.tests);
    }
    r.k = r.e;
}

// Sum over the `lanes`-lane group (every lane of the wave calls it).
__device__ __forceinline__ uint32_t group_sum(uint32_t v, uint32_t lanes)
{
    for (uint32_t o = 1; o < lanes; o <<= 1)
        v += (uint32_t)__shfl_xor((int)v, (int)o, 64);
    return v;
}

// Ray i (valid: the whole group) traced by its group of `lanes` lanes.
template <bool ANY, bool COUNT>
__device__ __forceinline__ void trace_split_lane(const DevFrame* __restrict__ fr,
                                                 const uint32_t* __restrict__ offsets,
                                                 const float4* __restrict__ trir,
                                                 const RaySoA& rays, const HitSoA& hits,
                                                 uint8_t* __restrict__ occluded,
                                                 uint32_t* __restrict__ counters, uint32_t coop,
                                                 uint32_t lanes, uint32_t i, bool valid)
{
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t j = lane & (lanes - 1u);
    const uint32_t g0 = lane - j;
    const uint64_t gmask = (lanes >= 64u ? ~0ull : ((1ull << lanes) - 1ull)) << g0;
    LaneRay r = {};
    r.id = -1;
    r.best = INFINITY;
    uint32_t cells = 0;
    Dda s = {};
    bool alive = false;
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    if (valid) {
        r.o0 = rays.ox[i];
        r.o1 = rays.oy[i];
        r.o2 = rays.oz[i];
        r.d0 = rays.dx[i];
        r.d1 = rays.dy[i];
        r.d2 = rays.dz[i];
        r.tmin = rays.tmin[i];
        r.tmax = rays.tmax[i];
        alive = grid_ok(fr) &&
                dda_setup(fr, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin, r.tmax, s) &&
                dda_skip(s, j, r.tmax);
    }
    float exit_t = 0.0f;
    if (alive) {
        load_cell(s, offsets, nx, ny, exit_t, r.bound, r.k, r.e);
        cells = 1;
    }
    bool winner = false;
    while (__ballot(alive) != 0) {
        const uint32_t t0 = r.tests;
        if (const uint64_t hv0 = coop_lanes_simple(alive, r.k, r.e, coop)) {
            for (uint64_t hv = hv0; hv != 0; hv &= hv - 1)
                coop_cell((uint32_t)__builtin_ctzll(hv), ANY, trir, r);
        }
        if (alive && r.k < r.e)
            lane_cell<ANY>(trir, r);
        // The group's earliest cell with a hit ends the ray; cells after it
        // (higher lanes of this round) are not part of the serial walk.
        const uint64_t hm = __ballot(alive && r.id >= 0) & gmask;
        if (hm != 0) {
            const uint32_t jw = (uint32_t)__builtin_ctzll(hm) - g0;
            if (alive && j > jw) {
                --cells;
                r.tests = t0;
            }
            winner = alive && j == jw;
            alive = false;
        } else if (alive) {
            if (dda_skip(s, lanes, r.tmax)) {
                load_cell(s, offsets, nx, ny, exit_t, r.bound, r.k, r.e);
                ++cells;
            } else {
                alive = false;
            }
        }
    }
    const bool group_hit = (__ballot(winner) & gmask) != 0;
    uint32_t sc = 0, st = 0;
    if (COUNT) {
        sc = group_sum(cells, lanes);
        st = group_sum(r.tests, lanes);
    }
    if (!valid)
        return;
    if (ANY) {
        if (j == 0)
            occluded[i] = group_hit ? 1 : 0;
    } else if (winner || (j == 0 && !group_hit)) {
        hits.id[i] = r.id;
        hits.t[i] = r.best;
        hits.u[i] = r.bu;
        hits.v[i] = r.bv;
    }
    if (COUNT && j == 0) {
        counters[2ull * i] = sc;
        counters[2ull * i + 1] = st;
    }
}

template <bool ANY, bool COUNT>
__global__ __launch_bounds__(256) void k_trace_split(TraceArgs a, RaySoA rays, HitSoA hits,
                                                     uint8_t* __restrict__ occluded,
                                                     uint32_t lanes)
{
    const uint32_t shift = (uint32_t)__builtin_ctz(lanes);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid;
    const uint32_t i = trace_ray(a, t >> shift, valid);
    trace_split_lane<ANY, COUNT>(a.fr, a.offsets, a.trir, rays, hits, occluded, a.counters,
                                 a.coop, lanes, i, valid);
}

// Batched multi-scene traversal with `lanes` lanes per ray (64 / lanes rays
// per wave): ray slot s is traced by threads [s * lanes, (s + 1) * lanes).  A
// wave covers 64 / lanes consecutive slots inside one 64-slot group, so one
// task, and the task's closest/any choice is uniform over the wave.
__global__ __launch_bounds__(256) void k_trace_batch_split(const BatchTask* __restrict__ tasks,
                                                           const uint32_t* __restrict__ map,
                                                           uint32_t total_slots, uint32_t coop,
                                                           uint32_t lanes)
{
    const uint32_t shift = (uint32_t)__builtin_ctz(lanes);
    const uint32_t slot = (blockIdx.x * blockDim.x + threadIdx.x) >> shift;
    const uint32_t group = __builtin_amdgcn_readfirstlane(slot >> 6);
    if ((group << 6) >= total_slots)
        return;   // whole waves only: total_slots is a multiple of 64
    const uint32_t task = __builtin_amdgcn_readfirstlane(map[group]);
    const BatchTask& T = tasks[task];
    bool valid;
    const uint32_t i = slot_ray(slot - T.slot_begin, T.image_width, T.n, valid);
    if (T.any)
        trace_split_lane<true, false>(T.fr, T.offsets, T.trir, T.rays, T.hits, T.occluded, nullptr,
                                      coop, lanes, i, valid);
    else
        trace_split_lane<false, false>(T.fr, T.offsets, T.trir, T.rays, T.hits, T.occluded,
                                       nullptr, coop, lanes, i, valid);
}

// ---------------------------------------------------------------------------
// Persistent waves with lane refill
// ---------------------------------------------------------------------------

// Each wave keeps its lanes busy by fetching new rays (one atomic per wave
// for all idle lanes, once at least `refill` lanes are idle or the wave has
// drained) and advances every lane by one work item per iteration -- one
// triangle test, or one cell (DDA step + cell range).  Lanes never wait for the
// longest ray of their wave and inactive ray slots cost one iteration.  Every
// wave exits once the slot counter passes n_slots and its lanes are done.
template <bool ANY, bool COUNT>
__global__ __launch_bounds__(256) void k_trace_persistent(TraceArgs a, RaySoA rays, HitSoA hits,
                                                          uint8_t* __restrict__ occluded,
                                                          uint32_t* __restrict__ slot_counter,
                                                          uint32_t n_slots, uint32_t refill)
{
    const DevFrame* __restrict__ fr = a.fr;
    const uint32_t* __restrict__ offsets = a.offsets;
    const float4* __restrict__ trir = a.trir;
    const bool ok = grid_ok(fr);
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t slot_limit =
        a.list ? (*a.list_count < n_slots ? *a.list_count : n_slots) : n_slots;

    int32_t ray = -1;         // ray index traced by this lane, -1 = idle
    bool exhausted = false;   // no slots left for this lane
    float o0 = 0, o1 = 0, o2 = 0, d0 = 0, d1 = 0, d2 = 0, tmin = 0, tmax = 0;
    Dda s = {};
    uint32_t k = 0, e = 0, cells = 0, tests = 0;
    bool loaded = false;
    float exit_t = 0.0f, bound = 0.0f;
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    float4 A = {}, B = {}, C = {};   // record k of the current cell, loaded ahead

    for (;;) {
        // ---- refill idle lanes -------------------------------------------
        const bool need = (ray < 0) && !exhausted;
        const uint64_t need_mask = __ballot(need);
        const uint64_t busy_mask = __ballot(ray >= 0);
        const uint32_t n_need = (uint32_t)__popcll(need_mask);
        if (need_mask != 0 && (n_need >= refill || busy_mask == 0)) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(need_mask);
            uint32_t base = 0;
            if (lane == leader)
                base = atomicAdd(slot_counter, n_need);
            base = __builtin_amdgcn_readlane(base, leader);
            if (need) {
                const uint32_t slot = base + lanes_below(need_mask);
                if (slot >= slot_limit) {
                    exhausted = true;
                } else {
                    bool valid;
                    const uint32_t ri = trace_ray(a, slot, valid);
                    if (valid) {
                        o0 = rays.ox[ri];
                        o1 = rays.oy[ri];
                        o2 = rays.oz[ri];
                        d0 = rays.dx[ri];
                        d1 = rays.dy[ri];
                        d2 = rays.dz[ri];
                        tmin = rays.tmin[ri];
                        tmax = rays.tmax[ri];
                        id = -1;
                        best = INFINITY;
                        bu = 0.0f;
                        bv = 0.0f;
                        cells = 0;
                        tests = 0;
                        loaded = false;
                        k = e = 0;
                        ray = (int32_t)ri;
                        if (!(ok && dda_setup(fr, o0, o1, o2, d0, d1, d2, tmin, tmax, s))) {
                            if (ANY) {
                                occluded[ri] = 0;
                            } else {
                                hits.id[ri] = -1;
                                hits.t[ri] = INFINITY;
                                hits.u[ri] = 0.0f;
                                hits.v[ri] = 0.0f;
                            }
                            if (COUNT) {
                                a.counters[2ull * ri] = 0;
                                a.counters[2ull * ri + 1] = 0;
                            }
                            ray = -1;
                        }
                    }
                }
            }
        }
        if (__ballot(ray >= 0) == 0) {
            if (__ballot(!exhausted) == 0)
                break;
            continue;
        }
        // ---- drained queue: finish dense cells with the whole wave -----------
        if (const uint64_t hv0 = coop_lanes(ray >= 0, exhausted, k, e, a.coop)) {
            LaneRay r = {o0, o1, o2, d0, d1, d2, tmin, tmax, bound, k, e, id, best, bu, bv, tests};
            for (uint64_t hv = hv0; hv != 0; hv &= hv - 1)
                coop_cell((uint32_t)__builtin_ctzll(hv), ANY, trir, r);
            k = r.k;
            id = r.id;
            best = r.best;
            bu = r.bu;
            bv = r.bv;
            tests = r.tests;
        }
        if (ray < 0)
            continue;

        // ---- one work item ---------------------------------------------------
        bool finish = false;
        if (k < e) {
            const float4 cA = A, cB = B, cC = C;
            ++k;
            if (k < e) {
                A = trir[3ull * k];
                B = trir[3ull * k + 1];
                C = trir[3ull * k + 2];
            }
            if (COUNT)
                ++tests;
            if (ANY) {
                if (any_test(cA, cB, cC, o0, o1, o2, d0, d1, d2, tmin, tmax)) {
                    id = 1;
                    finish = true;
                }
            } else {
                closest_test(cA, cB, cC, o0, o1, o2, d0, d1, d2, tmin, tmax, bound, id, best, bu,
                             bv);
            }
        } else {
            if (loaded) {
                if (id >= 0)   // closest: hit in this cell; any: occluded
                    finish = true;
                else if (exit_t > tmax || !dda_step(s))
                    finish = true;
            }
            if (!finish) {
                if (COUNT)
                    ++cells;
                exit_t = tmin_f(tmin_f(s.tx, s.ty), s.tz);
                const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
                k = offsets[ci];
                e = offsets[ci + 1];
                bound = exit_t * kExitScale;
                loaded = true;
                if (k < e) {
                    A = trir[3ull * k];
                    B = trir[3ull * k + 1];
                    C = trir[3ull * k + 2];
                }
            }
        }
        if (finish) {
            const uint32_t ri = (uint32_t)ray;
            if (ANY) {
                occluded[ri] = id >= 0 ? 1 : 0;
            } else {
                hits.id[ri] = id;
                hits.t[ri] = best;
                hits.u[ri] = bu;
                hits.v[ri] = bv;
            }
            if (COUNT) {
                a.counters[2ull * ri] = cells;
                a.counters[2ull * ri + 1] = tests;
            }
            ray = -1;
        }
    }
}

__global__ void k_fill_u32_one(uint32_t* __restrict__ p)
{
    if (threadIdx.x == 0)
        p[0] = 0u;
}

// ---------------------------------------------------------------------------
// Batched multi-scene traversal (device work queue)
// ---------------------------------------------------------------------------

// map[b] = index of the task owning slots [64b, 64b+64).
__global__ __launch_bounds__(256) void k_task_map(const BatchTask* __restrict__ tasks,
                                                  uint32_t* __restrict__ map)
{
    const BatchTask& t = tasks[blockIdx.x];
    const uint32_t b0 = t.slot_begin >> 6, nb = t.slots >> 6;
    for (uint32_t j = threadIdx.x; j < nb; j += blockDim.x)
        map[b0 + j] = blockIdx.x;
}

// Same per-lane algorithm as k_trace_persistent, but every lane carries its
// own task (scene grid, ray and output buffers, closest or any hit): a refill
// hands out global slots; slot s belongs to task map[s / 64].
__global__ __launch_bounds__(256) void k_trace_batch(const BatchTask* __restrict__ tasks,
                                                     const uint32_t* __restrict__ map,
                                                     uint32_t total_slots,
                                                     uint32_t* __restrict__ slot_counter,
                                                     uint32_t refill, uint32_t coop)
{
    const uint32_t lane = threadIdx.x & 63u;
    int32_t ray = -1;
    bool exhausted = false;
    uint32_t task = 0;
    bool any = false;
    const uint32_t* offsets = nullptr;
    const float4* trir = nullptr;
    uint32_t nx = 0, ny = 0;
    float o0 = 0, o1 = 0, o2 = 0, d0 = 0, d1 = 0, d2 = 0, tmin = 0, tmax = 0;
    Dda s = {};
    uint32_t k = 0, e = 0;
    bool loaded = false;
    float exit_t = 0.0f, bound = 0.0f;
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    float4 A = {}, B = {}, C = {};

    for (;;) {
        const bool need = (ray < 0) && !exhausted;
        const uint64_t need_mask = __ballot(need);
        const uint64_t busy_mask = __ballot(ray >= 0);
        const uint32_t n_need = (uint32_t)__popcll(need_mask);
        if (need_mask != 0 && (n_need >= refill || busy_mask == 0)) {
            const uint32_t leader = (uint32_t)__builtin_ctzll(need_mask);
            uint32_t base = 0;
            if (lane == leader)
                base = atomicAdd(slot_counter, n_need);
            base = __builtin_amdgcn_readlane(base, leader);
            if (need) {
                const uint32_t slot = base + lanes_below(need_mask);
                if (slot >= total_slots) {
                    exhausted = true;
                } else {
                    task = map[slot >> 6];
                    const BatchTask& T = tasks[task];
                    bool valid;
                    const uint32_t ri = slot_ray(slot - T.slot_begin, T.image_width, T.n, valid);
                    if (valid) {
                        const DevFrame* __restrict__ fr = T.fr;
                        any = T.any != 0;
                        offsets = T.offsets;
                        trir = T.trir;
                        nx = (uint32_t)fr->n[0];
                        ny = (uint32_t)fr->n[1];
                        o0 = T.rays.ox[ri];
                        o1 = T.rays.oy[ri];
                        o2 = T.rays.oz[ri];
                        d0 = T.rays.dx[ri];
                        d1 = T.rays.dy[ri];
                        d2 = T.rays.dz[ri];
                        tmin = T.rays.tmin[ri];
                        tmax = T.rays.tmax[ri];
                        id = -1;
                        best = INFINITY;
                        bu = 0.0f;
                        bv = 0.0f;
                        loaded = false;
                        k = e = 0;
                        ray = (int32_t)ri;
                        if (!(grid_ok(fr) && dda_setup(fr, o0, o1, o2, d0, d1, d2, tmin, tmax, s))) {
                            if (any) {
                                T.occluded[ri] = 0;
                            } else {
                                T.hits.id[ri] = -1;
                                T.hits.t[ri] = INFINITY;
                                T.hits.u[ri] = 0.0f;
                                T.hits.v[ri] = 0.0f;
                            }
                            ray = -1;
                        }
                    }
                }
            }
        }
        if (__ballot(ray >= 0) == 0) {
            if (__ballot(!exhausted) == 0)
                break;
            continue;
        }
        if (const uint64_t hv0 = coop_lanes(ray >= 0, exhausted, k, e, coop)) {
            uint32_t tests = 0;
            LaneRay r = {o0, o1, o2, d0, d1, d2, tmin, tmax, bound, k, e, id, best, bu, bv, tests};
            for (uint64_t hv = hv0; hv != 0; hv &= hv - 1) {
                const uint32_t l = (uint32_t)__builtin_ctzll(hv);
                coop_cell(l, bcast_u(any ? 1u : 0u, l) != 0, trir, r);
            }
            k = r.k;
            id = r.id;
            best = r.best;
            bu = r.bu;
            bv = r.bv;
        }
        if (ray < 0)
            continue;

        bool finish = false;
        if (k < e) {
            const float4 cA = A, cB = B, cC = C;
            ++k;
            if (k < e) {
                A = trir[3ull * k];
                B = trir[3ull * k + 1];
                C = trir[3ull * k + 2];
            }
            if (any) {
                if (any_test(cA, cB, cC, o0, o1, o2, d0, d1, d2, tmin, tmax)) {
                    id = 1;
                    finish = true;
                }
            } else {
                closest_test(cA, cB, cC, o0, o1, o2, d0, d1, d2, tmin, tmax, bound, id, best, bu,
                             bv);
            }
        } else {
            if (loaded) {
                if (id >= 0)   // closest: hit in this cell; any: occluded
                    finish = true;
                else if (exit_t > tmax || !dda_step(s))
                    finish = true;
            }
            if (!finish) {
                exit_t = tmin_f(tmin_f(s.tx, s.ty), s.tz);
                const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
                k = offsets[ci];
                e = offsets[ci + 1];
                bound = exit_t * kExitScale;
                loaded = true;
                if (k < e) {
                    A = trir[3ull * k];
                    B = trir[3ull * k + 1];
                    C = trir[3ull * k + 2];
                }
            }
        }
        if (finish) {
            const BatchTask& T = tasks[task];
            const uint32_t ri = (uint32_t)ray;
            if (any) {
                T.occluded[ri] = id >= 0 ? 1 : 0;
            } else {
                T.hits.id[ri] = id;
                T.hits.t[ri] = best;
                T.hits.u[ri] = bu;
                T.hits.v[ri] = bv;
            }
            ray = -1;
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
static uint32_t ray_grid(uint32_t n, uint32_t img_w, uint32_t block)
{
    return (uint32_t)((slot_count(n, img_w) + block - 1) / block);
}

// ---------------------------------------------------------------------------
// Two-level grid: one ray per lane
// ---------------------------------------------------------------------------
//
// The top DDA as in the one-level kernels; a top cell whose record carries a
// sub-grid resolution is walked by a DDA of its own from the point where the
// ray entered the cell (the clipped entry, then each top cell's exit), with
// hits bounded by the sub-cell's exit; other cells test their references
// bounded by the cell's exit.  The first (sub-)cell holding a hit ends a
// closest-hit ray, the first triangle in (tmin, tmax] an any-hit ray.
// Counters: top cells plus sub-cells visited, triangle tests.

// Tests records [b, e) of `rec`: closest (bound) or any.  True when the ray is
// done (a hit in this cell, or an occluder).
template <bool ANY>
__device__ __forceinline__ bool trace2_cell(const float4* __restrict__ rec, uint32_t b,
                                            uint32_t e, float o0, float o1, float o2, float d0,
                                            float d1, float d2, float tmin, float tmax,
                                            float bound, int32_t& id, float& best, float& bu,
                                            float& bv, uint32_t& tests)
{
    return test_records<ANY>(rec, b, e, o0, o1, o2, d0, d1, d2, tmin, tmax, bound, id, best, bu,
                             bv, tests);
}

__device__ __forceinline__ int32_t sub_vox(float val, float bmin, float inv, int32_t r)
{
    float x = (val - bmin) * inv;
    x = x > 0.0f ? x : 0.0f;
    const float hi = (float)(r - 1);
    x = x < hi ? x : hi;
    return (int32_t)x;
}

// Per-ray constants of the two-level walk.
struct Ray2 {
    float o0, o1, o2, d0, d1, d2, i0, i1, i2, tmin, tmax;
    bool n0, n1, n2;
};

// Sub-cell sizes and their reciprocals by resolution, staged in LDS by every
// block of a two-level kernel: sc = tab[16 a + r - 1], 1/sc = tab[48 + 16 a +
// r - 1] (the build evaluates cell_a / r and 1 / that, the frame's own
// operations, once per frame).
__device__ __forceinline__ void stage_sub_tab(const TraceArgs& a, float* tab)
{
    if (threadIdx.x < 96u)
        tab[threadIdx.x] = a.sub_tab[threadIdx.x];
    __syncthreads();
}

// DDA state inside a top cell's sub-grid: current sub-cell (and its index si
// in the sub-cell arrays, which a step along y or z changes by sy or sz; by
// +-1 along x), next boundary crossings and their increments.
struct Sub2 {
    uint32_t si, sy, sz;
    int32_t qx, qy, qz, ox, oy, oz;
    float tx, ty, tz, dtx, dty, dtz;
};

// Entry into the sub-grid of the top cell at s (record rec) at t_enter.
__device__ __forceinline__ void sub_setup(const DevFrame* __restrict__ fr, const float* tab,
                                          const Ray2& r, const Dda& s, uint2 rec, float t_enter,
                                          Sub2& q)
{
    const uint32_t base = rec.x & 0x7fffffffu;
    const int32_t r0 = (int32_t)(rec.y & 255u);
    const int32_t r1 = (int32_t)((rec.y >> 8) & 255u);
    const int32_t r2 = (int32_t)((rec.y >> 16) & 255u);
    const float sb0 = (float)s.cx * fr->cell[0] + fr->bmin[0];
    const float sb1 = (float)s.cy * fr->cell[1] + fr->bmin[1];
    const float sb2 = (float)s.cz * fr->cell[2] + fr->bmin[2];
    const float sc0 = tab[r0 - 1], sc1 = tab[16 + r1 - 1], sc2 = tab[32 + r2 - 1];
    const float si0 = tab[48 + r0 - 1], si1 = tab[64 + r1 - 1], si2 = tab[80 + r2 - 1];
    q.qx = sub_vox(r.o0 + r.d0 * t_enter, sb0, si0, r0);
    q.qy = sub_vox(r.o1 + r.d1 * t_enter, sb1, si1, r1);
    q.qz = sub_vox(r.o2 + r.d2 * t_enter, sb2, si2, r2);
    q.ox = r.n0 ? -1 : r0;
    q.oy = r.n1 ? -1 : r1;
    q.oz = r.n2 ? -1 : r2;
    q.tx = ((float)(q.qx + (r.n0 ? 0 : 1)) * sc0 + sb0 - r.o0) * r.i0;
    q.ty = ((float)(q.qy + (r.n1 ? 0 : 1)) * sc1 + sb1 - r.o1) * r.i1;
    q.tz = ((float)(q.qz + (r.n2 ? 0 : 1)) * sc2 + sb2 - r.o2) * r.i2;
    q.dtx = fabsf(sc0 * r.i0);
    q.dty = fabsf(sc1 * r.i1);
    q.dtz = fabsf(sc2 * r.i2);
    q.si = base + (uint32_t)q.qx +
           (uint32_t)r0 * ((uint32_t)q.qy + (uint32_t)r1 * (uint32_t)q.qz);
    q.sy = r.n1 ? 0u - (uint32_t)r0 : (uint32_t)r0;
    q.sz = r.n2 ? 0u - (uint32_t)(r0 * r1) : (uint32_t)(r0 * r1);
}

// One sub-grid step (the top DDA's tie order); false when it leaves the grid
// (q is then left as it was).  Branch-free: the crossed axis selects which
// coordinate, crossing and index change.
__device__ __forceinline__ bool sub_step(const Ray2& r, Sub2& q)
{
    const bool ax = q.tx < q.ty && q.tx < q.tz;
    const bool ay = !ax && q.ty < q.tz;
    const bool az = !ax && !ay;
    const int32_t dx = r.n0 ? -1 : 1;
    const int32_t qx = ax ? q.qx + dx : q.qx;
    const int32_t qy = ay ? q.qy + (r.n1 ? -1 : 1) : q.qy;
    const int32_t qz = az ? q.qz + (r.n2 ? -1 : 1) : q.qz;
    if ((ax && qx == q.ox) || (ay && qy == q.oy) || (az && qz == q.oz))
        return false;
    q.qx = qx;
    q.qy = qy;
    q.qz = qz;
    const float tx = q.tx + q.dtx, ty = q.ty + q.dty, tz = q.tz + q.dtz;
    q.tx = ax ? tx : q.tx;
    q.ty = ay ? ty : q.ty;
    q.tz = az ? tz : q.tz;
    q.si += ax ? (uint32_t)dx : (ay ? q.sy : q.sz);
    return true;
}

// One non-empty top cell (index ci, coordinates of s, record rec) of the
// two-level walk, entered at t_enter and left at exit_t: its sub-grid from t_enter when it has one and
// the second level is valid, else its own references bounded by exit_t.
// Counts the sub-cells visited and the tests.  True when the walk ends in this
// cell: a hit (closest), an occluder (any, *occ set) or a sub-cell whose exit
// lies beyond tmax.
template <bool ANY>
__device__ __forceinline__ bool top_cell2(const TraceArgs& a, const DevFrame* __restrict__ fr,
                                          const float* tab, bool sub_ok, const Ray2& r,
                                          const Dda& s, uint2 rec, uint32_t ci, float t_enter,
                                          float exit_t, bool& occ, int32_t& id, float& best,
                                          float& bu, float& bv, uint32_t& cells,
                                          uint32_t& tests)
{
    const bool has_sub = (rec.x >> 31) != 0u;
    if (!(sub_ok && has_sub)) {
        // the top cell's own list (also for a sub-grid cell when the second
        // level is not valid)
        const uint32_t kb = has_sub ? a.offsets[ci] : rec.x;
        const uint32_t ke = has_sub ? a.offsets[ci + 1] : rec.y;
        if (trace2_cell<ANY>(a.trir, kb, ke, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin, r.tmax,
                             exit_t * kExitScale, id, best, bu, bv, tests)) {
            occ = ANY;
            return true;
        }
        return false;
    }
    Sub2 q;
    sub_setup(fr, tab, r, s, rec, t_enter, q);
    uint32_t kb = a.sub_offsets[q.si], ke = a.sub_offsets[q.si + 1];
    for (;;) {
        ++cells;
        const float sexit = tmin_f(tmin_f(q.tx, q.ty), q.tz);
        // the next sub-cell and its offsets, loaded before this one's tests
        // (unconditionally: a step out of the sub-grid leaves q.si in place)
        const uint32_t cb = kb, ce = ke;
        const bool more = sub_step(r, q);
        kb = a.sub_offsets[q.si];
        ke = a.sub_offsets[q.si + 1];
        if (trace2_cell<ANY>(a.sub_trir, cb, ce, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin,
                             r.tmax, sexit * kExitScale, id, best, bu, bv, tests)) {
            occ = ANY;
            return true;
        }
        if (sexit > r.tmax || !more)
            return sexit > r.tmax;
    }
}

// Set-up of a two-level ray: constants, the top DDA and the clipped entry.
__device__ __forceinline__ bool ray2_setup(const DevFrame* __restrict__ fr, const RaySoA& rays,
                                           uint32_t i, Ray2& r, Dda& s, float& t_enter)
{
    r.o0 = rays.ox[i];
    r.o1 = rays.oy[i];
    r.o2 = rays.oz[i];
    r.d0 = rays.dx[i];
    r.d1 = rays.dy[i];
    r.d2 = rays.dz[i];
    r.tmin = rays.tmin[i];
    r.tmax = rays.tmax[i];
    if (!(grid_ok(fr) && dda_setup(fr, r.o0, r.o1, r.o2, r.d0, r.d1, r.d2, r.tmin, r.tmax, s)))
        return false;
    r.i0 = 1.0f / r.d0;
    r.i1 = 1.0f / r.d1;
    r.i2 = 1.0f / r.d2;
    // the clipped entry, as dda_setup computes it
    t_enter = r.tmin;
    t_enter = tmax_f(t_enter, tmin_f((fr->bmin[0] - r.o0) * r.i0, (fr->bmax[0] - r.o0) * r.i0));
    t_enter = tmax_f(t_enter, tmin_f((fr->bmin[1] - r.o1) * r.i1, (fr->bmax[1] - r.o1) * r.i1));
    t_enter = tmax_f(t_enter, tmin_f((fr->bmin[2] - r.o2) * r.i2, (fr->bmax[2] - r.o2) * r.i2));
    r.n0 = __builtin_signbit(r.d0);
    r.n1 = __builtin_signbit(r.d1);
    r.n2 = __builtin_signbit(r.d2);
    return true;
}

// The walk over the top grid goes item by item.  An item is a cell of the
// grid; for an empty cell with skip radius rr > 0 it is the cells ahead of it
// within rr on every axis (all empty), left in one step.  Each item's exit is
// its first boundary plane crossed, from the plane formula of the DDA set-up
// (t = ((float)k cell_a + bmin_a - o_a) / d_a), in the DDA's tie order.

// Skip radius of a cell record: {r, r} marks an empty cell.
__device__ __forceinline__ int32_t skip_radius(uint2 rec, bool& empty)
{
    empty = rec.x == rec.y;
    return empty && rec.x <= (uint32_t)kSkipMax ? (int32_t)rec.x : 0;
}

// Exit t of the item at the walk's cell s with radius rr; axis: the axis
// crossed.
__device__ __forceinline__ float item_exit(const DevFrame* __restrict__ fr, const Ray2& r,
                                           const Dda& s, int32_t rr, int& axis)
{
    const float t0 =
        ((float)(r.n0 ? s.cx - rr : s.cx + rr + 1) * fr->cell[0] + fr->bmin[0] - r.o0) * r.i0;
    const float t1 =
        ((float)(r.n1 ? s.cy - rr : s.cy + rr + 1) * fr->cell[1] + fr->bmin[1] - r.o1) * r.i1;
    const float t2 =
        ((float)(r.n2 ? s.cz - rr : s.cz + rr + 1) * fr->cell[2] + fr->bmin[2] - r.o2) * r.i2;
    if (t0 < t1 && t0 < t2) {
        axis = 0;
        return t0;
    }
    if (t1 < t2) {
        axis = 1;
        return t1;
    }
    axis = 2;
    return t2;
}

// v kept within the rr cells ahead of c (behind, for a negative direction).
__device__ __forceinline__ int32_t skip_cell(int32_t v, int32_t c, int32_t rr, bool neg)
{
    if (neg) {
        v = v < c ? v : c;
        return v > c - rr ? v : c - rr;
    }
    v = v > c ? v : c;
    return v < c + rr ? v : c + rr;
}

// Step to the next item across the current one's exit (axis, at t = e);
// false when that leaves the grid.  After a skip (rr > 0) the other two
// axes take the cells of the exit point, kept within the skipped cells: no
// axis ever moves backwards, and the crossed one moves by rr + 1.
__device__ __forceinline__ bool item_advance(const DevFrame* __restrict__ fr, const Ray2& r,
                                             Dda& s, int32_t rr, int axis, float e)
{
    if (axis == 0) {
        const int32_t c = r.n0 ? s.cx - rr - 1 : s.cx + rr + 1;
        if (c < 0 || c >= fr->n[0])
            return false;
        s.cx = c;
    } else if (axis == 1) {
        const int32_t c = r.n1 ? s.cy - rr - 1 : s.cy + rr + 1;
        if (c < 0 || c >= fr->n[1])
            return false;
        s.cy = c;
    } else {
        const int32_t c = r.n2 ? s.cz - rr - 1 : s.cz + rr + 1;
        if (c < 0 || c >= fr->n[2])
            return false;
        s.cz = c;
    }
    if (rr > 0) {
        if (axis != 0)
            s.cx = skip_cell(vox(r.o0 + r.d0 * e, fr->bmin[0], fr->inv[0], fr->n[0]), s.cx, rr,
                             r.n0);
        if (axis != 1)
            s.cy = skip_cell(vox(r.o1 + r.d1 * e, fr->bmin[1], fr->inv[1], fr->n[1]), s.cy, rr,
                             r.n1);
        if (axis != 2)
            s.cz = skip_cell(vox(r.o2 + r.d2 * e, fr->bmin[2], fr->inv[2], fr->n[2]), s.cz, rr,
                             r.n2);
    }
    return true;
}

template <bool ANY>
__global__ __launch_bounds__(128) void k_trace2(TraceArgs a, RaySoA rays, HitSoA hits,
                                                uint8_t* __restrict__ occluded)
{
    __shared__ float tab[96];
    stage_sub_tab(a, tab);
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid;
    const uint32_t i = trace_ray(a, slot, valid);
    if (!valid)
        return;
    const DevFrame* __restrict__ fr = a.fr;
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    bool occ = false;
    uint32_t cells = 0, tests = 0;
    // Over its capacity the second level's records are not complete: the top
    // grid alone then.
    const bool sub_ok = a.g2->status == 0;
    Ray2 r;
    Dda s;
    float t_enter;
    if (ray2_setup(fr, rays, i, r, s, t_enter)) {
        const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
        uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
        uint2 rec = a.cellrec[ci];
        for (;;) {
            ++cells;
            bool empty;
            const int32_t rr = skip_radius(rec, empty);
            int axis;
            const float e = item_exit(fr, r, s, rr, axis);
            // the next item and its record, loaded before this cell's tests
            Dda ns = s;
            const bool more = !(e > r.tmax) && item_advance(fr, r, ns, rr, axis, e);
            const uint32_t nci =
                (uint32_t)ns.cx + nx * ((uint32_t)ns.cy + ny * (uint32_t)ns.cz);
            // ns is inside the grid whether or not the walk goes on
            const uint2 nrec = a.cellrec[nci];
            if ((!empty && top_cell2<ANY>(a, fr, tab, sub_ok, r, s, rec, ci, t_enter, e, occ, id,
                                          best, bu, bv, cells, tests)) ||
                !more)
                break;
            t_enter = e;
            s = ns;
            ci = nci;
            rec = nrec;
        }
    }
    if (ANY) {
        occluded[i] = occ ? 1 : 0;
    } else {
        hits.id[i] = id;
        hits.t[i] = best;
        hits.u[i] = bu;
        hits.v[i] = bv;
    }
    if (a.counters) {
        a.counters[2ull * i] = cells;
        a.counters[2ull * i + 1] = tests;
    }
}

// `count` items of the walk, each after checking its exit against tmax (the
// serial walk's stop); t_enter receives the exit of the last item left.
__device__ __forceinline__ bool walk_skip(const TraceArgs& a, const DevFrame* __restrict__ fr,
                                          const Ray2& r, Dda& s, uint32_t count, float& t_enter)
{
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    for (uint32_t q = 0; q < count; ++q) {
        const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
        bool empty;
        const int32_t rr = skip_radius(a.cellrec[ci], empty);
        int axis;
        const float e = item_exit(fr, r, s, rr, axis);
        if (e > r.tmax || !item_advance(fr, r, s, rr, axis, e))
            return false;
        t_enter = e;
    }
    return true;
}

// Several lanes per ray (2..16, `lanes`): lane j of a ray's group walks the
// ray's top cells j, j + lanes, ... of the serial order, each one whole (its
// sub-grid included).  After every round the group's earliest cell that ends
// the walk (a hit, an occluder, a sub-cell exit beyond tmax) decides: its
// result is the serial result, and the cells after it in that round are not
// counted.  Latency of a long ray drops about `lanes`-fold; the sparse
// secondary stages are bound by it.
template <bool ANY>
__global__ __launch_bounds__(256) void k_trace2_split(TraceArgs a, RaySoA rays, HitSoA hits,
                                                      uint8_t* __restrict__ occluded,
                                                      uint32_t lanes)
{
    __shared__ float tab[96];
    stage_sub_tab(a, tab);
    const uint32_t shift = (uint32_t)__builtin_ctz(lanes);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid;
    const uint32_t i = trace_ray(a, t >> shift, valid);
    const DevFrame* __restrict__ fr = a.fr;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t j = lane & (lanes - 1u);
    const uint32_t g0 = lane - j;
    const uint64_t gmask = (lanes >= 64u ? ~0ull : ((1ull << lanes) - 1ull)) << g0;
    const bool sub_ok = a.g2->status == 0;
    const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    bool occ = false;
    uint32_t cells = 0, tests = 0;
    Ray2 r = {};
    Dda s = {};
    float t_enter = 0.0f;
    bool alive = valid && ray2_setup(fr, rays, i, r, s, t_enter) && walk_skip(a, fr, r, s, j, t_enter);
    bool winner = false;
    while (__ballot(alive) != 0) {
        const uint32_t c0 = cells, t0 = tests;
        bool stop = false;
        int32_t rr = 0;
        int axis = 0;
        float e = 0.0f;
        if (alive) {
            ++cells;
            const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
            const uint2 rec = a.cellrec[ci];
            bool empty;
            rr = skip_radius(rec, empty);
            e = item_exit(fr, r, s, rr, axis);
            stop = !empty && top_cell2<ANY>(a, fr, tab, sub_ok, r, s, rec, ci, t_enter, e, occ,
                                            id, best, bu, bv, cells, tests);
        }
        // The group's earliest cell that ends the walk ends the ray; cells
        // after it (higher lanes of this round) are not part of the serial walk.
        const uint64_t sm = __ballot(alive && stop) & gmask;
        if (sm != 0) {
            const uint32_t jw = (uint32_t)__builtin_ctzll(sm) - g0;
            if (alive && j > jw) {
                cells = c0;
                tests = t0;
            }
            winner = alive && j == jw && (ANY ? occ : id >= 0);
            alive = false;
        } else if (alive) {
            alive = !(e > r.tmax) && item_advance(fr, r, s, rr, axis, e);   // NaN tmax: no stop
            if (alive) {
                t_enter = e;
                alive = walk_skip(a, fr, r, s, lanes - 1u, t_enter);
            }
        }
    }
    const bool group_hit = (__ballot(winner) & gmask) != 0;
    uint32_t sc = 0, st = 0;
    if (a.counters) {
        sc = group_sum(cells, lanes);
        st = group_sum(tests, lanes);
    }
    if (!valid)
        return;
    if (ANY) {
        if (j == 0)
            occluded[i] = group_hit ? 1 : 0;
    } else if (winner || (j == 0 && !group_hit)) {
        hits.id[i] = winner ? id : -1;
        hits.t[i] = winner ? best : INFINITY;
        hits.u[i] = winner ? bu : 0.0f;
        hits.v[i] = winner ? bv : 0.0f;
    }
    if (a.counters && j == 0) {
        a.counters[2ull * i] = sc;
        a.counters[2ull * i + 1] = st;
    }
}

hipError_t launch_trace2(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                         uint8_t* occluded, uint32_t lanes, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    if (lanes <= 1) {
        const uint32_t g = ray_grid(a.n, a.image_width, 128);
        if (occluded)
            hipLaunchKernelGGL(k_trace2<true>, dim3(g), dim3(128), 0, s, a, rays, hits, occluded);
        else
            hipLaunchKernelGGL(k_trace2<false>, dim3(g), dim3(128), 0, s, a, rays, hits,
                               occluded);
        return hipGetLastError();
    }
    if (lanes > 16 || (lanes & (lanes - 1)) != 0)
        return hipErrorInvalidValue;
    const uint64_t threads = slot_count(a.n, a.image_width) * lanes;
    if (threads > 0xffffffffull - 256)
        return hipErrorInvalidValue;
    const dim3 g((uint32_t)((threads + 255) / 256));
    if (occluded)
        hipLaunchKernelGGL(k_trace2_split<true>, g, dim3(256), 0, s, a, rays, hits, occluded,
                           lanes);
    else
        hipLaunchKernelGGL(k_trace2_split<false>, g, dim3(256), 0, s, a, rays, hits, occluded,
                           lanes);
    return hipGetLastError();
}

hipError_t launch_trace_closest(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                                uint32_t block, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    const uint32_t g = ray_grid(a.n, a.image_width, block);
    if (a.counters && a.pool)
        hipLaunchKernelGGL((k_trace_closest<true, true>), dim3(g), dim3(block), 0, s, a, rays, hits);
    else if (a.counters)
        hipLaunchKernelGGL((k_trace_closest<true, false>), dim3(g), dim3(block), 0, s, a, rays,
                           hits);
    else if (a.pool)
        hipLaunchKernelGGL((k_trace_closest<false, true>), dim3(g), dim3(block), 0, s, a, rays,
                           hits);
    else
        hipLaunchKernelGGL((k_trace_closest<false, false>), dim3(g), dim3(block), 0, s, a, rays,
                           hits);
    return hipGetLastError();
}

hipError_t launch_trace_any(const TraceArgs& a, const RaySoA& rays, uint8_t* occluded,
                            uint32_t block, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    const uint32_t g = ray_grid(a.n, a.image_width, block);
    if (a.counters && a.pool)
        hipLaunchKernelGGL((k_trace_any<true, true>), dim3(g), dim3(block), 0, s, a, rays, occluded);
    else if (a.counters)
        hipLaunchKernelGGL((k_trace_any<true, false>), dim3(g), dim3(block), 0, s, a, rays,
                           occluded);
    else if (a.pool)
        hipLaunchKernelGGL((k_trace_any<false, true>), dim3(g), dim3(block), 0, s, a, rays,
                           occluded);
    else
        hipLaunchKernelGGL((k_trace_any<false, false>), dim3(g), dim3(block), 0, s, a, rays,
                           occluded);
    return hipGetLastError();
}

hipError_t launch_trace_split(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                              uint8_t* occluded, bool any, uint32_t lanes, uint32_t block,
                              hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    if (lanes < 2 || lanes > 16 || (lanes & (lanes - 1)) != 0)
        return hipErrorInvalidValue;
    const uint64_t threads = slot_count(a.n, a.image_width) * lanes;
    if (threads > 0xffffffffull - block)
        return hipErrorInvalidValue;
    const dim3 g((uint32_t)((threads + block - 1) / block));
    if (any) {
        if (a.counters)
            hipLaunchKernelGGL((k_trace_split<true, true>), g, dim3(block), 0, s, a, rays, hits,
                               occluded, lanes);
        else
            hipLaunchKernelGGL((k_trace_split<true, false>), g, dim3(block), 0, s, a, rays, hits,
                               occluded, lanes);
    } else {
        if (a.counters)
            hipLaunchKernelGGL((k_trace_split<false, true>), g, dim3(block), 0, s, a, rays, hits,
                               occluded, lanes);
        else
            hipLaunchKernelGGL((k_trace_split<false, false>), g, dim3(block), 0, s, a, rays, hits,
                               occluded, lanes);
    }
    return hipGetLastError();
}

// Resident blocks on the device for a persistent kernel at a block size.
template <class K>
static uint32_t resident_blocks(K kernel, uint32_t block)
{
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        (void)hipGetDevice(&dev);
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            cus <= 0)
            cus = 256;
    }
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, (int)block, 0) !=
            hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    return (uint32_t)cus * (uint32_t)per_cu;
}

hipError_t launch_trace_persistent(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                                   uint8_t* occluded, bool any, uint32_t* slot_counter,
                                   uint32_t block, uint32_t refill, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    const uint64_t n_slots = slot_count(a.n, a.image_width);
    if (n_slots > 0xffffffffull)
        return hipErrorInvalidValue;
    refill = refill < 1 ? 1 : (refill > 64 ? 64 : refill);
    hipLaunchKernelGGL(k_fill_u32_one, dim3(1), dim3(64), 0, s, slot_counter);
    const uint32_t need = (uint32_t)((n_slots + block - 1) / block);
#define GRT_LAUNCH_P(ANYV, CNTV)                                                                \
    do {                                                                                        \
        auto kern = k_trace_persistent<ANYV, CNTV>;                                             \
        uint32_t g = resident_blocks(kern, block);                                              \
        g = g < need ? g : need;                                                                \
        hipLaunchKernelGGL(kern, dim3(g), dim3(block), 0, s, a, rays, hits, occluded,           \
                           slot_counter, (uint32_t)n_slots, refill);                            \
    } while (0)
    if (any) {
        if (a.counters)
            GRT_LAUNCH_P(true, true);
        else
            GRT_LAUNCH_P(true, false);
    } else {
        if (a.counters)
            GRT_LAUNCH_P(false, true);
        else
            GRT_LAUNCH_P(false, false);
    }
#undef GRT_LAUNCH_P
    return hipGetLastError();
}

hipError_t launch_task_map(const BatchTask* tasks, uint32_t n_tasks, uint32_t* map, hipStream_t s)
{
    if (n_tasks == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_task_map, dim3(n_tasks), dim3(256), 0, s, tasks, map);
    return hipGetLastError();
}

hipError_t launch_trace_batch(const BatchTask* tasks, const uint32_t* map, uint32_t total_slots,
                              uint32_t* slot_counter, int mode, uint32_t block,
                              uint32_t refill, uint32_t lanes, uint32_t coop, hipStream_t s)
{
    if (total_slots == 0)
        return hipSuccess;
    if (mode == GRT_TRACE_SPLIT) {
        if (lanes < 2 || lanes > 16 || (lanes & (lanes - 1)) != 0)
            return hipErrorInvalidValue;
        const uint64_t threads = (uint64_t)total_slots * lanes;
        if (threads > 0xffffffffull - block)
            return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_trace_batch_split, dim3((uint32_t)((threads + block - 1) / block)),
                           dim3(block), 0, s, tasks, map, total_slots, coop, lanes);
        return hipGetLastError();
    }
    if (mode != GRT_TRACE_PERSISTENT) {
        const dim3 g((total_slots + block - 1) / block);
        if (mode == GRT_TRACE_POOLED)
            hipLaunchKernelGGL(k_trace_batch_simple<true>, g, dim3(block), 0, s, tasks, map,
                               total_slots, coop);
        else
            hipLaunchKernelGGL(k_trace_batch_simple<false>, g, dim3(block), 0, s, tasks, map,
                               total_slots, coop);
        return hipGetLastError();
    }
    refill = refill < 1 ? 1 : (refill > 64 ? 64 : refill);
    hipLaunchKernelGGL(k_fill_u32_one, dim3(1), dim3(64), 0, s, slot_counter);
    uint32_t g = resident_blocks(k_trace_batch, block);
    const uint32_t need = (total_slots + block - 1) / block;
    g = g < need ? g : need;
    hipLaunchKernelGGL(k_trace_batch, dim3(g), dim3(block), 0, s, tasks, map, total_slots,
                       slot_counter, refill, coop);
    return hipGetLastError();
}

}  // namespace grt
