// grt_mt.h -- Moller-Trumbore ray-triangle tests on the 48-byte triangle
// records, shared by the grid and kd-tree traversals so both evaluate the
// specification's binary32 operation sequence (RayTracing_kernel.cu:56-89, no
// determinant epsilon; the hit window of UniformGridAccStruct.cpp:362-378).
#pragma once

#include "grt_common.h"

namespace grt {

// ---------------------------------------------------------------------------
// Moller-Trumbore
// ---------------------------------------------------------------------------

// Split in two so the IEEE division runs only when some lane of the wave may
// hit.  mt_pre evaluates the numerators in the specification's order;
// mt_maybe is a conservative filter that returns false only when mt_exact is
// certainly a miss for the window (tmin, ub], ub an upper bound on accepted t
// (tmax, cell exit bound, best so far).  With r = RN(1/det), each of
// u = RN(un*r), v, t is within 2^-23 relative of un/det, vn/det, tn/det (no
// underflow inside the guarded magnitude ranges), so the margins 2^-40, 2^-16
// and 2^-20 keep every rejection on the miss side of the rounded comparison.
struct MtPre {
    float det, un, vn, tn;
};

__device__ __forceinline__ MtPre mt_pre(const float4 A, const float4 B, const float4 C, float o0,
                                        float o1, float o2, float d0, float d1, float d2)
{
    const float v0x = A.x, v0y = A.y, v0z = A.z;
    const float e1x = A.w, e1y = B.x, e1z = B.y;
    const float e2x = B.z, e2y = B.w, e2z = C.x;
    const float px = __builtin_fmaf(d1, e2z, -(d2 * e2y));
    const float py = __builtin_fmaf(d2, e2x, -(d0 * e2z));
    const float pz = __builtin_fmaf(d0, e2y, -(d1 * e2x));
    MtPre m;
    m.det = dot3(px, py, pz, e1x, e1y, e1z);
    const float tvx = o0 - v0x, tvy = o1 - v0y, tvz = o2 - v0z;
    const float qx = __builtin_fmaf(tvy, e1z, -(tvz * e1y));
    const float qy = __builtin_fmaf(tvz, e1x, -(tvx * e1z));
    const float qz = __builtin_fmaf(tvx, e1y, -(tvy * e1x));
    m.un = dot3(px, py, pz, tvx, tvy, tvz);
    m.vn = dot3(qx, qy, qz, d0, d1, d2);
    m.tn = dot3(qx, qy, qz, e2x, e2y, e2z);
    return m;
}

__device__ __forceinline__ bool mt_maybe(const MtPre& m, float tmin, float ub)
{
    const float ad = fabsf(m.det);
    if (!(ad >= 0x1p-100f && ad <= 0x1p100f))
        return true;
    const bool neg = m.det < 0.0f;
    const float us = neg ? -m.un : m.un;
    const float vs = neg ? -m.vn : m.vn;
    const float ts = neg ? -m.tn : m.tn;
    const float thr = ad * 0x1p-40f;
    if (us < -thr || vs < -thr)
        return false;
    if ((us + vs) > ad * (1.0f + 0x1p-16f))
        return false;
    if (ub >= 0x1p-20f && ub <= 0x1p40f && ts > (ub * ad) * (1.0f + 0x1p-20f))
        return false;
    if (tmin == 0.0f) {
        if (ts <= 0.0f)
            return false;
    } else if (tmin >= 0x1p-20f && tmin <= 0x1p20f && ts <= (tmin * ad) * (1.0f - 0x1p-20f)) {
        return false;
    }
    return true;
}

// The specification's division and inside test on the numerators.
__device__ __forceinline__ bool mt_exact(const MtPre& m, float& t, float& u, float& v)
{
    const float inv = 1.0f / m.det;
    u = m.un * inv;
    v = m.vn * inv;
    t = m.tn * inv;
    return (u >= 0.0f) && (v >= 0.0f) && ((u + v) <= 1.0f);
}

// Closest-hit candidate update for one record: the hit window of
// intersectTriangles -- (tmin, tmax], at most the cell exit bound, nearer than
// the best so far.  The triangle id is stored in the record's C.y.
__device__ __forceinline__ void closest_test(const float4 A, const float4 B, const float4 C,
                                             float o0, float o1, float o2, float d0, float d1,
                                             float d2, float tmin, float tmax, float bound,
                                             int32_t& id, float& best, float& bu, float& bv)
{
    const MtPre m = mt_pre(A, B, C, o0, o1, o2, d0, d1, d2);
    if (mt_maybe(m, tmin, tmin_f(tmin_f(tmax, bound), best))) {
        float t, u, v;
        if (mt_exact(m, t, u, v) && t > tmin && t <= tmax && t <= bound && t < best) {
            best = t;
            id = (int32_t)__float_as_uint(C.y);
            bu = u;
            bv = v;
        }
    }
}

__device__ __forceinline__ bool any_test(const float4 A, const float4 B, const float4 C, float o0,
                                         float o1, float o2, float d0, float d1, float d2,
                                         float tmin, float tmax)
{
    const MtPre m = mt_pre(A, B, C, o0, o1, o2, d0, d1, d2);
    if (!mt_maybe(m, tmin, tmax))
        return false;
    float t, u, v;
    return mt_exact(m, t, u, v) && t > tmin && t <= tmax;
}

// ---------------------------------------------------------------------------
// A list of records
// ---------------------------------------------------------------------------

// Read-only view in the constant address space: with a wave-uniform index the
// compiler reads it through the scalar unit (loads only).
typedef const __attribute__((address_space(4))) float* KF4;

__device__ __forceinline__ float4 kload4(KF4 p, uint64_t i)
{
    return make_float4(p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]);
}

__device__ __forceinline__ float4 rload(const float4* __restrict__ p, uint64_t i)
{
    return p[i];
}

__device__ __forceinline__ float4 rload(KF4 p, uint64_t i)
{
    return kload4(p, i);
}

// Tests the 48-byte records [b, e) of rec (a per-lane pointer, or the
// constant-space view for wave-uniform b, e): closest (hits bounded by
// `bound`) or any.  Counts the tests, an occluder included.  True when an
// any-hit ray found its occluder, or a closest-hit ray holds a hit.  Two
// record sets take turns: the next record loads while the current one is
// tested, unconditionally (the index stays within [b, e), so the last record
// may load twice), and nothing is copied between iterations -- a load under a
// per-lane condition would make every test wait for it at the join.
template <bool ANY, typename P>
__device__ __forceinline__ bool test_records(P rec, uint32_t b, uint32_t e, float o0, float o1,
                                             float o2, float d0, float d1, float d2, float tmin,
                                             float tmax, float bound, int32_t& id, float& best,
                                             float& bu, float& bv, uint32_t& tests)
{
    if (b >= e)
        return false;
    const uint32_t last = e - 1u;
    float4 A0 = rload(rec, 3ull * b), B0 = rload(rec, 3ull * b + 1),
           C0 = rload(rec, 3ull * b + 2);
    for (uint32_t k = b; k < e; k += 2) {
        const uint64_t q1 = 3ull * (k + 1u < last ? k + 1u : last);
        const float4 A1 = rload(rec, q1), B1 = rload(rec, q1 + 1), C1 = rload(rec, q1 + 2);
        ++tests;
        if (ANY) {
            if (any_test(A0, B0, C0, o0, o1, o2, d0, d1, d2, tmin, tmax))
                return true;
        } else {
            closest_test(A0, B0, C0, o0, o1, o2, d0, d1, d2, tmin, tmax, bound, id, best, bu, bv);
        }
        const uint64_t q2 = 3ull * (k + 2u < last ? k + 2u : last);
        A0 = rload(rec, q2);
        B0 = rload(rec, q2 + 1);
        C0 = rload(rec, q2 + 2);
        if (k + 1u < e) {
            ++tests;
            if (ANY) {
                if (any_test(A1, B1, C1, o0, o1, o2, d0, d1, d2, tmin, tmax))
                    return true;
            } else {
                closest_test(A1, B1, C1, o0, o1, o2, d0, d1, d2, tmin, tmax, bound, id, best, bu,
                             bv);
            }
        }
    }
    return !ANY && id >= 0;
}

}  // namespace grt
