// grt_kdtree.hip -- kd-tree traversal on gfx950 (wave64), one ray per lane.
//
// Reference semantics (KdTreeAccStruct.cpp:122-240):
//   - the ray is clipped to the tree's box (AabbIntersection::clipRay);
//   - descending, the plane distance d = (split - o[axis]) / dir[axis] picks
//     the far child alone when d < tnear, the near child alone when d > tfar,
//     else the near child with tfar = min(d, tfar) while the far child is
//     pushed with [max(d, tnear), tfar] (strict comparisons, so planes with
//     triangles in them are entered from both sides; the selects keep NaN
//     distances -- origin on the plane, dir[axis] = 0 -- on the whole
//     interval);
//   - a leaf's triangles are tested with the grid's hit window, the leaf's
//     tfar taking the place of the cell exit; the first leaf with a hit ends a
//     closest-hit ray, any triangle in (tmin, tmax] ends an any-hit ray;
//   - otherwise the traversal continues from the top of the stack.
// Node layout and records: grt_kdtree_build.cpp.  The stack holds at most one entry
// per tree level (entries from one descent path have distinct depths), and
// the builder caps the depth at kKdMaxDepth.
#include "grt_launch.h"
#include "grt_mt.h"

namespace grt {

__global__ __launch_bounds__(256) void k_kd_gather(const uint32_t* __restrict__ elements,
                                                   uint32_t n, const float4* __restrict__ tri,
                                                   float4* __restrict__ rec)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint64_t t = elements[i];
    rec[3ull * i] = tri[3 * t];
    rec[3ull * i + 1] = tri[3 * t + 1];
    rec[3ull * i + 2] = tri[3 * t + 2];
}

template <bool ANY>
__global__ __launch_bounds__(128) void k_kd_trace(KdArgs a, RaySoA rays, HitSoA hits,
                                                  uint8_t* __restrict__ occluded)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    bool valid;
    const uint32_t i = slot_ray(slot, a.image_width, a.n, valid);
    if (!valid)
        return;
    const float o0 = rays.ox[i], o1 = rays.oy[i], o2 = rays.oz[i];
    const float d0 = rays.dx[i], d1 = rays.dy[i], d2 = rays.dz[i];
    const float tmin = rays.tmin[i], tmax = rays.tmax[i];
    const float i0 = 1.0f / d0, i1 = 1.0f / d1, i2 = 1.0f / d2;
    float tn = tmin, tf = tmax;
    {
        const float p = (a.bmin[0] - o0) * i0, q = (a.bmax[0] - o0) * i0;
        tn = tmax_f(tn, tmin_f(p, q));
        tf = tmin_f(tf, tmax_f(p, q));
    }
    {
        const float p = (a.bmin[1] - o1) * i1, q = (a.bmax[1] - o1) * i1;
        tn = tmax_f(tn, tmin_f(p, q));
        tf = tmin_f(tf, tmax_f(p, q));
    }
    {
        const float p = (a.bmin[2] - o2) * i2, q = (a.bmax[2] - o2) * i2;
        tn = tmax_f(tn, tmin_f(p, q));
        tf = tmin_f(tf, tmax_f(p, q));
    }
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    bool occ = false;
    uint32_t leaves = 0, tests = 0;
    const bool n0 = __builtin_signbit(d0), n1 = __builtin_signbit(d1), n2 = __builtin_signbit(d2);
    if (tn <= tf) {
        uint32_t st_node[kKdMaxDepth];
        float st_tn[kKdMaxDepth], st_tf[kKdMaxDepth];
        uint32_t sp = 0, node = 0;
        for (;;) {
            uint2 nd = a.nodes[node];
            while (!(nd.y & 0x80000000u)) {
                const uint32_t axis = nd.y & 3u;
                const uint32_t left = nd.y >> 2;
                const float oa = axis == 0 ? o0 : (axis == 1 ? o1 : o2);
                const float ia = axis == 0 ? i0 : (axis == 1 ? i1 : i2);
                const bool neg = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
                const float dist = (__uint_as_float(nd.x) - oa) * ia;
                const uint32_t front = left + (neg ? 1u : 0u);
                const uint32_t back = left + (neg ? 0u : 1u);
                if (dist < tn) {
                    node = back;
                } else if (dist > tf) {
                    node = front;
                } else {
                    if (sp < kKdMaxDepth) {
                        st_node[sp] = back;
                        st_tn[sp] = dist > tn ? dist : tn;
                        st_tf[sp] = tf;
                        ++sp;
                    }
                    node = front;
                    tf = dist < tf ? dist : tf;
                }
                nd = a.nodes[node];
            }
            ++leaves;
            const uint32_t first = nd.x, end = nd.x + (nd.y & 0x7fffffffu);
            if (ANY) {
                for (uint32_t k = first; k < end; ++k) {
                    ++tests;
                    const float4 A = a.records[3ull * k], B = a.records[3ull * k + 1],
                                 C = a.records[3ull * k + 2];
                    if (any_test(A, B, C, o0, o1, o2, d0, d1, d2, tmin, tmax)) {
                        occ = true;
                        break;
                    }
                }
                if (occ)
                    break;
            } else {
                const float bound = tf * kExitScale;
                for (uint32_t k = first; k < end; ++k) {
                    ++tests;
                    const float4 A = a.records[3ull * k], B = a.records[3ull * k + 1],
                                 C = a.records[3ull * k + 2];
                    closest_test(A, B, C, o0, o1, o2, d0, d1, d2, tmin, tmax, bound, id, best, bu,
                                 bv);
                }
                if (id >= 0)
                    break;
            }
            if (sp == 0)
                break;
            --sp;
            node = st_node[sp];
            tn = st_tn[sp];
            tf = st_tf[sp];
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
        a.counters[2ull * i] = leaves;
        a.counters[2ull * i + 1] = tests;
    }
}

hipError_t launch_kd_gather(const uint32_t* elements, uint32_t n, const float4* tri, float4* rec,
                            hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_kd_gather, dim3(ceil_div(n, 256)), dim3(256), 0, s, elements, n, tri,
                       rec);
    return hipGetLastError();
}

hipError_t launch_kd_trace(const KdArgs& a, const RaySoA& rays, const HitSoA& hits,
                           uint8_t* occluded, uint32_t block, hipStream_t s)
{
    (void)block;
    const uint64_t slots = slot_count(a.n, a.image_width);
    const uint32_t nb = ceil_div(slots, 128);
    if (nb == 0)
        return hipSuccess;
    if (occluded)
        hipLaunchKernelGGL(k_kd_trace<true>, dim3(nb), dim3(128), 0, s, a, rays, hits, occluded);
    else
        hipLaunchKernelGGL(k_kd_trace<false>, dim3(nb), dim3(128), 0, s, a, rays, hits, occluded);
    return hipGetLastError();
}

}  // namespace grt
