// grt_instance.hip -- instance-level tracing (IAccStruct::traceNearestInstance,
// IAccStruct.cpp:26-73; Transform::inverseTransform, Transform.h:31-40).
//
// A world ray is taken into each instance's object space in turn: the origin
// through the inverse matrix, the direction through its 3x3 part without
// renormalising, so a hit's t is the same in both spaces.  Each instance's
// grid is traced with the interval cut to end just below the best hit so far,
// so only a strictly nearer hit replaces it and an earlier instance wins a tie
// (the reference passes its best distance into traceNearestGeometry).  The
// traversal itself is the single-scene kernel of the context's schedule.
#include "grt_launch.h"

namespace grt {

// The largest binary32 value below b (b finite).
__device__ __forceinline__ float float_below(float b)
{
    const int32_t bits = __float_as_int(b);
    if ((bits & 0x7fffffff) == 0)
        return -__int_as_float(1);   // below +-0: the negative denormal closest to 0
    return __int_as_float(bits > 0 ? bits - 1 : bits + 1);
}

__global__ __launch_bounds__(256) void k_instance_init(uint32_t n, HitSoA hits,
                                                       int32_t* __restrict__ instance)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    hits.id[i] = -1;
    hits.t[i] = INFINITY;
    hits.u[i] = 0.0f;
    hits.v[i] = 0.0f;
    instance[i] = -1;
}

// Object-space rays of one instance: o' = A o + a, d' = A d with
// x' = ((A00 x + A01 y) + A02 z) + a0; interval (tmin, min(tmax, below(best))].
__global__ __launch_bounds__(256) void k_instance_rays(uint32_t n, RaySoA in,
                                                       const float* __restrict__ best_t,
                                                       Affine inv, RaySoAOut out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const float* a = inv.m;
    const float ox = in.ox[i], oy = in.oy[i], oz = in.oz[i];
    const float dx = in.dx[i], dy = in.dy[i], dz = in.dz[i];
    out.ox[i] = ((a[0] * ox + a[1] * oy) + a[2] * oz) + a[3];
    out.oy[i] = ((a[4] * ox + a[5] * oy) + a[6] * oz) + a[7];
    out.oz[i] = ((a[8] * ox + a[9] * oy) + a[10] * oz) + a[11];
    out.dx[i] = (a[0] * dx + a[1] * dy) + a[2] * dz;
    out.dy[i] = (a[4] * dx + a[5] * dy) + a[6] * dz;
    out.dz[i] = (a[8] * dx + a[9] * dy) + a[10] * dz;
    out.tmin[i] = in.tmin[i];
    float tmax = in.tmax[i];
    const float b = best_t[i];
    if (b < INFINITY) {
        const float below = float_below(b);
        tmax = below < tmax ? below : tmax;
    }
    out.tmax[i] = tmax;
}

// Keep instance k's hit (strictly nearer by construction of its interval).
__global__ __launch_bounds__(256) void k_instance_merge(uint32_t n, HitSoAIn cand, HitSoA hits,
                                                        int32_t* __restrict__ instance, int32_t k)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t id = cand.id[i];
    if (id < 0)
        return;
    hits.id[i] = id;
    hits.t[i] = cand.t[i];
    hits.u[i] = cand.u[i];
    hits.v[i] = cand.v[i];
    instance[i] = k;
}

hipError_t launch_instance_init(uint32_t n, const HitSoA& hits, int32_t* instance, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_instance_init, dim3((n + 255) / 256), dim3(256), 0, s, n, hits, instance);
    return hipGetLastError();
}

hipError_t launch_instance_rays(uint32_t n, const RaySoA& in, const float* best_t,
                                const Affine& inv, const RaySoAOut& out, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_instance_rays, dim3((n + 255) / 256), dim3(256), 0, s, n, in, best_t, inv,
                       out);
    return hipGetLastError();
}

hipError_t launch_instance_merge(uint32_t n, const HitSoAIn& cand, const HitSoA& hits,
                                 int32_t* instance, int32_t k, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_instance_merge, dim3((n + 255) / 256), dim3(256), 0, s, n, cand, hits,
                       instance, k);
    return hipGetLastError();
}

}  // namespace grt
