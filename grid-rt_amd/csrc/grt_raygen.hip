// grt_raygen.hip -- ray generation on gfx950: primary rays from a pinhole
// camera, and shadow, mirror-reflection and ambient-occlusion rays from
// closest-hit records; plus the headlight shading of the reference CUDA tracer.
//
// Reference semantics:
//   primary rays        PinholeCamera.cpp:131-137, RayTracing_kernel.cu:145-161
//   shadow rays         shadow_FRAG.glsl:259-270, SimplePointLight.cpp:18-37
//   reflection rays     Sample.cpp:7-11,44-52, traversalIntersection_FRAG.glsl:285-296
//   shading normal      Sample.cpp:100-114, RayTracing_kernel.cu:92-102
//   headlight shading   RayTracing_kernel.cu:104-120
//   Phong shading       shading_FRAG.glsl:116-173, PhongMaterialColor.cpp:19-85,
//                       reflection blend GpuRenderer.cpp:233-297
// Every kernel is one output ray (or pixel) per lane, structure-of-arrays in
// and out, and follows the CPU oracle's operation order bit for bit.
#include "grt_launch.h"

namespace grt {

// ---------------------------------------------------------------------------
// Ray generation
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_gen_primary(DevCamera cam, RaySoAOut out)
{
    const uint32_t n = cam.width * cam.height;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t y = i / cam.width, x = i - y * cam.width;
    const float inv_w = 1.0f / (float)cam.width;
    const float inv_h = 1.0f / (float)cam.height;
    const float us = (float)x * inv_w;
    const float vs = (float)y * inv_h;
    out.ox[i] = cam.pos[0];
    out.oy[i] = cam.pos[1];
    out.oz[i] = cam.pos[2];
    out.dx[i] = (cam.base[0] + cam.nu[0] * us) + cam.nv[0] * vs;
    out.dy[i] = (cam.base[1] + cam.nu[1] * us) + cam.nv[1] * vs;
    out.dz[i] = (cam.base[2] + cam.nu[2] * us) + cam.nv[2] * vs;
    out.tmin[i] = kRayEpsilon;
    out.tmax[i] = INFINITY;
}

__device__ __forceinline__ void normalize3(float& x, float& y, float& z)
{
    const float inv = 1.0f / sqrtf(dot3(x, y, z, x, y, z));
    x = x * inv;
    y = y * inv;
    z = z * inv;
}

// Sample::computeShadingNormal with weights (1-(u+v), u, v), turned to face
// the incoming ray; hit point o + d*t.
__device__ __forceinline__ void hit_frame(const float* __restrict__ normals, int32_t id, float t,
                                          float u, float v, float o0, float o1, float o2,
                                          float d0, float d1, float d2, float h[3], float n[3])
{
    h[0] = o0 + d0 * t;
    h[1] = o1 + d1 * t;
    h[2] = o2 + d2 * t;
    const float* nn = normals + 9ull * (uint32_t)id;
    const float w = 1.0f - (u + v);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        n[k] = (nn[k] * w + nn[3 + k] * u) + nn[6 + k] * v;
    normalize3(n[0], n[1], n[2]);
    if (dot3(n[0], n[1], n[2], d0, d1, d2) > 0.0f) {
        n[0] = -n[0];
        n[1] = -n[1];
        n[2] = -n[2];
    }
}

__device__ __forceinline__ void put_inactive(const RaySoAOut& out, uint32_t i)
{
    out.ox[i] = 0.0f;
    out.oy[i] = 0.0f;
    out.oz[i] = 0.0f;
    out.dx[i] = 0.0f;
    out.dy[i] = 0.0f;
    out.dz[i] = 1.0f;
    out.tmin[i] = 1.0f;
    out.tmax[i] = -1.0f;
}

__global__ __launch_bounds__(256) void k_gen_shadow(uint32_t n, RaySoA in, HitSoAIn hits,
                                                    const float* __restrict__ normals, float lx,
                                                    float ly, float lz, float offset, RaySoAOut out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t id = hits.id[i];
    if (id < 0 || (uint32_t)id >= hits.n_tris) {
        put_inactive(out, i);
        return;
    }
    const float d0 = in.dx[i], d1 = in.dy[i], d2 = in.dz[i];
    float h[3], nn[3];
    hit_frame(normals, id, hits.t[i], hits.u[i], hits.v[i], in.ox[i], in.oy[i], in.oz[i], d0, d1,
              d2, h, nn);
    const float l0 = lx - h[0], l1 = ly - h[1], l2 = lz - h[2];
    if (dot3(nn[0], nn[1], nn[2], l0, l1, l2) <= 0.0f) {
        put_inactive(out, i);
        return;
    }
    const float dist = sqrtf(dot3(l0, l1, l2, l0, l1, l2));
    const float inv = 1.0f / dist;
    out.ox[i] = h[0] + nn[0] * offset;
    out.oy[i] = h[1] + nn[1] * offset;
    out.oz[i] = h[2] + nn[2] * offset;
    out.dx[i] = l0 * inv;
    out.dy[i] = l1 * inv;
    out.dz[i] = l2 * inv;
    out.tmin[i] = 0.0f;
    out.tmax[i] = dist;
}

__global__ __launch_bounds__(256) void k_gen_reflection(uint32_t n, RaySoA in, HitSoAIn hits,
                                                        const float* __restrict__ normals,
                                                        float offset, RaySoAOut out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t id = hits.id[i];
    if (id < 0 || (uint32_t)id >= hits.n_tris) {
        put_inactive(out, i);
        return;
    }
    const float d0 = in.dx[i], d1 = in.dy[i], d2 = in.dz[i];
    float h[3], nn[3];
    hit_frame(normals, id, hits.t[i], hits.u[i], hits.v[i], in.ox[i], in.oy[i], in.oz[i], d0, d1,
              d2, h, nn);
    const float proj = dot3(d0, d1, d2, nn[0], nn[1], nn[2]) * 2.0f;
    out.ox[i] = h[0] + nn[0] * offset;
    out.oy[i] = h[1] + nn[1] * offset;
    out.oz[i] = h[2] + nn[2] * offset;
    out.dx[i] = d0 - nn[0] * proj;
    out.dy[i] = d1 - nn[1] * proj;
    out.dz[i] = d2 - nn[2] * proj;
    out.tmin[i] = 0.0f;
    out.tmax[i] = INFINITY;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ float uniform24(uint64_t seed, uint64_t a, uint64_t b)
{
    const uint64_t h = mix64(seed ^ mix64(a * 0x9e3779b97f4a7c15ull + b));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

// sin/cos of 2*pi*turns: quadrant reduction in turns, Taylor polynomials.
__device__ __forceinline__ void sincos_turns(float turns, float& s, float& c)
{
    const float y = turns * 4.0f;
    const float q = rintf(y);
    const float r = y - q;
    const float a = r * 1.5707963267948966f;
    const float a2 = a * a;
    float sp = __builtin_fmaf(a2, 2.7557319e-6f, -1.9841270e-4f);
    sp = __builtin_fmaf(a2, sp, 8.3333333e-3f);
    sp = __builtin_fmaf(a2, sp, -1.6666667e-1f);
    sp = __builtin_fmaf(a2, sp, 1.0f);
    const float sa = a * sp;
    float cp = __builtin_fmaf(a2, 2.4801587e-5f, -1.3888889e-3f);
    cp = __builtin_fmaf(a2, cp, 4.1666667e-2f);
    cp = __builtin_fmaf(a2, cp, -0.5f);
    const float ca = __builtin_fmaf(a2, cp, 1.0f);
    const int qi = ((int)q) & 3;
    if (qi == 0) {
        s = sa;
        c = ca;
    } else if (qi == 1) {
        s = ca;
        c = -sa;
    } else if (qi == 2) {
        s = -sa;
        c = -ca;
    } else {
        s = -ca;
        c = sa;
    }
}

__global__ __launch_bounds__(256) void k_gen_ao(uint32_t n, uint32_t spp, uint64_t seed, RaySoA in,
                                                HitSoAIn hits, const float* __restrict__ normals,
                                                float offset, float ao_tmax, RaySoAOut out)
{
    const uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (j >= (uint64_t)n * spp)
        return;
    const uint32_t i = (uint32_t)(j / spp);
    const uint32_t k = (uint32_t)(j - (uint64_t)i * spp);
    const int32_t id = hits.id[i];
    if (id < 0 || (uint32_t)id >= hits.n_tris) {
        put_inactive(out, (uint32_t)j);
        return;
    }
    float h[3], nn[3];
    hit_frame(normals, id, hits.t[i], hits.u[i], hits.v[i], in.ox[i], in.oy[i], in.oz[i],
              in.dx[i], in.dy[i], in.dz[i], h, nn);
    const float sgn = copysignf(1.0f, nn[2]);
    const float a = -1.0f / (sgn + nn[2]);
    const float b = (nn[0] * nn[1]) * a;
    const float tb0 = 1.0f + ((sgn * nn[0]) * nn[0]) * a, tb1 = sgn * b, tb2 = -(sgn * nn[0]);
    const float bb0 = b, bb1 = sgn + (nn[1] * nn[1]) * a, bb2 = -nn[1];
    const float u1 = uniform24(seed, (uint64_t)i, 2ull * k);
    const float u2 = uniform24(seed, (uint64_t)i, 2ull * k + 1ull);
    const float r = sqrtf(u1);
    float sn, cs;
    sincos_turns(u2, sn, cs);
    const float lx = r * cs, ly = r * sn, lz = sqrtf(1.0f - u1);
    out.ox[j] = h[0] + nn[0] * offset;
    out.oy[j] = h[1] + nn[1] * offset;
    out.oz[j] = h[2] + nn[2] * offset;
    out.dx[j] = __builtin_fmaf(nn[0], lz, __builtin_fmaf(bb0, ly, tb0 * lx));
    out.dy[j] = __builtin_fmaf(nn[1], lz, __builtin_fmaf(bb1, ly, tb1 * lx));
    out.dz[j] = __builtin_fmaf(nn[2], lz, __builtin_fmaf(bb2, ly, tb2 * lx));
    out.tmin[j] = 0.0f;
    out.tmax[j] = ao_tmax;
}

// Headlight shading into RGBA8 (RayTracing_kernel.cu:92-120,262,274).
__device__ __forceinline__ uint32_t to_u8(float c)
{
    float x = c * 255.0f;
    x = x > 0.0f ? x : 0.0f;
    x = x < 255.0f ? x : 255.0f;
    return (uint32_t)x;
}

__global__ __launch_bounds__(256) void k_shade_headlight(uint32_t n, RaySoA in, HitSoAIn hits,
                                                         const float* __restrict__ normals,
                                                         uint32_t* __restrict__ rgba)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t id = hits.id[i];
    if (id < 0 || (uint32_t)id >= hits.n_tris) {
        rgba[i] = 0xffffffffu;
        return;
    }
    float dx = in.dx[i], dy = in.dy[i], dz = in.dz[i];
    normalize3(dx, dy, dz);
    const float* nn = normals + 9ull * (uint32_t)id;
    const float u = hits.u[i], v = hits.v[i];
    const float w = 1.0f - (u + v);
    float n0 = (nn[0] * w + nn[3] * u) + nn[6] * v;
    float n1 = (nn[1] * w + nn[4] * u) + nn[7] * v;
    float n2 = (nn[2] * w + nn[5] * u) + nn[8] * v;
    normalize3(n0, n1, n2);
    const float ndd = -dot3(n0, n1, n2, dx, dy, dz);
    const float r = (0.1f + (0.0f - 0.1f) * ndd) * 0.0f;
    const float g = (0.1f + (0.0f - 0.1f) * ndd) * 0.0f;
    const float b = (0.1f + (1.0f - 0.1f) * ndd) * 1.0f;
    rgba[i] = to_u8(r) | (to_u8(g) << 8) | (to_u8(b) << 16) | (255u << 24);
}

// ---------------------------------------------------------------------------
// Phong shading with shadow dimming and reflection accumulation
// ---------------------------------------------------------------------------
// x^e by square-and-multiply from the lowest bit of e (the specular power; an
// integer ladder is reproducible where a library powf is not).
__device__ __forceinline__ float pow_u32(float x, uint32_t e)
{
    float r = 1.0f, b = x;
    while (e) {
        if (e & 1u)
            r = r * b;
        e >>= 1;
        if (e)
            b = b * b;
    }
    return r;
}

// One pixel per lane.  The point is lit when N.L > 0 and its shadow ray is
// not occluded; lit: ambient + max(0, kd lc N.L) + s lc max(0, S.L)^e with
// S = normalise(reflect(d, N)); shadowed: ambient + max(0, kd lc N.L) x
// shadow_coeff.  depth 0 writes the colour (background for a miss); depth
// d > 0 adds weight x colour for a hit.
__global__ __launch_bounds__(256) void k_shade_phong(uint32_t n, RaySoA in, HitSoAIn hits,
                                                     const uint8_t* __restrict__ occluded,
                                                     const float* __restrict__ normals,
                                                     const float* __restrict__ diffuse,
                                                     DevPhong p, uint32_t depth, float weight,
                                                     float* __restrict__ rgb,
                                                     uint32_t* __restrict__ rgba)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const int32_t id = hits.id[i];
    float c0, c1, c2;
    if (id >= 0 && (uint32_t)id < hits.n_tris) {
        const float d0 = in.dx[i], d1 = in.dy[i], d2 = in.dz[i];
        float h[3], nn[3];
        hit_frame(normals, id, hits.t[i], hits.u[i], hits.v[i], in.ox[i], in.oy[i], in.oz[i],
                  d0, d1, d2, h, nn);
        const float l0 = p.light[0] - h[0], l1 = p.light[1] - h[1], l2 = p.light[2] - h[2];
        const bool lit = dot3(nn[0], nn[1], nn[2], l0, l1, l2) > 0.0f &&
                         !(occluded && occluded[i] != 0);
        const float dist = sqrtf(dot3(l0, l1, l2, l0, l1, l2));
        const float inv = 1.0f / dist;
        const float L0 = l0 * inv, L1 = l1 * inv, L2 = l2 * inv;
        const float proj = dot3(d0, d1, d2, nn[0], nn[1], nn[2]) * 2.0f;
        float s0 = d0 - nn[0] * proj, s1 = d1 - nn[1] * proj, s2 = d2 - nn[2] * proj;
        normalize3(s0, s1, s2);
        const float sdl = dot3(s0, s1, s2, L0, L1, L2);
        const float spec = (p.spec_exp > 0u && sdl > 0.0f) ? pow_u32(sdl, p.spec_exp) : 0.0f;
        const float ndl = dot3(nn[0], nn[1], nn[2], L0, L1, L2);
        float kd[3] = {1.0f, 1.0f, 1.0f};
        if (diffuse) {
            kd[0] = diffuse[3ull * (uint32_t)id];
            kd[1] = diffuse[3ull * (uint32_t)id + 1];
            kd[2] = diffuse[3ull * (uint32_t)id + 2];
        }
        float col[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            float dif = kd[c] * (p.light_color[c] * ndl);
            dif = dif > 0.0f ? dif : 0.0f;
            col[c] = lit ? (p.ambient[c] + dif) + p.specular[c] * (p.light_color[c] * spec)
                         : p.ambient[c] + dif * p.shadow_coeff;
        }
        if (depth == 0) {
            c0 = col[0];
            c1 = col[1];
            c2 = col[2];
        } else {
            c0 = rgb[3ull * i] + weight * col[0];
            c1 = rgb[3ull * i + 1] + weight * col[1];
            c2 = rgb[3ull * i + 2] + weight * col[2];
        }
    } else if (depth == 0) {
        c0 = p.background[0];
        c1 = p.background[1];
        c2 = p.background[2];
    } else {
        c0 = rgb[3ull * i];
        c1 = rgb[3ull * i + 1];
        c2 = rgb[3ull * i + 2];
    }
    rgb[3ull * i] = c0;
    rgb[3ull * i + 1] = c1;
    rgb[3ull * i + 2] = c2;
    if (rgba)
        rgba[i] = to_u8(c0) | (to_u8(c1) << 8) | (to_u8(c2) << 16) | (255u << 24);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
hipError_t launch_gen_primary(const DevCamera& cam, const RaySoAOut& out, hipStream_t s)
{
    const uint32_t n = cam.width * cam.height;
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gen_primary, dim3((n + 255) / 256), dim3(256), 0, s, cam, out);
    return hipGetLastError();
}

hipError_t launch_gen_shadow(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                             const float* normals, float lx, float ly, float lz, float offset,
                             const RaySoAOut& out, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gen_shadow, dim3((n + 255) / 256), dim3(256), 0, s, n, in, hits, normals,
                       lx, ly, lz, offset, out);
    return hipGetLastError();
}

hipError_t launch_gen_reflection(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                                 const float* normals, float offset, const RaySoAOut& out,
                                 hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gen_reflection, dim3((n + 255) / 256), dim3(256), 0, s, n, in, hits,
                       normals, offset, out);
    return hipGetLastError();
}

hipError_t launch_gen_ao(uint32_t n, uint32_t spp, uint64_t seed, const RaySoA& in,
                         const HitSoAIn& hits, const float* normals, float offset,
                         float ao_tmax, const RaySoAOut& out, hipStream_t s)
{
    const uint64_t total = (uint64_t)n * spp;
    if (total == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_gen_ao, dim3((uint32_t)((total + 255) / 256)), dim3(256), 0, s, n, spp,
                       seed, in, hits, normals, offset, ao_tmax, out);
    return hipGetLastError();
}

// Number of active rays (tmin <= tmax) added to *count: one atomic per wave.
__global__ __launch_bounds__(256) void k_count_active(const float* __restrict__ tmin,
                                                      const float* __restrict__ tmax, uint32_t n,
                                                      unsigned long long* __restrict__ count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool act = i < n && tmin[i] <= tmax[i];
    const uint64_t m = __ballot(act);
    if ((threadIdx.x & 63u) == 0 && m != 0)
        atomicAdd(count, (unsigned long long)__popcll(m));
}

hipError_t launch_count_active(const float* tmin, const float* tmax, uint32_t n,
                               unsigned long long* count, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_count_active, dim3((n + 255) / 256), dim3(256), 0, s, tmin, tmax, n,
                       count);
    return hipGetLastError();
}

hipError_t launch_shade_headlight(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                                  const float* normals, uint32_t* rgba, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_shade_headlight, dim3((n + 255) / 256), dim3(256), 0, s, n, in, hits,
                       normals, rgba);
    return hipGetLastError();
}

hipError_t launch_shade_phong(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                              const uint8_t* occluded, const float* normals,
                              const float* diffuse, const DevPhong& p, uint32_t depth,
                              float* rgb, uint32_t* rgba, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    float w = 1.0f;
    for (uint32_t k = 0; k < depth; ++k)
        w = w * p.reflect_coeff;
    hipLaunchKernelGGL(k_shade_phong, dim3((n + 255) / 256), dim3(256), 0, s, n, in, hits,
                       occluded, normals, diffuse, p, depth, w, rgb, rgba);
    return hipGetLastError();
}

}  // namespace grt

