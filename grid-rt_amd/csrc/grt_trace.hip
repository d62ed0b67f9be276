// grt_trace.hip -- one-ray-per-lane 3D-DDA traversal with Moller-Trumbore on
// gfx950, plus ray generation from a pinhole camera and from hit records.
//
// Reference semantics:
//   clip + DDA set-up   AabbIntersection.cpp:37-63, UniformGridAccStruct.cpp:118-149
//   DDA loop            UniformGridAccStruct.cpp:160-195 (single-axis step,
//                       ties to y then z)
//   cell loop           UniformGridAccStruct.cpp:362-378 (hits bounded by the
//                       current cell's exit)
//   Moller-Trumbore     RayTracing_kernel.cu:56-89 (no determinant epsilon)
//   any hit             shadow_FRAG.glsl:122-245
//   primary rays        PinholeCamera.cpp:131-137, RayTracing_kernel.cu:145-161
//   shadow/reflection   shadow_FRAG.glsl:259-270, Sample.cpp:7-52,95-114
//
// Data in HBM: rays and hits are structures of arrays (coalesced 4-byte lanes);
// the grid is offsets[m+1] + refs[R] (uint32); triangles are 48-byte records
// {v0.xyz, e1.x}{e1.yz, e2.xy}{e2.z, id, -, -} read with three 16-byte loads.
#include "grt_launch.h"

namespace grt {

// Lane -> ray index.  With image_width != 0 each wave64 covers an 8x8 pixel
// tile of a row-major image so neighbouring lanes trace neighbouring pixels.
__device__ __forceinline__ uint32_t ray_index(uint32_t img_w, uint32_t n, bool& valid)
{
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (img_w == 0) {
        valid = g < n;
        return g;
    }
    const uint32_t wave = g >> 6, lane = g & 63u;
    const uint32_t tiles_x = (img_w + 7u) >> 3;
    const uint32_t tx = wave % tiles_x, ty = wave / tiles_x;
    const uint32_t px = tx * 8u + (lane & 7u), py = ty * 8u + (lane >> 3);
    const uint32_t i = py * img_w + px;
    valid = (px < img_w) && (i < n);
    return i;
}

static uint32_t ray_grid(uint32_t n, uint32_t img_w, uint32_t block)
{
    if (img_w == 0)
        return (n + block - 1) / block;
    const uint32_t h = (n + img_w - 1) / img_w;
    const uint64_t threads = (uint64_t)((img_w + 7u) >> 3) * ((h + 7u) >> 3) * 64u;
    return (uint32_t)((threads + block - 1) / block);
}

// DDA state in registers.
struct Dda {
    int32_t cx, cy, cz;
    int32_t sx, sy, sz;
    int32_t ox, oy, oz;   // out-of-grid limits
    float tx, ty, tz;     // t at the next boundary per axis
    float dtx, dty, dtz;
};

// Ray::update + clipRay + DDA set-up.  Returns false when the ray misses the
// grid box (or its interval is empty).
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

// Moller-Trumbore on a packed 48-byte record; returns the inside test and
// writes t,u,v.
__device__ __forceinline__ bool mt_test(const float4* __restrict__ tri, uint32_t r, float o0,
                                        float o1, float o2, float d0, float d1, float d2,
                                        float& t, float& u, float& v)
{
    const float4 A = tri[3ull * r];
    const float4 B = tri[3ull * r + 1];
    const float4 C = tri[3ull * r + 2];
    const float v0x = A.x, v0y = A.y, v0z = A.z;
    const float e1x = A.w, e1y = B.x, e1z = B.y;
    const float e2x = B.z, e2y = B.w, e2z = C.x;
    const float px = __builtin_fmaf(d1, e2z, -(d2 * e2y));
    const float py = __builtin_fmaf(d2, e2x, -(d0 * e2z));
    const float pz = __builtin_fmaf(d0, e2y, -(d1 * e2x));
    const float det = dot3(px, py, pz, e1x, e1y, e1z);
    const float inv = 1.0f / det;
    const float tvx = o0 - v0x, tvy = o1 - v0y, tvz = o2 - v0z;
    const float qx = __builtin_fmaf(tvy, e1z, -(tvz * e1y));
    const float qy = __builtin_fmaf(tvz, e1x, -(tvx * e1z));
    const float qz = __builtin_fmaf(tvx, e1y, -(tvy * e1x));
    u = dot3(px, py, pz, tvx, tvy, tvz) * inv;
    v = dot3(qx, qy, qz, d0, d1, d2) * inv;
    t = dot3(qx, qy, qz, e2x, e2y, e2z) * inv;
    return (u >= 0.0f) && (v >= 0.0f) && ((u + v) <= 1.0f);
}

__device__ __forceinline__ bool grid_ok(const DevFrame* __restrict__ fr)
{
    return (fr->status & (kStatusFrameInvalid | kStatusCellsOver | kStatusRefsOver)) == 0;
}

template <bool COUNT>
__global__ __launch_bounds__(256) void k_trace_closest(TraceArgs a, RaySoA rays, HitSoA hits)
{
    bool valid;
    const uint32_t i = ray_index(a.image_width, a.n, valid);
    if (!valid)
        return;
    const DevFrame* __restrict__ fr = a.fr;
    const uint32_t* __restrict__ offsets = a.offsets;
    const uint32_t* __restrict__ refs = a.refs;
    const float4* __restrict__ tri = a.tri;
    const float o0 = rays.ox[i], o1 = rays.oy[i], o2 = rays.oz[i];
    const float d0 = rays.dx[i], d1 = rays.dy[i], d2 = rays.dz[i];
    const float tmin = rays.tmin[i], tmax = rays.tmax[i];
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    uint32_t cells = 0, tests = 0;
    Dda s;
    if (grid_ok(fr) && dda_setup(fr, o0, o1, o2, d0, d1, d2, tmin, tmax, s)) {
        const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
        for (;;) {
            if (COUNT)
                ++cells;
            const float exit_t = tmin_f(tmin_f(s.tx, s.ty), s.tz);
            const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
            const uint32_t b = offsets[ci], e = offsets[ci + 1];
            if (b < e) {
                const float bound = exit_t * kExitScale;
                for (uint32_t k = b; k < e; ++k) {
                    const uint32_t r = refs[k];
                    if (COUNT)
                        ++tests;
                    float t, u, v;
                    if (mt_test(tri, r, o0, o1, o2, d0, d1, d2, t, u, v) && t > tmin &&
                        t <= tmax && t <= bound && t < best) {
                        best = t;
                        id = (int32_t)r;
                        bu = u;
                        bv = v;
                    }
                }
                if (id >= 0)
                    break;
            }
            if (exit_t > tmax)
                break;
            if (!dda_step(s))
                break;
        }
    }
    hits.id[i] = id;
    hits.t[i] = best;
    hits.u[i] = bu;
    hits.v[i] = bv;
    if (COUNT) {
        a.counters[2ull * i] = cells;
        a.counters[2ull * i + 1] = tests;
    }
}

template <bool COUNT>
__global__ __launch_bounds__(256) void k_trace_any(TraceArgs a, RaySoA rays,
                                                   uint8_t* __restrict__ occluded)
{
    bool valid;
    const uint32_t i = ray_index(a.image_width, a.n, valid);
    if (!valid)
        return;
    const DevFrame* __restrict__ fr = a.fr;
    const uint32_t* __restrict__ offsets = a.offsets;
    const uint32_t* __restrict__ refs = a.refs;
    const float4* __restrict__ tri = a.tri;
    const float o0 = rays.ox[i], o1 = rays.oy[i], o2 = rays.oz[i];
    const float d0 = rays.dx[i], d1 = rays.dy[i], d2 = rays.dz[i];
    const float tmin = rays.tmin[i], tmax = rays.tmax[i];
    bool occ = false;
    uint32_t cells = 0, tests = 0;
    Dda s;
    if (grid_ok(fr) && dda_setup(fr, o0, o1, o2, d0, d1, d2, tmin, tmax, s)) {
        const uint32_t nx = (uint32_t)fr->n[0], ny = (uint32_t)fr->n[1];
        for (;;) {
            if (COUNT)
                ++cells;
            const float exit_t = tmin_f(tmin_f(s.tx, s.ty), s.tz);
            const uint32_t ci = (uint32_t)s.cx + nx * ((uint32_t)s.cy + ny * (uint32_t)s.cz);
            const uint32_t b = offsets[ci], e = offsets[ci + 1];
            for (uint32_t k = b; k < e; ++k) {
                if (COUNT)
                    ++tests;
                float t, u, v;
                if (mt_test(tri, refs[k], o0, o1, o2, d0, d1, d2, t, u, v) && t > tmin &&
                    t <= tmax) {
                    occ = true;
                    break;
                }
            }
            if (occ)
                break;
            if (exit_t > tmax)
                break;
            if (!dda_step(s))
                break;
        }
    }
    occluded[i] = occ ? 1 : 0;
    if (COUNT) {
        a.counters[2ull * i] = cells;
        a.counters[2ull * i + 1] = tests;
    }
}

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
// Host launchers
// ---------------------------------------------------------------------------
hipError_t launch_trace_closest(const TraceArgs& a, const RaySoA& rays, const HitSoA& hits,
                                uint32_t block, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    const uint32_t g = ray_grid(a.n, a.image_width, block);
    if (a.counters)
        hipLaunchKernelGGL(k_trace_closest<true>, dim3(g), dim3(block), 0, s, a, rays, hits);
    else
        hipLaunchKernelGGL(k_trace_closest<false>, dim3(g), dim3(block), 0, s, a, rays, hits);
    return hipGetLastError();
}

hipError_t launch_trace_any(const TraceArgs& a, const RaySoA& rays, uint8_t* occluded,
                            uint32_t block, hipStream_t s)
{
    if (a.n == 0)
        return hipSuccess;
    const uint32_t g = ray_grid(a.n, a.image_width, block);
    if (a.counters)
        hipLaunchKernelGGL(k_trace_any<true>, dim3(g), dim3(block), 0, s, a, rays, occluded);
    else
        hipLaunchKernelGGL(k_trace_any<false>, dim3(g), dim3(block), 0, s, a, rays, occluded);
    return hipGetLastError();
}

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

hipError_t launch_shade_headlight(uint32_t n, const RaySoA& in, const HitSoAIn& hits,
                                  const float* normals, uint32_t* rgba, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(k_shade_headlight, dim3((n + 255) / 256), dim3(256), 0, s, n, in, hits,
                       normals, rgba);
    return hipGetLastError();
}

}  // namespace grt
