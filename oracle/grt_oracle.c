/*
 * grt_oracle.c -- CPU oracle for the uniform-grid 3D-DDA + Moller-Trumbore path.
 *
 * TEST INFRASTRUCTURE ONLY (parity reference and CPU baseline); the product
 * library never links it.  Every function cites the reference file:line it
 * restates.  Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).
 *
 * Floating-point rules of the specification shared with the HIP kernels:
 *   - a fused multiply-add appears only as an explicit fmaf();
 *   - everything else is one IEEE-754 binary32 operation, evaluated in the
 *     order written (the file is compiled with -ffp-contract=off);
 *   - min/max are the ternaries a<b?a:b and a>b?a:b, so NaN handling is fixed;
 *   - division and sqrtf are correctly rounded.
 */
#include "grt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* Small helpers                                                             */
/* ------------------------------------------------------------------------ */
static inline float tmin_f(float a, float b) { return a < b ? a : b; }
static inline float tmax_f(float a, float b) { return a > b ? a : b; }

/* cross(a,b) with one fma per component. */
static inline void cross3(const float a[3], const float b[3], float c[3])
{
    c[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    c[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    c[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}

/* dot(a,b) = fma(a.z,b.z, fma(a.y,b.y, a.x*b.x)). */
static inline float dot3(const float a[3], const float b[3])
{
    return fmaf(a[2], b[2], fmaf(a[1], b[1], a[0] * b[0]));
}

double orc_now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* a2: grid frame                                                            */
/* ------------------------------------------------------------------------ */

/* Cube root used for the resolution factor.  The reference evaluates
 * powf(k*N/V, 1.0f/3.0f) (UniformGridAccStructBuilder.cpp:93); libm powf is
 * not reproducible across CPU and GPU math libraries, so the specification
 * fixes a cube root computed with basic IEEE-754 binary64 operations only:
 * x = m*2^e, m scaled into [0.5,4), eight Newton steps y <- (2y + m/y^2)/3
 * from y=1, then ldexp and rounding to binary32. */
float orc_cbrtf(float xf)
{
    double x = (double)xf;
    if (!(x > 0.0) || isinf(x))
        return xf;
    int e;
    double m = frexp(x, &e);
    int q = (e >= 0) ? (e / 3) : -((-e + 2) / 3);
    int r = e - 3 * q;
    m = ldexp(m, r);
    double y = 1.0;
    for (int i = 0; i < 8; ++i)
        y = (2.0 * y + m / (y * y)) / 3.0;
    return (float)ldexp(y, q);
}

/* Aabb::scaleBy(0.01f) (Aabb.cpp:78-91) followed by the resolution heuristic
 * (UniformGridAccStructBuilder.cpp:86-98) and setResolution
 * (UniformGridAccStruct.cpp:44-63).  lambda replaces the hard-coded k=6. */
int orc_grid_frame_from_box(const float bmin_raw[3], const float bmax_raw[3],
                            uint32_t n_tris, float lambda, orc_frame* f)
{
    if (n_tris == 0 || !(lambda > 0.0f))
        return -1;
    float diag[3];
    for (int d = 0; d < 3; ++d) {
        float mn = bmin_raw[d], mx = bmax_raw[d];
        float s = (mn == mx) ? 0.01f : fabsf(mx - mn) * 0.01f;
        f->bmin[d] = mn - s;
        f->bmax[d] = mx + s;
        diag[d] = f->bmax[d] - f->bmin[d];
    }
    float vol = (diag[0] * diag[1]) * diag[2];
    float x = (lambda * (float)n_tris) / vol;
    float fac = orc_cbrtf(x);
    uint64_t m = 1;
    for (int d = 0; d < 3; ++d) {
        float c = ceilf(diag[d] * fac);
        if (!(c < 1048576.0f))
            return -2;
        int32_t n = (int32_t)c;
        if (n < 1)
            n = 1;
        f->n[d] = n;
        m *= (uint64_t)n;
    }
    if (m > (1ull << 30))
        return -2;
    f->m = (uint32_t)m;
    for (int d = 0; d < 3; ++d) {
        f->cell[d] = diag[d] / (float)f->n[d];
        f->inv_cell[d] = 1.0f / f->cell[d];
    }
    return 0;
}

/* Aabb::buildFrom over all vertices (Aabb.cpp:11-53): exact min/max. */
int orc_grid_frame(const float* verts, uint32_t n_tris, float lambda, orc_frame* f)
{
    if (n_tris == 0)
        return -1;
    float mn[3] = {verts[0], verts[1], verts[2]};
    float mx[3] = {verts[0], verts[1], verts[2]};
    uint64_t nv = 3ull * n_tris;
    for (uint64_t i = 0; i < nv; ++i) {
        for (int d = 0; d < 3; ++d) {
            float v = verts[3 * i + d];
            mn[d] = v < mn[d] ? v : mn[d];
            mx[d] = v > mx[d] ? v : mx[d];
        }
    }
    return orc_grid_frame_from_box(mn, mx, n_tris, lambda, f);
}

/* ------------------------------------------------------------------------ */
/* a3-a5: grid build                                                         */
/* ------------------------------------------------------------------------ */

/* worldToVoxel (UniformGridAccStruct.cpp:92-98): binary32 (v-min)*invCell,
 * truncated; clamped to [0,n-1] as in countCellVP.glsl:31-35.  The clamp is
 * applied in float before truncation so NaN maps to cell 0. */
static inline int32_t vox(const orc_frame* f, float val, int d)
{
    float x = (val - f->bmin[d]) * f->inv_cell[d];
    x = x > 0.0f ? x : 0.0f;
    float hi = (float)(f->n[d] - 1);
    x = x < hi ? x : hi;
    return (int32_t)x;
}

/* Triangle AABB -> inclusive cell index box
 * (UniformGridAccStructBuilder.cpp:111-127). */
void orc_tri_cell_range(const orc_frame* f, const float* v9, int32_t lo[3], int32_t hi[3])
{
    for (int d = 0; d < 3; ++d) {
        float a = v9[d], b = v9[3 + d], c = v9[6 + d];
        float mn = a, mx = a;
        mn = b < mn ? b : mn;
        mx = b > mx ? b : mx;
        mn = c < mn ? c : mn;
        mx = c > mx ? c : mx;
        lo[d] = vox(f, mn, d);
        hi[d] = vox(f, mx, d);
    }
}

uint64_t orc_grid_count(const orc_frame* f, const float* verts, uint32_t n_tris, uint32_t* counts)
{
    memset(counts, 0, sizeof(uint32_t) * (size_t)f->m);
    uint64_t total = 0;
    const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
    for (uint32_t t = 0; t < n_tris; ++t) {
        int32_t lo[3], hi[3];
        orc_tri_cell_range(f, verts + 9ull * t, lo, hi);
        for (int32_t z = lo[2]; z <= hi[2]; ++z)
            for (int32_t y = lo[1]; y <= hi[1]; ++y)
                for (int32_t x = lo[0]; x <= hi[0]; ++x) {
                    uint32_t c = (uint32_t)x + nx * ((uint32_t)y + ny * (uint32_t)z);
                    counts[c] += 1;
                    total += 1;
                }
    }
    return total;
}

/* Exclusive scan of counts, then refs in triangle order: the reference pushes
 * triangle t into every cell of its AABB box while t ascends
 * (UniformGridAccStructBuilder.cpp:109-141), so each cell lists ascending ids.
 * Flattened like CudaRenderer.cpp:160-199 (start = offsets[c], count =
 * offsets[c+1]-offsets[c]) but with plain triangle ids instead of 3*id. */
void orc_grid_fill(const orc_frame* f, const float* verts, uint32_t n_tris,
                   const uint32_t* counts, uint32_t* offsets, uint32_t* refs)
{
    uint32_t acc = 0;
    for (uint32_t c = 0; c < f->m; ++c) {
        offsets[c] = acc;
        acc += counts[c];
    }
    offsets[f->m] = acc;
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)f->m);
    memcpy(cur, offsets, sizeof(uint32_t) * (size_t)f->m);
    const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
    for (uint32_t t = 0; t < n_tris; ++t) {
        int32_t lo[3], hi[3];
        orc_tri_cell_range(f, verts + 9ull * t, lo, hi);
        for (int32_t z = lo[2]; z <= hi[2]; ++z)
            for (int32_t y = lo[1]; y <= hi[1]; ++y)
                for (int32_t x = lo[0]; x <= hi[0]; ++x) {
                    uint32_t c = (uint32_t)x + nx * ((uint32_t)y + ny * (uint32_t)z);
                    refs[cur[c]++] = t;
                }
    }
    free(cur);
}

/* ------------------------------------------------------------------------ */
/* a6: camera                                                                */
/* ------------------------------------------------------------------------ */
static void normalize3(float v[3])
{
    float inv = 1.0f / sqrtf(dot3(v, v));
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

/* PinholeCamera::setLookAt + newFrame (PinholeCamera.cpp:23-55,91-103) with
 * the gluLookAt basis written out: f = normalize(center-eye),
 * X = normalize(f x up), Y = X x f, Z = -f; the camera looks along -Z. */
void orc_camera_lookat(const float eye[3], const float center[3], const float up[3],
                       float fovy_deg, float znear, uint32_t width, uint32_t height,
                       orc_camera* cam)
{
    float fw[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    normalize3(fw);
    float ax[3];
    cross3(fw, up, ax);
    normalize3(ax);
    float ay[3];
    cross3(ax, fw, ay);
    float az[3] = {-fw[0], -fw[1], -fw[2]};
    float sz = znear;
    float rad = (fovy_deg * 0.5f) * 0.017453292519943295f;
    float sy = (float)((double)sz * tan((double)rad));
    float sx = (sy * (float)width) / (float)height;
    float near_origin[3];
    for (int i = 0; i < 3; ++i) {
        near_origin[i] = ((eye[i] + ax[i] * -sx) + ay[i] * -sy) + az[i] * -sz;
        cam->pos[i] = eye[i];
        cam->nu[i] = ax[i] * (2.0f * sx);
        cam->nv[i] = ay[i] * (2.0f * sy);
        cam->base[i] = near_origin[i] - eye[i];
    }
    cam->width = width;
    cam->height = height;
}

/* PinholeCamera::getRayDirection (PinholeCamera.cpp:131-137) and the CUDA
 * kernel's ray setup (RayTracing_kernel.cu:145-161): pixel corners, row y
 * counted from the bottom, index y*W+x, direction not normalised.
 * tmin = rayEpsilon (IAccStruct.cpp:34), tmax = +inf. */
void orc_primary_rays(const orc_camera* cam, float* ox, float* oy, float* oz,
                      float* dx, float* dy, float* dz, float* tmin, float* tmax)
{
    const float inv_w = 1.0f / (float)cam->width;
    const float inv_h = 1.0f / (float)cam->height;
    for (uint32_t y = 0; y < cam->height; ++y) {
        for (uint32_t x = 0; x < cam->width; ++x) {
            size_t i = (size_t)y * cam->width + x;
            float us = (float)x * inv_w;
            float vs = (float)y * inv_h;
            ox[i] = cam->pos[0];
            oy[i] = cam->pos[1];
            oz[i] = cam->pos[2];
            dx[i] = (cam->base[0] + cam->nu[0] * us) + cam->nv[0] * vs;
            dy[i] = (cam->base[1] + cam->nu[1] * us) + cam->nv[1] * vs;
            dz[i] = (cam->base[2] + cam->nu[2] * us) + cam->nv[2] * vs;
            tmin[i] = ORC_RAY_EPSILON;
            tmax[i] = INFINITY;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a9: Moller-Trumbore                                                       */
/* ------------------------------------------------------------------------ */

/* Moller & Trumbore (1997) as in RayTracing_kernel.cu:56-89 and
 * traversalIntersection_FRAG.glsl:110-137: no determinant epsilon, edges
 * e1=v1-v0, e2=v2-v0, inside test u>=0, v>=0, u+v<=1.  det==0 gives inf/NaN
 * and the inside test then fails. */
static inline int mt_core(const float* v9, const float o[3], const float d[3],
                          float* t, float* u, float* v)
{
    const float e1[3] = {v9[3] - v9[0], v9[4] - v9[1], v9[5] - v9[2]};
    const float e2[3] = {v9[6] - v9[0], v9[7] - v9[1], v9[8] - v9[2]};
    float p[3];
    cross3(d, e2, p);
    const float det = dot3(p, e1);
    const float inv = 1.0f / det;
    const float tv[3] = {o[0] - v9[0], o[1] - v9[1], o[2] - v9[2]};
    float q[3];
    cross3(tv, e1, q);
    const float uu = dot3(p, tv) * inv;
    const float vv = dot3(q, d) * inv;
    const float tt = dot3(q, e2) * inv;
    *u = uu;
    *v = vv;
    *t = tt;
    return (uu >= 0.0f) && (vv >= 0.0f) && ((uu + vv) <= 1.0f);
}

int orc_mt(const float* v9, const float o[3], const float d[3], float tmin, float tmax,
           float* t, float* u, float* v)
{
    int in = mt_core(v9, o, d, t, u, v);
    return in && (*t > tmin) && (*t <= tmax);
}

/* ------------------------------------------------------------------------ */
/* a7+a8: clip and 3D-DDA                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t c[3], step[3], out[3];
    float tnext[3], tdelta[3];
} dda_state;

/* Ray::update + AabbIntersection::clipRay (Ray.h:26-37,
 * AabbIntersection.cpp:37-63) and the DDA set-up of traverse3ddda
 * (UniformGridAccStruct.cpp:118-149).  Returns 0 if the ray misses the box. */
static inline int dda_setup(const orc_frame* f, const float o[3], const float d[3],
                            float tmin, float tmax, dda_state* s)
{
    float inv[3];
    for (int i = 0; i < 3; ++i)
        inv[i] = 1.0f / d[i];
    float tn0 = tmin, tf0 = tmax;
    for (int i = 0; i < 3; ++i) {
        float a = (f->bmin[i] - o[i]) * inv[i];
        float b = (f->bmax[i] - o[i]) * inv[i];
        tn0 = tmax_f(tn0, tmin_f(a, b));
        tf0 = tmin_f(tf0, tmax_f(a, b));
    }
    if (tn0 > tf0)
        return 0;
    for (int i = 0; i < 3; ++i) {
        float p = o[i] + d[i] * tn0;
        s->c[i] = vox(f, p, i);
        int neg = signbit(d[i]) != 0;
        s->step[i] = neg ? -1 : 1;
        s->out[i] = neg ? -1 : f->n[i];
        float w = (float)(s->c[i] + (neg ? 0 : 1)) * f->cell[i] + f->bmin[i];
        s->tnext[i] = (w - o[i]) * inv[i];
        s->tdelta[i] = fabsf(f->cell[i] * inv[i]);
    }
    return 1;
}

/* Single-axis advance with the reference tie order
 * (UniformGridAccStruct.cpp:179-195).  Returns 0 when the ray leaves the grid. */
static inline int dda_step(dda_state* s)
{
    int a;
    if (s->tnext[0] < s->tnext[1] && s->tnext[0] < s->tnext[2])
        a = 0;
    else if (s->tnext[1] < s->tnext[2])
        a = 1;
    else
        a = 2;
    s->c[a] += s->step[a];
    if (s->c[a] == s->out[a])
        return 0;
    s->tnext[a] += s->tdelta[a];
    return 1;
}

/* Cell sequence of the DDA for one ray (test hook for the traversal order):
 * writes up to max_cells linear cell indices, returns the number visited. */
uint32_t orc_dda_cells(const orc_frame* f, const float o[3], const float d[3], float tmin,
                       float tmax, uint32_t* out, uint32_t max_cells)
{
    dda_state s;
    uint32_t k = 0;
    if (!dda_setup(f, o, d, tmin, tmax, &s))
        return 0;
    const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
    for (;;) {
        const uint32_t ci = (uint32_t)s.c[0] + nx * ((uint32_t)s.c[1] + ny * (uint32_t)s.c[2]);
        if (k < max_cells)
            out[k] = ci;
        ++k;
        const float exit_t = tmin_f(tmin_f(s.tnext[0], s.tnext[1]), s.tnext[2]);
        if (exit_t > tmax)
            break;
        if (!dda_step(&s))
            break;
    }
    return k;
}

static void trace_one_closest(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                              const float* verts, const float o[3], const float d[3],
                              float tmin, float tmax, int32_t* hid, float* ht, float* hu,
                              float* hv, orc_counters* cnt)
{
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    uint32_t cells = 0, tests = 0;
    dda_state s;
    if (dda_setup(f, o, d, tmin, tmax, &s)) {
        const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
        for (;;) {
            ++cells;
            const float exit_t = tmin_f(tmin_f(s.tnext[0], s.tnext[1]), s.tnext[2]);
            const uint32_t ci = (uint32_t)s.c[0] + nx * ((uint32_t)s.c[1] + ny * (uint32_t)s.c[2]);
            const uint32_t b = offsets[ci], e = offsets[ci + 1];
            if (b < e) {
                /* intersectTriangles (UniformGridAccStruct.cpp:362-378): hits are
                 * bounded by the exit of the current cell. */
                const float bound = exit_t * ORC_EXIT_SCALE;
                for (uint32_t k = b; k < e; ++k) {
                    const uint32_t tri = refs[k];
                    ++tests;
                    float t, u, v;
                    if (mt_core(verts + 9ull * tri, o, d, &t, &u, &v) && t > tmin &&
                        t <= tmax && t <= bound && t < best) {
                        best = t;
                        id = (int32_t)tri;
                        bu = u;
                        bv = v;
                    }
                }
                if (id >= 0)
                    break;
            }
            if (exit_t > tmax)
                break;
            if (!dda_step(&s))
                break;
        }
    }
    *hid = id;
    *ht = best;
    *hu = bu;
    *hv = bv;
    if (cnt) {
        cnt->cells = cells;
        cnt->tests = tests;
    }
}

static int trace_one_any(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                         const float* verts, const float o[3], const float d[3], float tmin,
                         float tmax, orc_counters* cnt)
{
    int occ = 0;
    uint32_t cells = 0, tests = 0;
    dda_state s;
    if (dda_setup(f, o, d, tmin, tmax, &s)) {
        const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
        for (;;) {
            ++cells;
            const float exit_t = tmin_f(tmin_f(s.tnext[0], s.tnext[1]), s.tnext[2]);
            const uint32_t ci = (uint32_t)s.c[0] + nx * ((uint32_t)s.c[1] + ny * (uint32_t)s.c[2]);
            const uint32_t b = offsets[ci], e = offsets[ci + 1];
            /* findAnyHit (shadow_FRAG.glsl:219-240): first hit in (tmin,tmax]
             * anywhere along the ray ends the search. */
            for (uint32_t k = b; k < e; ++k) {
                ++tests;
                float t, u, v;
                if (mt_core(verts + 9ull * refs[k], o, d, &t, &u, &v) && t > tmin && t <= tmax) {
                    occ = 1;
                    break;
                }
            }
            if (occ)
                break;
            if (exit_t > tmax)
                break;
            if (!dda_step(&s))
                break;
        }
    }
    if (cnt) {
        cnt->cells = cells;
        cnt->tests = tests;
    }
    return occ;
}

void orc_trace_closest(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                       const float* verts, uint32_t n_rays,
                       const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz,
                       const float* tmin, const float* tmax,
                       int32_t* hit_id, float* hit_t, float* hit_u, float* hit_v,
                       orc_counters* cnt, int n_threads, int chunk)
{
    if (chunk < 1)
        chunk = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        trace_one_closest(f, offsets, refs, verts, o, d, tmin[i], tmax[i], &hit_id[i], &hit_t[i],
                          &hit_u[i], &hit_v[i], cnt ? &cnt[i] : NULL);
    }
}

void orc_trace_any(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                   const float* verts, uint32_t n_rays,
                   const float* ox, const float* oy, const float* oz,
                   const float* dx, const float* dy, const float* dz,
                   const float* tmin, const float* tmax,
                   uint8_t* occluded, orc_counters* cnt, int n_threads, int chunk)
{
    if (chunk < 1)
        chunk = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        occluded[i] = (uint8_t)trace_one_any(f, offsets, refs, verts, o, d, tmin[i], tmax[i],
                                             cnt ? &cnt[i] : NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* Two-level grid                                                            */
/* ------------------------------------------------------------------------ */
/* A top cell c holding n_c >= min_refs references gets a uniform sub-grid of
 * r0 x r1 x r2 cells spanning exactly the cell: the top frame's rule applied
 * to the cell -- x = (lambda2 n_c) / ((cell0 cell1) cell2), f = cbrt(x),
 * r_a = ceil(cell_a f) clamped to [1, rmax] -- and none when r = (1,1,1).
 * Its frame: bmin_a = (float)c_a cell_a + bmin_a (the DDA's boundary formula),
 * cell'_a = cell_a / r_a, inv'_a = 1 / cell'_a.  A reference goes into every
 * sub-cell of its triangle's AABB box in that frame (clamp, truncate), with
 * ascending triangle ids per sub-cell.  cellsub[2c] = first sub-cell of c in
 * the global sub-cell order (0 without a sub-grid), cellsub[2c+1] = r0 |
 * r1 << 8 | r2 << 16 (0: no sub-grid); sub-cell (x, y, z) of c is
 * cellsub[2c] + x + r0 (y + r1 z).  An empty top cell has cellsub[2c+1] =
 * 0x80000000 | r, its skip radius: the largest r <= ORC_SKIP_MAX such that no
 * cell within r cells of it on every axis (Chebyshev distance; cells outside
 * the grid count as empty) holds a reference. */
#define ORC_SKIP_MAX 7
typedef struct {
    float bmin[3], cell[3], inv[3];
    int32_t r[3];
} sub_frame;

static inline int sub_frame_of(const orc_frame* f, uint32_t c, uint32_t packed, sub_frame* s)
{
    if (packed == 0 || (packed >> 31) != 0)
        return 0;
    const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
    const uint32_t cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    const uint32_t cc[3] = {cx, cy, cz};
    for (int a = 0; a < 3; ++a) {
        s->r[a] = (int32_t)((packed >> (8 * a)) & 0xffu);
        s->bmin[a] = (float)cc[a] * f->cell[a] + f->bmin[a];
        s->cell[a] = f->cell[a] / (float)s->r[a];
        s->inv[a] = 1.0f / s->cell[a];
    }
    return 1;
}

static inline int32_t sub_vox(const sub_frame* s, float val, int a)
{
    float x = (val - s->bmin[a]) * s->inv[a];
    x = x > 0.0f ? x : 0.0f;
    const float hi = (float)(s->r[a] - 1);
    x = x < hi ? x : hi;
    return (int32_t)x;
}

/* Skip radii of the empty cells: D = the Chebyshev distance to the nearest
 * occupied cell, capped at C = ORC_SKIP_MAX + 1, as three separable passes --
 * along x the distance to the nearest occupied cell of the row (a forward and
 * a backward sweep), then along y and z min over |k| < C of max(|k|, previous
 * pass at k) -- and r = D - 1. */
static void grid2_skip_radii(const orc_frame* f, const uint32_t* offsets, uint32_t* cellsub)
{
    const int32_t C = ORC_SKIP_MAX + 1;
    const int32_t n0 = f->n[0], n1 = f->n[1], n2 = f->n[2];
    const size_t m = f->m;
    uint8_t* d = (uint8_t*)malloc(m ? m : 1);
    uint8_t* e = (uint8_t*)malloc(m ? m : 1);
    for (int32_t z = 0; z < n2; ++z)
        for (int32_t y = 0; y < n1; ++y) {
            const size_t row = (size_t)n0 * ((size_t)y + (size_t)n1 * (size_t)z);
            int32_t last = -C;                  /* nearest occupied at or before x */
            for (int32_t x = 0; x < n0; ++x) {
                if (offsets[row + x + 1] != offsets[row + x])
                    last = x;
                const int32_t dist = x - last;
                d[row + x] = (uint8_t)(dist < C ? dist : C);
            }
            last = n0 + C;                      /* nearest occupied at or after x */
            for (int32_t x = n0 - 1; x >= 0; --x) {
                if (offsets[row + x + 1] != offsets[row + x])
                    last = x;
                const int32_t dist = last - x;
                if (dist < d[row + x])
                    d[row + x] = (uint8_t)dist;
            }
        }
    /* along y (d -> e), then z (e -> d) */
    for (int a = 1; a < 3; ++a) {
        const uint8_t* src = a == 1 ? d : e;
        uint8_t* dst = a == 1 ? e : d;
        const size_t stride = a == 1 ? (size_t)n0 : (size_t)n0 * (size_t)n1;
        const int32_t len = a == 1 ? n1 : n2;
        for (size_t c = 0; c < m; ++c) {
            const int32_t q = (int32_t)((c / stride) % (size_t)len);
            int32_t best = src[c];
            for (int32_t k = 1; k < C; ++k) {
                if (q - k >= 0) {
                    const int32_t v = src[c - (size_t)k * stride];
                    const int32_t cand = v > k ? v : k;
                    best = cand < best ? cand : best;
                }
                if (q + k < len) {
                    const int32_t v = src[c + (size_t)k * stride];
                    const int32_t cand = v > k ? v : k;
                    best = cand < best ? cand : best;
                }
            }
            dst[c] = (uint8_t)best;
        }
    }
    for (size_t c = 0; c < m; ++c)
        if (offsets[c + 1] == offsets[c])
            cellsub[2 * c + 1] = 0x80000000u | (uint32_t)(d[c] - 1);
    free(d);
    free(e);
}

uint64_t orc_grid2_plan(const orc_frame* f, const uint32_t* offsets, uint32_t min_refs,
                        float lambda2, uint32_t rmax, uint32_t* cellsub)
{
    const float vol = (f->cell[0] * f->cell[1]) * f->cell[2];
    uint64_t total = 0;
    for (uint32_t c = 0; c < f->m; ++c) {
        const uint32_t n = offsets[c + 1] - offsets[c];
        uint32_t packed = 0, m = 1;
        if (n >= min_refs && n > 0) {
            const float fac = orc_cbrtf((lambda2 * (float)n) / vol);
            uint32_t r[3];
            for (int a = 0; a < 3; ++a) {
                float x = ceilf(f->cell[a] * fac);
                x = x < 1.0f ? 1.0f : x;
                x = x > (float)rmax ? (float)rmax : x;
                r[a] = (uint32_t)x;
                m *= r[a];
            }
            if (m > 1)
                packed = r[0] | (r[1] << 8) | (r[2] << 16);
        }
        cellsub[2 * c] = packed ? (uint32_t)total : 0u;
        cellsub[2 * c + 1] = packed;
        if (packed)
            total += m;
    }
    grid2_skip_radii(f, offsets, cellsub);
    return total;
}

static inline void tri_sub_range(const sub_frame* s, const float* v9, int32_t lo[3],
                                 int32_t hi[3])
{
    for (int a = 0; a < 3; ++a) {
        float x = v9[a], y = v9[3 + a], z = v9[6 + a];
        float mn = x, mx = x;
        mn = y < mn ? y : mn;
        mx = y > mx ? y : mx;
        mn = z < mn ? z : mn;
        mx = z > mx ? z : mx;
        lo[a] = sub_vox(s, mn, a);
        hi[a] = sub_vox(s, mx, a);
    }
}

/* Sub-cell counts (n_sub entries) and their total. */
uint64_t orc_grid2_count(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                         const float* verts, const uint32_t* cellsub, uint32_t* counts,
                         uint64_t n_sub)
{
    memset(counts, 0, sizeof(uint32_t) * (size_t)n_sub);
    uint64_t total = 0;
    for (uint32_t c = 0; c < f->m; ++c) {
        sub_frame s;
        if (!sub_frame_of(f, c, cellsub[2 * c + 1], &s))
            continue;
        const uint32_t base = cellsub[2 * c];
        for (uint32_t k = offsets[c]; k < offsets[c + 1]; ++k) {
            int32_t lo[3], hi[3];
            tri_sub_range(&s, verts + 9ull * refs[k], lo, hi);
            for (int32_t z = lo[2]; z <= hi[2]; ++z)
                for (int32_t y = lo[1]; y <= hi[1]; ++y)
                    for (int32_t x = lo[0]; x <= hi[0]; ++x) {
                        counts[base + (uint32_t)x + (uint32_t)s.r[0] *
                                                        ((uint32_t)y + (uint32_t)s.r[1] * (uint32_t)z)]++;
                        ++total;
                    }
        }
    }
    return total;
}

void orc_grid2_fill(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                    const float* verts, const uint32_t* cellsub, const uint32_t* counts,
                    uint64_t n_sub, uint32_t* sub_offsets, uint32_t* sub_refs)
{
    uint32_t acc = 0;
    for (uint64_t q = 0; q < n_sub; ++q) {
        sub_offsets[q] = acc;
        acc += counts[q];
    }
    sub_offsets[n_sub] = acc;
    uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(n_sub ? n_sub : 1));
    memcpy(cur, sub_offsets, sizeof(uint32_t) * (size_t)n_sub);
    for (uint32_t c = 0; c < f->m; ++c) {
        sub_frame s;
        if (!sub_frame_of(f, c, cellsub[2 * c + 1], &s))
            continue;
        const uint32_t base = cellsub[2 * c];
        for (uint32_t k = offsets[c]; k < offsets[c + 1]; ++k) {
            int32_t lo[3], hi[3];
            tri_sub_range(&s, verts + 9ull * refs[k], lo, hi);
            for (int32_t z = lo[2]; z <= hi[2]; ++z)
                for (int32_t y = lo[1]; y <= hi[1]; ++y)
                    for (int32_t x = lo[0]; x <= hi[0]; ++x)
                        sub_refs[cur[base + (uint32_t)x + (uint32_t)s.r[0] *
                                                             ((uint32_t)y + (uint32_t)s.r[1] *
                                                                                (uint32_t)z)]++] = refs[k];
        }
    }
    free(cur);
}

/* The DDA inside one sub-grid from t_enter, with the ray's inverse direction:
 * the same set-up and single-axis step as the top level. */
typedef struct {
    int32_t c[3], step[3], out[3];
    float tnext[3], tdelta[3];
} sub_dda;

static inline void sub_dda_setup(const sub_frame* s, const float o[3], const float d[3],
                                 const float inv[3], float t_enter, sub_dda* q)
{
    for (int a = 0; a < 3; ++a) {
        const float p = o[a] + d[a] * t_enter;
        q->c[a] = sub_vox(s, p, a);
        const int neg = signbit(d[a]) != 0;
        q->step[a] = neg ? -1 : 1;
        q->out[a] = neg ? -1 : s->r[a];
        const float w = (float)(q->c[a] + (neg ? 0 : 1)) * s->cell[a] + s->bmin[a];
        q->tnext[a] = (w - o[a]) * inv[a];
        q->tdelta[a] = fabsf(s->cell[a] * inv[a]);
    }
}

static inline int sub_dda_step(sub_dda* q)
{
    int a;
    if (q->tnext[0] < q->tnext[1] && q->tnext[0] < q->tnext[2])
        a = 0;
    else if (q->tnext[1] < q->tnext[2])
        a = 1;
    else
        a = 2;
    q->c[a] += q->step[a];
    if (q->c[a] == q->out[a])
        return 0;
    q->tnext[a] += q->tdelta[a];
    return 1;
}

/* One ray through the two-level grid.  Closest hit: the first cell or
 * sub-cell holding a hit ends the ray, hits bounded by that (sub-)cell's exit;
 * any hit: the first triangle in (tmin, tmax].  Counters: top cells plus
 * sub-cells visited, triangle tests. */
static void grid2_trace_one(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                            const uint32_t* cellsub, const uint32_t* sub_offsets,
                            const uint32_t* sub_refs, const float* verts, const float o[3],
                            const float d[3], float tmin, float tmax, int any, int32_t* hid,
                            float* ht, float* hu, float* hv, uint8_t* occ, orc_counters* cnt)
{
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    int found_any = 0, done = 0;
    uint32_t cells = 0, tests = 0;
    dda_state st;
    float inv[3];
    for (int a = 0; a < 3; ++a)
        inv[a] = 1.0f / d[a];
    if (dda_setup(f, o, d, tmin, tmax, &st)) {
        float t_enter = tmin;
        for (int a = 0; a < 3; ++a) {   /* the clipped entry, as dda_setup computed it */
            const float p = (f->bmin[a] - o[a]) * inv[a], q = (f->bmax[a] - o[a]) * inv[a];
            t_enter = tmax_f(t_enter, tmin_f(p, q));
        }
        const uint32_t nx = (uint32_t)f->n[0], ny = (uint32_t)f->n[1];
        int neg[3];
        for (int a = 0; a < 3; ++a)
            neg[a] = signbit(d[a]) != 0;
        int32_t* c = st.c;
        /* The top walk goes item by item: a cell, or for an empty cell of skip
         * radius rr > 0 the cells within rr ahead of it on every axis, left in
         * one step.  An item's exit is its first boundary plane crossed, t =
         * ((float)k cell_a + bmin_a - o_a) / d_a, in the DDA's tie order (x
         * before y before z on ties).  After a skip the other two axes take
         * the cells of the exit point (vox), kept within the skipped cells. */
        for (;;) {
            ++cells;
            const uint32_t ci = (uint32_t)c[0] + nx * ((uint32_t)c[1] + ny * (uint32_t)c[2]);
            const int empty = offsets[ci + 1] == offsets[ci];
            const int32_t rr = empty ? (int32_t)(cellsub[2 * ci + 1] & 0xffu) : 0;
            float tp[3];
            for (int a = 0; a < 3; ++a) {
                const int32_t k = neg[a] ? c[a] - rr : c[a] + rr + 1;
                tp[a] = ((float)k * f->cell[a] + f->bmin[a] - o[a]) * inv[a];
            }
            const int ax = (tp[0] < tp[1] && tp[0] < tp[2]) ? 0 : (tp[1] < tp[2]) ? 1 : 2;
            const float exit_t = tp[ax];
            sub_frame sf;
            if (empty) {
                /* nothing to test */
            } else if (sub_frame_of(f, ci, cellsub[2 * ci + 1], &sf)) {
                const uint32_t base = cellsub[2 * ci];
                sub_dda q;
                sub_dda_setup(&sf, o, d, inv, t_enter, &q);
                for (;;) {
                    ++cells;
                    const float sexit = tmin_f(tmin_f(q.tnext[0], q.tnext[1]), q.tnext[2]);
                    const uint32_t si = base + (uint32_t)q.c[0] +
                                        (uint32_t)sf.r[0] * ((uint32_t)q.c[1] +
                                                             (uint32_t)sf.r[1] * (uint32_t)q.c[2]);
                    const uint32_t b = sub_offsets[si], e = sub_offsets[si + 1];
                    const float bound = sexit * ORC_EXIT_SCALE;
                    for (uint32_t k = b; k < e; ++k) {
                        const uint32_t tri = sub_refs[k];
                        ++tests;
                        float t, u, v;
                        if (!mt_core(verts + 9ull * tri, o, d, &t, &u, &v) || !(t > tmin) ||
                            !(t <= tmax))
                            continue;
                        if (any) {
                            found_any = 1;
                            break;
                        }
                        if (t <= bound && t < best) {
                            best = t;
                            id = (int32_t)tri;
                            bu = u;
                            bv = v;
                        }
                    }
                    if (found_any || id >= 0 || sexit > tmax) {
                        done = 1;
                        break;
                    }
                    if (!sub_dda_step(&q))
                        break;
                }
            } else {
                const uint32_t b = offsets[ci], e = offsets[ci + 1];
                const float bound = exit_t * ORC_EXIT_SCALE;
                for (uint32_t k = b; k < e; ++k) {
                    const uint32_t tri = refs[k];
                    ++tests;
                    float t, u, v;
                    if (!mt_core(verts + 9ull * tri, o, d, &t, &u, &v) || !(t > tmin) ||
                        !(t <= tmax))
                        continue;
                    if (any) {
                        found_any = 1;
                        break;
                    }
                    if (t <= bound && t < best) {
                        best = t;
                        id = (int32_t)tri;
                        bu = u;
                        bv = v;
                    }
                }
                if (found_any || id >= 0)
                    done = 1;
            }
            if (done || exit_t > tmax)
                break;
            t_enter = exit_t;
            const int32_t next = neg[ax] ? c[ax] - rr - 1 : c[ax] + rr + 1;
            if (next < 0 || next >= f->n[ax])
                break;
            c[ax] = next;
            if (rr > 0)
                for (int a = 0; a < 3; ++a) {
                    if (a == ax)
                        continue;
                    int32_t v = vox(f, o[a] + d[a] * exit_t, a);
                    if (neg[a]) {
                        v = v < c[a] ? v : c[a];
                        c[a] = v > c[a] - rr ? v : c[a] - rr;
                    } else {
                        v = v > c[a] ? v : c[a];
                        c[a] = v < c[a] + rr ? v : c[a] + rr;
                    }
                }
        }
    }
    if (any) {
        *occ = (uint8_t)found_any;
    } else {
        *hid = id;
        *ht = best;
        *hu = bu;
        *hv = bv;
    }
    if (cnt) {
        cnt->cells = cells;
        cnt->tests = tests;
    }
}

void orc_grid2_trace_closest(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                             const uint32_t* cellsub, const uint32_t* sub_offsets,
                             const uint32_t* sub_refs, const float* verts, uint32_t n_rays,
                             const float* ox, const float* oy, const float* oz,
                             const float* dx, const float* dy, const float* dz,
                             const float* tmin, const float* tmax,
                             int32_t* hit_id, float* hit_t, float* hit_u, float* hit_v,
                             orc_counters* cnt, int n_threads, int chunk)
{
    if (chunk < 1)
        chunk = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        grid2_trace_one(f, offsets, refs, cellsub, sub_offsets, sub_refs, verts, o, d, tmin[i],
                        tmax[i], 0, &hit_id[i], &hit_t[i], &hit_u[i], &hit_v[i], NULL,
                        cnt ? &cnt[i] : NULL);
    }
}

void orc_grid2_trace_any(const orc_frame* f, const uint32_t* offsets, const uint32_t* refs,
                         const uint32_t* cellsub, const uint32_t* sub_offsets,
                         const uint32_t* sub_refs, const float* verts, uint32_t n_rays,
                         const float* ox, const float* oy, const float* oz,
                         const float* dx, const float* dy, const float* dz,
                         const float* tmin, const float* tmax,
                         uint8_t* occluded, orc_counters* cnt, int n_threads, int chunk)
{
    if (chunk < 1)
        chunk = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        grid2_trace_one(f, offsets, refs, cellsub, sub_offsets, sub_refs, verts, o, d, tmin[i],
                        tmax[i], 1, NULL, NULL, NULL, NULL, &occluded[i], cnt ? &cnt[i] : NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* kd-tree traversal (the comparison accelerator)                            */
/* ------------------------------------------------------------------------ */
#define ORC_KD_MAX_DEPTH 64

/* KdTreeAccStruct::traceNearestGeometry / findLeaf (KdTreeAccStruct.cpp:
 * 122-201) over a flattened tree: node k is words [2k, 2k+1] -- inner
 * {split bits, (left << 2) | axis} with the right child at left + 1, leaf
 * {first element, 0x80000000 | count}; elements[] are triangle ids.  The ray
 * is clipped to box[0..5] (min xyz, max xyz) like the grid; d < tnear takes
 * the far child, d > tfar the near child, else the near child with
 * tfar = min(d, tfar) and the far child stacked with [max(d, tnear), tfar]
 * (selects as written, so a NaN d keeps the interval).  Closest hit: leaf
 * triangles tested with the grid's window, the leaf's tfar as the exit
 * bound; the first leaf with a hit ends the ray.  Any hit: (tmin, tmax]. */
static void kd_trace_one(const float box[6], const uint32_t* nodes, const uint32_t* elements,
                         const float* verts, const float o[3], const float d[3], float tmin,
                         float tmax, int any, int32_t* hid, float* ht, float* hu, float* hv,
                         uint8_t* occ, orc_counters* cnt)
{
    int32_t id = -1;
    float best = INFINITY, bu = 0.0f, bv = 0.0f;
    int hit_any = 0;
    uint32_t leaves = 0, tests = 0;
    float inv[3];
    for (int i = 0; i < 3; ++i)
        inv[i] = 1.0f / d[i];
    float tn = tmin, tf = tmax;
    for (int i = 0; i < 3; ++i) {
        const float a = (box[i] - o[i]) * inv[i];
        const float b = (box[3 + i] - o[i]) * inv[i];
        tn = tmax_f(tn, tmin_f(a, b));
        tf = tmin_f(tf, tmax_f(a, b));
    }
    if (tn <= tf) {
        uint32_t st_node[ORC_KD_MAX_DEPTH];
        float st_tn[ORC_KD_MAX_DEPTH], st_tf[ORC_KD_MAX_DEPTH];
        int sp = 0;
        uint32_t node = 0;
        for (;;) {
            while (!(nodes[2 * node + 1] & 0x80000000u)) {
                const uint32_t w = nodes[2 * node + 1];
                const int axis = (int)(w & 3u);
                const uint32_t left = w >> 2;
                float split;
                memcpy(&split, &nodes[2 * node], 4);
                const float dist = (split - o[axis]) * inv[axis];
                const int neg = signbit(d[axis]) != 0;
                const uint32_t front = left + (neg ? 1u : 0u);
                const uint32_t back = left + (neg ? 0u : 1u);
                if (dist < tn) {
                    node = back;
                } else if (dist > tf) {
                    node = front;
                } else {
                    if (sp < ORC_KD_MAX_DEPTH) {
                        st_node[sp] = back;
                        st_tn[sp] = dist > tn ? dist : tn;
                        st_tf[sp] = tf;
                        ++sp;
                    }
                    node = front;
                    tf = dist < tf ? dist : tf;
                }
            }
            ++leaves;
            const uint32_t first = nodes[2 * node];
            const uint32_t end = first + (nodes[2 * node + 1] & 0x7fffffffu);
            if (any) {
                for (uint32_t k = first; k < end; ++k) {
                    ++tests;
                    float t, u, v;
                    if (mt_core(verts + 9ull * elements[k], o, d, &t, &u, &v) && t > tmin &&
                        t <= tmax) {
                        hit_any = 1;
                        break;
                    }
                }
                if (hit_any)
                    break;
            } else {
                const float bound = tf * ORC_EXIT_SCALE;
                for (uint32_t k = first; k < end; ++k) {
                    const uint32_t tri = elements[k];
                    ++tests;
                    float t, u, v;
                    if (mt_core(verts + 9ull * tri, o, d, &t, &u, &v) && t > tmin && t <= tmax &&
                        t <= bound && t < best) {
                        best = t;
                        id = (int32_t)tri;
                        bu = u;
                        bv = v;
                    }
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
    if (any) {
        *occ = (uint8_t)hit_any;
    } else {
        *hid = id;
        *ht = best;
        *hu = bu;
        *hv = bv;
    }
    if (cnt) {
        cnt->cells = leaves;
        cnt->tests = tests;
    }
}

void orc_kd_trace_closest(const float box[6], const uint32_t* nodes, const uint32_t* elements,
                          const float* verts, uint32_t n_rays,
                          const float* ox, const float* oy, const float* oz,
                          const float* dx, const float* dy, const float* dz,
                          const float* tmin, const float* tmax,
                          int32_t* hit_id, float* hit_t, float* hit_u, float* hit_v,
                          orc_counters* cnt, int n_threads, int chunk)
{
    if (chunk < 1)
        chunk = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        kd_trace_one(box, nodes, elements, verts, o, d, tmin[i], tmax[i], 0, &hit_id[i],
                     &hit_t[i], &hit_u[i], &hit_v[i], NULL, cnt ? &cnt[i] : NULL);
    }
}

void orc_kd_trace_any(const float box[6], const uint32_t* nodes, const uint32_t* elements,
                      const float* verts, uint32_t n_rays,
                      const float* ox, const float* oy, const float* oz,
                      const float* dx, const float* dy, const float* dz,
                      const float* tmin, const float* tmax,
                      uint8_t* occluded, orc_counters* cnt, int n_threads, int chunk)
{
    if (chunk < 1)
        chunk = 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, chunk) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t i = 0; i < (int64_t)n_rays; ++i) {
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        kd_trace_one(box, nodes, elements, verts, o, d, tmin[i], tmax[i], 1, NULL, NULL, NULL,
                     NULL, &occluded[i], cnt ? &cnt[i] : NULL);
    }
}

/* ------------------------------------------------------------------------ */
/* a10-a12: secondary rays                                                   */
/* ------------------------------------------------------------------------ */

/* Sample::computeShadingNormal (Sample.cpp:100-114) with the CUDA weights
 * (RayTracing_kernel.cu:99-101): w0 = 1-(u+v). */
void orc_shading_normal(const float* n9, float u, float v, float out[3])
{
    const float w = 1.0f - (u + v);
    for (int i = 0; i < 3; ++i)
        out[i] = (n9[i] * w + n9[3 + i] * u) + n9[6 + i] * v;
    normalize3(out);
}

/* Hit point (Sample.cpp:95-98) and the shading normal turned to face the
 * incoming ray (surfaces are two-sided). */
static inline void hit_frame(const float o[3], const float d[3], float t, const float* n9,
                             float u, float v, float h[3], float nn[3])
{
    for (int i = 0; i < 3; ++i)
        h[i] = o[i] + d[i] * t;
    orc_shading_normal(n9, u, v, nn);
    if (dot3(nn, d) > 0.0f) {
        nn[0] = -nn[0];
        nn[1] = -nn[1];
        nn[2] = -nn[2];
    }
}

static inline void set_inactive(size_t i, float* ox, float* oy, float* oz, float* dx, float* dy,
                                float* dz, float* tn, float* tf)
{
    ox[i] = 0.0f;
    oy[i] = 0.0f;
    oz[i] = 0.0f;
    dx[i] = 0.0f;
    dy[i] = 0.0f;
    dz[i] = 1.0f;
    tn[i] = 1.0f;
    tf[i] = -1.0f;
}

/* Shadow rays (shadow_FRAG.glsl:259-270; SimplePointLight.cpp:18-37;
 * Sample.cpp:280-292): origin = hit + N*offset, direction = normalised vector
 * to the light, (0, distance to light]; no ray when N.L <= 0. */
void orc_gen_shadow(uint32_t n, const float* ox, const float* oy, const float* oz,
                    const float* dx, const float* dy, const float* dz,
                    const int32_t* hit_id, const float* hit_t, const float* hit_u,
                    const float* hit_v, const float* normals, const float light[3],
                    float offset,
                    float* sox, float* soy, float* soz, float* sdx, float* sdy, float* sdz,
                    float* stmin, float* stmax)
{
    for (size_t i = 0; i < n; ++i) {
        if (hit_id[i] < 0) {
            set_inactive(i, sox, soy, soz, sdx, sdy, sdz, stmin, stmax);
            continue;
        }
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        float h[3], nn[3];
        hit_frame(o, d, hit_t[i], normals + 9ull * (uint32_t)hit_id[i], hit_u[i], hit_v[i], h, nn);
        float l[3] = {light[0] - h[0], light[1] - h[1], light[2] - h[2]};
        if (dot3(nn, l) <= 0.0f) {
            set_inactive(i, sox, soy, soz, sdx, sdy, sdz, stmin, stmax);
            continue;
        }
        const float dist = sqrtf(dot3(l, l));
        const float inv = 1.0f / dist;
        sox[i] = h[0] + nn[0] * offset;
        soy[i] = h[1] + nn[1] * offset;
        soz[i] = h[2] + nn[2] * offset;
        sdx[i] = l[0] * inv;
        sdy[i] = l[1] * inv;
        sdz[i] = l[2] * inv;
        stmin[i] = 0.0f;
        stmax[i] = dist;
    }
}

/* Mirror reflection (Sample.cpp:7-11,44-52; traversalIntersection_FRAG.glsl:
 * 285-296): r = d - N*(2*(d.N)), origin = hit + N*offset, (0, +inf). */
void orc_gen_reflection(uint32_t n, const float* ox, const float* oy, const float* oz,
                        const float* dx, const float* dy, const float* dz,
                        const int32_t* hit_id, const float* hit_t, const float* hit_u,
                        const float* hit_v, const float* normals, float offset,
                        float* rox, float* roy, float* roz, float* rdx, float* rdy, float* rdz,
                        float* rtmin, float* rtmax)
{
    for (size_t i = 0; i < n; ++i) {
        if (hit_id[i] < 0) {
            set_inactive(i, rox, roy, roz, rdx, rdy, rdz, rtmin, rtmax);
            continue;
        }
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        float h[3], nn[3];
        hit_frame(o, d, hit_t[i], normals + 9ull * (uint32_t)hit_id[i], hit_u[i], hit_v[i], h, nn);
        const float proj = dot3(d, nn) * 2.0f;
        rox[i] = h[0] + nn[0] * offset;
        roy[i] = h[1] + nn[1] * offset;
        roz[i] = h[2] + nn[2] * offset;
        rdx[i] = d[0] - nn[0] * proj;
        rdy[i] = d[1] - nn[1] * proj;
        rdz[i] = d[2] - nn[2] * proj;
        rtmin[i] = 0.0f;
        rtmax[i] = INFINITY;
    }
}

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

float orc_uniform24(uint64_t seed, uint64_t a, uint64_t b)
{
    uint64_t h = mix64(seed ^ mix64(a * 0x9e3779b97f4a7c15ull + b));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

/* sin and cos of 2*pi*turns.  Quadrant reduction in turns (exact), then
 * Taylor polynomials of degree 9 (sin) and 8 (cos) on [-pi/4, pi/4]. */
void orc_sincos_turns(float turns, float* s_out, float* c_out)
{
    const float y = turns * 4.0f;
    const float q = rintf(y);
    const float r = y - q;
    const float a = r * 1.5707963267948966f;
    const float a2 = a * a;
    float sp = fmaf(a2, 2.7557319e-6f, -1.9841270e-4f);
    sp = fmaf(a2, sp, 8.3333333e-3f);
    sp = fmaf(a2, sp, -1.6666667e-1f);
    sp = fmaf(a2, sp, 1.0f);
    const float sa = a * sp;
    float cp = fmaf(a2, 2.4801587e-5f, -1.3888889e-3f);
    cp = fmaf(a2, cp, 4.1666667e-2f);
    cp = fmaf(a2, cp, -0.5f);
    const float ca = fmaf(a2, cp, 1.0f);
    const int qi = ((int)q) & 3;
    float s, c;
    switch (qi) {
    case 0: s = sa; c = ca; break;
    case 1: s = ca; c = -sa; break;
    case 2: s = -sa; c = -ca; break;
    default: s = -ca; c = sa; break;
    }
    *s_out = s;
    *c_out = c;
}

/* Cosine-weighted hemisphere directions (Malley's method) around the shading
 * normal, basis of Duff et al. (2017), origin = hit + N*offset, (0, ao_tmax].
 * Not in the reference; reuses the any-hit semantics of shadow rays. */
void orc_gen_ao(uint32_t n, uint32_t spp, uint64_t seed,
                const float* ox, const float* oy, const float* oz,
                const float* dx, const float* dy, const float* dz,
                const int32_t* hit_id, const float* hit_t, const float* hit_u,
                const float* hit_v, const float* normals, float offset, float ao_tmax,
                float* aox, float* aoy, float* aoz, float* adx, float* ady, float* adz,
                float* atmin, float* atmax)
{
    for (size_t i = 0; i < n; ++i) {
        if (hit_id[i] < 0) {
            for (uint32_t k = 0; k < spp; ++k)
                set_inactive(i * spp + k, aox, aoy, aoz, adx, ady, adz, atmin, atmax);
            continue;
        }
        const float o[3] = {ox[i], oy[i], oz[i]};
        const float d[3] = {dx[i], dy[i], dz[i]};
        float h[3], nn[3];
        hit_frame(o, d, hit_t[i], normals + 9ull * (uint32_t)hit_id[i], hit_u[i], hit_v[i], h, nn);
        const float sgn = copysignf(1.0f, nn[2]);
        const float a = -1.0f / (sgn + nn[2]);
        const float b = (nn[0] * nn[1]) * a;
        const float tb[3] = {1.0f + ((sgn * nn[0]) * nn[0]) * a, sgn * b, -(sgn * nn[0])};
        const float bb[3] = {b, sgn + (nn[1] * nn[1]) * a, -nn[1]};
        const float org[3] = {h[0] + nn[0] * offset, h[1] + nn[1] * offset, h[2] + nn[2] * offset};
        for (uint32_t k = 0; k < spp; ++k) {
            const size_t j = i * spp + k;
            const float u1 = orc_uniform24(seed, (uint64_t)i, 2ull * k);
            const float u2 = orc_uniform24(seed, (uint64_t)i, 2ull * k + 1ull);
            const float r = sqrtf(u1);
            float sn, cs;
            orc_sincos_turns(u2, &sn, &cs);
            const float lx = r * cs, ly = r * sn, lz = sqrtf(1.0f - u1);
            aox[j] = org[0];
            aoy[j] = org[1];
            aoz[j] = org[2];
            adx[j] = fmaf(nn[0], lz, fmaf(bb[0], ly, tb[0] * lx));
            ady[j] = fmaf(nn[1], lz, fmaf(bb[1], ly, tb[1] * lx));
            adz[j] = fmaf(nn[2], lz, fmaf(bb[2], ly, tb[2] * lx));
            atmin[j] = 0.0f;
            atmax[j] = ao_tmax;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Per-frame rigid transform                                                 */
/* ------------------------------------------------------------------------ */
void orc_transform(uint32_t n_tris, uint32_t tris_per_cluster, uint32_t n_cluster_tris,
                   const float* xf, const float* rest_verts, const float* rest_normals,
                   float* verts, float* normals)
{
    for (uint32_t t = 0; t < n_tris; ++t) {
        const float* rv = rest_verts + 9ull * t;
        const float* rn = rest_normals + 9ull * t;
        float* ov = verts + 9ull * t;
        float* on = normals + 9ull * t;
        if (t >= n_cluster_tris) {
            memcpy(ov, rv, 9 * sizeof(float));
            memcpy(on, rn, 9 * sizeof(float));
            continue;
        }
        const float* m = xf + 12ull * (t / tris_per_cluster);
        for (int k = 0; k < 3; ++k) {
            const float x = rv[3 * k], y = rv[3 * k + 1], z = rv[3 * k + 2];
            const float a = rn[3 * k], b = rn[3 * k + 1], c = rn[3 * k + 2];
            for (int r = 0; r < 3; ++r) {
                ov[3 * k + r] = fmaf(m[4 * r], x, fmaf(m[4 * r + 1], y, fmaf(m[4 * r + 2], z, m[4 * r + 3])));
                on[3 * k + r] = fmaf(m[4 * r], a, fmaf(m[4 * r + 1], b, m[4 * r + 2] * c));
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* FP64 brute force                                                          */
/* ------------------------------------------------------------------------ */
static inline int mt_f64(const float* v9, const double o[3], const double d[3],
                         double* t, double* u, double* v)
{
    const double v0[3] = {v9[0], v9[1], v9[2]};
    const double e1[3] = {(double)v9[3] - v0[0], (double)v9[4] - v0[1], (double)v9[5] - v0[2]};
    const double e2[3] = {(double)v9[6] - v0[0], (double)v9[7] - v0[1], (double)v9[8] - v0[2]};
    const double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2],
                         d[0] * e2[1] - d[1] * e2[0]};
    const double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (det == 0.0)
        return 0;
    const double tv[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
    const double q[3] = {tv[1] * e1[2] - tv[2] * e1[1], tv[2] * e1[0] - tv[0] * e1[2],
                         tv[0] * e1[1] - tv[1] * e1[0]};
    *u = (tv[0] * p[0] + tv[1] * p[1] + tv[2] * p[2]) / det;
    *v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) / det;
    *t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) / det;
    return (*u >= 0.0) && (*v >= 0.0) && (*u + *v <= 1.0);
}

/* Both brute-force passes take the rays in blocks of ORC_BRUTE_BLOCK: each
 * triangle is read once per block and tested against every ray of it, so a
 * scene far larger than the caches streams once per block instead of once per
 * ray.  Each ray still sees the triangles in ascending order, so the result is
 * that of the plain double loop. */
#define ORC_BRUTE_BLOCK 32

void orc_brute_closest_f64(const float* verts, uint32_t n_tris, uint32_t n_rays,
                           const float* ox, const float* oy, const float* oz,
                           const float* dx, const float* dy, const float* dz,
                           const float* tmin, const float* tmax,
                           int32_t* id, double* t1, double* t2, double* u, double* v,
                           int n_threads)
{
    const int64_t n_blocks = ((int64_t)n_rays + ORC_BRUTE_BLOCK - 1) / ORC_BRUTE_BLOCK;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t blk = 0; blk < n_blocks; ++blk) {
        const int64_t i0 = blk * ORC_BRUTE_BLOCK;
        const int nb = (int)((int64_t)n_rays - i0 < ORC_BRUTE_BLOCK ? (int64_t)n_rays - i0
                                                                     : ORC_BRUTE_BLOCK);
        double o[ORC_BRUTE_BLOCK][3], d[ORC_BRUTE_BLOCK][3], lo[ORC_BRUTE_BLOCK],
            hi[ORC_BRUTE_BLOCK];
        double b1[ORC_BRUTE_BLOCK], b2[ORC_BRUTE_BLOCK], bu[ORC_BRUTE_BLOCK], bv[ORC_BRUTE_BLOCK];
        int32_t bid[ORC_BRUTE_BLOCK];
        for (int j = 0; j < nb; ++j) {
            const int64_t i = i0 + j;
            o[j][0] = ox[i], o[j][1] = oy[i], o[j][2] = oz[i];
            d[j][0] = dx[i], d[j][1] = dy[i], d[j][2] = dz[i];
            lo[j] = tmin[i];
            hi[j] = tmax[i];
            b1[j] = b2[j] = INFINITY;
            bu[j] = bv[j] = 0.0;
            bid[j] = -1;
        }
        for (uint32_t k = 0; k < n_tris; ++k) {
            const float* tri = verts + 9ull * k;
            for (int j = 0; j < nb; ++j) {
                double t, uu, vv;
                if (!mt_f64(tri, o[j], d[j], &t, &uu, &vv) || !(t > lo[j]) || !(t <= hi[j]))
                    continue;
                if (t < b1[j]) {
                    b2[j] = b1[j];
                    b1[j] = t;
                    bid[j] = (int32_t)k;
                    bu[j] = uu;
                    bv[j] = vv;
                } else if (t < b2[j]) {
                    b2[j] = t;
                }
            }
        }
        for (int j = 0; j < nb; ++j) {
            const int64_t i = i0 + j;
            id[i] = bid[j];
            t1[i] = b1[j];
            t2[i] = b2[j];
            u[i] = bu[j];
            v[i] = bv[j];
        }
    }
}

void orc_brute_any_f64(const float* verts, uint32_t n_tris, uint32_t n_rays,
                       const float* ox, const float* oy, const float* oz,
                       const float* dx, const float* dy, const float* dz,
                       const float* tmin, const float* tmax,
                       uint8_t* occluded, double* t_near, int n_threads)
{
    const int64_t n_blocks = ((int64_t)n_rays + ORC_BRUTE_BLOCK - 1) / ORC_BRUTE_BLOCK;
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(n_threads > 1 ? n_threads : 1) if (n_threads > 1)
#endif
    for (int64_t blk = 0; blk < n_blocks; ++blk) {
        const int64_t i0 = blk * ORC_BRUTE_BLOCK;
        const int nb = (int)((int64_t)n_rays - i0 < ORC_BRUTE_BLOCK ? (int64_t)n_rays - i0
                                                                     : ORC_BRUTE_BLOCK);
        double o[ORC_BRUTE_BLOCK][3], d[ORC_BRUTE_BLOCK][3], lo[ORC_BRUTE_BLOCK],
            best[ORC_BRUTE_BLOCK];
        for (int j = 0; j < nb; ++j) {
            const int64_t i = i0 + j;
            o[j][0] = ox[i], o[j][1] = oy[i], o[j][2] = oz[i];
            d[j][0] = dx[i], d[j][1] = dy[i], d[j][2] = dz[i];
            lo[j] = tmin[i];
            best[j] = INFINITY;
        }
        for (uint32_t k = 0; k < n_tris; ++k) {
            const float* tri = verts + 9ull * k;
            for (int j = 0; j < nb; ++j) {
                double t, uu, vv;
                if (!mt_f64(tri, o[j], d[j], &t, &uu, &vv) || !(t > lo[j]))
                    continue;
                if (t < best[j])
                    best[j] = t;
            }
        }
        for (int j = 0; j < nb; ++j) {
            const int64_t i = i0 + j;
            const double hi = tmax[i];
            t_near[i] = best[j];
            occluded[i] = (uint8_t)(lo[j] <= hi && best[j] <= hi);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Headlight shading (RayTracing_kernel.cu:92-120)                           */
/* ------------------------------------------------------------------------ */
static inline uint32_t to_u8(float c)
{
    float x = c * 255.0f;
    x = x > 0.0f ? x : 0.0f;
    x = x < 255.0f ? x : 255.0f;
    return (uint32_t)x;
}

uint32_t orc_headlight_rgba8(const float d[3], const float* n9, int32_t id, float u, float v)
{
    if (id < 0)
        return 0xffffffffu;
    float dn[3] = {d[0], d[1], d[2]};
    normalize3(dn);
    float n[3];
    orc_shading_normal(n9, u, v, n);
    const float ndd = -dot3(n, dn);
    /* (ambient + (diffuse - ambient)*nDotD) * diffuse, ambient 0.1, diffuse (0,0,1) */
    const float r = (0.1f + (0.0f - 0.1f) * ndd) * 0.0f;
    const float g = (0.1f + (0.0f - 0.1f) * ndd) * 0.0f;
    const float b = (0.1f + (1.0f - 0.1f) * ndd) * 1.0f;
    return to_u8(r) | (to_u8(g) << 8) | (to_u8(b) << 16) | (255u << 24);
}

void orc_headlight_image(uint32_t n, const float* dx, const float* dy, const float* dz,
                         const float* normals9, const int32_t* id, const float* u,
                         const float* v, uint32_t* rgba)
{
    for (uint32_t i = 0; i < n; ++i) {
        const float d[3] = {dx[i], dy[i], dz[i]};
        const float* n9 = id[i] >= 0 ? normals9 + (size_t)id[i] * 9 : normals9;
        rgba[i] = orc_headlight_rgba8(d, n9, id[i], u[i], v[i]);
    }
}

/* ------------------------------------------------------------------------ */
/* Phong shading (shading_FRAG.glsl:116-173; PhongMaterialColor.cpp:19-85)   */
/* ------------------------------------------------------------------------ */
float orc_pow_u32(float x, uint32_t e)
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

/* Colour of one hit: hit point and shading normal facing the ray as for the
 * secondary rays; L = normalised vector to the light; the point is lit when
 * N.L > 0 and its shadow ray is not occluded (a point facing away from the
 * light gets no shadow ray and counts as shadowed).  Specular vector =
 * normalise(reflect(d, N)); specular = s x lc x max(0, S.L)^e; diffuse =
 * max(0, kd x lc x N.L); lit: ambient + diffuse + specular, shadowed:
 * ambient + diffuse x shadow_coeff. */
static void phong_colour(const float o[3], const float d[3], float t, const float* n9, float u,
                         float v, int shadowed, const float kd[3], const orc_phong* p,
                         float col[3])
{
    float h[3], nn[3];
    hit_frame(o, d, t, n9, u, v, h, nn);
    const float l[3] = {p->light[0] - h[0], p->light[1] - h[1], p->light[2] - h[2]};
    const int lit = dot3(nn, l) > 0.0f && !shadowed;
    const float dist = sqrtf(dot3(l, l));
    const float inv = 1.0f / dist;
    const float L[3] = {l[0] * inv, l[1] * inv, l[2] * inv};
    const float proj = dot3(d, nn) * 2.0f;
    float s[3] = {d[0] - nn[0] * proj, d[1] - nn[1] * proj, d[2] - nn[2] * proj};
    normalize3(s);
    const float sdl = dot3(s, L);
    const float spec = (p->spec_exp > 0 && sdl > 0.0f) ? orc_pow_u32(sdl, p->spec_exp) : 0.0f;
    const float ndl = dot3(nn, L);
    for (int c = 0; c < 3; ++c) {
        float dif = kd[c] * (p->light_color[c] * ndl);
        dif = dif > 0.0f ? dif : 0.0f;
        if (lit)
            col[c] = (p->ambient[c] + dif) + p->specular[c] * (p->light_color[c] * spec);
        else
            col[c] = p->ambient[c] + dif * p->shadow_coeff;
    }
}

void orc_shade_phong(uint32_t n, const float* ox, const float* oy, const float* oz,
                     const float* dx, const float* dy, const float* dz, const int32_t* hit_id,
                     const float* hit_t, const float* hit_u, const float* hit_v,
                     const uint8_t* occluded, const float* normals, const float* diffuse_rgb,
                     const orc_phong* p, uint32_t depth, float* rgb, uint32_t* rgba8)
{
    float w = 1.0f;
    for (uint32_t k = 0; k < depth; ++k)
        w = w * p->reflect_coeff;
    static const float white[3] = {1.0f, 1.0f, 1.0f};
    for (size_t i = 0; i < n; ++i) {
        float* out = rgb + 3 * i;
        if (hit_id[i] >= 0) {
            const uint32_t id = (uint32_t)hit_id[i];
            const float o[3] = {ox[i], oy[i], oz[i]};
            const float d[3] = {dx[i], dy[i], dz[i]};
            const float* kd = diffuse_rgb ? diffuse_rgb + 3ull * id : white;
            float col[3];
            phong_colour(o, d, hit_t[i], normals + 9ull * id, hit_u[i], hit_v[i],
                         occluded ? occluded[i] != 0 : 0, kd, p, col);
            for (int c = 0; c < 3; ++c)
                out[c] = depth == 0 ? col[c] : out[c] + w * col[c];
        } else if (depth == 0) {
            for (int c = 0; c < 3; ++c)
                out[c] = p->background[c];
        }
        if (rgba8)
            rgba8[i] = to_u8(out[0]) | (to_u8(out[1]) << 8) | (to_u8(out[2]) << 16) | (255u << 24);
    }
}
