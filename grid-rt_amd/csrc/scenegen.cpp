// scenegen.cpp -- seeded synthetic scenes (include/grt/grt_scenes.h).
#include <cmath>
#include <cstdint>
#include <cstring>

#include "grt/grt_scenes.h"

namespace {

// splitmix64 (Steele, Lea, Flood 2014): seeds the xoshiro state.
inline uint64_t splitmix64(uint64_t& x)
{
    uint64_t z = (x += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

// xoshiro256** 1.0 (Blackman & Vigna 2018).
struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed)
    {
        uint64_t x = seed;
        for (auto& w : s)
            w = splitmix64(x);
    }
    uint64_t next()
    {
        const uint64_t result = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return result;
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Box-Muller; one normal per call (the pair's second value is dropped so
    // the stream layout does not depend on call parity).
    double normal()
    {
        const double u1 = 1.0 - uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    void unit_vector(double v[3])
    {
        for (;;) {
            v[0] = normal();
            v[1] = normal();
            v[2] = normal();
            const double l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (l > 1e-12) {
                v[0] /= l;
                v[1] /= l;
                v[2] /= l;
                return;
            }
        }
    }
};

struct Writer {
    float* v;
    float* n;
    uint64_t t = 0;
    void tri(const double a[3], const double b[3], const double c[3], const double na[3],
             const double nb[3], const double nc[3])
    {
        if (v) {
            float* o = v + 9 * t;
            for (int k = 0; k < 3; ++k) {
                o[k] = (float)a[k];
                o[3 + k] = (float)b[k];
                o[6 + k] = (float)c[k];
            }
        }
        if (n) {
            float* o = n + 9 * t;
            for (int k = 0; k < 3; ++k) {
                o[k] = (float)na[k];
                o[3 + k] = (float)nb[k];
                o[6 + k] = (float)nc[k];
            }
        }
        ++t;
    }
};

// UV-sphere: 16 slices x 9 stacks, pole fans + 7 quad bands = 256 triangles.
void emit_sphere(Writer& w, double cx, double cy, double cz, double r)
{
    constexpr int S = 16, T = 9;
    auto vert = [&](int i, int j, double p[3], double nn[3]) {
        const double th = 3.141592653589793 * i / T;
        const double ph = 6.283185307179586 * (j % S) / S;
        nn[0] = std::sin(th) * std::cos(ph);
        nn[1] = std::cos(th);
        nn[2] = std::sin(th) * std::sin(ph);
        if (i == 0 || i == T) {
            nn[0] = 0.0;
            nn[2] = 0.0;
        }
        p[0] = cx + r * nn[0];
        p[1] = cy + r * nn[1];
        p[2] = cz + r * nn[2];
    };
    double a[3], b[3], c[3], d[3], na[3], nb[3], nc[3], nd[3];
    for (int j = 0; j < S; ++j) {
        vert(0, 0, a, na);
        vert(1, j + 1, b, nb);
        vert(1, j, c, nc);
        w.tri(a, b, c, na, nb, nc);
    }
    for (int i = 1; i < T - 1; ++i)
        for (int j = 0; j < S; ++j) {
            vert(i, j, a, na);
            vert(i + 1, j, b, nb);
            vert(i + 1, j + 1, c, nc);
            vert(i, j + 1, d, nd);
            w.tri(a, b, c, na, nb, nc);
            w.tri(a, c, d, na, nc, nd);
        }
    for (int j = 0; j < S; ++j) {
        vert(T, 0, a, na);
        vert(T - 1, j, b, nb);
        vert(T - 1, j + 1, c, nc);
        w.tri(a, b, c, na, nb, nc);
    }
}

}  // namespace

extern "C" {

uint32_t grt_gen_spheres(uint64_t seed, uint32_t n_spheres, float rmin, float rmax, float* verts,
                         float* normals)
{
    const uint64_t total = (uint64_t)n_spheres * GRT_SPHERE_TRIS;
    if (total > 0xffffffffull)
        return 0;
    if (!verts && !normals)
        return (uint32_t)total;
    Rng rng(seed);
    Writer w{verts, normals};
    for (uint32_t s = 0; s < n_spheres; ++s) {
        const double cx = rng.uniform(), cy = rng.uniform(), cz = rng.uniform();
        const double r = (double)rmin + ((double)rmax - (double)rmin) * rng.uniform();
        emit_sphere(w, cx, cy, cz, r);
    }
    return (uint32_t)total;
}

uint32_t grt_gen_sphere_clusters(uint64_t seed, uint32_t n_clusters, uint32_t spheres_per_cluster,
                                 float sigma, float rmin, float rmax, float* verts,
                                 float* normals)
{
    const uint64_t total = (uint64_t)n_clusters * spheres_per_cluster * GRT_SPHERE_TRIS;
    if (total > 0xffffffffull)
        return 0;
    if (!verts && !normals)
        return (uint32_t)total;
    Rng rng(seed);
    Writer w{verts, normals};
    for (uint32_t c = 0; c < n_clusters; ++c) {
        const double cx = rng.uniform(), cy = rng.uniform(), cz = rng.uniform();
        for (uint32_t s = 0; s < spheres_per_cluster; ++s) {
            const double x = cx + sigma * rng.normal();
            const double y = cy + sigma * rng.normal();
            const double z = cz + sigma * rng.normal();
            const double r = (double)rmin + ((double)rmax - (double)rmin) * rng.uniform();
            emit_sphere(w, x, y, z, r);
        }
    }
    return (uint32_t)total;
}

uint32_t grt_gen_stadium(uint64_t seed, uint32_t n_clusters, uint32_t tris_per_cluster,
                         uint32_t n_cluster_tris, float* verts, float* normals, float* centers)
{
    if ((uint64_t)n_clusters * tris_per_cluster < n_cluster_tris || n_cluster_tris > 0xfffffff0u)
        return 0;
    const uint32_t total = n_cluster_tris + 2;
    if (!verts && !normals && !centers)
        return total;
    Rng rng(seed);
    Writer w{verts, normals};
    const double mu = std::log(0.002), s = 0.5;
    uint32_t t = 0;
    for (uint32_t c = 0; c < n_clusters; ++c) {
        const double cc[3] = {rng.uniform(), rng.uniform(), rng.uniform()};
        const double sigma = 0.01 + 0.04 * rng.uniform();
        if (centers)
            for (int k = 0; k < 3; ++k)
                centers[3 * c + k] = (float)cc[k];
        for (uint32_t i = 0; i < tris_per_cluster && t < n_cluster_tris; ++i, ++t) {
            double v0[3], u1[3], u2[3];
            for (int k = 0; k < 3; ++k)
                v0[k] = cc[k] + sigma * rng.normal();
            rng.unit_vector(u1);
            rng.unit_vector(u2);
            const double l1 = std::exp(mu + s * rng.normal());
            const double l2 = std::exp(mu + s * rng.normal());
            double v1[3], v2[3];
            for (int k = 0; k < 3; ++k) {
                v1[k] = v0[k] + l1 * u1[k];
                v2[k] = v0[k] + l2 * u2[k];
            }
            const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
            const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
            double nn[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                            e1[0] * e2[1] - e1[1] * e2[0]};
            const double l = std::sqrt(nn[0] * nn[0] + nn[1] * nn[1] + nn[2] * nn[2]);
            if (l > 0.0)
                for (double& x : nn)
                    x /= l;
            w.tri(v0, v1, v2, nn, nn, nn);
        }
    }
    // Ground plane, 100x the unit cluster extent, below every cluster.
    const double y = -0.25, lo = -49.5, hi = 50.5;
    const double up[3] = {0.0, 1.0, 0.0};
    const double p0[3] = {lo, y, lo}, p1[3] = {hi, y, lo}, p2[3] = {hi, y, hi}, p3[3] = {lo, y, hi};
    w.tri(p0, p3, p2, up, up, up);
    w.tri(p0, p2, p1, up, up, up);
    return total;
}

void grt_gen_cluster_transforms(uint64_t seed, uint32_t frame, uint32_t n_clusters,
                                const float* centers, float max_translation, float* xf)
{
    Rng rng(seed ^ (0xa0761d6478bd642full * (uint64_t)(frame + 1)));
    for (uint32_t c = 0; c < n_clusters; ++c) {
        double ax[3];
        rng.unit_vector(ax);
        const double ang = 6.283185307179586 * rng.uniform();
        double tr[3];
        for (double& x : tr)
            x = max_translation * (2.0 * rng.uniform() - 1.0);
        // Rodrigues rotation matrix.
        const double cs = std::cos(ang), sn = std::sin(ang), oc = 1.0 - cs;
        const double x = ax[0], y = ax[1], z = ax[2];
        const double R[3][3] = {{cs + x * x * oc, x * y * oc - z * sn, x * z * oc + y * sn},
                                {y * x * oc + z * sn, cs + y * y * oc, y * z * oc - x * sn},
                                {z * x * oc - y * sn, z * y * oc + x * sn, cs + z * z * oc}};
        const double cc[3] = {centers[3 * c], centers[3 * c + 1], centers[3 * c + 2]};
        float* m = xf + 12ull * c;
        for (int r = 0; r < 3; ++r) {
            const double rc = R[r][0] * cc[0] + R[r][1] * cc[1] + R[r][2] * cc[2];
            m[4 * r] = (float)R[r][0];
            m[4 * r + 1] = (float)R[r][1];
            m[4 * r + 2] = (float)R[r][2];
            m[4 * r + 3] = (float)(cc[r] + tr[r] - rc);
        }
    }
}

double grt_gen_uniform(uint64_t seed, uint64_t index)
{
    uint64_t x = seed ^ (index * 0x9e3779b97f4a7c15ull);
    const uint64_t h = splitmix64(x);
    return (double)(h >> 11) * 0x1.0p-53;
}

}  // extern "C"
