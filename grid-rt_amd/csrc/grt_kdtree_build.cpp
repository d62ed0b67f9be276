// grt_kdtree_build.cpp -- the kd-tree comparison accelerator: a surface-area-
// heuristic kd-tree over a scene's triangles, built on the host and traced on
// the device (grt_kdtree.hip).
//
// Build (the reference's TriangleTreeBuilder, TriangleTreeBuilder.cpp:13-358,
// and the node layout of KdTreeAccStructBuilder::convertRawTree,
// KdTreeAccStructBuilder.cpp:39-97):
//   - the root box is the tight box of all vertices (Aabb::buildFrom);
//   - at every node each triangle is clipped to the node's box ("perfect
//     splits", AabbIntersection::clipTriangle); a triangle whose clipped part
//     is empty leaves the node;
//   - each axis contributes events at the clipped boxes' bounds -- start and
//     end, or one planar event for a box flat on that axis -- swept in order
//     of (position, end < planar < start);
//   - the cost of a plane is C_t + C_i (p_L n_L + p_R n_R) with C_t = 1,
//     C_i = 1.4, p the children's surface-area ratios, the triangles lying in
//     the plane put on the cheaper side, and a cut-off of empty space favoured
//     by the factor 0.8 + 0.2 x (extent of the non-empty side / node extent);
//   - a node becomes a leaf when the best cost is at least C_i n.
// Differences of this build: the clipping runs in binary64 and its box is
// rounded outwards to binary32, so a triangle that overlaps a node is never
// dropped from it by rounding; the depth is capped at kKdMaxDepth (a leaf is
// made there), which bounds the device traversal stack; subtrees are built on
// several host threads.
//
// Device layout: 8 bytes per node, nodes in breadth-first order with the two
// children of a node adjacent: {split bits, (left << 2) | axis} for inner
// nodes, {first element, 0x80000000 | count} for leaves; one 48-byte triangle
// record per leaf element, contiguous per leaf.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <future>
#include <memory>
#include <new>
#include <queue>
#include <vector>

#include "grt/grt.h"
#include "grt_internal.h"
#include "grt_launch.h"

using namespace grt;

namespace {

constexpr float kCostTraverse = 1.0f;
constexpr float kCostIntersect = 1.4f;

struct Box {
    float lo[3], hi[3];
};

float area(const Box& b)
{
    const float a = b.hi[0] - b.lo[0], c = b.hi[1] - b.lo[1], d = b.hi[2] - b.lo[2];
    return a * c + a * d + c * d;
}

// Sutherland-Hodgman clip of a triangle against a box in binary64; the box of
// the clipped polygon, rounded outwards to binary32 and intersected with the
// box.  False when nothing of the triangle lies in the box.
bool clip_triangle(const float* v9, const Box& box, Box& out)
{
    double poly[2][12][3];
    int n = 3, cur = 0;
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k)
            poly[0][i][k] = v9[3 * i + k];
    for (int plane = 0; plane < 6 && n > 0; ++plane) {
        const int k = plane % 3;
        const bool upper = plane >= 3;
        const double pos = upper ? box.hi[k] : box.lo[k];
        const double dir = upper ? -1.0 : 1.0;
        double(*in)[3] = poly[cur];
        double(*res)[3] = poly[cur ^ 1];
        int m = 0;
        for (int i = 0; i < n; ++i) {
            const double* a = in[i];
            const double* b = in[(i + 1) % n];
            const double da = dir * (a[k] - pos), db = dir * (b[k] - pos);
            if (da >= 0.0) {
                for (int c = 0; c < 3; ++c)
                    res[m][c] = a[c];
                ++m;
            }
            if ((da >= 0.0) != (db >= 0.0)) {
                const double t = da / (da - db);
                for (int c = 0; c < 3; ++c)
                    res[m][c] = a[c] + (b[c] - a[c]) * t;
                res[m][k] = pos;
                ++m;
            }
        }
        n = m;
        cur ^= 1;
    }
    if (n == 0)
        return false;
    for (int k = 0; k < 3; ++k) {
        double lo = poly[cur][0][k], hi = lo;
        for (int i = 1; i < n; ++i) {
            lo = std::min(lo, poly[cur][i][k]);
            hi = std::max(hi, poly[cur][i][k]);
        }
        float fl = (float)lo, fh = (float)hi;
        if ((double)fl > lo)
            fl = std::nextafter(fl, -INFINITY);
        if ((double)fh < hi)
            fh = std::nextafter(fh, INFINITY);
        out.lo[k] = std::max(fl, box.lo[k]);
        out.hi[k] = std::min(fh, box.hi[k]);
        if (!(out.lo[k] <= out.hi[k]))
            return false;
    }
    return true;
}

enum Side : uint8_t { kLeft, kRight };

struct Event {
    float pos;
    uint8_t type;   // 0 end, 1 planar, 2 start
    bool operator<(const Event& o) const
    {
        return pos < o.pos || (pos == o.pos && type < o.type);
    }
};

struct Split {
    float cost;
    Side side;
    int axis;
    float pos;
};

// Cost of the plane (axis k, position p) for nL triangles left of it, nP in
// it and nR right of it; the planar ones go to the cheaper side.
void sah(const Box& box, int k, float p, uint32_t nL, uint32_t nP, uint32_t nR, float& cost,
         Side& side)
{
    Box l = box, r = box;
    l.hi[k] = std::min(p, box.hi[k]);
    r.lo[k] = std::max(p, box.lo[k]);
    const float inv_area = 1.0f / area(box);
    const float pl = area(l) * inv_area, pr = area(r) * inv_area;
    const float inv_ext = 1.0f / (box.hi[k] - box.lo[k]);
    const float lext = (l.hi[k] - l.lo[k]) * inv_ext, rext = (r.hi[k] - r.lo[k]) * inv_ext;
    auto one = [&](uint32_t a, uint32_t b) {
        float c = kCostTraverse + kCostIntersect * (pl * (float)a + pr * (float)b);
        if (a == 0)
            c *= 0.8f + 0.2f * rext;
        else if (b == 0)
            c *= 0.8f + 0.2f * lext;
        return c;
    };
    const float cl = one(nL + nP, nR), cr = one(nL, nP + nR);
    if (cl <= cr) {
        cost = cl;
        side = kLeft;
    } else {
        cost = cr;
        side = kRight;
    }
}

struct RawNode {
    int axis = -1;
    float split = 0.0f;
    std::unique_ptr<RawNode> child[2];
    std::vector<uint32_t> tris;   // leaf
};

struct Builder {
    const float* verts = nullptr;
    uint32_t max_depth = kKdMaxDepth;

    std::unique_ptr<RawNode> build(const std::vector<uint32_t>& tris, const Box& box,
                                   uint32_t depth)
    {
        auto node = std::make_unique<RawNode>();
        const size_t n = tris.size();
        // Clip every triangle to the node's box.
        std::vector<Box> cb(n);
        std::vector<uint8_t> ok(n);
        uint32_t valid = 0;
        for (size_t i = 0; i < n; ++i) {
            ok[i] = clip_triangle(verts + 9ull * tris[i], box, cb[i]) ? 1 : 0;
            valid += ok[i];
        }
        Split best{INFINITY, kLeft, 0, 0.0f};
        if (depth < max_depth && valid > 0) {
            std::vector<Event> ev;
            ev.reserve(2 * valid);
            for (int k = 0; k < 3; ++k) {
                ev.clear();
                for (size_t i = 0; i < n; ++i) {
                    if (!ok[i])
                        continue;
                    if (cb[i].lo[k] == cb[i].hi[k]) {
                        ev.push_back(Event{cb[i].lo[k], 1});
                    } else {
                        ev.push_back(Event{cb[i].lo[k], 2});
                        ev.push_back(Event{cb[i].hi[k], 0});
                    }
                }
                std::sort(ev.begin(), ev.end());
                uint32_t nl = 0, nr = valid;
                for (size_t i = 0; i < ev.size();) {
                    const float p = ev[i].pos;
                    uint32_t ne = 0, np = 0, ns = 0;
                    while (i < ev.size() && ev[i].pos == p && ev[i].type == 0)
                        ++ne, ++i;
                    while (i < ev.size() && ev[i].pos == p && ev[i].type == 1)
                        ++np, ++i;
                    while (i < ev.size() && ev[i].pos == p && ev[i].type == 2)
                        ++ns, ++i;
                    nr -= np + ne;
                    float c;
                    Side s;
                    sah(box, k, p, nl, np, nr, c, s);
                    if (c < best.cost)
                        best = Split{c, s, k, p};
                    nl += ns + np;
                }
            }
        }
        if (!(best.cost < kCostIntersect * (float)valid)) {
            node->tris.reserve(valid);
            for (size_t i = 0; i < n; ++i)
                if (ok[i])
                    node->tris.push_back(tris[i]);
            return node;
        }
        // Partition: planar triangles on the plane go to the chosen side.
        std::vector<uint32_t> lt, rt;
        const int k = best.axis;
        const float p = best.pos;
        for (size_t i = 0; i < n; ++i) {
            if (!ok[i])
                continue;
            const float lo = cb[i].lo[k], hi = cb[i].hi[k];
            bool left, right;
            if (lo == hi) {
                left = lo < p || (lo == p && best.side == kLeft);
                right = lo > p || (lo == p && best.side == kRight);
            } else {
                left = lo < p;
                right = hi > p;
            }
            if (left)
                lt.push_back(tris[i]);
            if (right)
                rt.push_back(tris[i]);
        }
        std::vector<Box>().swap(cb);
        std::vector<uint8_t>().swap(ok);
        Box lb = box, rb = box;
        lb.hi[k] = std::min(p, box.hi[k]);
        rb.lo[k] = std::max(p, box.lo[k]);
        node->axis = k;
        node->split = p;
        if (depth < 6 && lt.size() + rt.size() > 20000) {
            auto fut = std::async(std::launch::async,
                                  [&]() { return build(lt, lb, depth + 1); });
            node->child[1] = build(rt, rb, depth + 1);
            node->child[0] = fut.get();
        } else {
            node->child[0] = build(lt, lb, depth + 1);
            node->child[1] = build(rt, rb, depth + 1);
        }
        return node;
    }
};

}  // namespace

struct grt_kdtree_t {
    grt_context_t* ctx = nullptr;
    grt_kdtree_info info{};
    std::vector<uint32_t> nodes_host;      // 2 words per node
    std::vector<uint32_t> elements_host;   // triangle id per leaf element
    DevMem nodes, records, elements;   // device copies (none for a host-only tree)
};

namespace {

// Build the tree of n_tris 9-float triangles into kd's host arrays and info.
int build_host_tree(const float* v, uint32_t nt, grt_kdtree_t* kd)
{
    Box root;
    for (int k = 0; k < 3; ++k) {
        root.lo[k] = INFINITY;
        root.hi[k] = -INFINITY;
    }
    for (size_t i = 0; i < (size_t)nt * 3; ++i)
        for (int k = 0; k < 3; ++k) {
            const float x = v[3 * i + k];
            if (!std::isfinite(x))
                return set_error(GRT_INVALID_PARAM, "kd-tree build: vertex %zu is not finite", i);
            root.lo[k] = std::min(root.lo[k], x);
            root.hi[k] = std::max(root.hi[k], x);
        }
    std::unique_ptr<RawNode> tree;
    try {
        std::vector<uint32_t> all(nt);
        for (uint32_t i = 0; i < nt; ++i)
            all[i] = i;
        Builder b;
        b.verts = v;
        tree = b.build(all, root, 0);
        // Breadth-first flattening, children adjacent.
        grt_kdtree_info& info = kd->info;
        for (int k = 0; k < 3; ++k) {
            info.bmin[k] = root.lo[k];
            info.bmax[k] = root.hi[k];
        }
        std::queue<std::pair<const RawNode*, uint32_t>> q;   // node, depth
        q.push({tree.get(), 0});
        while (!q.empty()) {
            const RawNode* nd = q.front().first;
            const uint32_t depth = q.front().second;
            q.pop();
            info.depth = std::max(info.depth, depth);
            if (nd->axis < 0) {
                const uint32_t first = (uint32_t)kd->elements_host.size();
                kd->elements_host.insert(kd->elements_host.end(), nd->tris.begin(),
                                         nd->tris.end());
                kd->nodes_host.push_back(first);
                kd->nodes_host.push_back(0x80000000u | (uint32_t)nd->tris.size());
                ++info.leaves;
                if (nd->tris.empty())
                    ++info.empty_leaves;
            } else {
                // The children go after every node already queued.
                const uint32_t self = (uint32_t)(kd->nodes_host.size() / 2);
                const uint32_t left = self + 1 + (uint32_t)q.size();
                uint32_t word0;
                std::memcpy(&word0, &nd->split, 4);
                kd->nodes_host.push_back(word0);
                kd->nodes_host.push_back((left << 2) | (uint32_t)nd->axis);
                q.push({nd->child[0].get(), depth + 1});
                q.push({nd->child[1].get(), depth + 1});
            }
        }
        info.nodes = (uint32_t)(kd->nodes_host.size() / 2);
        info.elements = (uint32_t)kd->elements_host.size();
    } catch (const std::bad_alloc&) {
        return set_error(GRT_NO_MEMORY, "kd-tree build: out of host memory");
    }
    return GRT_OK;
}

}  // namespace

extern "C" {

int grt_kdtree_build_host(const float* verts_host, uint32_t n_tris, grt_kdtree* out)
{
    if (!out || (!verts_host && n_tris))
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_build_host: NULL argument");
    *out = nullptr;
    if (n_tris == 0 || n_tris > (1u << 28))
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_build_host: n_tris=%u out of range",
                         n_tris);
    const auto t0 = std::chrono::steady_clock::now();
    grt_kdtree_t* kd = new (std::nothrow) grt_kdtree_t();
    if (!kd)
        return set_error(GRT_NO_MEMORY, "grt_kdtree_build_host: out of host memory");
    int rc = build_host_tree(verts_host, n_tris, kd);
    if (rc != GRT_OK) {
        delete kd;
        return rc;
    }
    kd->info.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = kd;
    return GRT_OK;
}

int grt_kdtree_build(grt_scene sc, grt_kdtree* out)
{
    if (!sc || !out)
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_build: NULL argument");
    *out = nullptr;
    if (sc->n_tris == 0)
        return set_error(GRT_NOT_READY, "grt_kdtree_build: empty scene");
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t nt = sc->n_tris;
    std::vector<float> v;
    try {
        v.resize((size_t)nt * 9);
    } catch (const std::bad_alloc&) {
        return set_error(GRT_NO_MEMORY, "grt_kdtree_build: out of host memory");
    }
    int rc = grt_scene_download_geometry(sc, v.data(), nullptr);
    if (rc != GRT_OK)
        return rc;
    grt_kdtree_t* kd = new (std::nothrow) grt_kdtree_t();
    if (!kd)
        return set_error(GRT_NO_MEMORY, "grt_kdtree_build: out of host memory");
    kd->ctx = sc->ctx;
    if ((rc = build_host_tree(v.data(), nt, kd)) != GRT_OK) {
        delete kd;
        return rc;
    }
    const grt_kdtree_info& info = kd->info;
    // Device copies: nodes, element ids, and the leaves' triangle records
    // gathered from the scene's packed records.
    hipStream_t s = sc->ctx->stream;
    const size_t ne = std::max<size_t>(1, info.elements);
    if ((rc = kd->nodes.reserve(kd->ctx, kd->nodes_host.size() * 4, false)) != GRT_OK ||
        (rc = kd->elements.reserve(kd->ctx, ne * 4, false)) != GRT_OK ||
        (rc = kd->records.reserve(kd->ctx, ne * 48, false)) != GRT_OK) {
        grt_kdtree_destroy(kd);
        return rc;
    }
    hipError_t e = hipMemcpyAsync(kd->nodes.ptr, kd->nodes_host.data(), kd->nodes_host.size() * 4,
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess && info.elements)
        e = hipMemcpyAsync(kd->elements.ptr, kd->elements_host.data(), (size_t)info.elements * 4,
                           hipMemcpyHostToDevice, s);
    if (e == hipSuccess && info.elements)
        e = launch_kd_gather(kd->elements.as<uint32_t>(), info.elements, sc->tri.as<float4>(),
                             kd->records.as<float4>(), s);
    if (e == hipSuccess)
        e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
        rc = hip_check(e, "grt_kdtree_build: upload");
        grt_kdtree_destroy(kd);
        return rc;
    }
    kd->info.build_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    *out = kd;
    return GRT_OK;
}

int grt_kdtree_destroy(grt_kdtree kd)
{
    if (!kd)
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_destroy: NULL tree");
    if (kd->ctx) {
        (void)hipStreamSynchronize(kd->ctx->stream);
        kd->nodes.release(kd->ctx);
        kd->elements.release(kd->ctx);
        kd->records.release(kd->ctx);
    }
    delete kd;
    return GRT_OK;
}

int grt_kdtree_info_get(grt_kdtree kd, grt_kdtree_info* out)
{
    if (!kd || !out)
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_info_get: NULL argument");
    *out = kd->info;
    return GRT_OK;
}

int grt_kdtree_download(grt_kdtree kd, uint32_t* nodes_host, uint32_t* elements_host)
{
    if (!kd)
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_download: NULL tree");
    if (nodes_host)
        std::memcpy(nodes_host, kd->nodes_host.data(), kd->nodes_host.size() * 4);
    if (elements_host && !kd->elements_host.empty())
        std::memcpy(elements_host, kd->elements_host.data(), kd->elements_host.size() * 4);
    return GRT_OK;
}

static int kd_trace(grt_kdtree kd, const grt_rays* rays, const grt_hits* hits, uint8_t* occluded,
                    uint32_t n, uint32_t image_width, uint32_t* counters, const char* what)
{
    if (!kd)
        return set_error(GRT_INVALID_PARAM, "%s: NULL tree", what);
    if (!kd->ctx)
        return set_error(GRT_NOT_READY, "%s: the tree was built on the host only", what);
    if (n == 0)
        return GRT_OK;
    const bool rays_ok = rays && rays->ox && rays->oy && rays->oz && rays->dx && rays->dy &&
                         rays->dz && rays->tmin && rays->tmax;
    const bool out_ok = hits ? (hits->id && hits->t && hits->u && hits->v) : occluded != nullptr;
    if (!rays_ok || !out_ok)
        return set_error(GRT_INVALID_PARAM, "%s: bad argument", what);
    if (slot_count(n, image_width) > 0xffffffffull)
        return set_error(GRT_INVALID_PARAM, "%s: too many rays", what);
    KdArgs a;
    for (int k = 0; k < 3; ++k) {
        a.bmin[k] = kd->info.bmin[k];
        a.bmax[k] = kd->info.bmax[k];
    }
    a.nodes = kd->nodes.as<uint2>();
    a.records = kd->records.as<float4>();
    a.n = n;
    a.image_width = image_width;
    a.counters = counters;
    const RaySoA r{rays->ox, rays->oy, rays->oz, rays->dx, rays->dy, rays->dz, rays->tmin,
                   rays->tmax};
    const HitSoA h = hits ? HitSoA{hits->id, hits->t, hits->u, hits->v}
                          : HitSoA{nullptr, nullptr, nullptr, nullptr};
    return hip_check(launch_kd_trace(a, r, h, occluded, kd->ctx->block_closest, kd->ctx->stream),
                     what);
}

int grt_kdtree_trace_closest(grt_kdtree kd, const grt_rays* rays, const grt_hits* hits,
                             uint32_t n, uint32_t image_width, uint32_t* counters_dev)
{
    if (!hits && n)
        return set_error(GRT_INVALID_PARAM, "grt_kdtree_trace_closest: hits is NULL");
    return kd_trace(kd, rays, hits, nullptr, n, image_width, counters_dev,
                    "grt_kdtree_trace_closest");
}

int grt_kdtree_trace_any(grt_kdtree kd, const grt_rays* rays, uint8_t* occluded_dev, uint32_t n,
                         uint32_t image_width, uint32_t* counters_dev)
{
    return kd_trace(kd, rays, nullptr, occluded_dev, n, image_width, counters_dev,
                    "grt_kdtree_trace_any");
}

}  // extern "C"
