// grt_mesh.cpp -- scene ingestion on the host, feeding grt_scene_upload_*:
//
//   primitive assembly   PrimitiveBuilder::end/updateTriangles/
//                        updateTriangleStrip/updatePolygonOrFan
//                        (PrimitiveBuilder.cpp:96-219), with the transformed
//                        vertex and normal of PrimitiveBuilder::addVertex
//                        (PrimitiveBuilder.cpp:71-88)
//   degenerate filter    TriAccel::buildFrom/valid (Triangle.cpp:11-63): a
//                        triangle is kept when the nine projection constants
//                        of its Wald accelerator are finite; kept triangles
//                        get consecutive ids (PrimitiveBuilder.cpp:120-128)
//   OBJ faces            GpuObjLoader.cpp:138-151: a face of k corners becomes
//                        the fan (c0, ci, ci+1); missing normals are (0,0,0);
//                        usemtl selects the MTL material's Kd as the
//                        triangles' diffuse colour
//   procedural scenes    rtdb::loadSphere (rtdb.cpp:142-192), a unit cube, and
//                        the layouts of rtdb.cpp:9-57,215-287 (triangle, logo,
//                        evenly spaced triangles, the mixed scene, 125 copies
//                        of a model), the teapot given as a mesh
//
// Everything here is host code; the result is 9 floats of vertices, 9 of
// normals and 3 of diffuse colour per triangle.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <sstream>
#include <string>
#include <vector>

#include "grt/grt.h"
#include "grt_internal.h"

using grt::set_error;

struct grt_mesh_t {
    std::vector<float> verts;      // 9 per kept triangle
    std::vector<float> normals;    // 9 per kept triangle
    std::vector<float> diffuse;    // 3 per kept triangle
    std::vector<uint32_t> source;  // assembled-triangle index of each kept triangle
    uint32_t assembled = 0;        // triangles assembled so far, kept or dropped
    // immediate mode
    int prim = 0;
    bool open = false;
    std::vector<float> pv, pn;     // transformed vertices / normals of the open primitive
    float normal[3] = {0.0f, 0.0f, 1.0f};
    float color[3] = {1.0f, 1.0f, 1.0f};
    float xf[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    float nxf[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};   // inverse transpose of xf's 3x3 part
};

namespace {

// TriAccel::buildFrom + valid (Triangle.cpp:11-63) in binary32: the edges b =
// v2 - v0, c = v1 - v0, the normal c x b, the dominant axis k (x if |nx|
// beats both others, else y if |ny| > |nz|, else z) and the nine constants of
// the projected test.  A triangle of zero area, or one whose constants
// overflow, is invalid.
bool tri_valid(const float* v)
{
    const float v0[3] = {v[0], v[1], v[2]};
    float b[3], c[3], n[3];
    for (int d = 0; d < 3; ++d) {
        b[d] = v[6 + d] - v0[d];
        c[d] = v[3 + d] - v0[d];
    }
    n[0] = c[1] * b[2] - c[2] * b[1];
    n[1] = c[2] * b[0] - c[0] * b[2];
    n[2] = c[0] * b[1] - c[1] * b[0];
    int k;
    if (std::fabs(n[0]) > std::fabs(n[1]) && std::fabs(n[0]) > std::fabs(n[2]))
        k = 0;
    else if (std::fabs(n[1]) > std::fabs(n[2]))
        k = 1;
    else
        k = 2;
    const int u = (k + 1) % 3, w = (k + 2) % 3;
    float a[9];
    a[0] = n[u] / n[k];
    a[1] = n[w] / n[k];
    const float s = 1.0f / n[k];
    const float nl[3] = {n[0] * s, n[1] * s, n[2] * s};
    a[2] = (v0[0] * nl[0] + v0[1] * nl[1]) + v0[2] * nl[2];
    const float inv = 1.0f / (b[u] * c[w] - b[w] * c[u]);
    a[3] = -b[w] * inv;
    a[4] = b[u] * inv;
    a[5] = (b[w] * v0[u] - b[u] * v0[w]) * inv;
    a[6] = c[w] * inv;
    a[7] = -c[u] * inv;
    a[8] = (c[u] * v0[w] - c[w] * v0[u]) * inv;
    for (float x : a)
        if (!std::isfinite(x))
            return false;
    return true;
}

void push_triangle(grt_mesh_t* m, const float* p0, const float* p1, const float* p2,
                   const float* n0, const float* n1, const float* n2, const float* kd,
                   bool filter)
{
    float v[9], n[9];
    const float* P[3] = {p0, p1, p2};
    const float* N[3] = {n0, n1, n2};
    for (int k = 0; k < 3; ++k)
        for (int d = 0; d < 3; ++d) {
            v[3 * k + d] = P[k][d];
            n[3 * k + d] = N[k][d];
        }
    const uint32_t index = m->assembled++;
    if (filter && !tri_valid(v))
        return;
    m->verts.insert(m->verts.end(), v, v + 9);
    m->normals.insert(m->normals.end(), n, n + 9);
    m->diffuse.insert(m->diffuse.end(), kd, kd + 3);
    m->source.push_back(index);
}

// Primitive assembly of the open primitive (PrimitiveBuilder.cpp:96-219).
void assemble(grt_mesh_t* m)
{
    const size_t nv = m->pv.size() / 3;
    auto V = [&](size_t i) { return &m->pv[3 * i]; };
    auto N = [&](size_t i) { return &m->pn[3 * i]; };
    switch (m->prim) {
    case GRT_TRIANGLES:
        for (size_t i = 0; i + 2 < nv; i += 3)
            push_triangle(m, V(i), V(i + 1), V(i + 2), N(i), N(i + 1), N(i + 2), m->color, true);
        break;
    case GRT_TRIANGLE_STRIP:
        // every other triangle reversed as (i+2, i+1, i)
        for (size_t i = 0; i + 2 < nv; ++i) {
            if (i & 1)
                push_triangle(m, V(i + 2), V(i + 1), V(i), N(i + 2), N(i + 1), N(i), m->color,
                              true);
            else
                push_triangle(m, V(i), V(i + 1), V(i + 2), N(i), N(i + 1), N(i + 2), m->color,
                              true);
        }
        break;
    default:   // fan and polygon: (0, i, i+1)
        for (size_t i = 1; i + 1 < nv; ++i)
            push_triangle(m, V(0), V(i), V(i + 1), N(0), N(i), N(i + 1), m->color, true);
        break;
    }
}

bool mesh_ok(grt_mesh m, const char* who)
{
    if (!m) {
        set_error(GRT_INVALID_PARAM, "%s: NULL mesh", who);
        return false;
    }
    return true;
}

// Inverse transpose of a 3x3 matrix (row-major), in binary64.
bool inverse_transpose(const float* a3x4, float out[9])
{
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            m[r][c] = a3x4[4 * r + c];
    const double c00 = m[1][1] * m[2][2] - m[1][2] * m[2][1];
    const double c01 = m[1][2] * m[2][0] - m[1][0] * m[2][2];
    const double c02 = m[1][0] * m[2][1] - m[1][1] * m[2][0];
    const double det = m[0][0] * c00 + m[0][1] * c01 + m[0][2] * c02;
    if (!(std::fabs(det) > 0.0) || !std::isfinite(det))
        return false;
    // cofactor matrix / det = inverse transpose
    const double cof[3][3] = {
        {c00, c01, c02},
        {m[0][2] * m[2][1] - m[0][1] * m[2][2], m[0][0] * m[2][2] - m[0][2] * m[2][0],
         m[0][1] * m[2][0] - m[0][0] * m[2][1]},
        {m[0][1] * m[1][2] - m[0][2] * m[1][1], m[0][2] * m[1][0] - m[0][0] * m[1][2],
         m[0][0] * m[1][1] - m[0][1] * m[1][0]}};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            out[3 * r + c] = (float)(cof[r][c] / det);
    return true;
}

std::string dir_of(const std::string& path)
{
    const size_t s = path.find_last_of('/');
    return s == std::string::npos ? std::string() : path.substr(0, s + 1);
}

// Kd of every material of an MTL file (newmtl / Kd).
void read_mtl(const std::string& path, std::map<std::string, std::vector<float>>& mats)
{
    std::ifstream in(path);
    std::string line, cur;
    while (std::getline(in, line)) {
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#')
            continue;
        if (key == "newmtl") {
            ls >> cur;
            mats[cur] = {1.0f, 1.0f, 1.0f};
        } else if (key == "Kd" && !cur.empty()) {
            float r = 1, g = 1, b = 1;
            ls >> r >> g >> b;
            mats[cur] = {r, g, b};
        }
    }
}

// One face corner "v", "v/t", "v//n" or "v/t/n"; indices are 1-based, or
// negative relative to the elements read so far; 0 = absent.
bool parse_corner(const std::string& s, long nv, long nn, long& vi, long& ni)
{
    vi = 0;
    ni = 0;
    const char* p = s.c_str();
    char* end = nullptr;
    errno = 0;
    long a = std::strtol(p, &end, 10);
    if (end == p || errno)
        return false;
    vi = a < 0 ? nv + a + 1 : a;
    if (*end == '/') {
        p = end + 1;
        if (*p != '/') {
            std::strtol(p, &end, 10);   // texture coordinate, unused
            p = end;
        }
        if (*p == '/') {
            ++p;
            long c = std::strtol(p, &end, 10);
            if (end != p)
                ni = c < 0 ? nn + c + 1 : c;
        }
    }
    return true;
}

}  // namespace

extern "C" {

int grt_mesh_create(grt_mesh* out)
{
    if (!out)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_create: NULL argument");
    *out = new (std::nothrow) grt_mesh_t();
    return *out ? GRT_OK : set_error(GRT_NO_MEMORY, "grt_mesh_create: out of host memory");
}

int grt_mesh_destroy(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_destroy"))
        return GRT_INVALID_PARAM;
    delete m;
    return GRT_OK;
}

int grt_mesh_clear(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_clear"))
        return GRT_INVALID_PARAM;
    m->verts.clear();
    m->normals.clear();
    m->diffuse.clear();
    m->source.clear();
    m->assembled = 0;
    m->open = false;
    m->pv.clear();
    m->pn.clear();
    return GRT_OK;
}

int grt_mesh_set_transform(grt_mesh m, const float xf[12])
{
    if (!mesh_ok(m, "grt_mesh_set_transform"))
        return GRT_INVALID_PARAM;
    static const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    const float* a = xf ? xf : id;
    float nxf[9];
    if (!inverse_transpose(a, nxf))
        return set_error(GRT_INVALID_PARAM, "grt_mesh_set_transform: singular matrix");
    std::memcpy(m->xf, a, sizeof(m->xf));
    std::memcpy(m->nxf, nxf, sizeof(m->nxf));
    return GRT_OK;
}

int grt_mesh_set_color(grt_mesh m, float r, float g, float b)
{
    if (!mesh_ok(m, "grt_mesh_set_color"))
        return GRT_INVALID_PARAM;
    m->color[0] = r;
    m->color[1] = g;
    m->color[2] = b;
    return GRT_OK;
}

int grt_mesh_begin(grt_mesh m, int primitive)
{
    if (!mesh_ok(m, "grt_mesh_begin"))
        return GRT_INVALID_PARAM;
    if (m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_begin: a primitive is already open");
    if (primitive != GRT_TRIANGLES && primitive != GRT_TRIANGLE_STRIP &&
        primitive != GRT_TRIANGLE_FAN && primitive != GRT_POLYGON)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_begin: unknown primitive 0x%x", primitive);
    m->prim = primitive;
    m->open = true;
    m->pv.clear();
    m->pn.clear();
    return GRT_OK;
}

int grt_mesh_normal(grt_mesh m, float x, float y, float z)
{
    if (!mesh_ok(m, "grt_mesh_normal"))
        return GRT_INVALID_PARAM;
    m->normal[0] = x;
    m->normal[1] = y;
    m->normal[2] = z;
    return GRT_OK;
}

// PrimitiveBuilder::addVertex: the vertex through the transform, the current
// normal through its inverse transpose, normalised.
int grt_mesh_vertex(grt_mesh m, float x, float y, float z)
{
    if (!mesh_ok(m, "grt_mesh_vertex"))
        return GRT_INVALID_PARAM;
    if (!m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_vertex: no open primitive");
    const float* a = m->xf;
    const float p[3] = {((a[0] * x + a[1] * y) + a[2] * z) + a[3],
                        ((a[4] * x + a[5] * y) + a[6] * z) + a[7],
                        ((a[8] * x + a[9] * y) + a[10] * z) + a[11]};
    const float* q = m->nxf;
    const float* n = m->normal;
    float t[3] = {(q[0] * n[0] + q[1] * n[1]) + q[2] * n[2],
                  (q[3] * n[0] + q[4] * n[1]) + q[5] * n[2],
                  (q[6] * n[0] + q[7] * n[1]) + q[8] * n[2]};
    const float len2 = (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
    if (len2 > 0.0f) {
        const float inv = 1.0f / std::sqrt(len2);
        t[0] = t[0] * inv;
        t[1] = t[1] * inv;
        t[2] = t[2] * inv;
    }
    m->pv.insert(m->pv.end(), p, p + 3);
    m->pn.insert(m->pn.end(), t, t + 3);
    return GRT_OK;
}

int grt_mesh_end(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_end"))
        return GRT_INVALID_PARAM;
    if (!m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_end: no open primitive");
    assemble(m);
    m->open = false;
    m->pv.clear();
    m->pn.clear();
    return GRT_OK;
}

namespace {

int load_obj(grt_mesh m, const char* path, uint32_t flags);

}  // namespace

int grt_mesh_load_obj(grt_mesh m, const char* path, uint32_t flags)
{
    if (!mesh_ok(m, "grt_mesh_load_obj") || !path)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_load_obj: NULL argument");
    // All or nothing: a file refused part way leaves the mesh as it was.
    const size_t kept = m->source.size();
    const uint32_t assembled = m->assembled;
    const int rc = load_obj(m, path, flags);
    if (rc != GRT_OK) {
        m->verts.resize(9 * kept);
        m->normals.resize(9 * kept);
        m->diffuse.resize(3 * kept);
        m->source.resize(kept);
        m->assembled = assembled;
    }
    return rc;
}

namespace {

int load_obj(grt_mesh m, const char* path, uint32_t flags)
{
    std::ifstream in(path);
    if (!in)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_load_obj: cannot open %s", path);
    const bool filter = (flags & GRT_MESH_DROP_DEGENERATE) != 0;
    std::vector<float> v, vn;
    std::map<std::string, std::vector<float>> mats;
    const float white[3] = {1.0f, 1.0f, 1.0f};
    const float zero[3] = {0.0f, 0.0f, 0.0f};
    std::vector<float> kd(white, white + 3);
    std::string line;
    uint64_t lineno = 0;
    std::vector<long> cv, cn;
    while (std::getline(in, line)) {
        ++lineno;
        std::istringstream ls(line);
        std::string key;
        if (!(ls >> key) || key[0] == '#')
            continue;
        if (key == "v" || key == "vn") {
            float a = 0, b = 0, c = 0;
            if (!(ls >> a >> b >> c))
                return set_error(GRT_INVALID_PARAM, "grt_mesh_load_obj: %s:%llu: bad %s", path,
                                 (unsigned long long)lineno, key.c_str());
            std::vector<float>& dst = key == "v" ? v : vn;
            dst.push_back(a);
            dst.push_back(b);
            dst.push_back(c);
        } else if (key == "f") {
            cv.clear();
            cn.clear();
            std::string tok;
            const long nv = (long)(v.size() / 3), nn = (long)(vn.size() / 3);
            while (ls >> tok) {
                long vi, ni;
                if (!parse_corner(tok, nv, nn, vi, ni) || vi < 0 || vi > nv || ni < 0 || ni > nn)
                    return set_error(GRT_INVALID_PARAM,
                                     "grt_mesh_load_obj: %s:%llu: bad face corner '%s'", path,
                                     (unsigned long long)lineno, tok.c_str());
                cv.push_back(vi);
                cn.push_back(ni);
            }
            // Absent indices give (0,0,0), as GpuObjLoader::addElem does.
            auto P = [&](size_t i) { return cv[i] > 0 ? &v[3 * (size_t)(cv[i] - 1)] : zero; };
            auto N = [&](size_t i) { return cn[i] > 0 ? &vn[3 * (size_t)(cn[i] - 1)] : zero; };
            for (size_t i = 1; i + 1 < cv.size(); ++i)
                push_triangle(m, P(0), P(i), P(i + 1), N(0), N(i), N(i + 1), kd.data(), filter);
        } else if (key == "mtllib") {
            std::string name;
            ls >> name;
            read_mtl(dir_of(path) + name, mats);
        } else if (key == "usemtl") {
            std::string name;
            ls >> name;
            auto it = mats.find(name);
            kd = it != mats.end() ? it->second : std::vector<float>(white, white + 3);
        }
    }
    return GRT_OK;
}

}  // namespace

// rtdb::loadSphere (rtdb.cpp:142-192): stacks/2 strips of slices+1 vertex
// pairs, latitude angles j * 2pi / (slices/2), longitudes i * pi / slices,
// unit radius, normal = position; the strips' pole triangles have zero area
// and are dropped by the filter.
int grt_mesh_add_sphere(grt_mesh m, uint32_t slices, uint32_t stacks)
{
    if (!mesh_ok(m, "grt_mesh_add_sphere"))
        return GRT_INVALID_PARAM;
    if (slices < 2 || stacks < 2 || m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_add_sphere: bad arguments");
    const double two_pi = 6.283185307179586, pi = 3.141592653589793;
    for (uint32_t j = 0; j < stacks / 2; ++j) {
        const double t1 = j * two_pi / (slices / 2), t2 = (j + 1) * two_pi / (slices / 2);
        grt_mesh_begin(m, GRT_TRIANGLE_STRIP);
        for (uint32_t i = 0; i <= slices; ++i) {
            const double t3 = i * pi / slices;
            for (double t : {t1, t2}) {
                const float e[3] = {(float)(std::cos(t) * std::cos(t3)), (float)std::sin(t),
                                    (float)(std::cos(t) * std::sin(t3))};
                grt_mesh_normal(m, e[0], e[1], e[2]);
                grt_mesh_vertex(m, e[0], e[1], e[2]);
            }
        }
        grt_mesh_end(m);
    }
    return GRT_OK;
}

// The cube [-0.5, 0.5]^3 (the corners of rtdb's Cube.h), two triangles per
// face, counter-clockwise seen from outside, face normals.
int grt_mesh_add_cube(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_add_cube"))
        return GRT_INVALID_PARAM;
    if (m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_add_cube: a primitive is open");
    for (int axis = 0; axis < 3; ++axis)
        for (int side = 0; side < 2; ++side) {
            const float s = side ? 0.5f : -0.5f;
            const int u = (axis + 1) % 3, w = (axis + 2) % 3;
            float n[3] = {0, 0, 0};
            n[axis] = side ? 1.0f : -1.0f;
            // corners in the (u, w) plane, ordered so that (u x w) points along +axis
            float c[4][3];
            const float q[4][2] = {{-0.5f, -0.5f}, {0.5f, -0.5f}, {0.5f, 0.5f}, {-0.5f, 0.5f}};
            for (int k = 0; k < 4; ++k) {
                c[k][axis] = s;
                c[k][u] = side ? q[k][0] : q[3 - k][0];
                c[k][w] = side ? q[k][1] : q[3 - k][1];
            }
            grt_mesh_normal(m, n[0], n[1], n[2]);
            grt_mesh_begin(m, GRT_POLYGON);
            for (int k = 0; k < 4; ++k)
                grt_mesh_vertex(m, c[k][0], c[k][1], c[k][2]);
            grt_mesh_end(m);
        }
    return GRT_OK;
}

// rtdb::loadTriangle (rtdb.cpp:9-18): (0,0,0), (1,0,0), (0,1,0), normal +z.
int grt_mesh_add_triangle(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_add_triangle"))
        return GRT_INVALID_PARAM;
    if (m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_add_triangle: a primitive is open");
    grt_mesh_normal(m, 0.0f, 0.0f, 1.0f);
    grt_mesh_begin(m, GRT_TRIANGLES);
    grt_mesh_vertex(m, 0.0f, 0.0f, 0.0f);
    grt_mesh_vertex(m, 1.0f, 0.0f, 0.0f);
    grt_mesh_vertex(m, 0.0f, 1.0f, 0.0f);
    return grt_mesh_end(m);
}

// rtdb::loadLogo (rtdb.cpp:20-57): three triangles in the z = 0 plane, normal
// +z.  The reference coloured each vertex; colour is per triangle here, so
// the logo takes the current colour.
int grt_mesh_add_logo(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_add_logo"))
        return GRT_INVALID_PARAM;
    if (m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_add_logo: a primitive is open");
    static const float p[9][2] = {{0.0f, 0.0f}, {1.0f, 0.0f}, {0.0f, 1.0f},
                                  {2.0f, 0.0f}, {3.0f, 0.0f}, {3.0f, 1.0f},
                                  {1.5f, 0.0f}, {2.5f, 1.0f}, {0.5f, 1.0f}};
    grt_mesh_normal(m, 0.0f, 0.0f, 1.0f);
    grt_mesh_begin(m, GRT_TRIANGLES);
    for (const auto& q : p)
        grt_mesh_vertex(m, q[0], q[1], 0.0f);
    return grt_mesh_end(m);
}

// Append model's triangles through m's current transform, as if each had been
// added as a GRT_TRIANGLES primitive: vertices as points, normals through the
// inverse transpose and renormalised, the degenerate filter applied again,
// each triangle keeping its diffuse colour.
int grt_mesh_add_mesh(grt_mesh m, grt_mesh model)
{
    if (!mesh_ok(m, "grt_mesh_add_mesh") || !mesh_ok(model, "grt_mesh_add_mesh"))
        return GRT_INVALID_PARAM;
    if (m->open)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_add_mesh: a primitive is open");
    if (m == model)
        return set_error(GRT_INVALID_PARAM, "grt_mesh_add_mesh: a mesh cannot add itself");
    const size_t nt = model->source.size();
    const float saved_color[3] = {m->color[0], m->color[1], m->color[2]};
    const float saved_normal[3] = {m->normal[0], m->normal[1], m->normal[2]};
    for (size_t t = 0; t < nt; ++t) {
        const float* v = &model->verts[9 * t];
        const float* n = &model->normals[9 * t];
        const float* kd = &model->diffuse[3 * t];
        m->color[0] = kd[0];
        m->color[1] = kd[1];
        m->color[2] = kd[2];
        grt_mesh_begin(m, GRT_TRIANGLES);
        for (int k = 0; k < 3; ++k) {
            grt_mesh_normal(m, n[3 * k], n[3 * k + 1], n[3 * k + 2]);
            grt_mesh_vertex(m, v[3 * k], v[3 * k + 1], v[3 * k + 2]);
        }
        grt_mesh_end(m);
    }
    std::memcpy(m->color, saved_color, sizeof(saved_color));
    std::memcpy(m->normal, saved_normal, sizeof(saved_normal));
    return GRT_OK;
}

}  // extern "C"

namespace {

// The reference's MatrixStack::scale/translate post-multiply the current
// matrix (MatrixStack.cpp:26-38): the new operation applies to a vertex first.
void affine_mul(const float a[12], const float b[12], float out[12])
{
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            out[4 * r + c] = (a[4 * r] * b[c] + a[4 * r + 1] * b[4 + c]) + a[4 * r + 2] * b[8 + c];
        out[4 * r + 3] = ((a[4 * r] * b[3] + a[4 * r + 1] * b[7]) + a[4 * r + 2] * b[11]) +
                         a[4 * r + 3];
    }
}

void translated(const float base[12], float x, float y, float z, float out[12])
{
    const float t[12] = {1, 0, 0, x, 0, 1, 0, y, 0, 0, 1, z};
    affine_mul(base, t, out);
}

void scaled(const float base[12], float s, float out[12])
{
    const float t[12] = {s, 0, 0, 0, 0, s, 0, 0, 0, 0, s, 0};
    affine_mul(base, t, out);
}

// Run `add` with m's transform set to xf, then restore the transform.
template <class F>
int with_transform(grt_mesh m, const float xf[12], F add)
{
    float saved[12];
    std::memcpy(saved, m->xf, sizeof(saved));
    int rc = grt_mesh_set_transform(m, xf);
    if (rc == GRT_OK)
        rc = add();
    const int rc2 = grt_mesh_set_transform(m, saved);
    return rc != GRT_OK ? rc : rc2;
}

}  // namespace

extern "C" {

// rtdb::loadEvenSpaced (rtdb.cpp:215-233): the unit triangle at x, y, z in
// {0, 1.5, 3} (27 copies, x fastest).
int grt_mesh_add_even_spaced(grt_mesh m)
{
    if (!mesh_ok(m, "grt_mesh_add_even_spaced"))
        return GRT_INVALID_PARAM;
    const float base[12] = {m->xf[0], m->xf[1], m->xf[2],  m->xf[3],
                            m->xf[4], m->xf[5], m->xf[6],  m->xf[7],
                            m->xf[8], m->xf[9], m->xf[10], m->xf[11]};
    for (int z = 0; z < 3; ++z)
        for (int y = 0; y < 3; ++y)
            for (int x = 0; x < 3; ++x) {
                float xf[12];
                translated(base, 1.5f * (float)x, 1.5f * (float)y, 1.5f * (float)z, xf);
                int rc = with_transform(m, xf, [&]() { return grt_mesh_add_triangle(m); });
                if (rc != GRT_OK)
                    return rc;
            }
    return GRT_OK;
}

// rtdb::loadMixed (rtdb.cpp:235-266): the logo at (1, 2, 2.5); unit cubes at
// (0,5,0), (5,5,0), (5,5,5), (0,5,5); and `model` (the reference's teapot)
// scaled by 5 with offsets (0,0,0), (5,0,0), (5,0,5), (0,0,5) before the
// scale.  The translations accumulate as in the reference.
int grt_mesh_add_mixed(grt_mesh m, grt_mesh model)
{
    if (!mesh_ok(m, "grt_mesh_add_mixed") || !mesh_ok(model, "grt_mesh_add_mixed"))
        return GRT_INVALID_PARAM;
    const float base[12] = {m->xf[0], m->xf[1], m->xf[2],  m->xf[3],
                            m->xf[4], m->xf[5], m->xf[6],  m->xf[7],
                            m->xf[8], m->xf[9], m->xf[10], m->xf[11]};
    float xf[12], next[12];
    translated(base, 1.0f, 2.0f, 2.5f, xf);
    int rc = with_transform(m, xf, [&]() { return grt_mesh_add_logo(m); });
    static const float steps[4][3] = {{0, 5, 0}, {5, 0, 0}, {0, 0, 5}, {-5, 0, 0}};
    std::memcpy(xf, base, sizeof(xf));
    for (int k = 0; k < 4 && rc == GRT_OK; ++k) {
        translated(xf, steps[k][0], steps[k][1], steps[k][2], next);
        std::memcpy(xf, next, sizeof(xf));
        rc = with_transform(m, xf, [&]() { return grt_mesh_add_cube(m); });
    }
    static const float tsteps[4][3] = {{0, 0, 0}, {5, 0, 0}, {0, 0, 5}, {-5, 0, 0}};
    scaled(base, 5.0f, xf);
    for (int k = 0; k < 4 && rc == GRT_OK; ++k) {
        translated(xf, tsteps[k][0], tsteps[k][1], tsteps[k][2], next);
        std::memcpy(xf, next, sizeof(xf));
        rc = with_transform(m, xf, [&]() { return grt_mesh_add_mesh(m, model); });
    }
    return rc;
}

// rtdb::loadTeapots125 (rtdb.cpp:268-287): 125 copies of `model` (the
// reference's teapot, 125 x 2256 = 282k triangles there), each scaled by 5
// after a translation by (2x, 2y, 2z), x, y, z in 0..4 (x fastest).
int grt_mesh_add_teapots125(grt_mesh m, grt_mesh model)
{
    if (!mesh_ok(m, "grt_mesh_add_teapots125") || !mesh_ok(model, "grt_mesh_add_teapots125"))
        return GRT_INVALID_PARAM;
    const float base[12] = {m->xf[0], m->xf[1], m->xf[2],  m->xf[3],
                            m->xf[4], m->xf[5], m->xf[6],  m->xf[7],
                            m->xf[8], m->xf[9], m->xf[10], m->xf[11]};
    float s5[12];
    scaled(base, 5.0f, s5);
    for (int z = 0; z < 5; ++z)
        for (int y = 0; y < 5; ++y)
            for (int x = 0; x < 5; ++x) {
                float xf[12];
                translated(s5, (float)(x * 2), (float)(y * 2), (float)(z * 2), xf);
                int rc = with_transform(m, xf, [&]() { return grt_mesh_add_mesh(m, model); });
                if (rc != GRT_OK)
                    return rc;
            }
    return GRT_OK;
}

int grt_triangle_valid(const float v9[9])
{
    return v9 && tri_valid(v9) ? 1 : 0;
}

int grt_mesh_info(grt_mesh m, uint32_t* n_tris, uint32_t* n_dropped)
{
    if (!mesh_ok(m, "grt_mesh_info"))
        return GRT_INVALID_PARAM;
    const uint32_t kept = (uint32_t)m->source.size();
    if (n_tris)
        *n_tris = kept;
    if (n_dropped)
        *n_dropped = m->assembled - kept;
    return GRT_OK;
}

int grt_mesh_copy(grt_mesh m, float* verts9, float* normals9, float* diffuse3,
                  uint32_t* source_index)
{
    if (!mesh_ok(m, "grt_mesh_copy"))
        return GRT_INVALID_PARAM;
    if (verts9 && !m->verts.empty())
        std::memcpy(verts9, m->verts.data(), m->verts.size() * 4);
    if (normals9 && !m->normals.empty())
        std::memcpy(normals9, m->normals.data(), m->normals.size() * 4);
    if (diffuse3 && !m->diffuse.empty())
        std::memcpy(diffuse3, m->diffuse.data(), m->diffuse.size() * 4);
    if (source_index && !m->source.empty())
        std::memcpy(source_index, m->source.data(), m->source.size() * 4);
    return GRT_OK;
}

int grt_scene_upload_mesh(grt_scene sc, grt_mesh m)
{
    if (!sc || !mesh_ok(m, "grt_scene_upload_mesh"))
        return set_error(GRT_INVALID_PARAM, "grt_scene_upload_mesh: NULL argument");
    const uint32_t n = (uint32_t)m->source.size();
    if (n == 0)
        return set_error(GRT_INVALID_PARAM, "grt_scene_upload_mesh: empty mesh");
    int rc = grt_scene_upload_triangles(sc, m->verts.data(), n);
    if (rc != GRT_OK)
        return rc;
    return grt_scene_upload_normals(sc, m->normals.data(), n);
}

}  // extern "C"
