// rt_compat.cpp -- the reference's plugin classes for the traced path
// (include/rt/IPlugin.h, IRenderer.h, ICamera.h, IAccStruct.h,
// IAccStructBuilder.h; rtp::PinholeCamera; rtc::CudaRenderer) over grt_*.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <limits>

#include "grt/rt_compat.h"
#include "grt_internal.h"

namespace grt {

void check(int status)
{
    if (status != GRT_OK)
        throw Error(status, grt_last_error());
}

static void hip_ok(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        check(hip_check(e, what));
}

// ---- records ------------------------------------------------------------

void Ray::update()   // Ray.h:26-37
{
    invDir.set(1.0f / dir.x, 1.0f / dir.y, 1.0f / dir.z);
    const float d[3] = {dir.x, dir.y, dir.z};
    for (int k = 0; k < 3; ++k) {
        dirSignBits[k] = std::signbit(d[k]) ? 1 : 0;
        dirSigns[k] = dirSignBits[k] ? -1 : 1;
    }
}

void Geometry::clear()
{
    accStruct = ref_ptr<IAccStruct>();
    triDesc.clear();
    vertices.clear();
    normals.clear();
}

// ---- plugin bases: default implementations do nothing (IPlugin.cpp) -----

void IPlugin::newFrame() {}
void IPlugin::receiveParameter(int, void*) {}
IPlugin::~IPlugin() = default;

void IRenderer::render() {}

void ICamera::translate(float, float, float) {}
void ICamera::rotate(float, uint32_t) {}
void ICamera::setLookAt(float, float, float, float, float, float, float, float, float) {}
void ICamera::setPerspective(float, float, float) {}
void ICamera::setViewport(uint32_t, uint32_t) {}
void ICamera::getViewport(uint32_t& w, uint32_t& h) { w = h = 0; }
void ICamera::getRayOrigin(vec3f&, float, float) {}
void ICamera::getRayDirection(vec3f&, float, float) {}

void IAccStruct::clear() {}
void IAccStruct::setBoundingBox(const Aabb& bbox) { bbox_ = bbox; }
const Aabb& IAccStruct::getBoundingBox() const { return bbox_; }
void IAccStruct::traceNearestInstance(const std::vector<Instance>&, Sample&) {}
void IAccStruct::traceNearestGeometry(const Instance&, Ray&, Hit&) {}
bool IAccStruct::traceAnyInstance(const std::vector<Instance>&, Sample&) { return false; }
bool IAccStruct::traceAnyGeometry(const Instance&, Ray&, Hit&) { return false; }

void IAccStructBuilder::buildGeometry(Geometry*) {}
IAccStruct* IAccStructBuilder::buildInstance(const std::vector<Instance>) { return nullptr; }

// ---- PinholeCamera (PinholeCamera.cpp:6-151) -----------------------------

PinholeCamera::PinholeCamera()
{
    const float id[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    std::memcpy(axis_, id, sizeof(axis_));
}

void PinholeCamera::newFrame()
{
    const vec3f& t = continuousTranslation;
    if (t.x != 0.0f || t.y != 0.0f || t.z != 0.0f)
        translate(t.x, t.y, t.z);
    if (!dirty_)
        return;
    dirty_ = false;
    const grt_camera c = toGrtCamera();
    baseDir_.set(c.base[0], c.base[1], c.base[2]);
    nearU_.set(c.nu[0], c.nu[1], c.nu[2]);
    nearV_.set(c.nv[0], c.nv[1], c.nv[2]);
    invWidth_ = 1.0f / (float)width_;
    invHeight_ = 1.0f / (float)height_;
}

grt_camera PinholeCamera::toGrtCamera() const
{
    grt_camera c;
    const float pos[3] = {position_.x, position_.y, position_.z};
    camera_from_basis(pos, axis_[0], axis_[1], axis_[2], fovy_, zNear_, width_, height_, &c);
    return c;
}

void PinholeCamera::translate(float x, float y, float z)   // PinholeCamera.cpp:57-63
{
    position_.x += (axis_[0][0] * x + axis_[1][0] * y) + axis_[2][0] * z;
    position_.y += (axis_[0][1] * x + axis_[1][1] * y) + axis_[2][1] * z;
    position_.z += (axis_[0][2] * x + axis_[1][2] * y) + axis_[2][2] * z;
    dirty_ = true;
}

void PinholeCamera::rotate(float radians, uint32_t axis)   // PinholeCamera.cpp:65-89
{
    if (axis > RT_AXIS_Z)
        return;
    // Rodrigues rotation of the two other axes about axis_[axis], in double.
    const double k[3] = {axis_[axis][0], axis_[axis][1], axis_[axis][2]};
    const double c = std::cos((double)radians), s = std::sin((double)radians);
    for (uint32_t a = 0; a < 3; ++a) {
        if (a == axis)
            continue;
        const double v[3] = {axis_[a][0], axis_[a][1], axis_[a][2]};
        const double kv = k[0] * v[0] + k[1] * v[1] + k[2] * v[2];
        const double kxv[3] = {k[1] * v[2] - k[2] * v[1], k[2] * v[0] - k[0] * v[2],
                               k[0] * v[1] - k[1] * v[0]};
        for (int i = 0; i < 3; ++i)
            axis_[a][i] = (float)(v[i] * c + kxv[i] * s + k[i] * kv * (1.0 - c));
    }
    dirty_ = true;
}

void PinholeCamera::setLookAt(float eyeX, float eyeY, float eyeZ, float centerX, float centerY,
                              float centerZ, float upX, float upY, float upZ)
{
    position_.set(eyeX, eyeY, eyeZ);
    const float eye[3] = {eyeX, eyeY, eyeZ}, center[3] = {centerX, centerY, centerZ},
                up[3] = {upX, upY, upZ};
    lookat_basis(eye, center, up, axis_[0], axis_[1], axis_[2]);
    dirty_ = true;
}

void PinholeCamera::setPerspective(float fovY, float zNear, float zFar)
{
    fovy_ = fovY;
    zNear_ = zNear;
    zFar_ = zFar;
    dirty_ = true;
}

void PinholeCamera::setViewport(uint32_t width, uint32_t height)
{
    width_ = width;
    height_ = height;
    dirty_ = true;
}

void PinholeCamera::getViewport(uint32_t& width, uint32_t& height)
{
    width = width_;
    height = height_;
}

void PinholeCamera::getRayOrigin(vec3f& origin, float, float) { origin = position_; }

void PinholeCamera::getRayDirection(vec3f& dir, float x, float y)   // PinholeCamera.cpp:131-137
{
    const float us = x * invWidth_, vs = y * invHeight_;
    dir.set((baseDir_.x + nearU_.x * us) + nearV_.x * vs,
            (baseDir_.y + nearU_.y * us) + nearV_.y * vs,
            (baseDir_.z + nearU_.z * us) + nearV_.z * vs);
}

// ---- HipGridAccStruct ------------------------------------------------------

HipGridAccStruct::HipGridAccStruct(grt_context ctx) : ctx_(ctx)
{
    check(grt_scene_create(ctx_, &scene_));
}

HipGridAccStruct::~HipGridAccStruct()
{
    if (one_)
        (void)hipFree(one_);
    if (scene_)
        grt_scene_destroy(scene_);
}

void HipGridAccStruct::clear()
{
    if (scene_)
        grt_scene_destroy(scene_);
    scene_ = nullptr;
    check(grt_scene_create(ctx_, &scene_));
}

grt_grid_info HipGridAccStruct::gridInfo() const
{
    grt_grid_info info;
    check(grt_grid_info_get(scene_, &info));
    return info;
}

void HipGridAccStruct::traceBatch(const grt_rays& rays, const grt_hits* hits, uint8_t* occluded,
                                  uint32_t n, uint32_t image_width)
{
    if (hits)
        check(grt_trace_closest(scene_, &rays, hits, n, image_width, nullptr));
    else
        check(grt_trace_any(scene_, &rays, occluded, n, image_width, nullptr));
}

void HipGridAccStruct::traceNearest(const grt_rays& rays, const grt_hits& hits, uint32_t n,
                                    uint32_t image_width)
{
    traceBatch(rays, &hits, nullptr, n, image_width);
}

void HipGridAccStruct::traceAny(const grt_rays& rays, uint8_t* occluded, uint32_t n,
                                uint32_t image_width)
{
    traceBatch(rays, nullptr, occluded, n, image_width);
}

// One ray through the batch path: 8 ray floats + 4 hit words + 1 flag byte.
void HipGridAccStruct::traceOne(Ray& ray, bool any, Hit& hit, bool& found)
{
    if (!one_)
        hip_ok(hipMalloc(&one_, 64), "per-ray scratch");
    float* f = static_cast<float*>(one_);
    const float in[8] = {ray.orig.x, ray.orig.y, ray.orig.z, ray.dir.x,
                         ray.dir.y,  ray.dir.z,  ray.tnear,  ray.tfar};
    grt_rays r = {f, f + 1, f + 2, f + 3, f + 4, f + 5, f + 6, f + 7};
    hipStream_t s = ctx_->stream;
    hip_ok(hipMemcpyAsync(f, in, sizeof(in), hipMemcpyHostToDevice, s), "per-ray upload");
    if (any) {
        uint8_t* occ = reinterpret_cast<uint8_t*>(f + 12);
        traceBatch(r, nullptr, occ, 1, 0);
        uint8_t o = 0;
        hip_ok(hipMemcpyAsync(&o, occ, 1, hipMemcpyDeviceToHost, s), "per-ray result");
        check(grt_synchronize(ctx_));
        found = o != 0;
        return;
    }
    grt_hits h = {reinterpret_cast<int32_t*>(f + 8), f + 9, f + 10, f + 11};
    traceBatch(r, &h, nullptr, 1, 0);
    float out[4];
    hip_ok(hipMemcpyAsync(out, f + 8, sizeof(out), hipMemcpyDeviceToHost, s), "per-ray result");
    check(grt_synchronize(ctx_));
    int32_t id;
    std::memcpy(&id, &out[0], 4);
    found = id >= 0;
    if (found) {
        hit.triangleId = (uint32_t)id;
        hit.distance = out[1];
        hit.v1Coord = out[2];
        hit.v2Coord = out[3];
        hit.v0Coord = 1.0f - (out[2] + out[3]);
    }
}

void HipGridAccStruct::traceNearestGeometry(const Instance& instance, Ray& ray, Hit& hit)
{
    Hit h;
    bool found = false;
    traceOne(ray, false, h, found);
    if (found && h.distance < hit.distance) {   // IAccStruct.cpp:83-87
        h.instance = &instance;
        hit = h;
    }
}

bool HipGridAccStruct::traceAnyGeometry(const Instance&, Ray& ray, Hit& hit)
{
    bool found = false;
    traceOne(ray, true, hit, found);
    return found;
}

// The ray in an instance's object space (Transform::inverseTransform,
// Transform.h:31-40): the origin through the inverse matrix, the direction
// through its 3x3 part, not renormalised, so t is shared with world space.
static Ray object_ray(const Instance& inst, const Ray& world)
{
    float a[12];
    check(grt_affine_inverse(inst.transform, a));
    const vec3f& o = world.orig;
    const vec3f& d = world.dir;
    Ray r = world;
    r.orig.set(((a[0] * o.x + a[1] * o.y) + a[2] * o.z) + a[3],
               ((a[4] * o.x + a[5] * o.y) + a[6] * o.z) + a[7],
               ((a[8] * o.x + a[9] * o.y) + a[10] * o.z) + a[11]);
    r.dir.set((a[0] * d.x + a[1] * d.y) + a[2] * d.z, (a[4] * d.x + a[5] * d.y) + a[6] * d.z,
              (a[8] * d.x + a[9] * d.y) + a[10] * d.z);
    r.update();
    return r;
}

// IAccStruct.cpp:26-73: rays start at tnear = 2e-4 with tfar = +inf; each
// instance's geometry is traced with the ray in its object space and keeps
// the earlier hit unless it finds a strictly nearer one.  Batches of rays go
// through grt_trace_instances instead.
void HipGridAccStruct::traceNearestInstance(const std::vector<Instance>& instances,
                                            Sample& sample)
{
    sample.ray.tnear = 2e-4f;
    sample.ray.tfar = std::numeric_limits<float>::infinity();
    sample.ray.update();
    sample.hit.instance = nullptr;
    sample.hit.distance = std::numeric_limits<float>::max();
    for (const Instance& inst : instances) {
        if (!inst.geometry || !inst.geometry->accStruct)
            throw Error(GRT_NOT_READY, "traceNearestInstance: instance without a built geometry");
        Ray local = object_ray(inst, sample.ray);
        inst.geometry->accStruct->traceNearestGeometry(inst, local, sample.hit);
    }
}

bool HipGridAccStruct::traceAnyInstance(const std::vector<Instance>& instances, Sample& sample)
{
    for (const Instance& inst : instances) {
        if (!inst.geometry || !inst.geometry->accStruct)
            throw Error(GRT_NOT_READY, "traceAnyInstance: instance without a built geometry");
        Ray local = object_ray(inst, sample.ray);
        if (inst.geometry->accStruct->traceAnyGeometry(inst, local, sample.hit))
            return true;
    }
    return false;
}

// ---- HipGridAccStructBuilder ----------------------------------------------

HipGridAccStructBuilder::HipGridAccStructBuilder(grt_context ctx) : ctx_(ctx) {}

void HipGridAccStructBuilder::receiveParameter(int paramId, void* paramValue)
{
    if (paramId == GRT_PARAM_LAMBDA && paramValue) {
        const float lam = *static_cast<const float*>(paramValue);
        if (!(lam > 0.0f))
            throw Error(GRT_INVALID_PARAM, "lambda must be positive");
        lambda_ = lam;
    }
}

// Expand a geometry's indexed triangles into 9 floats of vertices (and of
// normals when every vertex has one), as CudaRenderer.cpp:86-147 did to float4,
// and upload them into the accelerator's scene.
static void upload_geometry(Geometry* g, grt_scene scene)
{
    if (!g)
        throw Error(GRT_INVALID_PARAM, "buildGeometry: NULL geometry");
    const size_t nt = g->triDesc.size();
    if (nt == 0 || nt > 0xffffffffu / 9)
        throw Error(GRT_INVALID_PARAM, "buildGeometry: triangle count out of range");
    const bool with_normals = g->normals.size() == g->vertices.size() && !g->normals.empty();
    // Expand the indexed triangles (CudaRenderer.cpp:86-147 did the same, to float4).
    std::vector<float> v(nt * 9), n(with_normals ? nt * 9 : 0);
    for (size_t t = 0; t < nt; ++t)
        for (uint32_t k = 0; k < 3; ++k) {
            const uint32_t vi = g->triDesc[t].v[k];
            if (vi >= g->vertices.size())
                throw Error(GRT_INVALID_PARAM, "buildGeometry: vertex index out of range");
            const vec3f& p = g->vertices[vi];
            v[t * 9 + k * 3 + 0] = p.x;
            v[t * 9 + k * 3 + 1] = p.y;
            v[t * 9 + k * 3 + 2] = p.z;
            if (with_normals) {
                const vec3f& q = g->normals[vi];
                n[t * 9 + k * 3 + 0] = q.x;
                n[t * 9 + k * 3 + 1] = q.y;
                n[t * 9 + k * 3 + 2] = q.z;
            }
        }
    check(grt_scene_upload_triangles(scene, v.data(), (uint32_t)nt));
    if (with_normals)
        check(grt_scene_upload_normals(scene, n.data(), (uint32_t)nt));
}

void HipGridAccStructBuilder::buildGeometry(Geometry* g)
{
    if (!g)
        throw Error(GRT_INVALID_PARAM, "buildGeometry: NULL geometry");
    HipGridAccStruct* acc = dynamic_cast<HipGridAccStruct*>(g->accStruct.get());
    if (!acc || dynamic_cast<HipKdTreeAccStruct*>(acc)) {
        acc = new HipGridAccStruct(ctx_);
        g->accStruct = acc;
    }
    upload_geometry(g, acc->scene_);
    check(grt_grid_build(acc->scene_, lambda_));
    const grt_grid_info info = acc->gridInfo();
    Aabb box;
    box.minv.set(info.bmin[0], info.bmin[1], info.bmin[2]);
    box.maxv.set(info.bmax[0], info.bmax[1], info.bmax[2]);
    acc->setBoundingBox(box);
}

// ---- HipKdTreeAccStruct / HipKdTreeAccStructBuilder --------------------------

HipKdTreeAccStruct::~HipKdTreeAccStruct()
{
    if (tree_)
        grt_kdtree_destroy(tree_);
}

void HipKdTreeAccStruct::clear()
{
    if (tree_)
        grt_kdtree_destroy(tree_);
    tree_ = nullptr;
    HipGridAccStruct::clear();
}

grt_kdtree_info HipKdTreeAccStruct::treeInfo() const
{
    if (!tree_)
        throw Error(GRT_NOT_READY, "HipKdTreeAccStruct: no tree built");
    grt_kdtree_info info;
    check(grt_kdtree_info_get(tree_, &info));
    return info;
}

void HipKdTreeAccStruct::traceBatch(const grt_rays& rays, const grt_hits* hits, uint8_t* occluded,
                                    uint32_t n, uint32_t image_width)
{
    if (!tree_)
        throw Error(GRT_NOT_READY, "HipKdTreeAccStruct: no tree built");
    if (hits)
        check(grt_kdtree_trace_closest(tree_, &rays, hits, n, image_width, nullptr));
    else
        check(grt_kdtree_trace_any(tree_, &rays, occluded, n, image_width, nullptr));
}

HipKdTreeAccStructBuilder::HipKdTreeAccStructBuilder(grt_context ctx) : ctx_(ctx) {}

void HipKdTreeAccStructBuilder::buildGeometry(Geometry* g)
{
    if (!g)
        throw Error(GRT_INVALID_PARAM, "buildGeometry: NULL geometry");
    HipKdTreeAccStruct* acc = dynamic_cast<HipKdTreeAccStruct*>(g->accStruct.get());
    if (!acc) {
        acc = new HipKdTreeAccStruct(ctx_);
        g->accStruct = acc;
    }
    upload_geometry(g, acc->scene_);
    if (acc->tree_)
        grt_kdtree_destroy(acc->tree_);
    acc->tree_ = nullptr;
    check(grt_kdtree_build(acc->scene_, &acc->tree_));
    const grt_kdtree_info info = acc->treeInfo();
    Aabb box;
    box.minv.set(info.bmin[0], info.bmin[1], info.bmin[2]);
    box.maxv.set(info.bmax[0], info.bmax[1], info.bmax[2]);
    acc->setBoundingBox(box);
}

// ---- HipRenderer -----------------------------------------------------------

HipRenderer::HipRenderer(grt_context ctx, PinholeCamera* camera, Geometry* geometry)
    : ctx_(ctx), camera_(camera), geometry_(geometry)
{
    if (!camera || !geometry)
        throw Error(GRT_INVALID_PARAM, "HipRenderer: camera and geometry are required");
}

HipRenderer::~HipRenderer()
{
    if (buf_) {
        grt_synchronize(ctx_);
        (void)hipFree(buf_);
    }
}

void HipRenderer::receiveParameter(int paramId, void* paramValue)
{
    if (paramId == GRT_PARAM_BLOCK_SIZE && paramValue) {
        const uint32_t b = *static_cast<const uint32_t*>(paramValue);
        check(grt_set_block_size(ctx_, b, b));
    }
}

void HipRenderer::reserve(uint32_t n)
{
    if (n <= cap_)
        return;
    if (buf_) {
        check(grt_synchronize(ctx_));
        (void)hipFree(buf_);
        buf_ = nullptr;
        cap_ = 0;
    }
    hip_ok(hipMalloc(&buf_, (size_t)n * 4 * 13), "frame buffers");
    cap_ = n;
    float* f = static_cast<float*>(buf_);
    const size_t N = n;
    rays_ = {f, f + N, f + 2 * N, f + 3 * N, f + 4 * N, f + 5 * N, f + 6 * N, f + 7 * N};
    hits_ = {reinterpret_cast<int32_t*>(f + 8 * N), f + 9 * N, f + 10 * N, f + 11 * N};
    frame_ = reinterpret_cast<uint32_t*>(f + 12 * N);
}

void HipRenderer::newFrame()   // CudaRenderer.cpp:9-31
{
    camera_->newFrame();
    cam_ = camera_->toGrtCamera();
    if (cam_.width == 0 || cam_.height == 0)
        throw Error(GRT_INVALID_PARAM, "HipRenderer: empty viewport");
    reserve(cam_.width * cam_.height);
}

void HipRenderer::render()   // CudaRenderer.cpp:33-63 + RayTracing.cu:227-253
{
    HipGridAccStruct* acc = dynamic_cast<HipGridAccStruct*>(geometry_->accStruct.get());
    if (!acc)
        throw Error(GRT_NOT_READY, "HipRenderer: geometry has no HipGridAccStruct");
    if (!frame_)
        newFrame();
    const uint32_t n = cam_.width * cam_.height;
    check(grt_gen_primary(ctx_, &cam_, &rays_));
    acc->traceNearest(rays_, hits_, n, cam_.width);
    check(grt_shade_headlight(acc->scene(), &rays_, &hits_, n, frame_));
}

void HipRenderer::readFrameBuffer(std::vector<uint32_t>& rgba) const
{
    check(grt_synchronize(ctx_));
    const uint32_t n = cam_.width * cam_.height;
    rgba.resize(n);
    hip_ok(hipMemcpy(rgba.data(), frame_, (size_t)n * 4, hipMemcpyDeviceToHost), "read frame");
}

void HipRenderer::readHits(std::vector<int32_t>& id, std::vector<float>& t,
                           std::vector<float>& u, std::vector<float>& v) const
{
    check(grt_synchronize(ctx_));
    const size_t n = (size_t)cam_.width * cam_.height;
    id.resize(n);
    t.resize(n);
    u.resize(n);
    v.resize(n);
    hip_ok(hipMemcpy(id.data(), hits_.id, n * 4, hipMemcpyDeviceToHost), "read hits");
    hip_ok(hipMemcpy(t.data(), hits_.t, n * 4, hipMemcpyDeviceToHost), "read hits");
    hip_ok(hipMemcpy(u.data(), hits_.u, n * 4, hipMemcpyDeviceToHost), "read hits");
    hip_ok(hipMemcpy(v.data(), hits_.v, n * 4, hipMemcpyDeviceToHost), "read hits");
}

// ---------------------------------------------------------------------------
// PrimitiveBuilder
// ---------------------------------------------------------------------------
void PrimitiveBuilder::begin(int primitiveType, Geometry* geometry)   // PrimitiveBuilder.cpp:18-23
{
    if (!geometry)
        throw Error(GRT_INVALID_PARAM, "PrimitiveBuilder::begin: NULL geometry");
    if (primitiveType != RT_TRIANGLES && primitiveType != RT_TRIANGLE_STRIP &&
        primitiveType != RT_TRIANGLE_FAN && primitiveType != RT_POLYGON)
        throw Error(GRT_INVALID_PARAM, "PrimitiveBuilder::begin: unknown primitive type");
    type_ = primitiveType;
    geometry_ = geometry;
    start_ = geometry->vertices.size();
}

void PrimitiveBuilder::setMatrix(const float xf[12])
{
    static const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    const float* a = xf ? xf : id;
    double m[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            m[r][c] = a[4 * r + c];
    double cof[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) {
            const int r1 = (r + 1) % 3, r2 = (r + 2) % 3, c1 = (c + 1) % 3, c2 = (c + 2) % 3;
            cof[r][c] = m[r1][c1] * m[r2][c2] - m[r1][c2] * m[r2][c1];
        }
    const double det = m[0][0] * cof[0][0] + m[0][1] * cof[0][1] + m[0][2] * cof[0][2];
    if (!(det != 0.0) || !std::isfinite(det))
        throw Error(GRT_INVALID_PARAM, "PrimitiveBuilder::setMatrix: singular matrix");
    std::memcpy(xf_, a, sizeof(xf_));
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            nxf_[3 * r + c] = (float)(cof[r][c] / det);   // inverse transpose
}

void PrimitiveBuilder::setNormal(float x, float y, float z)
{
    normal_.set(x, y, z);
}

void PrimitiveBuilder::addVertex(float x, float y, float z)   // PrimitiveBuilder.cpp:71-88
{
    if (!geometry_)
        throw Error(GRT_INVALID_PARAM, "PrimitiveBuilder::addVertex: no primitive begun");
    const float* a = xf_;
    geometry_->vertices.emplace_back(((a[0] * x + a[1] * y) + a[2] * z) + a[3],
                                     ((a[4] * x + a[5] * y) + a[6] * z) + a[7],
                                     ((a[8] * x + a[9] * y) + a[10] * z) + a[11]);
    const float* q = nxf_;
    const vec3f& n = normal_;
    float t[3] = {(q[0] * n.x + q[1] * n.y) + q[2] * n.z, (q[3] * n.x + q[4] * n.y) + q[5] * n.z,
                  (q[6] * n.x + q[7] * n.y) + q[8] * n.z};
    const float len2 = (t[0] * t[0] + t[1] * t[1]) + t[2] * t[2];
    if (len2 > 0.0f) {
        const float inv = 1.0f / std::sqrt(len2);
        t[0] = t[0] * inv;
        t[1] = t[1] * inv;
        t[2] = t[2] * inv;
    }
    geometry_->normals.emplace_back(t[0], t[1], t[2]);
}

void PrimitiveBuilder::store(uint32_t a, uint32_t b, uint32_t c)   // storeValidTriangle
{
    const std::vector<vec3f>& v = geometry_->vertices;
    const float tri[9] = {v[a].x, v[a].y, v[a].z, v[b].x, v[b].y, v[b].z, v[c].x, v[c].y, v[c].z};
    if (!grt_triangle_valid(tri))
        return;
    TriDesc d;
    d.v[0] = a;
    d.v[1] = b;
    d.v[2] = c;
    geometry_->triDesc.push_back(d);
}

void PrimitiveBuilder::end()   // PrimitiveBuilder.cpp:96-219
{
    if (!geometry_)
        throw Error(GRT_INVALID_PARAM, "PrimitiveBuilder::end: no primitive begun");
    const size_t n = geometry_->vertices.size() - start_;
    const uint32_t s = (uint32_t)start_;
    if (type_ == RT_TRIANGLES) {
        for (size_t i = 0; i + 2 < n; i += 3)
            store(s + (uint32_t)i, s + (uint32_t)i + 1, s + (uint32_t)i + 2);
    } else if (type_ == RT_TRIANGLE_STRIP) {
        for (size_t i = 0; i + 2 < n; ++i) {
            const uint32_t k = s + (uint32_t)i;
            if (i & 1)
                store(k + 2, k + 1, k);
            else
                store(k, k + 1, k + 2);
        }
    } else {
        for (size_t i = 1; i + 1 < n; ++i)
            store(s, s + (uint32_t)i, s + (uint32_t)i + 1);
    }
    geometry_ = nullptr;
}

// ---------------------------------------------------------------------------
// HipGridRenderer
// ---------------------------------------------------------------------------
HipGridRenderer::HipGridRenderer(grt_context ctx, PinholeCamera* camera, Geometry* geometry,
                                 HipGridAccStructBuilder* builder)
    : ctx_(ctx), camera_(camera), geometry_(geometry), builder_(builder)
{
    if (!camera || !geometry || !builder)
        throw Error(GRT_INVALID_PARAM, "HipGridRenderer: camera, geometry and builder are required");
    grt_phong_defaults(&phong_);
}

HipGridRenderer::~HipGridRenderer()
{
    if (buf_ || diffuse_)
        grt_synchronize(ctx_);
    if (buf_)
        (void)hipFree(buf_);
    if (diffuse_)
        (void)hipFree(diffuse_);
}

void HipGridRenderer::receiveParameter(int paramId, void* paramValue)
{
    if (paramId == GRT_PARAM_BLOCK_SIZE && paramValue) {
        const uint32_t b = *static_cast<const uint32_t*>(paramValue);
        check(grt_set_block_size(ctx_, b, b));
    }
}

void HipGridRenderer::rebuildGrid(bool recomputeResolution)
{
    (void)recomputeResolution;
    need_rebuild_ = true;
}

void HipGridRenderer::setDiffuseColors(const std::vector<float>& rgb)
{
    if (!rgb.empty() && rgb.size() != geometry_->triDesc.size() * 3)
        throw Error(GRT_INVALID_PARAM, "setDiffuseColors: need 3 floats per triangle");
    if (diffuse_) {
        check(grt_synchronize(ctx_));
        (void)hipFree(diffuse_);
        diffuse_ = nullptr;
        diffuse_n_ = 0;
    }
    if (rgb.empty())
        return;
    hip_ok(hipMalloc(&diffuse_, rgb.size() * 4), "diffuse colours");
    hip_ok(hipMemcpy(diffuse_, rgb.data(), rgb.size() * 4, hipMemcpyHostToDevice),
           "diffuse colours");
    diffuse_n_ = rgb.size() / 3;
}

// Per pixel: primary, shadow and bounce rays (8 floats each), two hit records
// (4 words each), rgb (3 floats), the RGBA8 pixel, and the occlusion byte.
void HipGridRenderer::reserve(uint32_t n)
{
    if (n <= cap_)
        return;
    if (buf_) {
        check(grt_synchronize(ctx_));
        (void)hipFree(buf_);
        buf_ = nullptr;
        cap_ = 0;
    }
    const size_t N = n;
    hip_ok(hipMalloc(&buf_, N * 4 * 36 + N), "frame buffers");
    cap_ = n;
    float* f = static_cast<float*>(buf_);
    auto rays_at = [&](size_t k) {
        float* p = f + k * N;
        return grt_rays{p, p + N, p + 2 * N, p + 3 * N, p + 4 * N, p + 5 * N, p + 6 * N,
                        p + 7 * N};
    };
    rays_ = rays_at(0);
    shadow_ = rays_at(8);
    bounce_ = rays_at(16);
    hits_ = {reinterpret_cast<int32_t*>(f + 24 * N), f + 25 * N, f + 26 * N, f + 27 * N};
    bounce_hits_ = {reinterpret_cast<int32_t*>(f + 28 * N), f + 29 * N, f + 30 * N, f + 31 * N};
    rgb_ = f + 32 * N;
    frame_ = reinterpret_cast<uint32_t*>(f + 35 * N);
    occluded_ = reinterpret_cast<uint8_t*>(f + 36 * N);
}

void HipGridRenderer::newFrame()   // GpuRenderer.cpp:43-92
{
    camera_->newFrame();
    cam_ = camera_->toGrtCamera();
    if (cam_.width == 0 || cam_.height == 0)
        throw Error(GRT_INVALID_PARAM, "HipGridRenderer: empty viewport");
    reserve(cam_.width * cam_.height);
}

void HipGridRenderer::render()   // GpuRenderer.cpp:94-297
{
    if (!geometry_->accStruct || animation_ || need_rebuild_) {
        builder_->buildGeometry(geometry_.get());
        need_rebuild_ = false;
    }
    HipGridAccStruct* acc = dynamic_cast<HipGridAccStruct*>(geometry_->accStruct.get());
    if (!acc)
        throw Error(GRT_NOT_READY, "HipGridRenderer: geometry has no HipGridAccStruct");
    if (diffuse_ && diffuse_n_ != geometry_->triDesc.size())
        throw Error(GRT_INVALID_PARAM, "HipGridRenderer: diffuse colours do not match the geometry");
    if (!frame_)
        newFrame();
    const uint32_t n = cam_.width * cam_.height;
    grt_scene sc = acc->scene();
    check(grt_gen_primary(ctx_, &cam_, &rays_));
    acc->traceNearest(rays_, hits_, n, cam_.width);
    check(grt_gen_shadow(sc, &rays_, &hits_, n, phong_.light, 1e-3f, &shadow_));
    acc->traceAny(shadow_, occluded_, n, cam_.width);
    check(grt_shade_phong(sc, &rays_, &hits_, occluded_, n, diffuse_, &phong_, 0, rgb_,
                          reflection_ ? nullptr : frame_));
    if (!reflection_)
        return;
    check(grt_gen_reflection(sc, &rays_, &hits_, n, 1e-3f, &bounce_));
    acc->traceNearest(bounce_, bounce_hits_, n, cam_.width);
    check(grt_gen_shadow(sc, &bounce_, &bounce_hits_, n, phong_.light, 1e-3f, &shadow_));
    acc->traceAny(shadow_, occluded_, n, cam_.width);
    check(grt_shade_phong(sc, &bounce_, &bounce_hits_, occluded_, n, diffuse_, &phong_, 1, rgb_,
                          frame_));
}

void HipGridRenderer::readFrameBuffer(std::vector<uint32_t>& rgba) const
{
    check(grt_synchronize(ctx_));
    const size_t n = (size_t)cam_.width * cam_.height;
    rgba.resize(n);
    hip_ok(hipMemcpy(rgba.data(), frame_, n * 4, hipMemcpyDeviceToHost), "read frame");
}

void HipGridRenderer::readRgb(std::vector<float>& rgb) const
{
    check(grt_synchronize(ctx_));
    const size_t n = (size_t)cam_.width * cam_.height;
    rgb.resize(3 * n);
    hip_ok(hipMemcpy(rgb.data(), rgb_, n * 12, hipMemcpyDeviceToHost), "read rgb");
}

}  // namespace grt
