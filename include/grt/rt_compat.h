// rt_compat.h -- the reference's C++ plugin interface for the traced path, over
// the grt_* C ABI.
//
// The reference's renderer, camera and acceleration-structure plugins derive
// from rt::IPlugin (include/rt/IPlugin.h:9-18) and are held by intrusive
// reference-counted pointers (include/rt/Plugins.h:15-25).  The classes below
// keep those names, virtual methods and argument meanings so that code written
// against rt::IRenderer / rt::ICamera / rt::IAccStructBuilder / rt::IAccStruct
// moves over by a namespace change:
//
//   rt::IPlugin            include/rt/IPlugin.h:9-18          -> grt::IPlugin
//   rt::IRenderer          include/rt/IRenderer.h:8-13        -> grt::IRenderer
//   rt::ICamera            include/rt/ICamera.h:10-30         -> grt::ICamera
//   rtp::PinholeCamera     include/rtp/PinholeCamera.h:9-74   -> grt::PinholeCamera
//   rt::IAccStructBuilder  include/rt/IAccStructBuilder.h:12-18 -> grt::IAccStructBuilder
//   rt::IAccStruct         include/rt/IAccStruct.h:10-32      -> grt::IAccStruct
//   rt::Geometry / Instance / Ray / Hit / Aabb / Sample        -> same names in grt::
//
// The HIP implementations are HipGridAccStructBuilder (the device grid build in
// place of rtp::UniformGridAccStructBuilder), HipGridAccStruct (device
// traversal in place of rtp::UniformGridAccStruct) and HipRenderer (the frame
// loop of rtc::CudaRenderer, src/rtcuda/CudaRenderer.cpp:9-209).  Per-ray
// calls (traceNearestGeometry, traceAnyGeometry) are kept for compatibility;
// the batch calls (traceNearest, traceAny over device structure-of-arrays
// buffers) are the throughput path.  Errors throw grt::Error (carrying the
// GRT_* status) where the reference aborted the process.  Instance transforms
// (include/rt/Instance.h:13-20) are not part of this path: an Instance here is a
// geometry and its box.
#ifndef GRT_RT_COMPAT_H
#define GRT_RT_COMPAT_H

#include <atomic>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "grt/grt.h"

namespace grt {

// RT_* codes (include/rt/common.h:19-26).
enum : uint32_t { RT_AXIS_X = 0, RT_AXIS_Y = 1, RT_AXIS_Z = 2 };

// Parameter ids for IPlugin::receiveParameter.
enum : int {
    GRT_PARAM_LAMBDA = 0x4001,       // float*: grid resolution factor (builder)
    GRT_PARAM_BLOCK_SIZE = 0x4002,   // uint32_t*: traversal block size (renderer)
};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// Throws grt::Error for a non-zero grt_* status.
void check(int status);

struct vec3f {
    float x = 0, y = 0, z = 0;
    vec3f() = default;
    vec3f(float x_, float y_, float z_) : x(x_), y(y_), z(z_) {}
    void set(float x_, float y_, float z_) { x = x_; y = y_; z = z_; }
};

// vr::RefCounted / vr::ref_ptr (intrusive counting, include/rt/Plugins.h:15-25).
class RefCounted {
public:
    void ref() const { count_.fetch_add(1, std::memory_order_relaxed); }
    void unref() const
    {
        if (count_.fetch_sub(1, std::memory_order_acq_rel) == 1)
            delete this;
    }
    int refCount() const { return count_.load(std::memory_order_relaxed); }

protected:
    RefCounted() = default;
    RefCounted(const RefCounted&) = delete;
    RefCounted& operator=(const RefCounted&) = delete;
    virtual ~RefCounted() = default;

private:
    mutable std::atomic<int> count_{0};
};

template <class T>
class ref_ptr {
public:
    ref_ptr() = default;
    ref_ptr(T* p) : p_(p) { if (p_) p_->ref(); }
    ref_ptr(const ref_ptr& o) : p_(o.p_) { if (p_) p_->ref(); }
    ref_ptr(ref_ptr&& o) noexcept : p_(o.p_) { o.p_ = nullptr; }
    ~ref_ptr() { if (p_) p_->unref(); }
    ref_ptr& operator=(ref_ptr o) { std::swap(p_, o.p_); return *this; }
    T* get() const { return p_; }
    T* operator->() const { return p_; }
    T& operator*() const { return *p_; }
    explicit operator bool() const { return p_ != nullptr; }

private:
    T* p_ = nullptr;
};

// include/rt/Aabb.h:9-45.
struct Aabb {
    vec3f minv, maxv;
};

// include/rt/Ray.h:9-37.  update() fills invDir and the direction signs.
struct Ray {
    void update();
    vec3f orig;
    vec3f dir;
    float tnear = 0.0f;
    float tfar = 0.0f;
    vec3f invDir;
    int32_t dirSigns[3] = {1, 1, 1};      // -1 negative, 1 otherwise
    int32_t dirSignBits[3] = {0, 0, 0};   // 1 negative, 0 otherwise
};

struct Instance;

// include/rt/Hit.h:9-17.  v0Coord = 1 - (u + v), v1Coord = u, v2Coord = v
// (RayTriIntersection.cpp:96-99); distance = t.
struct Hit {
    uint32_t triangleId = 0;
    float v0Coord = 0, v1Coord = 0, v2Coord = 0;
    float distance = 0;
    const Instance* instance = nullptr;
};

// The ray and hit fields of rt::Sample (include/rt/Sample.h:74-82).
struct Sample {
    Ray ray;
    Hit hit;
};

class IAccStruct;

// include/rt/Geometry.h:10-32: indexed triangles; vertex k of triangle t is
// vertices[triDesc[t].v[k]].  normals is indexed like vertices (may be empty).
struct TriDesc {
    uint32_t v[3];
};

class Geometry : public RefCounted {
public:
    void clear();
    const vec3f& getVertex(uint32_t t, uint32_t k) const { return vertices[triDesc[t].v[k]]; }
    const vec3f& getNormal(uint32_t t, uint32_t k) const { return normals[triDesc[t].v[k]]; }

    ref_ptr<IAccStruct> accStruct;
    std::vector<TriDesc> triDesc;
    std::vector<vec3f> vertices;
    std::vector<vec3f> normals;
};

// include/rt/Instance.h:13-20: a geometry placed by a row-major 3x4
// object-to-world matrix (identity by default).
struct Instance {
    Geometry* geometry = nullptr;
    Aabb bbox;
    float transform[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
};

// include/rt/IPlugin.h:9-18.
class IPlugin : public RefCounted {
public:
    virtual void newFrame();
    virtual void receiveParameter(int paramId, void* paramValue);

protected:
    ~IPlugin() override;
};

// include/rt/IRenderer.h:8-13.
class IRenderer : public IPlugin {
public:
    virtual void render();
};

// include/rt/ICamera.h:10-30.
class ICamera : public IPlugin {
public:
    virtual void translate(float x, float y, float z);
    virtual void rotate(float radians, uint32_t axis);
    virtual void setLookAt(float eyeX, float eyeY, float eyeZ, float centerX, float centerY,
                           float centerZ, float upX, float upY, float upZ);
    virtual void setPerspective(float fovY, float zNear, float zFar);
    virtual void setViewport(uint32_t width, uint32_t height);
    virtual void getViewport(uint32_t& width, uint32_t& height);
    virtual void getRayOrigin(vec3f& origin, float x, float y);
    virtual void getRayDirection(vec3f& dir, float x, float y);

    vec3f continuousTranslation;
};

// include/rtp/PinholeCamera.h:9-74, PinholeCamera.cpp:6-151.  setLookAt takes
// the gluLookAt basis; rotate turns the basis about one of its own axes.
class PinholeCamera : public ICamera {
public:
    PinholeCamera();
    void newFrame() override;
    void translate(float x, float y, float z) override;
    void rotate(float radians, uint32_t axis) override;
    void setLookAt(float eyeX, float eyeY, float eyeZ, float centerX, float centerY,
                   float centerZ, float upX, float upY, float upZ) override;
    void setPerspective(float fovY, float zNear, float zFar) override;
    void setViewport(uint32_t width, uint32_t height) override;
    void getViewport(uint32_t& width, uint32_t& height) override;
    void getRayOrigin(vec3f& origin, float x, float y) override;
    void getRayDirection(vec3f& dir, float x, float y) override;

    const vec3f& getPosition() const { return position_; }
    const vec3f& getBaseDir() const { return baseDir_; }
    const vec3f& getNearU() const { return nearU_; }
    const vec3f& getNearV() const { return nearV_; }
    // The frame's camera in the form the C ABI takes (after newFrame).
    grt_camera toGrtCamera() const;

private:
    bool dirty_ = true;
    vec3f position_;
    float axis_[3][3];   // X, Y, Z
    float fovy_ = 60.0f, zNear_ = 1.0f, zFar_ = 1000.0f;
    uint32_t width_ = 400, height_ = 300;
    vec3f baseDir_, nearU_, nearV_;
    float invWidth_ = 0, invHeight_ = 0;
};

// include/rt/IAccStruct.h:10-32.
class IAccStruct : public IPlugin {
public:
    virtual void clear();
    virtual void setBoundingBox(const Aabb& bbox);
    virtual const Aabb& getBoundingBox() const;
    virtual void traceNearestInstance(const std::vector<Instance>& instances, Sample& sample);
    virtual void traceNearestGeometry(const Instance& instance, Ray& ray, Hit& hit);
    virtual bool traceAnyInstance(const std::vector<Instance>& instances, Sample& sample);
    virtual bool traceAnyGeometry(const Instance& instance, Ray& ray, Hit& hit);

protected:
    Aabb bbox_;
};

// include/rt/IAccStructBuilder.h:12-18.
class IAccStructBuilder : public IPlugin {
public:
    virtual void buildGeometry(Geometry* geometry);
    virtual IAccStruct* buildInstance(const std::vector<Instance> instances);
};

// The device grid of one geometry (rtp::UniformGridAccStruct,
// UniformGridAccStruct.cpp:30-42,113-196,362-378, traced on the GPU).
class HipGridAccStruct : public IAccStruct {
public:
    explicit HipGridAccStruct(grt_context ctx);

    void clear() override;
    // Per-ray calls: one launch each.  traceNearestGeometry keeps an earlier,
    // closer hit.distance as the reference does (IAccStruct.cpp:83-87).
    void traceNearestGeometry(const Instance& instance, Ray& ray, Hit& hit) override;
    bool traceAnyGeometry(const Instance& instance, Ray& ray, Hit& hit) override;
    void traceNearestInstance(const std::vector<Instance>& instances, Sample& sample) override;
    bool traceAnyInstance(const std::vector<Instance>& instances, Sample& sample) override;

    // Batch calls over device buffers (the throughput path).
    void traceNearest(const grt_rays& rays, const grt_hits& hits, uint32_t n,
                      uint32_t image_width = 0);
    void traceAny(const grt_rays& rays, uint8_t* occluded, uint32_t n, uint32_t image_width = 0);

    grt_scene scene() const { return scene_; }
    grt_grid_info gridInfo() const;

protected:
    ~HipGridAccStruct() override;
    // The batch trace every call above goes through: closest hit into hits,
    // or any hit into occluded when hits is NULL.
    virtual void traceBatch(const grt_rays& rays, const grt_hits* hits, uint8_t* occluded,
                            uint32_t n, uint32_t image_width);
    grt_context ctx_;
    grt_scene scene_ = nullptr;

private:
    friend class HipGridAccStructBuilder;
    friend class HipKdTreeAccStructBuilder;
    void traceOne(Ray& ray, bool any, Hit& hit, bool& found);
    void* one_ = nullptr;   // device scratch for the per-ray calls
};

// The kd-tree of one geometry (rtp::KdTreeAccStruct, KdTreeAccStruct.cpp:
// 53-240): the same calls as HipGridAccStruct, traced through a surface-area-
// heuristic kd-tree (grt_kdtree_*) instead of the grid.
class HipKdTreeAccStruct : public HipGridAccStruct {
public:
    explicit HipKdTreeAccStruct(grt_context ctx) : HipGridAccStruct(ctx) {}
    void clear() override;
    grt_kdtree tree() const { return tree_; }
    grt_kdtree_info treeInfo() const;

protected:
    ~HipKdTreeAccStruct() override;
    void traceBatch(const grt_rays& rays, const grt_hits* hits, uint8_t* occluded, uint32_t n,
                    uint32_t image_width) override;

private:
    friend class HipKdTreeAccStructBuilder;
    grt_kdtree tree_ = nullptr;
};

// Device grid build (UniformGridAccStructBuilder::buildGeometry,
// UniformGridAccStructBuilder.cpp:58-149).  receiveParameter(GRT_PARAM_LAMBDA,
// float*) sets lambda (default 5); building a geometry that already holds a
// HipGridAccStruct reuses its device buffers (the per-frame rebuild).
class HipGridAccStructBuilder : public IAccStructBuilder {
public:
    explicit HipGridAccStructBuilder(grt_context ctx);
    void receiveParameter(int paramId, void* paramValue) override;
    void buildGeometry(Geometry* geometry) override;
    float lambda() const { return lambda_; }

private:
    grt_context ctx_;
    float lambda_ = 5.0f;
};

// kd-tree build (KdTreeAccStructBuilder::buildGeometry,
// KdTreeAccStructBuilder.cpp:21-28, with TriangleTreeBuilder's SAH build):
// the geometry's accStruct becomes a HipKdTreeAccStruct.
class HipKdTreeAccStructBuilder : public IAccStructBuilder {
public:
    explicit HipKdTreeAccStructBuilder(grt_context ctx);
    void buildGeometry(Geometry* geometry) override;

private:
    grt_context ctx_;
};

// rtc::CudaRenderer (CudaRenderer.cpp:9-209): newFrame reads the camera,
// render traces one primary ray per pixel and writes the headlight-shaded
// RGBA8 frame (RayTracing_kernel.cu:104-120) into frameBuffer().
class HipRenderer : public IRenderer {
public:
    HipRenderer(grt_context ctx, PinholeCamera* camera, Geometry* geometry);
    void newFrame() override;
    void render() override;
    void receiveParameter(int paramId, void* paramValue) override;

    // Device buffers of the last frame: width*height RGBA8 pixels and hit records.
    const uint32_t* frameBuffer() const { return frame_; }
    const grt_hits& hits() const { return hits_; }
    void readFrameBuffer(std::vector<uint32_t>& rgba) const;
    void readHits(std::vector<int32_t>& id, std::vector<float>& t, std::vector<float>& u,
                  std::vector<float>& v) const;

protected:
    ~HipRenderer() override;

private:
    void reserve(uint32_t n);
    grt_context ctx_;
    ref_ptr<PinholeCamera> camera_;
    ref_ptr<Geometry> geometry_;
    grt_camera cam_{};
    void* buf_ = nullptr;
    uint32_t cap_ = 0;
    grt_rays rays_{};
    grt_hits hits_{};
    uint32_t* frame_ = nullptr;
};

// rt::PrimitiveBuilder (PrimitiveBuilder.cpp:5-219), as rt::Context drove it
// (beginPrimitive / setNormal / addVertex / endPrimitive, Context.cpp:285-343):
// addVertex appends the vertex through the current matrix, and the current
// normal through its inverse transpose, normalised, to the geometry; end()
// assembles the primitive's triangles into triDesc (triangles; strips with
// every other triangle reversed; fans and polygons around the first vertex)
// and keeps those that pass the degenerate filter (grt_triangle_valid), so
// the kept triangles get consecutive ids.
enum : int {
    RT_TRIANGLES = GRT_TRIANGLES,
    RT_TRIANGLE_STRIP = GRT_TRIANGLE_STRIP,
    RT_TRIANGLE_FAN = GRT_TRIANGLE_FAN,
    RT_POLYGON = GRT_POLYGON
};

class PrimitiveBuilder {
public:
    void begin(int primitiveType, Geometry* geometry);
    // Row-major 3x4 affine matrix (Transform::setMatrix); NULL: identity.
    void setMatrix(const float xf[12]);
    void setNormal(float x, float y, float z);
    void addVertex(float x, float y, float z);
    void end();

private:
    void store(uint32_t a, uint32_t b, uint32_t c);
    int type_ = RT_TRIANGLES;
    Geometry* geometry_ = nullptr;
    size_t start_ = 0;
    float xf_[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    float nxf_[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    vec3f normal_{0.0f, 0.0f, 1.0f};
};

// rtgl::GpuRenderer (GpuRenderer.h:18-71, GpuRenderer.cpp:43-297,663-703):
// render() rebuilds the grid when needed (every frame with animation
// enabled, as the reference did after each keyframe upload), traces one
// primary ray per pixel, runs the shadow pass and the Phong shading pass
// (shading_FRAG.glsl), and with reflections enabled (ENABLE_REFLECTION,
// GpuRenderer.cpp:24) adds one mirror bounce with its own shadow pass.  The
// frame is RGBA8 in frameBuffer(), converted from the float colours in rgb().
// Shadow and reflection rays start 1e-3 off the surface along the normal
// (shadow_FRAG.glsl:267-268).
class HipGridRenderer : public IRenderer {
public:
    HipGridRenderer(grt_context ctx, PinholeCamera* camera, Geometry* geometry,
                    HipGridAccStructBuilder* builder);
    void newFrame() override;
    void render() override;
    void receiveParameter(int paramId, void* paramValue) override;

    void setAnimationEnabled(bool enabled) { animation_ = enabled; }
    void setReflectionEnabled(bool enabled) { reflection_ = enabled; }
    // Rebuild the geometry's grid with the builder before the next frame.  The
    // resolution always follows the current box and lambda, so
    // recomputeResolution = false gives the same grid as true.
    void rebuildGrid(bool recomputeResolution);
    // Shading parameters (grt_phong_defaults at construction).
    grt_phong& phong() { return phong_; }
    // Per-triangle diffuse colours, 3 floats per triangle (the reference's
    // material texture); empty = white.
    void setDiffuseColors(const std::vector<float>& rgb);

    const uint32_t* frameBuffer() const { return frame_; }
    const float* rgb() const { return rgb_; }
    void readFrameBuffer(std::vector<uint32_t>& rgba) const;
    void readRgb(std::vector<float>& rgb) const;

protected:
    ~HipGridRenderer() override;

private:
    void reserve(uint32_t n);
    grt_context ctx_;
    ref_ptr<PinholeCamera> camera_;
    ref_ptr<Geometry> geometry_;
    ref_ptr<HipGridAccStructBuilder> builder_;
    grt_camera cam_{};
    grt_phong phong_{};
    bool animation_ = false;
    bool reflection_ = false;
    bool need_rebuild_ = false;
    void* buf_ = nullptr;
    uint32_t cap_ = 0;
    grt_rays rays_{}, shadow_{}, bounce_{};
    grt_hits hits_{}, bounce_hits_{};
    uint8_t* occluded_ = nullptr;
    float* rgb_ = nullptr;
    uint32_t* frame_ = nullptr;
    float* diffuse_ = nullptr;
    size_t diffuse_n_ = 0;
};

}  // namespace grt

#endif  // GRT_RT_COMPAT_H
