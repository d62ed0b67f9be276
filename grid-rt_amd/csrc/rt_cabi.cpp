// rt_cabi.cpp -- the reference tracer's flat C ABI (src/rtcuda/RayTracing.h:6-44)
// over the grt_* API.  The reference kept its inputs in file-static device
// buffers bound to textures (RayTracing.cu:16-28,106-170); here the transfers
// are kept on the host until grtRayTrace, which converts them once into a
// grt_scene (three float4 per triangle -> nine floats; {start, count} cells and
// 3*id references -> offsets[m+1] and plain ids) and then renders from device
// memory for every later frame until something is transferred again.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "grt/grt.h"
#include "grt/rt_cabi.h"
#include "grt_internal.h"

namespace {

struct CabiState {
    std::mutex mu;
    grt_context ctx = nullptr;
    grt_scene scene = nullptr;
    uint32_t* frame = nullptr;     // registered RGBA8 buffer (caller's)
    grt_camera cam{};
    bool cam_set = false;
    // Host copies of the transfers, consumed by the next grtRayTrace.
    std::vector<float> verts, normals;            // 9 floats per triangle
    std::vector<int32_t> cell_ptrs, cell_ids;
    float bmin[3] = {0, 0, 0}, bmax[3] = {0, 0, 0}, cell[3] = {0, 0, 0};
    int32_t n[3] = {0, 0, 0};
    bool geom_dirty = false, grid_dirty = false;
    bool have_verts = false, have_normals = false, have_grid_params = false;
    // Per-frame device buffers: 8 ray arrays + 4 hit arrays, cap rays each.
    void* buf = nullptr;
    uint32_t cap = 0;
    grt_rays rays{};
    grt_hits hits{};
    uint32_t last_n = 0;
};

CabiState& state()
{
    static CabiState s;
    return s;
}

int need_ctx(CabiState& s)
{
    if (!s.ctx)
        return grt::set_error(GRT_NOT_READY, "rtInitContext has not been called");
    return GRT_OK;
}

// 12 floats per triangle (three float4) -> 9.
int unpack_float4x3(const float* src, uint32_t size, std::vector<float>& dst, const char* what)
{
    if (!src && size)
        return grt::set_error(GRT_INVALID_PARAM, "%s: NULL data", what);
    if (size % 12 != 0)
        return grt::set_error(GRT_INVALID_PARAM, "%s: size %u is not 12 floats per triangle",
                              what, size);
    const uint32_t nt = size / 12;
    dst.resize((size_t)nt * 9);
    for (uint32_t t = 0; t < nt; ++t)
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c)
                dst[(size_t)t * 9 + k * 3 + c] = src[(size_t)t * 12 + k * 4 + c];
    return GRT_OK;
}

int release_buffers(CabiState& s)
{
    if (s.buf)
        (void)hipFree(s.buf);
    s.buf = nullptr;
    s.cap = 0;
    return GRT_OK;
}

int reserve_rays(CabiState& s, uint32_t n)
{
    if (n <= s.cap)
        return GRT_OK;
    release_buffers(s);
    const size_t bytes = (size_t)n * 4 * 12;
    if (hipMalloc(&s.buf, bytes) != hipSuccess) {
        s.buf = nullptr;
        return grt::set_error(GRT_NO_MEMORY, "grtRayTrace: cannot allocate %zu bytes", bytes);
    }
    s.cap = n;
    float* f = static_cast<float*>(s.buf);
    const size_t N = n;
    s.rays = {f, f + N, f + 2 * N, f + 3 * N, f + 4 * N, f + 5 * N, f + 6 * N, f + 7 * N};
    s.hits = {reinterpret_cast<int32_t*>(f + 8 * N), f + 9 * N, f + 10 * N, f + 11 * N};
    return GRT_OK;
}

// Flattened {start, count} cells + 3*id references -> offsets / plain ids, in
// cell order (the cells need not be stored contiguously in cellTriangleIds).
int convert_grid(const CabiState& s, uint32_t n_tris, std::vector<uint32_t>& offsets,
                 std::vector<uint32_t>& refs)
{
    const uint64_t m = (uint64_t)s.n[0] * (uint64_t)s.n[1] * (uint64_t)s.n[2];
    if (s.cell_ptrs.size() != 2 * m)
        return grt::set_error(GRT_INVALID_PARAM,
                              "grtRayTrace: %zu cell pointers for %llu cells (want 2 per cell)",
                              s.cell_ptrs.size(), (unsigned long long)m);
    offsets.resize(m + 1);
    refs.clear();
    refs.reserve(s.cell_ids.size());
    const uint64_t nids = s.cell_ids.size();
    for (uint64_t c = 0; c < m; ++c) {
        offsets[c] = (uint32_t)refs.size();
        const int64_t start = s.cell_ptrs[2 * c], count = s.cell_ptrs[2 * c + 1];
        if (start < 0 || count < 0 || (uint64_t)(start + count) > nids)
            return grt::set_error(GRT_INVALID_PARAM,
                                  "grtRayTrace: cell %llu = {%lld, %lld} outside %llu ids",
                                  (unsigned long long)c, (long long)start, (long long)count,
                                  (unsigned long long)nids);
        for (int64_t k = 0; k < count; ++k) {
            const int32_t id3 = s.cell_ids[start + k];
            if (id3 < 0 || id3 % 3 != 0 || (uint32_t)(id3 / 3) >= n_tris)
                return grt::set_error(GRT_INVALID_PARAM,
                                      "grtRayTrace: triangle id %d is not 3*t for t < %u", id3,
                                      n_tris);
            refs.push_back((uint32_t)(id3 / 3));
        }
    }
    offsets[m] = (uint32_t)refs.size();
    return GRT_OK;
}

int sync_scene(CabiState& s)
{
    int rc;
    if (!s.have_verts || !s.have_grid_params || s.cell_ptrs.empty())
        return grt::set_error(GRT_NOT_READY,
                              "grtRayTrace: vertices, cell pointers and grid parameters are "
                              "required");
    const uint32_t nt = (uint32_t)(s.verts.size() / 9);
    if (s.have_normals && s.normals.size() != s.verts.size())
        return grt::set_error(GRT_INVALID_PARAM, "grtRayTrace: %zu normals for %u triangles",
                              s.normals.size() / 9, nt);
    if (!s.scene && (rc = grt_scene_create(s.ctx, &s.scene)) != GRT_OK)
        return rc;
    if (s.geom_dirty) {
        if ((rc = grt_scene_upload_triangles(s.scene, s.verts.data(), nt)) != GRT_OK)
            return rc;
        if (s.have_normals &&
            (rc = grt_scene_upload_normals(s.scene, s.normals.data(), nt)) != GRT_OK)
            return rc;
        s.geom_dirty = false;
        s.grid_dirty = true;   // the 48-byte records are gathered at grid upload
    }
    if (s.grid_dirty) {
        std::vector<uint32_t> offsets, refs;
        if ((rc = convert_grid(s, nt, offsets, refs)) != GRT_OK)
            return rc;
        grt_grid_info info{};
        for (int d = 0; d < 3; ++d) {
            info.bmin[d] = s.bmin[d];
            info.bmax[d] = s.bmax[d];
            info.cell[d] = s.cell[d];
            info.inv_cell[d] = 1.0f / s.cell[d];   // RayTracing.cu:193-195
            info.n[d] = s.n[d];
        }
        if ((rc = grt_grid_upload(s.scene, &info, offsets.data(), refs.data())) != GRT_OK)
            return rc;
        s.grid_dirty = false;
    }
    return GRT_OK;
}

}  // namespace

extern "C" {

bool rtInitContext(void)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (s.ctx)
        return true;
    if (grt_init(0, &s.ctx) != GRT_OK) {
        s.ctx = nullptr;
        return false;
    }
    return true;
}

int rtShutdownContext(void)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.ctx)
        return GRT_OK;
    grt_synchronize(s.ctx);
    release_buffers(s);
    if (s.scene)
        grt_scene_destroy(s.scene);
    grt_shutdown(s.ctx);
    s.scene = nullptr;
    s.ctx = nullptr;
    s.frame = nullptr;
    s.cam_set = false;
    s.verts.clear();
    s.normals.clear();
    s.cell_ptrs.clear();
    s.cell_ids.clear();
    s.have_verts = s.have_normals = s.have_grid_params = false;
    s.geom_dirty = s.grid_dirty = false;
    s.last_n = 0;
    return GRT_OK;
}

const char* grtLastError(void) { return grt_last_error(); }

int grtRegisterBuffer(uint32_t* rgba8_dev)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!rgba8_dev)
        return grt::set_error(GRT_INVALID_PARAM, "grtRegisterBuffer: NULL buffer");
    s.frame = rgba8_dev;
    return GRT_OK;
}

int grtUnregisterBuffer(uint32_t* rgba8_dev)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (rgba8_dev == s.frame)
        s.frame = nullptr;
    return GRT_OK;
}

int grtSetScreenSize(uint32_t width, uint32_t height)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (width == 0 || height == 0 || (uint64_t)width * height > 0xffffffffull / 64)
        return grt::set_error(GRT_INVALID_PARAM, "grtSetScreenSize: %ux%u", width, height);
    s.cam.width = width;
    s.cam.height = height;
    return GRT_OK;
}

int grtSetCameraParameters(float posX, float posY, float posZ, float baseDirX, float baseDirY,
                           float baseDirZ, float uNearX, float uNearY, float uNearZ,
                           float vNearX, float vNearY, float vNearZ)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    const float v[12] = {posX,   posY,   posZ,   baseDirX, baseDirY, baseDirZ,
                         uNearX, uNearY, uNearZ, vNearX,   vNearY,   vNearZ};
    std::memcpy(s.cam.pos, v, 12);
    std::memcpy(s.cam.base, v + 3, 12);
    std::memcpy(s.cam.nu, v + 6, 12);
    std::memcpy(s.cam.nv, v + 9, 12);
    s.cam_set = true;
    return GRT_OK;
}

int grtTransferTriangleVertices(const float* vertices, uint32_t size)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = unpack_float4x3(vertices, size, s.verts, "grtTransferTriangleVertices");
    if (rc != GRT_OK)
        return rc;
    s.have_verts = true;
    s.geom_dirty = true;
    return GRT_OK;
}

int grtTransferTriangleNormals(const float* normals, uint32_t size)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = unpack_float4x3(normals, size, s.normals, "grtTransferTriangleNormals");
    if (rc != GRT_OK)
        return rc;
    s.have_normals = true;
    s.geom_dirty = true;
    return GRT_OK;
}

int grtTransferCellPointers(const int32_t* cellPointers, uint32_t size)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if ((!cellPointers && size) || size % 2 != 0)
        return grt::set_error(GRT_INVALID_PARAM,
                              "grtTransferCellPointers: want {start, count} pairs, got %u ints",
                              size);
    s.cell_ptrs.assign(cellPointers, cellPointers + size);
    s.grid_dirty = true;
    return GRT_OK;
}

int grtTransferCellTriangleIds(const int32_t* cellTriangleIds, uint32_t size)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!cellTriangleIds && size)
        return grt::set_error(GRT_INVALID_PARAM, "grtTransferCellTriangleIds: NULL data");
    s.cell_ids.assign(cellTriangleIds, cellTriangleIds + size);
    s.grid_dirty = true;
    return GRT_OK;
}

int grtSetBoxMin(float x, float y, float z)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    s.bmin[0] = x;
    s.bmin[1] = y;
    s.bmin[2] = z;
    s.grid_dirty = true;
    return GRT_OK;
}

int grtSetBoxMax(float x, float y, float z)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    s.bmax[0] = x;
    s.bmax[1] = y;
    s.bmax[2] = z;
    s.grid_dirty = true;
    return GRT_OK;
}

int grtSetGridParameters(float cellSizeX, float cellSizeY, float cellSizeZ, int32_t numCellsX,
                         int32_t numCellsY, int32_t numCellsZ)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    const float c[3] = {cellSizeX, cellSizeY, cellSizeZ};
    const int32_t n[3] = {numCellsX, numCellsY, numCellsZ};
    for (int d = 0; d < 3; ++d)
        if (!(c[d] > 0.0f) || n[d] < 1)
            return grt::set_error(GRT_INVALID_PARAM,
                                  "grtSetGridParameters: axis %d cell %g, %d cells", d, c[d],
                                  n[d]);
    for (int d = 0; d < 3; ++d) {
        s.cell[d] = c[d];
        s.n[d] = n[d];
    }
    s.have_grid_params = true;
    s.grid_dirty = true;
    return GRT_OK;
}

int grtSetBlockSize(uint32_t width, uint32_t height)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = need_ctx(s);
    if (rc != GRT_OK)
        return rc;
    const uint64_t b = (uint64_t)width * height;
    if (b != 64 && b != 128 && b != 192 && b != 256)
        return GRT_OK;   // the reference ignored the block size altogether
    return grt_set_block_size(s.ctx, (uint32_t)b, (uint32_t)b);
}

int grtRayTrace(void)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    int rc = need_ctx(s);
    if (rc != GRT_OK)
        return rc;
    if (!s.frame)
        return grt::set_error(GRT_NOT_READY, "grtRayTrace: no frame buffer registered");
    if (!s.cam_set || s.cam.width == 0)
        return grt::set_error(GRT_NOT_READY, "grtRayTrace: camera or screen size not set");
    if (!s.have_normals)
        return grt::set_error(GRT_NOT_READY, "grtRayTrace: headlight shading needs normals");
    const uint32_t n = s.cam.width * s.cam.height;
    // The frame buffer must hold n pixels inside its allocation.
    hipDeviceptr_t base = nullptr;
    size_t alloc = 0;
    if (hipMemGetAddressRange(&base, &alloc, (hipDeviceptr_t)s.frame) != hipSuccess ||
        (size_t)((char*)s.frame - (char*)base) + (size_t)n * 4 > alloc) {
        (void)hipGetLastError();
        return grt::set_error(GRT_INVALID_PARAM,
                              "grtRayTrace: the registered buffer does not hold %ux%u pixels "
                              "of device memory", s.cam.width, s.cam.height);
    }
    if ((rc = sync_scene(s)) != GRT_OK)
        return rc;
    if ((rc = reserve_rays(s, n)) != GRT_OK)
        return rc;
    if ((rc = grt_gen_primary(s.ctx, &s.cam, &s.rays)) != GRT_OK)
        return rc;
    if ((rc = grt_trace_closest(s.scene, &s.rays, &s.hits, n, s.cam.width, nullptr)) != GRT_OK)
        return rc;
    if ((rc = grt_shade_headlight(s.scene, &s.rays, &s.hits, n, s.frame)) != GRT_OK)
        return rc;
    s.last_n = n;
    return grt_synchronize(s.ctx);
}

int grtGetHitRecords(int32_t* id, float* t, float* u, float* v)
{
    CabiState& s = state();
    std::lock_guard<std::mutex> g(s.mu);
    if (!s.ctx || s.last_n == 0)
        return grt::set_error(GRT_NOT_READY, "grtGetHitRecords: no frame traced yet");
    const size_t bytes = (size_t)s.last_n * 4;
    void* dst[4] = {id, t, u, v};
    const void* src[4] = {s.hits.id, s.hits.t, s.hits.u, s.hits.v};
    for (int k = 0; k < 4; ++k) {
        if (!dst[k])
            continue;
        const hipError_t e = hipMemcpy(dst[k], src[k], bytes, hipMemcpyDeviceToHost);
        if (e != hipSuccess)
            return grt::hip_check(e, "grtGetHitRecords");
    }
    return GRT_OK;
}

}  // extern "C"
