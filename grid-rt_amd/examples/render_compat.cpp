// render_compat.cpp -- a frame rendered through the reference-style plugin
// classes (grt/rt_compat.h), the way rtc::CudaRenderer was driven
// (CudaRenderer.cpp:9-209, Canvas.cpp:174-179): build the geometry's grid with
// the builder plugin, set up a PinholeCamera, newFrame + render.  It also
// traces a few pixels one ray at a time (traceNearestInstance) and checks them
// against the frame.
//
//   render_compat SCENE OUT WIDTH HEIGHT LAMBDA EYE_X EYE_Y EYE_Z AT_X AT_Y AT_Z [kd]
//
// SCENE: uint32 n, then 9n floats of vertices and 9n floats of normals.
// OUT:   int32 id[w*h], float t[w*h], u[w*h], v[w*h], uint32 rgba[w*h],
//        then the grid's n[3] (int32) and bmin[3], bmax[3] (float).
// With "kd" the geometry is built by the kd-tree builder plugin instead
// (KdTreeAccStructBuilder), and the trailer holds n = {-1,-1,-1} and the
// tree's box.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "grt/rt_compat.h"

static int fail(const char* msg)
{
    std::fprintf(stderr, "render_compat: %s\n", msg);
    return 1;
}

int main(int argc, char** argv)
{
    if (argc != 12 && !(argc == 13 && std::strcmp(argv[12], "kd") == 0))
        return fail("usage: SCENE OUT WIDTH HEIGHT LAMBDA EYE(3) AT(3) [kd]");
    const bool kd = argc == 13;
    const uint32_t width = (uint32_t)std::atoi(argv[3]), height = (uint32_t)std::atoi(argv[4]);
    float lambda = (float)std::atof(argv[5]);
    float eye[3], at[3];
    for (int k = 0; k < 3; ++k) {
        eye[k] = (float)std::atof(argv[6 + k]);
        at[k] = (float)std::atof(argv[9 + k]);
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f)
        return fail("cannot open scene");
    uint32_t n = 0;
    if (std::fread(&n, 4, 1, f) != 1 || n == 0)
        return fail("bad scene header");
    std::vector<float> v(9 * (size_t)n), nv(9 * (size_t)n);
    if (std::fread(v.data(), 4, v.size(), f) != v.size() ||
        std::fread(nv.data(), 4, nv.size(), f) != nv.size())
        return fail("short scene file");
    std::fclose(f);

    grt_context ctx = nullptr;
    if (grt_init(0, &ctx) != GRT_OK)
        return fail(grt_last_error());
    int rc = 0;
    try {
        grt::ref_ptr<grt::Geometry> geom(new grt::Geometry);
        geom->vertices.resize(3 * (size_t)n);
        geom->normals.resize(3 * (size_t)n);
        geom->triDesc.resize(n);
        for (uint32_t t = 0; t < n; ++t)
            for (uint32_t k = 0; k < 3; ++k) {
                const size_t i = 3 * (size_t)t + k;
                geom->vertices[i].set(v[3 * i], v[3 * i + 1], v[3 * i + 2]);
                geom->normals[i].set(nv[3 * i], nv[3 * i + 1], nv[3 * i + 2]);
                geom->triDesc[t].v[k] = (uint32_t)i;
            }
        if (kd) {
            grt::ref_ptr<grt::HipKdTreeAccStructBuilder> builder(
                new grt::HipKdTreeAccStructBuilder(ctx));
            builder->buildGeometry(geom.get());
        } else {
            grt::ref_ptr<grt::HipGridAccStructBuilder> builder(
                new grt::HipGridAccStructBuilder(ctx));
            builder->receiveParameter(grt::GRT_PARAM_LAMBDA, &lambda);
            builder->buildGeometry(geom.get());
        }

        grt::ref_ptr<grt::PinholeCamera> cam(new grt::PinholeCamera);
        cam->setLookAt(eye[0], eye[1], eye[2], at[0], at[1], at[2], 0.0f, 1.0f, 0.0f);
        cam->setPerspective(60.0f, 1.0f, 1000.0f);
        cam->setViewport(width, height);

        grt::ref_ptr<grt::HipRenderer> renderer(new grt::HipRenderer(ctx, cam.get(), geom.get()));
        renderer->newFrame();
        renderer->render();
        std::vector<int32_t> id;
        std::vector<float> t, u, w;
        std::vector<uint32_t> rgba;
        renderer->readHits(id, t, u, w);
        renderer->readFrameBuffer(rgba);

        // A few pixels again, one ray at a time (IAccStruct::traceNearestInstance).
        std::vector<grt::Instance> inst(1);
        inst[0].geometry = geom.get();
        const uint32_t probes[4][2] = {{width / 2, height / 2}, {0, 0}, {width / 3, height / 4},
                                       {width - 1, height - 1}};
        for (const auto& p : probes) {
            grt::Sample s;
            cam->getRayOrigin(s.ray.orig, (float)p[0], (float)p[1]);
            cam->getRayDirection(s.ray.dir, (float)p[0], (float)p[1]);
            geom->accStruct->traceNearestInstance(inst, s);
            const size_t i = (size_t)p[1] * width + p[0];
            const bool hit = s.hit.instance != nullptr;
            if (hit != (id[i] >= 0) ||
                (hit && ((int32_t)s.hit.triangleId != id[i] || s.hit.distance != t[i] ||
                         s.hit.v1Coord != u[i] || s.hit.v2Coord != w[i]))) {
                std::fprintf(stderr, "render_compat: per-ray trace differs at pixel %u,%u\n",
                             p[0], p[1]);
                rc = 2;
            }
        }
        grt_grid_info info;
        if (kd) {
            const grt_kdtree_info ki =
                static_cast<grt::HipKdTreeAccStruct*>(geom->accStruct.get())->treeInfo();
            for (int k = 0; k < 3; ++k) {
                info.n[k] = -1;
                info.bmin[k] = ki.bmin[k];
                info.bmax[k] = ki.bmax[k];
            }
        } else {
            info = static_cast<grt::HipGridAccStruct*>(geom->accStruct.get())->gridInfo();
        }
        FILE* o = std::fopen(argv[2], "wb");
        if (!o)
            throw grt::Error(GRT_INVALID_PARAM, "cannot open output");
        const size_t np = (size_t)width * height;
        std::fwrite(id.data(), 4, np, o);
        std::fwrite(t.data(), 4, np, o);
        std::fwrite(u.data(), 4, np, o);
        std::fwrite(w.data(), 4, np, o);
        std::fwrite(rgba.data(), 4, np, o);
        std::fwrite(info.n, 4, 3, o);
        std::fwrite(info.bmin, 4, 3, o);
        std::fwrite(info.bmax, 4, 3, o);
        std::fclose(o);
    } catch (const grt::Error& e) {
        std::fprintf(stderr, "render_compat: error %d: %s\n", e.code(), e.what());
        rc = 3;
    }
    grt_shutdown(ctx);
    if (rc == 0)
        std::printf("render_compat: %ux%u frame written\n", width, height);
    return rc;
}
