// render_phong.cpp -- a Phong frame through the reference-style renderer
// plugin grt::HipGridRenderer (the counterpart of rtgl::GpuRenderer,
// GpuRenderer.cpp:43-297): builder plugin, PinholeCamera, reflections on,
// newFrame + render.  A second frame after rebuildGrid() must be identical.
//
//   render_phong SCENE OUT WIDTH HEIGHT LAMBDA EYE_X EYE_Y EYE_Z AT_X AT_Y AT_Z [COLOURS]
//
// SCENE:   uint32 n, then 9n floats of vertices and 9n floats of normals.
// COLOURS: 3n floats, the diffuse colour of each triangle (default white).
// OUT:     float rgb[3*w*h], then uint32 rgba[w*h].
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "grt/rt_compat.h"

static int fail(const char* msg)
{
    std::fprintf(stderr, "render_phong: %s\n", msg);
    return 1;
}

static bool read_floats(const char* path, std::vector<float>& out, size_t skip_words)
{
    FILE* f = std::fopen(path, "rb");
    if (!f)
        return false;
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, (long)(4 * skip_words), SEEK_SET);
    out.resize(bytes > 0 ? (size_t)bytes / 4 - skip_words : 0);
    const bool ok = std::fread(out.data(), 4, out.size(), f) == out.size();
    std::fclose(f);
    return ok;
}

int main(int argc, char** argv)
{
    if (argc != 12 && argc != 13)
        return fail("usage: SCENE OUT WIDTH HEIGHT LAMBDA EYE(3) AT(3) [COLOURS]");
    const uint32_t width = (uint32_t)std::atoi(argv[3]), height = (uint32_t)std::atoi(argv[4]);
    float lambda = (float)std::atof(argv[5]);
    float eye[3], at[3];
    for (int k = 0; k < 3; ++k) {
        eye[k] = (float)std::atof(argv[6 + k]);
        at[k] = (float)std::atof(argv[9 + k]);
    }
    std::vector<float> scene;
    if (!read_floats(argv[1], scene, 1))
        return fail("cannot read scene");
    const size_t n = scene.size() / 18;
    if (n == 0 || scene.size() != 18 * n)
        return fail("bad scene file");
    std::vector<float> colours;
    if (argc == 13 && (!read_floats(argv[12], colours, 0) || colours.size() != 3 * n))
        return fail("bad colour file");

    grt_context ctx = nullptr;
    if (grt_init(0, &ctx) != GRT_OK)
        return fail(grt_last_error());
    int rc = 0;
    try {
        grt::ref_ptr<grt::Geometry> geom(new grt::Geometry);
        geom->vertices.resize(3 * n);
        geom->normals.resize(3 * n);
        geom->triDesc.resize(n);
        for (size_t i = 0; i < 3 * n; ++i) {
            geom->vertices[i].set(scene[3 * i], scene[3 * i + 1], scene[3 * i + 2]);
            geom->normals[i].set(scene[9 * n + 3 * i], scene[9 * n + 3 * i + 1],
                                 scene[9 * n + 3 * i + 2]);
            geom->triDesc[i / 3].v[i % 3] = (uint32_t)i;
        }
        grt::ref_ptr<grt::HipGridAccStructBuilder> builder(new grt::HipGridAccStructBuilder(ctx));
        builder->receiveParameter(grt::GRT_PARAM_LAMBDA, &lambda);

        grt::ref_ptr<grt::PinholeCamera> cam(new grt::PinholeCamera);
        cam->setLookAt(eye[0], eye[1], eye[2], at[0], at[1], at[2], 0.0f, 1.0f, 0.0f);
        cam->setPerspective(60.0f, 1.0f, 1000.0f);
        cam->setViewport(width, height);

        grt::ref_ptr<grt::HipGridRenderer> r(
            new grt::HipGridRenderer(ctx, cam.get(), geom.get(), builder.get()));
        r->setReflectionEnabled(true);
        if (!colours.empty())
            r->setDiffuseColors(colours);
        r->newFrame();
        r->render();   // builds the grid on the first frame
        std::vector<float> rgb;
        std::vector<uint32_t> rgba, again;
        r->readRgb(rgb);
        r->readFrameBuffer(rgba);
        r->rebuildGrid(true);
        r->newFrame();
        r->render();
        r->readFrameBuffer(again);
        if (again != rgba) {
            std::fprintf(stderr, "render_phong: frame after rebuildGrid differs\n");
            rc = 2;
        }
        FILE* o = std::fopen(argv[2], "wb");
        if (!o)
            throw grt::Error(GRT_INVALID_PARAM, "cannot open output");
        std::fwrite(rgb.data(), 4, rgb.size(), o);
        std::fwrite(rgba.data(), 4, rgba.size(), o);
        std::fclose(o);
    } catch (const grt::Error& e) {
        std::fprintf(stderr, "render_phong: error %d: %s\n", e.code(), e.what());
        rc = 3;
    }
    grt_shutdown(ctx);
    if (rc == 0)
        std::printf("render_phong: %ux%u frame written\n", width, height);
    return rc;
}
