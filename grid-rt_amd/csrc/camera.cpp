// camera.cpp -- PinholeCamera set-up on the host (PinholeCamera.cpp:23-55,
// 91-118).  The reference derives its orientation through vrbase's
// makeLookAt/getRotation (PinholeCamera.cpp:99-101), which is not vendored; the
// basis here is the gluLookAt one written out: f = normalize(center - eye),
// X = normalize(f x up), Y = X x f, Z = -f, and the camera looks along -Z.
#include <cmath>

#include "grt_internal.h"

namespace grt {

static inline float dot3h(const float a[3], const float b[3])
{
    return std::fmaf(a[2], b[2], std::fmaf(a[1], b[1], a[0] * b[0]));
}

static inline void cross3h(const float a[3], const float b[3], float c[3])
{
    c[0] = std::fmaf(a[1], b[2], -(a[2] * b[1]));
    c[1] = std::fmaf(a[2], b[0], -(a[0] * b[2]));
    c[2] = std::fmaf(a[0], b[1], -(a[1] * b[0]));
}

static inline void normalize3h(float v[3])
{
    const float inv = 1.0f / std::sqrt(dot3h(v, v));
    v[0] = v[0] * inv;
    v[1] = v[1] * inv;
    v[2] = v[2] * inv;
}

void lookat_basis(const float eye[3], const float center[3], const float up[3], float ax[3],
                  float ay[3], float az[3])
{
    float fw[3] = {center[0] - eye[0], center[1] - eye[1], center[2] - eye[2]};
    normalize3h(fw);
    cross3h(fw, up, ax);
    normalize3h(ax);
    cross3h(ax, fw, ay);
    az[0] = -fw[0];
    az[1] = -fw[1];
    az[2] = -fw[2];
}

void camera_from_basis(const float pos[3], const float ax[3], const float ay[3], const float az[3],
                       float fovy_deg, float znear, uint32_t width, uint32_t height,
                       grt_camera* cam)
{
    // newFrame: near plane at zNear, half extents sy = zNear*tan(fovy/2), sx = sy*W/H.
    const float sz = znear;
    const float rad = (fovy_deg * 0.5f) * 0.017453292519943295f;
    const float sy = (float)((double)sz * std::tan((double)rad));
    const float sx = (sy * (float)width) / (float)height;
    for (int i = 0; i < 3; ++i) {
        const float near_origin = ((pos[i] + ax[i] * -sx) + ay[i] * -sy) + az[i] * -sz;
        cam->pos[i] = pos[i];
        cam->nu[i] = ax[i] * (2.0f * sx);
        cam->nv[i] = ay[i] * (2.0f * sy);
        cam->base[i] = near_origin - pos[i];
    }
    cam->width = width;
    cam->height = height;
}

void camera_lookat(const float eye[3], const float center[3], const float up[3], float fovy_deg,
                   float znear, uint32_t width, uint32_t height, grt_camera* cam)
{
    float ax[3], ay[3], az[3];
    lookat_basis(eye, center, up, ax, ay, az);
    camera_from_basis(eye, ax, ay, az, fovy_deg, znear, width, height, cam);
}

}  // namespace grt
