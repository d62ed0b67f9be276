/*
 * rt_cabi.h -- the reference tracer's flat C ABI, served by the HIP path.
 *
 * The reference's GPU tracer is driven through eleven extern "C" calls on
 * file-static state (src/rtcuda/RayTracing.h:6-44, implemented in
 * src/rtcuda/RayTracing.cu:41-255): the host builds the grid, expands
 * triangles to three float4 per triangle, flattens the cells into
 * {start, count} pairs and triangle ids pre-multiplied by 3
 * (src/rtcuda/CudaRenderer.cpp:65-209), and cudaRayTrace() renders one
 * headlight-shaded RGBA8 frame into a registered pixel buffer.  Each call below
 * replaces the reference call named beside it, with the same argument order and
 * meaning, so a caller changes only the prefix:
 *
 *   rtInitContext            -> rtInitContext           (RayTracing.h:8)
 *   cudaRegisterBuffer       -> grtRegisterBuffer       (RayTracing.h:11)
 *   cudaUnregisterBuffer     -> grtUnregisterBuffer     (RayTracing.h:14)
 *   cudaSetScreenSize        -> grtSetScreenSize        (RayTracing.h:17)
 *   cudaSetCameraParameters  -> grtSetCameraParameters  (RayTracing.h:20-23)
 *   cudaTransferTriangleVertices -> grtTransferTriangleVertices (RayTracing.h:27)
 *   cudaTransferTriangleNormals  -> grtTransferTriangleNormals  (RayTracing.h:28)
 *   cudaTransferCellPointers     -> grtTransferCellPointers     (RayTracing.h:31)
 *   cudaTransferCellTriangleIds  -> grtTransferCellTriangleIds  (RayTracing.h:32)
 *   cudaSetBoxMin / cudaSetBoxMax -> grtSetBoxMin / grtSetBoxMax (RayTracing.h:33-34)
 *   cudaSetGridParameters    -> grtSetGridParameters    (RayTracing.h:35-36)
 *   cudaSetBlockSize         -> grtSetBlockSize         (RayTracing.h:39)
 *   cudaRayTrace             -> grtRayTrace             (RayTracing.h:42)
 *
 * Differences, all additive:
 *   - The pixel buffer is device memory the caller owns (width*height uint32,
 *     RGBA8, row y = 0 at the bottom as in the reference) instead of a GL pixel
 *     buffer object id; there is no GL interop.
 *   - Every call returns a status (GRT_OK = 0 or a GRT_* code from grt.h, the
 *     message in grtLastError()) instead of aborting the process; a caller that
 *     ignores the result keeps the reference's calling pattern.
 *   - Device memory is released by rtShutdownContext (the reference never freed
 *     it) and is replaced, not leaked, when data is transferred again.
 *   - grtGetHitRecords exposes the frame's hit records (triangle id, t, u, v).
 * The state is process-wide, as in the reference, and guarded by a mutex.
 */
#ifndef GRT_RT_CABI_H
#define GRT_RT_CABI_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RayTracing.h:8 / RayTracing.cu:43-53: open device 0.  False on failure. */
bool rtInitContext(void);
/* No counterpart (the reference leaked its buffers): free all device state. */
int rtShutdownContext(void);
const char* grtLastError(void);

/* RayTracing.h:11-14: the frame buffer, width*height RGBA8 pixels of device
 * memory owned by the caller.  Unregistering a buffer that is not the
 * registered one does nothing (RayTracing.cu:65-73). */
int grtRegisterBuffer(uint32_t* rgba8_dev);
int grtUnregisterBuffer(uint32_t* rgba8_dev);

/* RayTracing.h:17 / RayTracing.cu:75-82. */
int grtSetScreenSize(uint32_t width, uint32_t height);
/* RayTracing.h:20-23 / RayTracing.cu:84-104: PinholeCamera position, base
 * direction and near-plane U, V vectors (PinholeCamera.cpp:41-54). */
int grtSetCameraParameters(float posX, float posY, float posZ,
                           float baseDirX, float baseDirY, float baseDirZ,
                           float uNearX, float uNearY, float uNearZ,
                           float vNearX, float vNearY, float vNearZ);

/* RayTracing.h:27-28 / RayTracing.cu:106-138: `size` floats, 12 per triangle
 * (three float4 vertices or normals, w ignored; CudaRenderer.cpp:86-147). */
int grtTransferTriangleVertices(const float* vertices, uint32_t size);
int grtTransferTriangleNormals(const float* normals, uint32_t size);

/* RayTracing.h:31-32 / RayTracing.cu:140-170: `size` int32 each.
 * cellPointers: {start, count} per cell, x fastest then y then z
 * (CudaRenderer.cpp:160-196); cellTriangleIds: 3 * triangle index
 * (CudaRenderer.cpp:190). */
int grtTransferCellPointers(const int32_t* cellPointers, uint32_t size);
int grtTransferCellTriangleIds(const int32_t* cellTriangleIds, uint32_t size);

/* RayTracing.h:33-36 / RayTracing.cu:172-219: the grid's (already padded)
 * box, cell size and cells per axis; the inverse cell size is 1/cellSize. */
int grtSetBoxMin(float x, float y, float z);
int grtSetBoxMax(float x, float y, float z);
int grtSetGridParameters(float cellSizeX, float cellSizeY, float cellSizeZ,
                         int32_t numCellsX, int32_t numCellsY, int32_t numCellsZ);

/* RayTracing.h:39 (a no-op in RayTracing.cu:221-225): width*height is taken
 * as the traversal kernels' block size when it is 64, 128, 192 or 256. */
int grtSetBlockSize(uint32_t width, uint32_t height);

/* RayTracing.h:42 / RayTracing.cu:227-253 and kernel rayTrace
 * (RayTracing_kernel.cu:140-275): one primary ray through each pixel corner,
 * closest hit, headlight shading (RayTracing_kernel.cu:104-120), written to the
 * registered buffer; returns when the frame is complete. */
int grtRayTrace(void);

/* No counterpart: copy the last frame's hit records to the host (NULL skips a
 * field), width*height entries each, id -1 on a miss. */
int grtGetHitRecords(int32_t* id, float* t, float* u, float* v);

#ifdef __cplusplus
}
#endif

#endif /* GRT_RT_CABI_H */
