/*
 * vhm_io -- host-side companions of the fusion path, in libvhm.so: the
 * reference demo's synthetic scene and binary PLY writer (src/main.cpp) and
 * its dataset loader (include/tsdfmc/dataset_io.h), with a PNG decoder that
 * runs on Linux (the reference decodes through Windows WIC only).
 *
 * None of these functions touches the GPU; they run where the reference
 * runs them, on the host, and feed or drain the C-ABI of vhm.h.
 *
 * Conventions as vhm.h: int status (VHM_OK == 0); on failure
 * vhm_io_last_error() returns the calling thread's message, which is the
 * text of the exception the reference throws in the same situation.
 */
#ifndef VHM_IO_H_
#define VHM_IO_H_

#include <stddef.h>
#include <stdint.h>

#include "vhm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* File missing, unreadable, malformed or in an unsupported format. */
#define VHM_ERR_IO 7

/* The calling thread's last vhm_io error message ("" if none). */
const char* vhm_io_last_error(void);

/* ---------------------------------------------------- synthetic demo --- */

/* Intrinsics of the synthetic demo for a width x height image
 * (runSyntheticDemo, main.cpp:404-410). */
void vhm_demo_intrinsics(int32_t width, int32_t height, vhm_intrinsics* K);

/* Camera pose of demo frame `frame_index` of `frame_count`: a circle of
 * radius 1.15 m looking at (0.08, -0.02, 0) (main.cpp:424-432), through
 * makeLookAtPose (main.cpp:177-187). */
void vhm_demo_pose(int32_t frame_index, int32_t frame_count, vhm_pose* T_wc);

/* traceScene (main.cpp:255-272): the demo scene, a sphere and a box.
 * Returns 1 and the ray parameter of the nearest hit, or 0 on a miss. */
int vhm_demo_trace(const float origin[3], const float direction[3], float* hit_t);

/* The demo's noise generator: one std::mt19937 seeded with `seed` (the demo
 * uses 42, main.cpp:414) that persists across frames. */
typedef struct vhm_demo_renderer vhm_demo_renderer;
int vhm_demo_renderer_create(uint32_t seed, vhm_demo_renderer** out);
void vhm_demo_renderer_destroy(vhm_demo_renderer* r);

/* renderSyntheticDepth (main.cpp:295-335): width*height depths in metres
 * (0 = no measurement) into depth_out.  With VHM_DEMO_NOISE the frame
 * carries the demo's patterned bias, Gaussian noise (sigma 2.5 mm) and 3 %
 * dropout, drawing from r exactly as the reference does; without it the
 * frame is the noise-free ray-cast depth and r is not advanced. */
#define VHM_DEMO_NOISE 1u
int vhm_demo_render(vhm_demo_renderer* r, const vhm_intrinsics* K, const vhm_pose* T_wc, int32_t frame_index,
                    uint32_t flags, float* depth_out);

/* ------------------------------------------------------------- PLY --- */

/* writeMeshAsPly (main.cpp:337-398): binary little-endian, unindexed
 * (three vertices per face, each with the face's unit normal), degenerate
 * triangles (|(v1-v0) x (v2-v0)| <= 1e-8) dropped.  triangles: n x 9 floats
 * in host memory.  *n_written (may be NULL) receives the faces written. */
int vhm_write_ply(const char* path, const float* triangles, size_t n_triangles, size_t* n_written);

/* --------------------------------------------------------- datasets --- */

/* listDatasetFrameIds (dataset_io.h:146-163): the all-digit subdirectory
 * names of root as integers, ascending.  Writes min(n, capacity) ids. */
int vhm_dataset_frame_ids(const char* root, int32_t* ids, size_t capacity, size_t* n);

/* loadDatasetIntrinsics (dataset_io.h:380-394): `path` (normally
 * <root>/intrinsic.txt, else the first <root>/<id>/params.txt) holds
 * fx fy cx cy [width height] [depth_scale] (dataset_io.h:92-133); a file
 * without a resolution takes it from the first <id>/depth.png. */
int vhm_dataset_intrinsics(const char* path, vhm_intrinsics* K, float* depth_scale);

/* loadPoseFromFile (dataset_io.h:165-196): 7 values (quaternion w x y z,
 * translation) or a row-major 3x4 / 4x4 matrix (12 or 16 values). */
int vhm_load_pose(const char* path, vhm_pose* T_wc);

/* quaternionToMatrix (math_types.h:158-178), row-major into R. */
void vhm_quaternion_to_matrix(float w, float x, float y, float z, float R[9]);

/* Size of a PNG image (completeIntrinsicsResolutionFromDepth,
 * dataset_io.h:241-276). */
int vhm_png_size(const char* path, int32_t* width, int32_t* height);

/* loadDepthPng (dataset_io.h:290-356): decodes a grayscale PNG to 16-bit
 * samples; depth = sample * depth_scale, 0 where the sample is 0 or the
 * product is <= 0.  The image must be K->width x K->height.  depth_out:
 * width*height floats; *valid receives the number of valid depths. */
int vhm_load_depth_png(const char* path, const vhm_intrinsics* K, float depth_scale, float* depth_out,
                       size_t* valid);

/* The same decode without the scaling: raw 16-bit samples, for
 * vhm_integrate_u16 (which applies the scale on the GPU with the same
 * rule).  *valid counts the samples that give a valid depth at
 * depth_scale. */
int vhm_load_depth_png_u16(const char* path, const vhm_intrinsics* K, float depth_scale, uint16_t* raw_out,
                           size_t* valid);

/* loadDatasetDepthFrame (dataset_io.h:396-409): <root>/<frame_id>/depth.png
 * and <root>/<frame_id>/<pose_filename>. */
int vhm_load_dataset_frame(const char* root, int32_t frame_id, const char* pose_filename,
                           const vhm_intrinsics* K, float depth_scale, vhm_pose* T_wc, float* depth_out,
                           size_t* valid);

#ifdef __cplusplus
}
#endif

#endif /* VHM_IO_H_ */
