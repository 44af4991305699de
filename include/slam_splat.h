/* slam_splat.h -- C ABI of the surfel splat prediction and fill-in on MI355X
 * (gfx950): the producers of the model maps that the odometry's
 * initICPModel / initRGBModel consume.
 *
 * Replaces IndexMap::combinedPredict (IndexMap.cpp:243-345; shaders
 * splat.vert, combo_splat.frag) and FillIn::vertex / normal / image
 * (FillIn.cpp:21-198; fill_vertex.frag, fill_normal.frag, fill_rgb.frag) of
 * the reference.  Part of libslam_odom.so; status codes, frame views and
 * memory kinds are those of slam_odom.h.
 *
 * Model layout (the reference's Vertex, 48 bytes per surfel, Vertex.h):
 *   float[12] = { x, y, z, confidence,  colour, unused, init_time, last_time,
 *                 nx, ny, nz, radius }
 * with the colour a float holding (r << 16 | g << 8 | b) and positions and
 * normals in the global frame.
 *
 * Outputs of a prediction, width x height, row 0 first:
 *   SPLAT_IMAGE   uint8 RGBA  (r, g, b, 255) of the nearest surfel, 0 elsewhere
 *   SPLAT_VERTEX  float RGBA  camera-frame vertex and confidence, 0 elsewhere
 *   SPLAT_NORMAL  float RGBA  camera-frame normal and radius, 0 elsewhere
 *   SPLAT_TIME    uint16      init_time of the nearest surfel, 0 elsewhere
 * Outputs of a fill-in: SPLAT_FILL_IMAGE (uint8 RGBA), SPLAT_FILL_VERTEX and
 * SPLAT_FILL_NORMAL (float RGBA).
 */
#ifndef SLAM_SPLAT_H
#define SLAM_SPLAT_H

#include "slam_odom.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPLAT_IMAGE 0
#define SPLAT_VERTEX 1
#define SPLAT_NORMAL 2
#define SPLAT_TIME 3
#define SPLAT_FILL_IMAGE 4
#define SPLAT_FILL_VERTEX 5
#define SPLAT_FILL_NORMAL 6

#define SPLAT_FLOATS 12          /* floats per surfel */
#define SPLAT_MAX_POINT_SIZE 256 /* point sizes are clamped to [1, this] pixels */

typedef struct splat_ctx *splat_handle;

/* The prediction and fill-in buffers of one IndexMap target (the frame
 * loop's ACTIVE and INACTIVE predictions are two handles) and FillIn's
 * textures (IndexMap.cpp:21-110, FillIn.cpp:21-54).  max_host_surfels sizes
 * the device staging for models passed in host memory (0: device models
 * only).  All device memory is allocated here. */
int splat_create(int width, int height, float cx, float cy, float fx, float fy, int max_host_surfels,
                 splat_handle *out);
int splat_destroy(splat_handle h);

/* Run the handle's work on this stream (a hipStream_t; NULL restores the
 * handle's own), e.g. odom_get_stream's, so that a prediction, its fill-in
 * and the odometry's initICPModel are ordered without host waits. */
int splat_set_stream(splat_handle h, void *stream);

/* IndexMap::combinedPredict(pose, model, depthCutoff, confThreshold, time,
 * maxTime, timeDelta, type), IndexMap.cpp:243-345: pose is the 4x4
 * camera-to-global transform (row-major, host); model holds n surfels
 * (SPLAT_FLOATS floats each) in host or device memory (ODOM_MEM_*).
 * Enqueued on the handle's stream; a device model must stay valid until
 * the stream reaches the call. */
int splat_combined_predict(splat_handle h, const float pose[16], const float *model, int n, int mem,
                           float depth_cutoff, float conf_threshold, int time, int max_time, int time_delta);

/* FillIn::vertex(vertexTex, depth, passthrough_geometry),
 * FillIn::normal(normalTex, depth, passthrough_geometry) and
 * FillIn::image(imageTex, rgb, passthrough_image) on the handle's last
 * prediction (FillIn.cpp:56-198): where the prediction has no surfel (vertex
 * and normal: z == 0; image: r + g + b == 0) or passthrough is set, the
 * output comes from the depth frame (ODOM_FMT_DEPTH_U16) or the colour frame
 * (ODOM_FMT_RGB8 / RGBA8). */
int splat_fill_in(splat_handle h, const odom_frame_view *depth, const odom_frame_view *rgb,
                  int passthrough_geometry, int passthrough_image);

/* Device pointer of an output (SPLAT_*), valid for the handle's life; and a
 * synchronous copy of it to host memory (tightly packed). */
int splat_get_buffer(splat_handle h, int which, void **dev_ptr);
int splat_read(splat_handle h, int which, void *host_dst);
int splat_synchronize(splat_handle h);

/* Measurement: reps predictions of a device model with HIP events around
 * each kernel on the handle's stream; average microseconds per launch of the
 * raster pass, the resolve pass and the fill-in (fill_depth NULL: no
 * fill-in timed), and the number of fragments drawn by the last prediction
 * (fragments that passed the surfel disk test). */
int splat_profile(splat_handle h, const float pose[16], const float *model_dev, int n, float depth_cutoff,
                  float conf_threshold, int time, int max_time, int time_delta, const odom_frame_view *fill_depth,
                  const odom_frame_view *fill_rgb, int reps, float us[3], long long *fragments);

#ifdef __cplusplus
}
#endif

#endif /* SLAM_SPLAT_H */
