/* slam_surfel.h -- C ABI of the surfel model's fusion on MI355X (gfx950):
 * initialisation from the first frame, the index map, data association and
 * update, and the cleaning of unstable surfels.
 *
 * Replaces Model::initialise, Model::fuse and Model::clean (Model.cpp:208-599;
 * shaders vertex_feedback.vert/.geom, init_unstable.vert, data.vert/.geom/
 * .frag, update.vert, copy_unstable.vert/.geom) and IndexMap::predictIndices
 * (IndexMap.cpp:151-210; index_map.vert/.frag) of the reference, with
 * IndexMap::FACTOR = 1 as in the reference.  The deformation-graph part of
 * the cleaning (copy_unstable.vert:103-207, active only with graph nodes) is
 * not included.  Part of libslam_odom.so; status codes, frame views and
 * memory kinds are those of slam_odom.h, and the surfel layout is the one of
 * slam_splat.h (12 floats), so the model feeds splat_combined_predict.
 *
 * Depth frames are uint16 millimetres (ODOM_FMT_DEPTH_U16) and are turned
 * into metres as ComputePack's METRIC pass does (depth_metric.frag: 0 outside
 * [300, depth_cutoff * 1000]); colour frames are RGB8.
 */
#ifndef SLAM_SURFEL_H
#define SLAM_SURFEL_H

#include "slam_odom.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The model, its index map and the fusion scratch on one device. */
typedef struct surfel_ctx *surfel_handle;

/* The model buffers (two of max_surfels surfels, Model.cpp:26-80), the index
 * map (IndexMap.cpp:21-110 at FACTOR 1) and the fusion scratch.  All device
 * memory is allocated here; max_surfels + width * height must stay below
 * 2^23. */
int surfel_create(int width, int height, float cx, float cy, float fx, float fy, int max_surfels,
                  surfel_handle *out);
int surfel_destroy(surfel_handle h);
int surfel_set_stream(surfel_handle h, void *stream);

/* Model::initialise(FeedbackBuffer raw, filtered) of the first frame
 * (elastic_fusion_file.cpp:353-360): a surfel per pixel with depth in
 * (0, max_depth], in the camera frame.  Synchronous (the count is read
 * back, as glGetQueryObjectuiv does). */
int surfel_initialise(surfel_handle h, const odom_frame_view *rgb, const odom_frame_view *raw_depth,
                      const odom_frame_view *filtered_depth, float depth_cutoff, int time, float max_depth);

/* IndexMap::predictIndices(pose, time, model, max_depth, time_delta). */
int surfel_predict_indices(surfel_handle h, const float pose[16], int time, float max_depth, int time_delta);

/* Model::fuse(time, rgb, raw, filtered, depth_cutoff, conf, weighting) at
 * this pose (Model.cpp:336-487): data association against the current index
 * map, the update of matched surfels, and the list of new points that the
 * next surfel_clean appends. */
int surfel_fuse(surfel_handle h, const float pose[16], int time, const odom_frame_view *rgb,
                const odom_frame_view *raw_depth, const odom_frame_view *filtered_depth, float depth_cutoff,
                float max_depth, float weighting);

/* Model::clean(time, conf_threshold, graph = {}, time_delta, ...) at this
 * pose (Model.cpp:489-599) against the current index map: keeps, in order,
 * the surfels and then the new points that pass copy_unstable's tests.
 * Synchronous (the count is read back). */
int surfel_clean(surfel_handle h, const float pose[16], int time, float conf_threshold, int time_delta);

/* Model::lastCount(); the current model buffer (device, count * 12
 * floats, valid until the next initialise or clean); a copy to the host;
 * and a model set from host or device memory (ODOM_MEM_*). */
int surfel_count(surfel_handle h, int *count);
int surfel_get_model(surfel_handle h, const float **dev_ptr);
int surfel_read_model(surfel_handle h, float *host_dst, int max_surfels, int *count);
int surfel_set_model(surfel_handle h, const float *model, int n, int mem);

/* The index map of the last surfel_predict_indices (index uint32, vertex and
 * confidence, colour and time, normal and radius; width * height each) and
 * the new points of the last surfel_fuse (12 floats each). */
int surfel_read_index_map(surfel_handle h, uint32_t *index, float *vert_conf, float *color_time, float *norm_rad);
int surfel_read_fresh(surfel_handle h, float *host_dst, int max_points, int *count);

/* Measurement: microseconds per call of predict_indices, fuse and clean
 * (HIP events on the handle's stream around one frame's sequence
 * predict_indices -> fuse -> predict_indices -> clean, reps times from the
 * same starting model, which is restored before each rep; afterwards the
 * model is the one update's result).  Allocates a copy of the starting model
 * for the call's duration. */
int surfel_profile(surfel_handle h, const float pose[16], int time, const odom_frame_view *rgb,
                   const odom_frame_view *raw_depth, const odom_frame_view *filtered_depth, float depth_cutoff,
                   float max_depth, float weighting, float conf_threshold, int time_delta, int reps, float us[4]);

#ifdef __cplusplus
}
#endif

#endif /* SLAM_SURFEL_H */
