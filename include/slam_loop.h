/* slam_loop.h -- the reference's frame loop without loop closure, native, over
 * the odometry, splat and surfel C ABIs (slam_odom.h, slam_splat.h,
 * slam_surfel.h).
 *
 * Replaces the per-frame body of elastic_fusion_file.cpp:335-420 and
 * 551-602 with closeLoops off:
 *   bilateral filter (ComputePack FILTER);
 *   first frame: Model::initialise and initFirstRGB;
 *   later frames: frame-to-model tracking against the filled-in prediction
 *   (initICPModel, initRGBModel, initICP, initRGB,
 *   getIncrementalTransformation), the velocity weighting, and, when the
 *   previous call's ICP error was below 1e-4, predictIndices, fuse,
 *   predictIndices, clean;
 *   every frame: the prediction and fill-in for the next one (predict()).
 * The prediction the reference makes right after tracking (:399) feeds only
 * the loop-closure search and is not made.  All work runs on the odometry's
 * stream; the call waits where the reference reads results on the host (the
 * pose and the surfel count).  Part of libslam_odom.so.
 */
#ifndef SLAM_LOOP_H
#define SLAM_LOOP_H

#include "slam_odom.h"
#include "slam_splat.h"
#include "slam_surfel.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct efloop_params {
    float depth_cutoff;   /* depthCutoff: bilateral and metric window (m), ef_file.cfg:15 */
    float max_depth;      /* maxDepthProcessed (m), ef_file.cfg:16 */
    float confidence;     /* ef_file.cfg:7 */
    int time_delta;       /* ef_file.cfg:13 */
    float icp_weight;     /* icpWeight */
    int so3;              /* the SO3 pre-alignment of frame-to-model tracking */
    int max_surfels;      /* model capacity */
    int use_graph;        /* replay the tracking call from a hipGraph */
    int autotune;         /* after the first tracking call, time the step kernels' launch
                             shapes on its inputs and keep the fastest (odom_autotune).  The
                             shapes set the order of the block sums, and the choice follows
                             timing, so poses can differ between runs in their last bits;
                             0 keeps the fixed default shapes (reproducible runs). */
} efloop_params;

typedef struct efloop_ctx *efloop_handle;

/* The reference configs' values (ef_file.cfg): 3 m, 20 m, 10, 200, icpWeight
 * 10, SO3 on; 4M surfels; graphs and autotuning on. */
int efloop_default_params(efloop_params *p);
int efloop_create(int width, int height, float cx, float cy, float fx, float fy, const efloop_params *p,
                  efloop_handle *out);
int efloop_destroy(efloop_handle h);

/* One frame: depth ODOM_FMT_DEPTH_U16 (mm), colour ODOM_FMT_RGB8, host or
 * device.  pose_out: the tracked camera-to-model pose, 4x4 row-major. */
int efloop_process(efloop_handle h, const odom_frame_view *rgb, const odom_frame_view *depth, float pose_out[16]);

/* The loop's components (owned by the loop), its frame count and the last
 * fusion weighting. */
int efloop_get_handles(efloop_handle h, odom_handle *odom, splat_handle *prediction, surfel_handle *model);
int efloop_get_state(efloop_handle h, int *tick, float *last_weighting, float *last_icp_error);

#ifdef __cplusplus
}
#endif

#endif /* SLAM_LOOP_H */
