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
                             timing, so poses can then differ between runs in their last
                             bits; 0 (the default) keeps the library's table of shapes and
                             reproducible runs. */
} efloop_params;

typedef struct efloop_ctx *efloop_handle;

/* The reference configs' values (ef_file.cfg): 3 m, 20 m, 10, 200, icpWeight
 * 10, SO3 on; 4M surfels; graphs on, autotuning off. */
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

/* ---- n sequences in lockstep (no counterpart in the reference, whose frame
 * loop tracks one sequence): every sequence has its own loop (model,
 * prediction, odometry handle and stream) and the same per-frame work as
 * efloop_process, except that the n tracking calls of a frame are one batched
 * call (odom_track_batch_policy, under `pol`), and, with a network set, the
 * learned start's input and forward pass (slam_cnn.h) run once for all n
 * sequences.  A sequence's results are bit-identical to the same sequence
 * run as a batch of one.  Autotuning is not done (the shapes are the
 * library's table). ---- */
typedef struct efloop_batch_s *efloop_batch;

/* Per sequence, since creation. */
typedef struct efloop_counters {
    int frames_tracked;        /* tracking calls */
    long long gn_iterations;   /* Gauss-Newton iterations of the kept refinements, summed */
    long long alt_iterations;  /* Gauss-Newton iterations of refinements judged away, summed */
    int predictions;           /* tracking calls started from the network's prediction */
    int predictions_kept;      /* ... whose refinement the judge kept */
} efloop_counters;

int efloop_batch_create(int n, int width, int height, float cx, float cy, float fx, float fy,
                        const efloop_params *p, const odom_policy *pol, efloop_batch *out);
int efloop_batch_destroy(efloop_batch b);
/* The learned start: the warm-start network's parameters (host fp32, the
 * layout of cnn_param_layout) and output scales.  640x480 sequences only.
 * Without it each call starts as `pol` says. */
int efloop_batch_set_net(efloop_batch b, const float *params, size_t nparams, float t_scale, float r_scale);
/* One frame of every sequence: rgb[i], depth[i] as in efloop_process;
 * poses_out: n row-major 4x4 poses. */
int efloop_batch_process(efloop_batch b, const odom_frame_view *rgb, const odom_frame_view *depth,
                         float *poses_out);
int efloop_batch_get_loop(efloop_batch b, int i, efloop_handle *out);
int efloop_batch_get_counters(efloop_batch b, int i, efloop_counters *out);

#ifdef __cplusplus
}
#endif

#endif /* SLAM_LOOP_H */
