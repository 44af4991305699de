/*
 * slam_odom.h -- C ABI of the MI355X (gfx950) dense RGB-D odometry library
 * (libslam_odom.so).
 *
 * Drop-in for the reference's src/odom path (siw-engineering/slam, an
 * ElasticFusion fork): one handle is one RGBDOdometryef object
 * (src/odom/RGBDOdometryef.h:28-131).  The GL textures of the reference
 * (GPUTexture*, mapped through CUDA-GL interop) are replaced by frame views:
 * a pointer, a row pitch, a pixel format and whether the pointer is host or
 * device memory.  Every entry point returns an int status (ODOM_OK = 0).
 *
 * Threading: a handle owns one HIP stream and is not re-entrant; distinct
 * handles may be used from distinct host threads.  Inputs are borrowed only
 * for the duration of the call.  Calls that read results back (the pose, the
 * statistics, the surfel count of slam_surfel.h) poll the stream until they
 * arrive, keeping the calling thread busy meanwhile, rather than sleeping
 * until the runtime's completion interrupt.
 */
#ifndef SLAM_ODOM_H
#define SLAM_ODOM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define ODOM_OK 0
#define ODOM_ERR_ARG (-1)      /* bad argument (null, size or format mismatch) */
#define ODOM_ERR_HIP (-2)      /* a HIP runtime call failed */
#define ODOM_ERR_STATE (-3)    /* call order violated */
#define ODOM_ERR_NODEV (-4)    /* no gfx950 device */

/* pixel formats of odom_frame_view (the reference's texture formats,
 * elastic_fusion_file.cpp:301-303, FillIn.cpp:29-42) */
#define ODOM_FMT_DEPTH_U16 1   /* uint16 depth, 1000 units per metre */
#define ODOM_FMT_RGB8 2        /* 3 bytes per pixel */
#define ODOM_FMT_RGBA8 3       /* 4 bytes per pixel (alpha ignored) */
#define ODOM_FMT_RGBA32F 4     /* float4 per pixel: vertex (x,y,z,w) or normal */

#define ODOM_MEM_HOST 0
#define ODOM_MEM_DEVICE 1

typedef struct odom_frame_view {
    const void *ptr;
    int pitch_bytes;           /* row pitch; 0 = tightly packed */
    int width, height;
    int format;                /* ODOM_FMT_* */
    int mem;                   /* ODOM_MEM_* */
} odom_frame_view;

typedef struct odom_params {
    int width, height;
    float cx, cy, fx, fy;
    float dist_thresh;         /* default 0.10 m (RGBDOdometryef.h:35) */
    float angle_thresh;        /* default sin(20 deg) (RGBDOdometryef.h:36) */
} odom_params;

/* Reported per call (RGBDOdometryef.h:62-70 plus flags). */
typedef struct odom_stats {
    float last_icp_error, last_icp_count;
    float last_rgb_error, last_rgb_count;
    float last_so3_error, last_so3_count;
    int so3_iterations;
    int gn_iterations;         /* Gauss-Newton iterations run */
    int degenerate;            /* last solve had no information (identity update) */
    int reverted;              /* the 0.3 m revert was taken */
    int seed_used;             /* odom_track_seeded / odom_track_policy: the last call kept the prediction */
    int seed_calls, seed_kept; /* predicted calls so far, and how many kept the prediction */
    int alt_iterations;        /* a judged policy call: Gauss-Newton iterations of the refinement not kept */
} odom_stats;

typedef struct odom_ctx *odom_handle;

/* Defaults of the RGBDOdometryef constructor (RGBDOdometryef.h:32-36). */
int odom_default_params(odom_params *p, int width, int height, float cx, float cy, float fx, float fy);

/* RGBDOdometryef::RGBDOdometryef, RGBDOdometryef.cpp:21-111.  Allocates every
 * device buffer once. */
int odom_create(const odom_params *p, odom_handle *out);
/* RGBDOdometryef::~RGBDOdometryef, RGBDOdometryef.cpp:113-116 */
int odom_destroy(odom_handle h);

/* The depth pre-filter of the reference's frame loop: depth_bilateral.frag:30-76
 * as run by ComputePack::compute (elastic_fusion_file.cpp:342-346).  dst is a
 * width*height uint16 buffer in memory space dst_mem. */
int odom_bilateral_depth(odom_handle h, const odom_frame_view *raw, float max_depth_m, void *dst, int dst_mem);

/* RGBDOdometryef::initICP(GPUTexture *filteredDepth, float depthCutoff),
 * RGBDOdometryef.cpp:118-142. */
int odom_init_icp_depth(odom_handle h, const odom_frame_view *filtered_depth, float depth_cutoff);
/* Bilateral pre-filter fused with initICP(filteredDepth): the pair of calls
 * ComputePack FILTER + initICP of elastic_fusion_file.cpp:346, 370. */
int odom_init_icp_raw_depth(odom_handle h, const odom_frame_view *raw_depth, float max_depth_m,
                            float depth_cutoff);
/* RGBDOdometryef::initICP(GPUTexture *v, GPUTexture *n, float), RGBDOdometryef.cpp:144-167 */
int odom_init_icp_maps(odom_handle h, const odom_frame_view *vertices, const odom_frame_view *normals,
                       float depth_cutoff);
/* RGBDOdometryef::initICPModel, RGBDOdometryef.cpp:169-206.  pose: row-major
 * 4x4 camera-to-world of the model prediction. */
int odom_init_icp_model(odom_handle h, const odom_frame_view *vertices, const odom_frame_view *normals,
                        float depth_cutoff, const float pose[16]);
/* initICPModel with the handle's current pose estimate as modelPose (the
 * currPose the reference's frame loop passes, elastic_fusion_file.cpp:366-368),
 * read on the device: lets a sequence be enqueued with no host round trip. */
int odom_init_icp_model_current(odom_handle h, const odom_frame_view *vertices, const odom_frame_view *normals,
                                float depth_cutoff);
/* RGBDOdometryef::initRGB, RGBDOdometryef.cpp:243-247 */
int odom_init_rgb(odom_handle h, const odom_frame_view *rgb);
/* RGBDOdometryef::initRGBModel, RGBDOdometryef.cpp:237-241 */
int odom_init_rgb_model(odom_handle h, const odom_frame_view *rgb);
/* RGBDOdometryef::initFirstRGB, RGBDOdometryef.cpp:249-265 */
int odom_init_first_rgb(odom_handle h, const odom_frame_view *rgb);

/* RGBDOdometryef::getIncrementalTransformation, RGBDOdometryef.cpp:267-595.
 * trans[3] and rot[9] (row-major) are read as the prior and overwritten with
 * the result.  Synchronous. */
int odom_track(odom_handle h, float trans[3], float rot[9], int rgb_only, float icp_weight, int pyramid,
               int fast_odom, int so3);
/* Asynchronous form: the prior is the handle's device-resident pose (the
 * result of the previous call, or odom_set_pose); nothing is read back. */
int odom_track_async(odom_handle h, int rgb_only, float icp_weight, int pyramid, int fast_odom, int so3);
int odom_set_pose(odom_handle h, const float trans[3], const float rot[9]);
int odom_get_pose(odom_handle h, float trans[3], float rot[9]);
int odom_synchronize(odom_handle h);
/* Replay each distinct tracking call (options, schedule and image parity)
 * from a hipGraph captured on its first use; 0 launches kernel by kernel. */
int odom_set_graph(odom_handle h, int enable);
/* No counterpart in the reference.  The SO3 pre-alignment loop
 * (RGBDOdometryef.cpp:294-380) runs as one launch whose workgroups meet at a
 * grid barrier after every iteration (when the pairs' workgroups fit on the
 * device at once, else one launch per iteration).  spin_limit: polls before a
 * barrier wait gives up (default 1 << 20, about a second); a call in which
 * one gave up makes odom_get_pose and odom_get_stats return ODOM_ERR_HIP
 * until odom_init_first_rgb; 0 gives up at the first wait (the failure path,
 * for tests).  spin_limit < 0 selects one launch per iteration.  Both forms give the same results, bit for bit.  A batch takes
 * its first handle's setting when it is created. */
int odom_set_so3_loop(odom_handle h, int spin_limit);
/* Record every call's resulting pose ([R|t], 12 floats row-major) into a
 * device ring of `capacity` entries, slot = call index % capacity; the call
 * index restarts at 0.  capacity 0 turns the log off. */
int odom_set_pose_log(odom_handle h, float *dev_ptr, int capacity);

/* RGBDOdometryef::getCovariance, RGBDOdometryef.cpp:597-600: lastA^-1 */
int odom_get_covariance(odom_handle h, double cov[36]);
/* public lastA / lastb (RGBDOdometryef.h:69-70); A row-major */
int odom_get_last_system(odom_handle h, double A[36], double b[6]);
int odom_get_stats(odom_handle h, odom_stats *s);

/* ---- kernel-level entry points: the launchers of src/odom/utils.cuh:62-175,
 * run on the handle's current pyramids with explicit constants.  Used for
 * parity checks and autotuning. ---- */

/* icpStep (reduce.cu:418-490) at `level`: out[29] = 27 JtJ/Jtr products,
 * sum r^2, inliers (JtJJtrSE3 order, types.cuh:85-96); mask (host, optional,
 * width_l*height_l bytes) gets 1 per inlier. */
int odom_icp_step(odom_handle h, int level, const float Rcurr[9], const float tcurr[3],
                  const float Rprev_inv[9], const float tprev[3], double out[29], uint8_t *mask);
/* computeDerivativeImages (utils.cu:608-638) on all levels of the next image. */
int odom_compute_derivatives(odom_handle h);
/* computeRgbResidual (reduce.cu:869-935): count and int32 sum of diff^2;
 * mask gets 1 per valid correspondence. */
int odom_rgb_residual(odom_handle h, int level, const float krkinv[9], const float kt[3], int32_t *count,
                      int32_t *sigma_sum, uint8_t *mask);
/* rgbStep (reduce.cu:626-678) with the correspondences of krkinv/kt */
int odom_rgb_step(odom_handle h, int level, float sigma, const float krkinv[9], const float kt[3],
                  double out[29]);
/* so3Step (reduce.cu:1082-1140) on level 2 of lastNextImage/nextImage */
int odom_so3_step(odom_handle h, const float H[9], const float kinv[9], const float krlr[9], double out[11],
                  uint8_t *mask);
/* Copy an internal buffer to host memory.  Names: depth (uint16), vmap_curr,
 * nmap_curr, vmap_g_prev, nmap_g_prev (3 planes float), last_depth,
 * next_depth (float), last_image, next_image, last_next_image (uint8), dxy
 * (int16 pairs). */
int odom_read_buffer(odom_handle h, const char *name, int level, void *dst, size_t bytes);

/* Launch shapes of the two Gauss-Newton step kernels (the role of the
 * reference's GPUConfig.h table, GPUConfig.h:7-149).  kernel 0 = ICP /
 * association pass, 1 = photometric pass; threads 256 or 512, 1 to 4 pixels
 * per thread per tile. */
int odom_set_step_shape(odom_handle h, int kernel, int level, int threads, int pixels_per_thread);
int odom_get_step_shape(odom_handle h, int kernel, int level, int *threads, int *pixels_per_thread);
/* One getIncrementalTransformation on the current inputs and pose with a
 * HIP event pair around every Gauss-Newton step launch on the handle's
 * stream: mean microseconds of pass A (ICP + photometric association) and
 * pass B (photometric products) per pyramid level, the number of launches
 * per level, and the whole call's duration. */
int odom_profile_track(odom_handle h, int rgb_only, float icp_weight, int pyramid, int fast_odom, int so3,
                       float us_a[3], float us_b[3], int n_a[3], int n_b[3], float *us_total);
/* Mean microseconds of the two step kernels per level as they run inside a
 * tracking call (ICP/association pass us_a, photometric pass us_b): after a
 * completed call with the same options, each level's last iteration is
 * relaunched reps times back to back, as one graph replay between one event
 * pair (so the host's launch cost leaves no gaps between the kernels). */
int odom_profile_steps(odom_handle h, int rgb_only, float icp_weight, int reps, float us_a[3], float us_b[3]);
/* Times the supported shapes on the handle's current inputs and keeps the
 * fastest per kernel and level. */
int odom_autotune(odom_handle h, int rgb_only, float icp_weight);

/* The loop-closure acceptance test of the reference's frame loop
 * (elastic_fusion_file.cpp:474-497) on the last tracking call of this
 * handle, run as model-to-model registration (odom_init_icp_model with the
 * inactive model's maps, odom_init_icp_maps with the active model's, so3 off):
 * ok = every diagonal entry of getCovariance() <= cov_thresh and
 * lastICPCount > icp_count_thresh and lastICPError < icp_err_thresh.  A
 * singular last system (no covariance) is not accepted.  Synchronous. */
int odom_loop_check(odom_handle h, double cov_thresh, float icp_count_thresh, float icp_err_thresh, int *ok);

/* ---- tracking policies: the start, the iteration schedule and the
 * convergence test of a call, and the judge of a learned start (no
 * counterpart in the reference, whose getIncrementalTransformation always
 * starts from the previous pose, optionally rotated by the SO3
 * pre-alignment, and runs a fixed 10/5/4 schedule, RGBDOdometryef.cpp:
 * 382-420).  With the default policy odom_track_policy(h, NULL, ...) is
 * getIncrementalTransformation(rgbOnly, icpWeight, pyramid, !fastOdom, so3). ---- */
#define ODOM_START_PREVIOUS 0  /* the previous pose (the reference's start) */
#define ODOM_START_VELOCITY 1  /* constant velocity: the previous pose moved by the last call's increment */
#define ODOM_JUDGE_NONE 0      /* a prediction is kept */
#define ODOM_JUDGE_COARSE 1    /* the prediction and the constant-velocity start are both refined through
                                  the level-2 iterations; the better one continues through levels 1 and 0 */
#define ODOM_JUDGE_FULL 2      /* both are refined through every level; the better one is the result */
typedef struct odom_policy {
    int iters[3];              /* Gauss-Newton iterations at levels 0, 1, 2 (at most 63 each) */
    float conv_t, conv_r;      /* a level stops after a solve whose update is below both (m, rad);
                                  conv_t = 0: every scheduled iteration runs */
    int start;                 /* ODOM_START_* when there is no prediction */
    int so3;                   /* the SO3 pre-alignment (only with ODOM_START_PREVIOUS and no prediction) */
    int judge;                 /* ODOM_JUDGE_*, for a prediction */
    float judge_tol;           /* the prediction falls back when its refinement's ICP residual
                                  (lastICPError) is more than judge_tol above, or its inlier count
                                  more than judge_tol below, the constant-velocity start's
                                  (photometric error and count for RGB-only calls) */
} odom_policy;
/* iters 10/5/4, no convergence test, start previous, so3 on, judge full, tol 0.01 */
int odom_default_policy(odom_policy *p);
/* getIncrementalTransformation under a policy, asynchronously (the prior is
 * the handle's device-resident pose, as odom_track_async).  pred: a device
 * prediction of the camera motion since the previous pose (6 floats, t /
 * t_scale and rotation vector / r_scale, slam_cnn.h) or NULL.  With a
 * prediction and a judge, the constant-velocity start is refined beside it
 * in the same launches (the handle's second state) and the judge keeps the
 * better; odom_get_stats reports whether the prediction was kept and the
 * iterations of both refinements. */
int odom_track_policy(odom_handle h, const float *pred, float t_scale, float r_scale, const odom_policy *pol,
                      int rgb_only, float icp_weight);

/* ---- batches: many frame pairs (handles, of any size) tracked by the same
 * launches.  No counterpart in the reference, whose RGBDOdometryef tracks one
 * pair per object and call (RGBDOdometryef.cpp:267-595); this is the batched
 * form of that call.  Every launch covers every pair: the ICP / association
 * and photometric passes of one Gauss-Newton iteration are one launch each
 * over all pairs, with blocks shared out in proportion to each pair's pixel
 * count.  The handles keep their own inputs (init* calls on each handle),
 * state and results (odom_get_pose / odom_get_stats per handle).  All
 * handles must share dist_thresh and angle_thresh. ---- */
typedef struct odom_batch_s *odom_batch;
int odom_batch_create(const odom_handle *handles, int n, odom_batch *out);
int odom_batch_destroy(odom_batch b);
/* getIncrementalTransformation on every pair of the batch, asynchronously on
 * the batch's stream; each handle's stream waits for it, so per-handle calls
 * made afterwards see the results. */
int odom_track_batch(odom_batch b, int rgb_only, float icp_weight, int pyramid, int fast_odom, int so3);
int odom_batch_synchronize(odom_batch b);
/* Blocks per step launch over the whole batch (default: 4 per compute unit),
 * shared out by pixel count, at most 256 per pair.  0: every pair gets the
 * block count of a single-pair call (its sums are then bit-identical to the
 * pair tracked alone). */
int odom_batch_set_blocks(odom_batch b, int target_blocks);
int odom_batch_get_blocks(odom_batch b, int kernel, int level, int *total_blocks);
int odom_batch_set_step_shape(odom_batch b, int kernel, int level, int threads, int pixels_per_thread);
int odom_batch_get_step_shape(odom_batch b, int kernel, int level, int *threads, int *pixels_per_thread);
/* As odom_autotune for the batch's launches; needs a completed batched call
 * on the current inputs. */
int odom_batch_autotune(odom_batch b, int rgb_only, float icp_weight);
/* 1 (default): each iteration's solve runs once per pair in a kernel of its
 * own before the ICP pass; 0: every block of the ICP pass repeats it in its
 * prologue (the single-pair form). */
int odom_batch_set_split_solve(odom_batch b, int enable);
/* Capture each distinct call (options and image parities) as a hipGraph on
 * first use and replay it afterwards. */
int odom_batch_set_graph(odom_batch b, int enable);
/* As odom_profile_track, for one batched call. */
int odom_batch_profile_track(odom_batch b, int rgb_only, float icp_weight, int pyramid, int fast_odom, int so3,
                             float us_a[3], float us_b[3], int n_a[3], int n_b[3], float *us_total);
/* The batch's HIP stream (hipStream_t): a caller enqueues the warm-start
 * network (slam_cnn.h) there between odom_batch_warmstart_input and
 * odom_track_batch_policy. */
int odom_batch_get_stream(odom_batch b, void **stream);
/* The warm-start network input of every pair (640x480 handles only): bf16
 * [n][60][80][4] at device pointer x, sample i from handle i, on the batch's
 * stream after each handle's preceding work. */
int odom_batch_warmstart_input(odom_batch b, void *x);
/* odom_track_policy on every pair of the batch; pred: 6 floats per handle
 * (device) or NULL. */
int odom_track_batch_policy(odom_batch b, const float *pred, float t_scale, float r_scale, const odom_policy *pol,
                            int rgb_only, float icp_weight);

/* ---- learned warm start (no counterpart in the reference, whose tracker
 * always starts from the previous pose, RGBDOdometryef.cpp:267-401) ---- */
/* The handle's HIP stream (hipStream_t), so a caller can order its own work
 * (e.g. the warm-start network, slam_cnn.h) with the handle's. */
int odom_get_stream(odom_handle h, void **stream);
/* Enqueue the warm-start network input (bf16 [1][60][80][4], device pointer
 * x) from the handle's level-2 model depth and intensity and current depth
 * and intensity (640x480 frames only). */
int odom_init_warmstart_input(odom_handle h, void *x);
/* getIncrementalTransformation from a predicted start: pred (device, 6
 * floats) is the camera motion since the previous pose as (t / t_scale,
 * rotation vector / r_scale); the solve starts at that motion with the
 * schedule iters[3] (iterations at levels 0, 1, 2) and no SO3 pass.  With
 * select, the prediction is kept only if the level-2 ICP association from it
 * is not worse (inliers, mean squared residual) than from the previous pose;
 * otherwise the call starts from the previous pose.  Asynchronous. */
int odom_track_seeded(odom_handle h, const float *pred, float t_scale, float r_scale, const int iters[3],
                      int rgb_only, float icp_weight, int select);

/* ---- an entry point that allocates nothing.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under hipGraph stream capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It is odom_icp_step on caller-provided maps: the pass's
 * descriptor, state record, block partials and FP64 totals live in one
 * file-scope device array of ODOM_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends, and the call makes no hipMalloc/hipFree.  Every call on
 * a GPU shares that GPU's array (its contents are rewritten by each call
 * before they are read): calls that overlap in time on one GPU race. ---- */
#define ODOM_NOALLOC_SCRATCH_BYTES 55296
#define ODOM_NOALLOC_MAX_DIM 8192
/* icpStep (reduce.cu:418-490, the JtJJtrSE3 sums of types.cuh:85-96) over
 * device maps in the reference's layout (createVMap/createNMap, utils.cu:
 * 109-204): planar float [3][height][width] (x, y, z planes, row pitch =
 * width), NaN x = invalid; vcur/ncur in the current camera frame, vprev/nprev
 * in the global frame.  fx..cy: the intrinsics of the maps' level.  out[29]
 * as odom_icp_step.  width and height at most ODOM_NOALLOC_MAX_DIM
 * (ODOM_ERR_ARG beyond).  Work runs on `stream` (a hipStream_t, 0 = the null
 * stream); the call returns when out is written. */
int odom_icp_step_noalloc(const float *vcur, const float *ncur, const float *vprev, const float *nprev,
                          int width, int height, float fx, float fy, float cx, float cy,
                          const float Rcurr[9], const float tcurr[3], const float Rprev_inv[9],
                          const float tprev[3], float dist_thresh, float angle_thresh, double out[29],
                          void *stream);

/* The measured HBM copy peak the benchmark reports beside the 8 TB/s
 * specification (no counterpart in the reference): `reps` device copies of
 * `bytes` (a multiple of 16, 16-byte aligned pointers) from src to dst on
 * `stream` after one untimed copy; ms_per_copy gets the mean duration (HIP
 * events).  Bandwidth = 2 * bytes / duration. */
int odom_stream_copy(const void *src, void *dst, size_t bytes, int reps, float *ms_per_copy, void *stream);

/* Device name / arch of the current device, for logs. */
int odom_device_info(char *name, int name_len, char *arch, int arch_len);

#ifdef __cplusplus
}
#endif
#endif
