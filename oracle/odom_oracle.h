/*
 * odom_oracle.h -- CPU oracle of the dense RGB-D odometry path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path in slam_amd/csrc.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never calls it.
 *
 * It restates, in plain C, the algorithm of the reference's src/odom path
 * (siw-engineering/slam, an ElasticFusion fork).  Every function cites the
 * reference file:line it follows.  Per-pixel arithmetic is IEEE binary32 in
 * the order the reference writes it, with no FMA contraction
 * (-ffp-contract=off), correctly rounded division and square root, and
 * round-half-even float->int conversion that saturates and maps NaN to 0
 * (the semantics of __float2int_rn).  Sums of per-pixel products are taken in
 * FP64; the 6x6 / 3x3 solves and the SE(3) update run in FP64.
 *
 * The *_f64 step variants run the association and products in binary64 as
 * well ("all-FP64 run") and are used to measure the error of the mixed
 * precision device path.
 *
 * Layouts match the reference's DeviceArray2D usage:
 *   vertex / normal maps: planar, 3 planes of w*h floats (x, then y, then z);
 *   a pixel is invalid iff its x value is NaN (utils.cu:130).
 *   Model maps (RGBA32F textures) are interleaved float4 per pixel.
 */
#ifndef ODOM_ORACLE_H
#define ODOM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float fx, fy, cx, cy; } orc_intr;

/* CameraModel::operator()(level), Camera.h:14-18: f and c divided by 2^level. */
orc_intr orc_level_intr(orc_intr k, int level);

/* Deterministic exp for x <= 0 shared (as an algorithm) with the device path:
 * Cody-Waite reduction by ln2 and a degree-7 Taylor polynomial.  Returns 0
 * for x < -87. */
float orc_bexpf(float x);
/* threads of the oracle's OpenMP row loops; the host's online processors */
void orc_set_threads(int n);
int orc_get_threads(void);
int orc_hardware_concurrency(void);

/* a1: depth_bilateral.frag:30-76 */
void orc_bilateral_u16(const uint16_t *src, uint16_t *dst, int w, int h, float maxD);
/* a2: utils.cu:57-107 (pyrDownGaussKernel, sigma_color 30) */
void orc_pyrdown_u16(const uint16_t *src, int w, int h, uint16_t *dst);
/* a3: utils.cu:109-149 */
void orc_vmap(const uint16_t *depth, int w, int h, orc_intr k, float cutoff, float *vmap);
/* a4: utils.cu:151-204 */
void orc_nmap(const float *vmap, int w, int h, float *nmap);
/* a5: utils.cu:270-330 */
void orc_copy_maps(const float *v4, const float *n4, int w, int h, float *vmap, float *nmap);
/* a6: utils.cu:365-444; dst is (w/2)x(h/2) */
void orc_resize_map(const float *src, int w, int h, float *dst, int normalize);
/* a7: utils.cu:206-268, in place */
void orc_transform_maps(float *vmap, float *nmap, int w, int h, const float R[9], const float t[3]);
/* a8: utils.cu:526-546 */
void orc_vertices_to_depth(const float *v4, int w, int h, float cutoff, float *depth);
/* a9: utils.cu:332-363, 446-468 */
void orc_pyrdown_f32(const float *src, int w, int h, float *dst);
/* a10: utils.cu:550-577 (weights 0.114/0.299/0.587 on bytes 0/1/2) */
void orc_intensity(const uint8_t *rgb, int w, int h, int channels, uint8_t *dst);
/* a11: utils.cu:470-524 */
void orc_pyrdown_u8(const uint8_t *src, int w, int h, uint8_t *dst);
/* a12: utils.cu:579-638 */
void orc_derivative(const uint8_t *src, int w, int h, int16_t *dx, int16_t *dy);
/* a13: utils.cu:640-673; cloud is 3 planes */
void orc_point_cloud(const float *depth, int w, int h, orc_intr k, float *cloud);

/* reduce.cu:75-81 */
typedef struct {
    int16_t zero_x, zero_y, one_x, one_y;
    float diff;
    int32_t valid;
} orc_dataterm;

/* a14: reduce.cu:257-490.  out[0..26] = upper-triangular products of the
 * row [J | r] in JtJJtrSE3 order (types.cuh:85-93), out[27] = sum r^2,
 * out[28] = inlier count.  mask (nullable) gets 1 per inlier pixel. */
void orc_icp_step(const float *vmap_curr, const float *nmap_curr,
                  const float *vmap_g_prev, const float *nmap_g_prev,
                  int w, int h, orc_intr k,
                  const float Rcurr[9], const float tcurr[3],
                  const float Rprev_inv[9], const float tprev[3],
                  float distThres, float angleThres, double out[29], uint8_t *mask);
void orc_icp_step_f64(const float *vmap_curr, const float *nmap_curr,
                      const float *vmap_g_prev, const float *nmap_g_prev,
                      int w, int h, orc_intr k,
                      const double Rcurr[9], const double tcurr[3],
                      const double Rprev_inv[9], const double tprev[3],
                      double distThres, double angleThres, double out[29], uint8_t *mask);

/* a15: reduce.cu:739-935.  corres (nullable) gets one DataTerm per pixel;
 * count and the int32 (wrapping) sum of diff^2 are returned. */
void orc_rgb_residual(float minScale, const int16_t *dIdx, const int16_t *dIdy,
                      const float *lastDepth, const float *nextDepth,
                      const uint8_t *lastImage, const uint8_t *nextImage,
                      int w, int h, float maxDepthDelta,
                      const float krkinv[9], const float kt[3],
                      orc_dataterm *corres, int32_t *count, int32_t *sigma);
void orc_rgb_residual_f64(float minScale, const int16_t *dIdx, const int16_t *dIdy,
                          const float *lastDepth, const float *nextDepth,
                          const uint8_t *lastImage, const uint8_t *nextImage,
                          int w, int h, double maxDepthDelta,
                          const double krkinv[9], const double kt[3],
                          orc_dataterm *corres, int32_t *count, int32_t *sigma);

/* a16: reduce.cu:494-678; the cloud is recomputed from lastDepth (a13). */
void orc_rgb_step(const orc_dataterm *corres, int w, int h, float sigma,
                  const float *lastDepth, orc_intr k,
                  const int16_t *dIdx, const int16_t *dIdy, float sobelScale, double out[29]);
void orc_rgb_step_f64(const orc_dataterm *corres, int w, int h, double sigma,
                      const float *lastDepth, orc_intr k,
                      const int16_t *dIdx, const int16_t *dIdy, double sobelScale, double out[29]);

/* a17: reduce.cu:937-1140.  out[0..8] JtJJtrSO3 order, out[9] sum r^2,
 * out[10] count. */
void orc_so3_step(const uint8_t *lastImage, const uint8_t *nextImage, int w, int h,
                  const float H[9], const float kinv[9], const float krlr[9],
                  double out[11], uint8_t *mask);
void orc_so3_step_f64(const uint8_t *lastImage, const uint8_t *nextImage, int w, int h,
                      const double H[9], const double kinv[9], const double krlr[9],
                      double out[11], uint8_t *mask);

/* Unpack of the 29 / 11 packed sums into a symmetric A and b
 * (reduce.cu:475-489, 1125-1139). */
void orc_unpack_se3(const double in[29], double A[36], double b[6]);
void orc_unpack_so3(const double in[11], double A[9], double b[3]);

/* Symmetric solve with diagonal pivoting and a pseudo-inverse of D, the
 * semantics of Eigen's LDLT::solve as called at RGBDOdometryef.cpp:366,544.
 * Returns the number of pivots treated as zero. */
int orc_ldlt_solve(int n, const double *A, const double *b, double *x);
/* utils.h:16-52 */
void orc_rodrigues(const double r[3], double R[9]);

/* The tracker (RGBDOdometryef.{h,cpp}).  Buffers and cross-call state as in
 * the reference (ctor :21-111). */
typedef struct orc_odom orc_odom;

orc_odom *orc_create(int w, int h, float cx, float cy, float fx, float fy,
                     float distThresh, float angleThresh);
void orc_destroy(orc_odom *o);
/* precision: 0 = FP32 per pixel (mirror of the device path), 1 = all FP64 */
void orc_set_precision(orc_odom *o, int fp64);
void orc_init_icp_depth(orc_odom *o, const uint16_t *depth, float depthCutoff);
void orc_init_icp_maps(orc_odom *o, const float *v4, const float *n4, float depthCutoff);
void orc_init_icp_model(orc_odom *o, const float *v4, const float *n4, float depthCutoff,
                        const float pose[16]);
void orc_init_rgb(orc_odom *o, const uint8_t *rgb, int channels);
void orc_init_rgb_model(orc_odom *o, const uint8_t *rgb, int channels);
void orc_init_first_rgb(orc_odom *o, const uint8_t *rgb, int channels);
/* RGBDOdometryef.cpp:267-595.  t and R (row-major) in/out. */
void orc_track(orc_odom *o, float t[3], float R[9], int rgbOnly, float icpWeight,
               int pyramid, int fastOdom, int so3);
/* The same with an explicit schedule (iterations per level, 0..2) and an
 * optional seeded increment `seed` (row-major 4x4, the resultRt convention:
 * current = prev * seed^-1); NULL for none.  No counterpart in the reference:
 * the warm-start form. */
void orc_track_ex(orc_odom *o, float t[3], float R[9], int rgbOnly, float icpWeight, const int iters[3],
                  int so3, const double *seed);
/* orc_track_ex over levels hi..lo (2..0 is the whole call) with the
 * convergence policy of the library's odom_track_policy (no counterpart in
 * the reference, whose schedule is fixed): a level ends after a solve whose
 * update has |dt| < convT and |dw| < convR (convT 0: off).  resultRt_out
 * (optional): the increment at the end (identity after the 0.3 m revert);
 * iters_done (optional): solves run.  With lo > 0 the run stops after level
 * lo without the revert or the image swap, t/R = the current estimate, and
 * a later run continues from resultRt_out as its seed. */
void orc_track_run(orc_odom *o, float t[3], float R[9], int rgbOnly, float icpWeight, const int iters[3],
                   int so3, const double *seed, float convT, float convR, int hi, int lo, double resultRt_out[16],
                   int *iters_done);
/* stats[0..5] = lastICPError, lastICPCount, lastRGBError, lastRGBCount,
 * lastSO3Error, lastSO3Count; stats[6] = number of SO3 iterations run. */
void orc_get_stats(const orc_odom *o, float stats[7]);
void orc_get_last_system(const orc_odom *o, double A[36], double b[6]);
/* rotation found by the SO3 pre-alignment of the last track call (:294-380) */
void orc_get_so3_rotation(const orc_odom *o, double R[9]);
/* RGBDOdometryef.cpp:597-600: lastA^-1 (Gauss-Jordan with partial pivoting) */
int orc_get_covariance(const orc_odom *o, double cov[36]);
/* access to internal pyramids for tests; level 0..2 */
const float *orc_buf_f32(const orc_odom *o, const char *name, int level);
const uint8_t *orc_buf_u8(const orc_odom *o, const char *name, int level);
const int16_t *orc_buf_s16(const orc_odom *o, const char *name, int level);
const uint16_t *orc_buf_u16(const orc_odom *o, const char *name, int level);

#ifdef __cplusplus
}
#endif
#endif
