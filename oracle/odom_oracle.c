/*
 * odom_oracle.c -- CPU oracle of the dense RGB-D odometry path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * A restatement in C of the reference (siw-engineering/slam) files
 *   src/odom/RGBDOdometryef.{h,cpp}, src/odom/utils.h, src/odom/utils.cu,
 *   src/odom/reduce.cu, src/cuda/types.cuh, src/cuda/operators.cuh,
 *   src/sensors/Camera.h, src/gl/shaders/depth_bilateral.frag.
 * Each function cites the lines it follows.  Build with -ffp-contract=off
 * (oracle/Makefile): the per-pixel expressions are evaluated in the
 * reference's written order with one rounding per operation.
 *
 * Choices where the reference is not IEEE-defined (its nvcc build used
 * --ftz --prec-div=false --prec-sqrt=false and FMA contraction, and the GL
 * shader's exp has driver-defined precision): correctly rounded / and sqrt,
 * 1/sqrt in place of rsqrtf, no contraction, orc_bexpf() in place of GLSL
 * exp, GLSL round() as round-half-away-from-zero, NaN->uchar as 0.
 */
#include "odom_oracle.h"

#include <float.h>
#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <unistd.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Threads of the OpenMP row loops (the CPU baseline's 1-core and all-core
 * runs), and the host's online processors. */
void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

int orc_hardware_concurrency(void)
{
    return (int)sysconf(_SC_NPROCESSORS_ONLN);
}

static const float ORC_NAN = NAN;

orc_intr orc_level_intr(orc_intr k, int level)
{
    /* Camera.h:14-18: int div = 1 << level; f / div (float / int) */
    const int div = 1 << level;
    orc_intr o = {k.fx / (float)div, k.fy / (float)div, k.cx / (float)div, k.cy / (float)div};
    return o;
}

/* __float2int_rn semantics: round half to even, saturate, NaN -> 0.  The
 * saturation bound only has to exceed any image dimension. */
static inline int f2i_rn_f(float f)
{
    if (!(f == f)) return 0;
    float r = rintf(f);
    if (r > 1.0e9f) r = 1.0e9f;
    if (r < -1.0e9f) r = -1.0e9f;
    return (int)r;
}

static inline int f2i_rn_d(double f)
{
    if (!(f == f)) return 0;
    double r = rint(f);
    if (r > 1.0e9) r = 1.0e9;
    if (r < -1.0e9) r = -1.0e9;
    return (int)r;
}

float orc_bexpf(float x)
{
    if (!(x >= -87.0f)) return 0.0f;
    if (x > 0.0f) x = 0.0f;
    const float kf = rintf(x * 1.44269504f);
    float r = x - kf * 0.693145751953125f; /* ln2 high part, 9 trailing zero bits */
    r = r - kf * 1.42860677e-06f;           /* ln2 low part */
    float p = 1.98412698e-04f;              /* 1/5040 */
    p = p * r + 1.38888889e-03f;            /* 1/720 */
    p = p * r + 8.33333333e-03f;            /* 1/120 */
    p = p * r + 4.16666667e-02f;            /* 1/24 */
    p = p * r + 1.66666667e-01f;            /* 1/6 */
    p = p * r + 0.5f;
    p = p * r + 1.0f;
    p = p * r + 1.0f;
    const int kk = (int)kf; /* in [-126, 0] */
    union { uint32_t u; float f; } sc;
    sc.u = (uint32_t)(kk + 127) << 23;
    return p * sc.f;
}

/* ---------------------------------------------------------------- a1 */
/* depth_bilateral.frag:30-76.  The centre texel is sampled at (x, y)
 * exactly; neighbours are not validity-tested.  The weight
 * exp(-(space2*0.024691358 + color2*0.000555556)) (:67) is evaluated as the
 * product of its two factors, each from orc_bexpf: a 13x13 spatial table and
 * a range table over |value - tmp|.  Range factors for |value - tmp| >= 380
 * (below 1.6e-35) are taken as 0: they cannot change the binary32 sums,
 * whose centre term is 1. */
#define BILAT_RANGE_TAPS 380

void orc_bilateral_u16(const uint16_t *src, uint16_t *dst, int w, int h, float maxD)
{
    const uint32_t maxv = (uint32_t)(maxD * 1000.0f);
    float ES[169], EC[BILAT_RANGE_TAPS + 1];
    for (int dy = -6; dy <= 6; ++dy)
        for (int dx = -6; dx <= 6; ++dx)
            ES[(dy + 6) * 13 + dx + 6] = orc_bexpf(-((float)(dx * dx + dy * dy) * 0.024691358f));
    for (int a = 0; a < BILAT_RANGE_TAPS; ++a) EC[a] = orc_bexpf(-((float)(a * a) * 0.000555556f));
    EC[BILAT_RANGE_TAPS] = 0.0f;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const uint32_t value = src[(size_t)y * w + x];
            if (value > maxv || value < 300u) {
                dst[(size_t)y * w + x] = 0;
                continue;
            }
            const int tx = (x - 6 + 13 < w) ? x - 6 + 13 : w;
            const int ty = (y - 6 + 13 < h) ? y - 6 + 13 : h;
            float sum1 = 0.0f, sum2 = 0.0f;
            for (int cy = (y - 6 > 0 ? y - 6 : 0); cy < ty; ++cy) {
                for (int cx = (x - 6 > 0 ? x - 6 : 0); cx < tx; ++cx) {
                    const uint32_t tmp = src[(size_t)cy * w + cx];
                    int a = (int)value - (int)tmp;
                    a = a < 0 ? -a : a;
                    if (a > BILAT_RANGE_TAPS) a = BILAT_RANGE_TAPS;
                    const float weight = ES[(cy - y + 6) * 13 + (cx - x + 6)] * EC[a];
                    sum1 += (float)tmp * weight;
                    sum2 += weight;
                }
            }
            dst[(size_t)y * w + x] = (uint16_t)(uint32_t)roundf(sum1 / sum2);
        }
    }
}

/* ---------------------------------------------------------------- a2 */
/* utils.cu:57-94 pyrDownGaussKernel with sigma_color = 30 (utils.cu:103) */
void orc_pyrdown_u16(const uint16_t *src, int w, int h, uint16_t *dst)
{
    const int dw = w / 2, dh = h / 2;
    static const float wts[3] = {0.375f, 0.25f, 0.0625f};
#pragma omp parallel for schedule(static)
    for (int y = 0; y < dh; ++y) {
        for (int x = 0; x < dw; ++x) {
            const int center = src[(size_t)(2 * y) * w + 2 * x];
            const int x_mi = (2 * x - 2 > 0 ? 2 * x - 2 : 0) - 2 * x;
            const int y_mi = (2 * y - 2 > 0 ? 2 * y - 2 : 0) - 2 * y;
            const int x_ma = (w < 2 * x + 3 ? w : 2 * x + 3) - 2 * x;
            const int y_ma = (h < 2 * y + 3 ? h : 2 * y + 3) - 2 * y;
            float sum = 0.0f, wall = 0.0f;
            for (int yi = y_mi; yi < y_ma; ++yi)
                for (int xi = x_mi; xi < x_ma; ++xi) {
                    const int val = src[(size_t)(2 * y + yi) * w + 2 * x + xi];
                    if ((float)abs(val - center) < 3.0f * 30.0f) {
                        sum += (float)val * wts[abs(xi)] * wts[abs(yi)];
                        wall += wts[abs(xi)] * wts[abs(yi)];
                    }
                }
            dst[(size_t)y * dw + x] = (uint16_t)(int)(sum / wall);
        }
    }
}

/* ---------------------------------------------------------------- a3 */
/* utils.cu:109-149; fx_inv = 1.f / fx on the host (utils.cu:147).  Invalid
 * pixels get NaN in all three planes (the reference writes the x plane only
 * and every consumer tests x). */
void orc_vmap(const uint16_t *depth, int w, int h, orc_intr k, float cutoff, float *vmap)
{
    const size_t plane = (size_t)w * h;
    const float fx_inv = 1.f / k.fx, fy_inv = 1.f / k.fy;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < h; ++v)
        for (int u = 0; u < w; ++u) {
            const size_t i = (size_t)v * w + u;
            const float z = (float)depth[i] / 1000.f;
            if (z != 0 && z < cutoff) {
                vmap[i] = z * ((float)u - k.cx) * fx_inv;
                vmap[i + plane] = z * ((float)v - k.cy) * fy_inv;
                vmap[i + 2 * plane] = z;
            } else {
                vmap[i] = vmap[i + plane] = vmap[i + 2 * plane] = ORC_NAN;
            }
        }
}

static inline void normalized3(float v[3])
{
    /* operators.cuh:81-85 with 1/sqrtf in place of rsqrtf */
    const float rn = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] = v[0] * rn; v[1] = v[1] * rn; v[2] = v[2] * rn;
}

/* ---------------------------------------------------------------- a4 */
/* utils.cu:151-188 */
void orc_nmap(const float *vmap, int w, int h, float *nmap)
{
    const size_t plane = (size_t)w * h;
#pragma omp parallel for schedule(static)
    for (int v = 0; v < h; ++v)
        for (int u = 0; u < w; ++u) {
            const size_t i = (size_t)v * w + u;
            if (u == w - 1 || v == h - 1) {
                nmap[i] = nmap[i + plane] = nmap[i + 2 * plane] = ORC_NAN;
                continue;
            }
            const size_t i01 = i + 1, i10 = i + w;
            if (!isnan(vmap[i]) && !isnan(vmap[i01]) && !isnan(vmap[i10])) {
                const float a[3] = {vmap[i01] - vmap[i], vmap[i01 + plane] - vmap[i + plane],
                                    vmap[i01 + 2 * plane] - vmap[i + 2 * plane]};
                const float b[3] = {vmap[i10] - vmap[i], vmap[i10 + plane] - vmap[i + plane],
                                    vmap[i10 + 2 * plane] - vmap[i + 2 * plane]};
                float r[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                              a[0] * b[1] - a[1] * b[0]};
                normalized3(r);
                nmap[i] = r[0]; nmap[i + plane] = r[1]; nmap[i + 2 * plane] = r[2];
            } else {
                nmap[i] = nmap[i + plane] = nmap[i + 2 * plane] = ORC_NAN;
            }
        }
}

/* ---------------------------------------------------------------- a5 */
/* utils.cu:270-310: RGBA32F -> planar; invalid iff vertex z == 0 */
void orc_copy_maps(const float *v4, const float *n4, int w, int h, float *vmap, float *nmap)
{
    const size_t plane = (size_t)w * h;
    for (size_t i = 0; i < plane; ++i) {
        const int ok = !(v4[4 * i + 2] == 0);
        for (int c = 0; c < 3; ++c) {
            vmap[i + c * plane] = ok ? v4[4 * i + c] : ORC_NAN;
            nmap[i + c * plane] = ok ? n4[4 * i + c] : ORC_NAN;
        }
    }
}

/* ---------------------------------------------------------------- a6 */
/* utils.cu:365-416 resizeMapKernel<normalize> */
void orc_resize_map(const float *src, int w, int h, float *dst, int normalize)
{
    const int dw = w / 2, dh = h / 2;
    const size_t sp = (size_t)w * h, dp = (size_t)dw * dh;
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            const size_t o = (size_t)y * dw + x;
            const size_t a = (size_t)(2 * y) * w + 2 * x, b = a + 1, c = a + w, d = c + 1;
            if (isnan(src[a]) || isnan(src[b]) || isnan(src[c]) || isnan(src[d])) {
                dst[o] = dst[o + dp] = dst[o + 2 * dp] = ORC_NAN;
                continue;
            }
            float n[3];
            for (int p = 0; p < 3; ++p) {
                const size_t off = p * sp;
                n[p] = (src[a + off] + src[b + off] + src[c + off] + src[d + off]) / 4.0f;
            }
            if (normalize) normalized3(n);
            dst[o] = n[0]; dst[o + dp] = n[1]; dst[o + 2 * dp] = n[2];
        }
}

/* ---------------------------------------------------------------- a7 */
/* utils.cu:206-248: v' = R v + t, n' = R n, NaN passes through */
void orc_transform_maps(float *vmap, float *nmap, int w, int h, const float R[9], const float t[3])
{
    const size_t plane = (size_t)w * h;
    for (size_t i = 0; i < plane; ++i) {
        if (!isnan(vmap[i])) {
            const float v[3] = {vmap[i], vmap[i + plane], vmap[i + 2 * plane]};
            for (int r = 0; r < 3; ++r)
                vmap[i + r * plane] = R[3 * r] * v[0] + R[3 * r + 1] * v[1] + R[3 * r + 2] * v[2] + t[r];
        } else {
            vmap[i] = vmap[i + plane] = vmap[i + 2 * plane] = ORC_NAN;
        }
        if (!isnan(nmap[i])) {
            const float n[3] = {nmap[i], nmap[i + plane], nmap[i + 2 * plane]};
            for (int r = 0; r < 3; ++r)
                nmap[i + r * plane] = R[3 * r] * n[0] + R[3 * r + 1] * n[1] + R[3 * r + 2] * n[2];
        } else {
            nmap[i] = nmap[i + plane] = nmap[i + 2 * plane] = ORC_NAN;
        }
    }
}

/* ---------------------------------------------------------------- a8 */
/* utils.cu:526-537 */
void orc_vertices_to_depth(const float *v4, int w, int h, float cutoff, float *depth)
{
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        const float z = v4[4 * i + 2];
        depth[i] = (z > cutoff || z <= 0) ? ORC_NAN : z;
    }
}

static const float GAUSS25[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6,
                                  4, 16, 24, 16, 4, 1, 4, 6, 4, 1};

/* ---------------------------------------------------------------- a9 */
/* utils.cu:332-363: loop bound min(2x+3, cols-1) and the reversed weight
 * index (ty-cy-1)*5+(tx-cx-1) are kept. */
void orc_pyrdown_f32(const float *src, int w, int h, float *dst)
{
    const int dw = w / 2, dh = h / 2;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            const int tx = (2 * x + 3 < w - 1) ? 2 * x + 3 : w - 1;
            const int ty = (2 * y + 3 < h - 1) ? 2 * y + 3 : h - 1;
            float sum = 0.0f;
            int count = 0;
            for (int cy = (2 * y - 2 > 0 ? 2 * y - 2 : 0); cy < ty; ++cy)
                for (int cx = (2 * x - 2 > 0 ? 2 * x - 2 : 0); cx < tx; ++cx) {
                    const float v = src[(size_t)cy * w + cx];
                    if (!isnan(v)) {
                        const float g = GAUSS25[(ty - cy - 1) * 5 + (tx - cx - 1)];
                        sum += v * g;
                        count = (int)((float)count + g);
                    }
                }
            dst[(size_t)y * dw + x] = sum / (float)count;
        }
}

/* ---------------------------------------------------------------- a10 */
/* utils.cu:550-563 */
void orc_intensity(const uint8_t *rgb, int w, int h, int channels, uint8_t *dst)
{
    const size_t n = (size_t)w * h;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t *p = rgb + i * (size_t)channels;
        const int value = (int)((float)p[0] * 0.114f + (float)p[1] * 0.299f + (float)p[2] * 0.587f);
        dst[i] = (uint8_t)value;
    }
}

/* ---------------------------------------------------------------- a11 */
/* utils.cu:470-500; count == 0 gives 0 (NaN -> uchar) */
void orc_pyrdown_u8(const uint8_t *src, int w, int h, uint8_t *dst)
{
    const int dw = w / 2, dh = h / 2;
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            const int tx = (2 * x + 3 < w - 1) ? 2 * x + 3 : w - 1;
            const int ty = (2 * y + 3 < h - 1) ? 2 * y + 3 : h - 1;
            float sum = 0.0f;
            int count = 0;
            for (int cy = (2 * y - 2 > 0 ? 2 * y - 2 : 0); cy < ty; ++cy)
                for (int cx = (2 * x - 2 > 0 ? 2 * x - 2 : 0); cx < tx; ++cx) {
                    const uint8_t v = src[(size_t)cy * w + cx];
                    if (v > 0) {
                        const float g = GAUSS25[(ty - cy - 1) * 5 + (tx - cx - 1)];
                        sum += (float)v * g;
                        count = (int)((float)count + g);
                    }
                }
            const float q = sum / (float)count;
            dst[(size_t)y * dw + x] = (count == 0) ? 0 : (uint8_t)q;
        }
}

/* ---------------------------------------------------------------- a12 */
/* utils.cu:582-606 with the kernels of utils.cu:614-620 */
static const float GSX[9] = {0.52201f, 0.00000f, -0.52201f, 0.79451f, -0.00000f, -0.79451f,
                             0.52201f, 0.00000f, -0.52201f};
static const float GSY[9] = {0.52201f, 0.79451f, 0.52201f, 0.00000f, 0.00000f, 0.00000f,
                             -0.52201f, -0.79451f, -0.52201f};

void orc_derivative(const uint8_t *src, int w, int h, int16_t *dx, int16_t *dy)
{
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            float dxv = 0.0f, dyv = 0.0f;
            int ki = 8;
            for (int j = (y - 1 > 0 ? y - 1 : 0); j <= (y + 1 < h - 1 ? y + 1 : h - 1); j++)
                for (int i = (x - 1 > 0 ? x - 1 : 0); i <= (x + 1 < w - 1 ? x + 1 : w - 1); i++) {
                    const float s = (float)src[(size_t)j * w + i];
                    dxv += s * GSX[ki];
                    dyv += s * GSY[ki];
                    --ki;
                }
            dx[(size_t)y * w + x] = (int16_t)dxv;
            dy[(size_t)y * w + x] = (int16_t)dyv;
        }
}

/* ---------------------------------------------------------------- a13 */
/* utils.cu:640-658 */
void orc_point_cloud(const float *depth, int w, int h, orc_intr k, float *cloud)
{
    const size_t plane = (size_t)w * h;
    const float invFx = 1.0f / k.fx, invFy = 1.0f / k.fy;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x) {
            const size_t i = (size_t)y * w + x;
            const float z = depth[i];
            cloud[i] = ((float)x - k.cx) * z * invFx;
            cloud[i + plane] = ((float)y - k.cy) * z * invFy;
            cloud[i + 2 * plane] = z;
        }
}

/* ------------------------------------------------- steps (a14 - a17) */
#define REAL float
#define FN(name) name##_f32
#define F2I f2i_rn_f
#define SQRT sqrtf
#define FABS fabsf
#include "odom_oracle_steps.inc"
#undef REAL
#undef FN
#undef F2I
#undef SQRT
#undef FABS

#define REAL double
#define FN(name) name##_f64
#define F2I f2i_rn_d
#define SQRT sqrt
#define FABS fabs
#include "odom_oracle_steps.inc"
#undef REAL
#undef FN
#undef F2I
#undef SQRT
#undef FABS

void orc_icp_step(const float *vc, const float *nc, const float *vp, const float *np, int w, int h,
                  orc_intr k, const float Rc[9], const float tc[3], const float Rpi[9],
                  const float tp[3], float distThres, float angleThres, double out[29], uint8_t *mask)
{
    icp_step_f32(vc, nc, vp, np, w, h, k, Rc, tc, Rpi, tp, distThres, angleThres, out, mask);
}

void orc_icp_step_f64(const float *vc, const float *nc, const float *vp, const float *np, int w,
                      int h, orc_intr k, const double Rc[9], const double tc[3],
                      const double Rpi[9], const double tp[3], double distThres,
                      double angleThres, double out[29], uint8_t *mask)
{
    icp_step_f64(vc, nc, vp, np, w, h, k, Rc, tc, Rpi, tp, distThres, angleThres, out, mask);
}

void orc_rgb_residual(float minScale, const int16_t *dIdx, const int16_t *dIdy,
                      const float *lastDepth, const float *nextDepth, const uint8_t *lastImage,
                      const uint8_t *nextImage, int w, int h, float maxDepthDelta,
                      const float krkinv[9], const float kt[3], orc_dataterm *corres,
                      int32_t *count, int32_t *sigma)
{
    rgb_residual_f32(minScale, dIdx, dIdy, lastDepth, nextDepth, lastImage, nextImage, w, h,
                     maxDepthDelta, krkinv, kt, corres, count, sigma);
}

void orc_rgb_residual_f64(float minScale, const int16_t *dIdx, const int16_t *dIdy,
                          const float *lastDepth, const float *nextDepth,
                          const uint8_t *lastImage, const uint8_t *nextImage, int w, int h,
                          double maxDepthDelta, const double krkinv[9], const double kt[3],
                          orc_dataterm *corres, int32_t *count, int32_t *sigma)
{
    rgb_residual_f64(minScale, dIdx, dIdy, lastDepth, nextDepth, lastImage, nextImage, w, h,
                     maxDepthDelta, krkinv, kt, corres, count, sigma);
}

void orc_rgb_step(const orc_dataterm *corres, int w, int h, float sigma, const float *lastDepth,
                  orc_intr k, const int16_t *dIdx, const int16_t *dIdy, float sobelScale,
                  double out[29])
{
    rgb_step_f32(corres, w, h, sigma, lastDepth, k, dIdx, dIdy, sobelScale, out);
}

void orc_rgb_step_f64(const orc_dataterm *corres, int w, int h, double sigma,
                      const float *lastDepth, orc_intr k, const int16_t *dIdx,
                      const int16_t *dIdy, double sobelScale, double out[29])
{
    rgb_step_f64(corres, w, h, sigma, lastDepth, k, dIdx, dIdy, sobelScale, out);
}

void orc_so3_step(const uint8_t *last, const uint8_t *next, int w, int h, const float H[9],
                  const float kinv[9], const float krlr[9], double out[11], uint8_t *mask)
{
    so3_step_f32(last, next, w, h, H, kinv, krlr, out, mask);
}

void orc_so3_step_f64(const uint8_t *last, const uint8_t *next, int w, int h, const double H[9],
                      const double kinv[9], const double krlr[9], double out[11], uint8_t *mask)
{
    so3_step_f64(last, next, w, h, H, kinv, krlr, out, mask);
}

/* ---------------------------------------------------------- a18 unpack */
/* reduce.cu:475-489 */
void orc_unpack_se3(const double in[29], double A[36], double b[6])
{
    int s = 0;
    for (int i = 0; i < 6; ++i)
        for (int j = i; j < 7; ++j) {
            const double v = in[s++];
            if (j == 6) b[i] = v;
            else A[j * 6 + i] = A[i * 6 + j] = v;
        }
}

/* reduce.cu:1125-1139 */
void orc_unpack_so3(const double in[11], double A[9], double b[3])
{
    int s = 0;
    for (int i = 0; i < 3; ++i)
        for (int j = i; j < 4; ++j) {
            const double v = in[s++];
            if (j == 3) b[i] = v;
            else A[j * 3 + i] = A[i * 3 + j] = v;
        }
}

/* --------------------------------------------------------- a20 solve */
/* LDL^T with symmetric diagonal pivoting (largest remaining |diagonal|) and
 * a pseudo-inverse of D (pivots with |d| <= DBL_MIN give a zero component):
 * the documented behaviour of Eigen's LDLT::solve used at
 * RGBDOdometryef.cpp:366, 544, 550, 556. */
int orc_ldlt_solve(int n, const double *Ain, const double *b, double *x)
{
    double A[36];
    int perm[6];
    double y[6];
    int zeros = 0;
    for (int i = 0; i < n * n; ++i) A[i] = Ain[i];
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(A[k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (fabs(A[i * n + i]) > best) { best = fabs(A[i * n + i]); p = i; }
        if (p != k) {
            for (int j = 0; j < n; ++j) { double t = A[k * n + j]; A[k * n + j] = A[p * n + j]; A[p * n + j] = t; }
            for (int i = 0; i < n; ++i) { double t = A[i * n + k]; A[i * n + k] = A[i * n + p]; A[i * n + p] = t; }
            int t = perm[k]; perm[k] = perm[p]; perm[p] = t;
        }
        const double d = A[k * n + k];
        if (fabs(d) <= DBL_MIN) {
            for (int i = k + 1; i < n; ++i) A[i * n + k] = 0.0;
            continue;
        }
        for (int i = k + 1; i < n; ++i) A[i * n + k] = A[i * n + k] / d;
        for (int i = k + 1; i < n; ++i)
            for (int j = k + 1; j <= i; ++j) {
                A[i * n + j] = A[i * n + j] - A[i * n + k] * A[j * n + k] * d;
                A[j * n + i] = A[i * n + j];
            }
    }
    for (int i = 0; i < n; ++i) y[i] = b[perm[i]];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) y[i] = y[i] - A[i * n + j] * y[j];
    for (int i = 0; i < n; ++i) {
        const double d = A[i * n + i];
        if (fabs(d) > DBL_MIN) y[i] = y[i] / d;
        else { y[i] = 0.0; ++zeros; }
    }
    for (int i = n - 1; i >= 0; --i)
        for (int j = i + 1; j < n; ++j) y[i] = y[i] - A[j * n + i] * y[j];
    for (int i = 0; i < n; ++i) x[perm[i]] = y[i];
    return zeros;
}

/* utils.h:16-52 */
void orc_rodrigues(const double src[3], double R[9])
{
    for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
    double rx = src[0], ry = src[1], rz = src[2];
    const double theta = sqrt(rx * rx + ry * ry + rz * rz);
    if (theta >= DBL_EPSILON) {
        const double I[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        const double c = cos(theta), s = sin(theta), c1 = 1. - c;
        const double itheta = theta ? 1. / theta : 0.;
        rx *= itheta; ry *= itheta; rz *= itheta;
        const double rrt[9] = {rx * rx, rx * ry, rx * rz, rx * ry, ry * ry, ry * rz, rx * rz, ry * rz, rz * rz};
        const double rx_[9] = {0, -rz, ry, rz, 0, -rx, -ry, rx, 0};
        for (int k = 0; k < 9; ++k) R[k] = c * I[k] + c1 * rrt[k] + s * rx_[k];
    }
}

/* ---------------------------------------------------- small matrix ops */
static void mm3d(const double *a, const double *b, double *o)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

static void mm3f(const float *a, const float *b, float *o)
{
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            o[i * 3 + j] = a[i * 3] * b[j] + a[i * 3 + 1] * b[3 + j] + a[i * 3 + 2] * b[6 + j];
}

static void mm4d(const double *a, const double *b, double *o)
{
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            o[i * 4 + j] = a[i * 4] * b[j] + a[i * 4 + 1] * b[4 + j] + a[i * 4 + 2] * b[8 + j] +
                           a[i * 4 + 3] * b[12 + j];
}

/* 3x3 float inverse by cofactors (Rprev.inverse(), RGBDOdometryef.cpp:386) */
static void inv3f(const float *m, float *o)
{
    const float c0 = m[4] * m[8] - m[5] * m[7];
    const float c1 = m[5] * m[6] - m[3] * m[8];
    const float c2 = m[3] * m[7] - m[4] * m[6];
    const float det = m[0] * c0 + m[1] * c1 + m[2] * c2;
    const float id = 1.0f / det;
    o[0] = c0 * id; o[1] = (m[2] * m[7] - m[1] * m[8]) * id; o[2] = (m[1] * m[5] - m[2] * m[4]) * id;
    o[3] = c1 * id; o[4] = (m[0] * m[8] - m[2] * m[6]) * id; o[5] = (m[2] * m[3] - m[0] * m[5]) * id;
    o[6] = c2 * id; o[7] = (m[1] * m[6] - m[0] * m[7]) * id; o[8] = (m[0] * m[4] - m[1] * m[3]) * id;
}

/* K and K^-1 of a level in double, K^-1 in closed form */
static void kmat(orc_intr k, double K[9], double Ki[9])
{
    const double fx = k.fx, fy = k.fy, cx = k.cx, cy = k.cy;
    K[0] = fx; K[1] = 0; K[2] = cx; K[3] = 0; K[4] = fy; K[5] = cy; K[6] = 0; K[7] = 0; K[8] = 1;
    Ki[0] = 1.0 / fx; Ki[1] = 0; Ki[2] = -cx / fx;
    Ki[3] = 0; Ki[4] = 1.0 / fy; Ki[5] = -cy / fy;
    Ki[6] = 0; Ki[7] = 0; Ki[8] = 1;
}

/* ------------------------------------------------------------ tracker */
#define NPYR 3

struct orc_odom {
    int w, h;
    orc_intr intr;
    float distThres, angleThres;
    int fp64;
    int pw[NPYR], ph[NPYR];
    uint16_t *depth_tmp[NPYR];
    float *vmaps_tmp, *nmaps_tmp;
    float *vmap_g_prev[NPYR], *nmap_g_prev[NPYR], *vmap_curr[NPYR], *nmap_curr[NPYR];
    float *lastDepth[NPYR], *nextDepth[NPYR];
    uint8_t *lastImage[NPYR], *nextImage[NPYR], *lastNextImage[NPYR];
    int16_t *dIdx[NPYR], *dIdy[NPYR];
    orc_dataterm *corres[NPYR];
    float lastICPError, lastICPCount, lastRGBError, lastRGBCount, lastSO3Error, lastSO3Count;
    int so3_iters;
    double so3R[9];                 /* resultR after the SO3 loop (identity when off) */
    double lastA[36], lastb[6];
};

/* RGBDOdometryef.cpp:21-111 */
orc_odom *orc_create(int w, int h, float cx, float cy, float fx, float fy, float distThresh,
                     float angleThresh)
{
    orc_odom *o = (orc_odom *)calloc(1, sizeof(orc_odom));
    o->w = w; o->h = h;
    o->intr.fx = fx; o->intr.fy = fy; o->intr.cx = cx; o->intr.cy = cy;
    o->distThres = distThresh; o->angleThres = angleThresh;
    o->lastICPError = 0; o->lastICPCount = (float)(w * h);
    o->lastRGBError = 0; o->lastRGBCount = (float)(w * h);
    o->lastSO3Error = 0; o->lastSO3Count = (float)(w * h);
    for (int l = 0; l < NPYR; ++l) {
        const int pw = w >> l, ph = h >> l;
        const size_t n = (size_t)pw * ph;
        o->pw[l] = pw; o->ph[l] = ph;
        o->depth_tmp[l] = (uint16_t *)calloc(n, 2);
        o->vmap_g_prev[l] = (float *)calloc(3 * n, 4);
        o->nmap_g_prev[l] = (float *)calloc(3 * n, 4);
        o->vmap_curr[l] = (float *)calloc(3 * n, 4);
        o->nmap_curr[l] = (float *)calloc(3 * n, 4);
        o->lastDepth[l] = (float *)calloc(n, 4);
        o->nextDepth[l] = (float *)calloc(n, 4);
        o->lastImage[l] = (uint8_t *)calloc(n, 1);
        o->nextImage[l] = (uint8_t *)calloc(n, 1);
        o->lastNextImage[l] = (uint8_t *)calloc(n, 1);
        o->dIdx[l] = (int16_t *)calloc(n, 2);
        o->dIdy[l] = (int16_t *)calloc(n, 2);
        o->corres[l] = (orc_dataterm *)calloc(n, sizeof(orc_dataterm));
    }
    o->vmaps_tmp = (float *)calloc((size_t)4 * w * h, 4);
    o->nmaps_tmp = (float *)calloc((size_t)4 * w * h, 4);
    return o;
}

void orc_destroy(orc_odom *o)
{
    if (!o) return;
    for (int l = 0; l < NPYR; ++l) {
        free(o->depth_tmp[l]); free(o->vmap_g_prev[l]); free(o->nmap_g_prev[l]);
        free(o->vmap_curr[l]); free(o->nmap_curr[l]); free(o->lastDepth[l]); free(o->nextDepth[l]);
        free(o->lastImage[l]); free(o->nextImage[l]); free(o->lastNextImage[l]);
        free(o->dIdx[l]); free(o->dIdy[l]); free(o->corres[l]);
    }
    free(o->vmaps_tmp); free(o->nmaps_tmp);
    free(o);
}

void orc_set_precision(orc_odom *o, int fp64) { o->fp64 = fp64; }

/* RGBDOdometryef.cpp:118-142 */
void orc_init_icp_depth(orc_odom *o, const uint16_t *depth, float depthCutoff)
{
    memcpy(o->depth_tmp[0], depth, (size_t)o->w * o->h * 2);
    for (int l = 1; l < NPYR; ++l)
        orc_pyrdown_u16(o->depth_tmp[l - 1], o->pw[l - 1], o->ph[l - 1], o->depth_tmp[l]);
    for (int l = 0; l < NPYR; ++l) {
        orc_vmap(o->depth_tmp[l], o->pw[l], o->ph[l], orc_level_intr(o->intr, l), depthCutoff, o->vmap_curr[l]);
        orc_nmap(o->vmap_curr[l], o->pw[l], o->ph[l], o->nmap_curr[l]);
    }
}

/* RGBDOdometryef.cpp:144-167 */
void orc_init_icp_maps(orc_odom *o, const float *v4, const float *n4, float depthCutoff)
{
    (void)depthCutoff;
    memcpy(o->vmaps_tmp, v4, (size_t)16 * o->w * o->h);
    memcpy(o->nmaps_tmp, n4, (size_t)16 * o->w * o->h);
    orc_copy_maps(o->vmaps_tmp, o->nmaps_tmp, o->w, o->h, o->vmap_curr[0], o->nmap_curr[0]);
    for (int l = 1; l < NPYR; ++l) {
        orc_resize_map(o->vmap_curr[l - 1], o->pw[l - 1], o->ph[l - 1], o->vmap_curr[l], 0);
        orc_resize_map(o->nmap_curr[l - 1], o->pw[l - 1], o->ph[l - 1], o->nmap_curr[l], 1);
    }
}

/* RGBDOdometryef.cpp:169-206; pose is row-major 4x4 */
void orc_init_icp_model(orc_odom *o, const float *v4, const float *n4, float depthCutoff,
                        const float pose[16])
{
    (void)depthCutoff;
    memcpy(o->vmaps_tmp, v4, (size_t)16 * o->w * o->h);
    memcpy(o->nmaps_tmp, n4, (size_t)16 * o->w * o->h);
    orc_copy_maps(o->vmaps_tmp, o->nmaps_tmp, o->w, o->h, o->vmap_g_prev[0], o->nmap_g_prev[0]);
    for (int l = 1; l < NPYR; ++l) {
        orc_resize_map(o->vmap_g_prev[l - 1], o->pw[l - 1], o->ph[l - 1], o->vmap_g_prev[l], 0);
        orc_resize_map(o->nmap_g_prev[l - 1], o->pw[l - 1], o->ph[l - 1], o->nmap_g_prev[l], 1);
    }
    const float R[9] = {pose[0], pose[1], pose[2], pose[4], pose[5], pose[6], pose[8], pose[9], pose[10]};
    const float t[3] = {pose[3], pose[7], pose[11]};
    for (int l = 0; l < NPYR; ++l)
        orc_transform_maps(o->vmap_g_prev[l], o->nmap_g_prev[l], o->pw[l], o->ph[l], R, t);
}

/* RGBDOdometryef.cpp:208-235 (maxDepthRGB = 6.0, :37) */
static void populate_rgbd(orc_odom *o, const uint8_t *rgb, int channels, float **dstDepth,
                          uint8_t **dstImage)
{
    orc_vertices_to_depth(o->vmaps_tmp, o->w, o->h, 6.0f, dstDepth[0]);
    for (int l = 0; l + 1 < NPYR; ++l)
        orc_pyrdown_f32(dstDepth[l], o->pw[l], o->ph[l], dstDepth[l + 1]);
    orc_intensity(rgb, o->w, o->h, channels, dstImage[0]);
    for (int l = 0; l + 1 < NPYR; ++l)
        orc_pyrdown_u8(dstImage[l], o->pw[l], o->ph[l], dstImage[l + 1]);
}

void orc_init_rgb_model(orc_odom *o, const uint8_t *rgb, int channels)
{
    populate_rgbd(o, rgb, channels, o->lastDepth, o->lastImage);
}

void orc_init_rgb(orc_odom *o, const uint8_t *rgb, int channels)
{
    populate_rgbd(o, rgb, channels, o->nextDepth, o->nextImage);
}

/* RGBDOdometryef.cpp:249-265 */
void orc_init_first_rgb(orc_odom *o, const uint8_t *rgb, int channels)
{
    orc_intensity(rgb, o->w, o->h, channels, o->lastNextImage[0]);
    for (int l = 0; l + 1 < NPYR; ++l)
        orc_pyrdown_u8(o->lastNextImage[l], o->pw[l], o->ph[l], o->lastNextImage[l + 1]);
}

static void d2f(const double *a, float *b, int n)
{
    for (int i = 0; i < n; ++i) b[i] = (float)a[i];
}

/* RGBDOdometryef.cpp:267-595 */
/* currentT = [Rprev|tprev] * resultRt^-1 with resultRt taken to float
 * (RGBDOdometryef.cpp:563-575) */
static void pose_from_result(const double resultRt[16], const float Rprev[9], const float tprev[3], float Rcurr[9],
                             float tcurr[3])
{
    float Rf[9], tf[3], Ri[9], ti[3];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) Rf[a * 3 + b] = (float)resultRt[a * 4 + b];
        tf[a] = (float)resultRt[a * 4 + 3];
    }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Ri[a * 3 + b] = Rf[b * 3 + a];
    for (int a = 0; a < 3; ++a) ti[a] = -(Ri[a * 3] * tf[0] + Ri[a * 3 + 1] * tf[1] + Ri[a * 3 + 2] * tf[2]);
    mm3f(Rprev, Ri, Rcurr);
    for (int a = 0; a < 3; ++a)
        tcurr[a] = Rprev[a * 3] * ti[0] + Rprev[a * 3 + 1] * ti[1] + Rprev[a * 3 + 2] * ti[2] + tprev[a];
}

void orc_track(orc_odom *o, float t[3], float R[9], int rgbOnly, float icpWeight, int pyramid,
               int fastOdom, int so3)
{
    const int iters[NPYR] = {fastOdom ? 3 : 10, pyramid ? 5 : 0, pyramid ? 4 : 0};
    orc_track_ex(o, t, R, rgbOnly, icpWeight, iters, so3, NULL);
}

void orc_track_ex(orc_odom *o, float t[3], float R[9], int rgbOnly, float icpWeight, const int iters[3],
                  int so3, const double *seed)
{
    orc_track_run(o, t, R, rgbOnly, icpWeight, iters, so3, seed, 0.0f, 0.0f, NPYR - 1, 0, NULL, NULL);
}

void orc_track_run(orc_odom *o, float t[3], float R[9], int rgbOnly, float icpWeight, const int iters[3],
                   int so3, const double *seed, float convT, float convR, int hi, int lo, double resultRt_out[16],
                   int *iters_done)
{
    const int icp = !rgbOnly && icpWeight > 0;
    const int rgb = rgbOnly || icpWeight < 100;
    const float sobelScale = (float)(1.0 / pow(2.0, 3)); /* sobelSize 3, :34-35 */
    const float maxDepthDeltaRGB = 0.07f;                   /* :36 */
    const float minGrad[NPYR] = {5, 3, 1};                  /* :107-110 */
    float Rprev[9], tprev[3], Rcurr[9], tcurr[3];
    memcpy(Rprev, R, sizeof Rprev); memcpy(tprev, t, sizeof tprev);
    memcpy(Rcurr, R, sizeof Rcurr); memcpy(tcurr, t, sizeof tcurr);

    if (rgb)
        for (int l = 0; l < NPYR; ++l)
            orc_derivative(o->nextImage[l], o->pw[l], o->ph[l], o->dIdx[l], o->dIdy[l]);

    double resultR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    o->so3_iters = 0;
    if (hi != NPYR - 1) so3 = 0;    /* the SO3 pre-alignment precedes the coarsest level */
    if (so3) { /* :294-380 */
        const int lvl = 2;
        float R_lr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        double K[9], Ki[9];
        kmat(orc_level_intr(o->intr, lvl), K, Ki);
        float lastError = FLT_MAX / 2, lastCount = FLT_MAX / 2;
        double lastResultR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        for (int i = 0; i < 10; ++i) {
            double KR[9], Hd[9];
            mm3d(K, resultR, KR);
            mm3d(KR, Ki, Hd);
            double sums[11];
            if (o->fp64) {
                orc_so3_step_f64(o->lastNextImage[lvl], o->nextImage[lvl], o->pw[lvl], o->ph[lvl],
                                 Hd, Ki, KR, sums, NULL);
            } else {
                float Hf[9], Kif[9], KRf[9];
                d2f(Hd, Hf, 9); d2f(Ki, Kif, 9); d2f(KR, KRf, 9);
                orc_so3_step(o->lastNextImage[lvl], o->nextImage[lvl], o->pw[lvl], o->ph[lvl],
                             Hf, Kif, KRf, sums, NULL);
            }
            o->so3_iters++;
            const float res0 = (float)sums[9], res1 = (float)sums[10];
            o->lastSO3Error = sqrtf(res0) / res1;
            o->lastSO3Count = res1;
            if (o->lastSO3Error < lastError && lastCount == o->lastSO3Count) break;
            else if (o->lastSO3Error > lastError + 0.001f) {
                o->lastSO3Error = lastError;
                o->lastSO3Count = lastCount;
                memcpy(resultR, lastResultR, sizeof resultR);
                break;
            }
            lastError = o->lastSO3Error;
            lastCount = o->lastSO3Count;
            memcpy(lastResultR, resultR, sizeof resultR);
            double A3[9], b3[3], delta[3], rot[9];
            orc_unpack_so3(sums, A3, b3);
            orc_ldlt_solve(3, A3, b3, delta);
            orc_rodrigues(delta, rot);
            float rotf[9], nR[9];
            d2f(rot, rotf, 9);
            mm3f(rotf, R_lr, nR);
            memcpy(R_lr, nR, sizeof nR);
            for (int k = 0; k < 9; ++k) resultR[k] = (double)R_lr[k];
        }
    }

    memcpy(o->so3R, resultR, sizeof resultR);

    int iterations[NPYR];
    for (int l = 0; l < NPYR; ++l) iterations[l] = iters[l];

    float Rprev_inv[9];
    inv3f(Rprev, Rprev_inv);

    double resultRt[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    if (so3)
        for (int x = 0; x < 3; ++x)
            for (int y = 0; y < 3; ++y) resultRt[x * 4 + y] = resultR[x * 3 + y];
    if (seed) {
        /* a seeded increment (no counterpart in the reference): the first
         * iteration already starts from it, for the ICP term too */
        memcpy(resultRt, seed, sizeof resultRt);
        pose_from_result(resultRt, Rprev, tprev, Rcurr, tcurr);
    }

    float residual[2] = {0.0f, 0.0f};
    int ndone = 0;
    for (int i = hi; i >= lo; --i) {
        const orc_intr ki = orc_level_intr(o->intr, i);
        const int pw = o->pw[i], ph = o->ph[i];
        double K[9], Ki[9];
        kmat(ki, K, Ki);
        o->lastRGBError = FLT_MAX;
        for (int j = 0; j < iterations[i]; ++j) {
            /* Rt = resultRt^-1 as the rigid inverse [R^T | -R^T t] */
            double Rinv[9], tinv[3];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) Rinv[a * 3 + b] = resultRt[b * 4 + a];
            for (int a = 0; a < 3; ++a)
                tinv[a] = -(Rinv[a * 3] * resultRt[3] + Rinv[a * 3 + 1] * resultRt[7] +
                            Rinv[a * 3 + 2] * resultRt[11]);
            double KR[9], KRKi[9], Kt[3];
            mm3d(K, Rinv, KR);
            mm3d(KR, Ki, KRKi);
            for (int a = 0; a < 3; ++a) Kt[a] = K[a * 3] * tinv[0] + K[a * 3 + 1] * tinv[1] + K[a * 3 + 2] * tinv[2];

            int32_t sigma = 0, rgbSize = 0;
            const float minScale = (float)(pow(minGrad[i], 2.0) / pow(sobelScale, 2.0));
            if (rgb) {
                if (o->fp64) {
                    orc_rgb_residual_f64(minScale, o->dIdx[i], o->dIdy[i], o->lastDepth[i], o->nextDepth[i],
                                         o->lastImage[i], o->nextImage[i], pw, ph, (double)maxDepthDeltaRGB,
                                         KRKi, Kt, o->corres[i], &rgbSize, &sigma);
                } else {
                    float KRKif[9], Ktf[3];
                    d2f(KRKi, KRKif, 9); d2f(Kt, Ktf, 3);
                    orc_rgb_residual(minScale, o->dIdx[i], o->dIdy[i], o->lastDepth[i], o->nextDepth[i],
                                     o->lastImage[i], o->nextImage[i], pw, ph, maxDepthDeltaRGB, KRKif,
                                     Ktf, o->corres[i], &rgbSize, &sigma);
                }
            }
            /* :457-458, with the precedence of the original expression */
            float sigmaVal = (float)sqrt((double)(((float)sigma / (float)rgbSize == 0) ? 1 : rgbSize));
            const float rgbError = (float)(sqrt((double)sigma) / (double)(rgbSize == 0 ? 1 : rgbSize));
            if (rgbOnly && rgbError > o->lastRGBError) break;
            o->lastRGBError = rgbError;
            o->lastRGBCount = (float)rgbSize;
            if (rgbOnly) sigmaVal = -1;

            double sicp[29] = {0}, srgb[29] = {0};
            if (icp) {
                if (o->fp64) {
                    double Rc[9], tc[3], Rpi[9], tp[3];
                    for (int a = 0; a < 9; ++a) { Rc[a] = Rcurr[a]; Rpi[a] = Rprev_inv[a]; }
                    for (int a = 0; a < 3; ++a) { tc[a] = tcurr[a]; tp[a] = tprev[a]; }
                    orc_icp_step_f64(o->vmap_curr[i], o->nmap_curr[i], o->vmap_g_prev[i], o->nmap_g_prev[i],
                                     pw, ph, ki, Rc, tc, Rpi, tp, (double)o->distThres,
                                     (double)o->angleThres, sicp, NULL);
                } else {
                    orc_icp_step(o->vmap_curr[i], o->nmap_curr[i], o->vmap_g_prev[i], o->nmap_g_prev[i], pw,
                                 ph, ki, Rcurr, tcurr, Rprev_inv, tprev, o->distThres, o->angleThres, sicp,
                                 NULL);
                }
                residual[0] = (float)sicp[27];
                residual[1] = (float)sicp[28];
            }
            o->lastICPError = sqrtf(residual[0]) / residual[1];
            o->lastICPCount = residual[1];
            if (rgb) {
                if (o->fp64)
                    orc_rgb_step_f64(o->corres[i], pw, ph, (double)sigmaVal, o->lastDepth[i], ki, o->dIdx[i],
                                     o->dIdy[i], (double)sobelScale, srgb);
                else
                    orc_rgb_step(o->corres[i], pw, ph, sigmaVal, o->lastDepth[i], ki, o->dIdx[i], o->dIdy[i],
                                 sobelScale, srgb);
            }
            double Ai[36], bi[6], Ar[36], br[6], x[6];
            orc_unpack_se3(sicp, Ai, bi);
            orc_unpack_se3(srgb, Ar, br);
            if (icp && rgb) {
                const double wgt = icpWeight;
                for (int a = 0; a < 36; ++a) o->lastA[a] = Ar[a] + wgt * wgt * Ai[a];
                for (int a = 0; a < 6; ++a) o->lastb[a] = br[a] + wgt * bi[a];
            } else if (icp) {
                memcpy(o->lastA, Ai, sizeof Ai); memcpy(o->lastb, bi, sizeof bi);
            } else {
                memcpy(o->lastA, Ar, sizeof Ar); memcpy(o->lastb, br, sizeof br);
            }
            orc_ldlt_solve(6, o->lastA, o->lastb, x);

            /* computeUpdateSE3, utils.h:54-74 */
            double rot[9], upd[16] = {0}, nrt[16];
            orc_rodrigues(x + 3, rot);
            for (int a = 0; a < 3; ++a) {
                for (int b = 0; b < 3; ++b) upd[a * 4 + b] = rot[a * 3 + b];
                upd[a * 4 + 3] = x[a];
            }
            upd[15] = 1.0;
            mm4d(upd, resultRt, nrt);
            memcpy(resultRt, nrt, sizeof nrt);
            /* rgbOdom.inverse(): [R^T | -R^T t]; currentT = [Rprev|tprev] * inv (:567-575) */
            pose_from_result(resultRt, Rprev, tprev, Rcurr, tcurr);
            ++ndone;
            /* the convergence policy (no counterpart in the reference, whose
             * schedule is fixed, :382-384, 420): an update below both
             * thresholds ends the level */
            if (convT > 0.0f) {
                const double nt = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
                const double nr = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
                if (nt < (double)convT && nr < (double)convR) break;
            }
        }
    }
    if (iters_done) *iters_done = ndone;
    if (lo > 0) {
        /* a partial run (the coarse levels of a judged call): no revert,
         * no image swap; the increment continues in a later run */
        if (resultRt_out) memcpy(resultRt_out, resultRt, sizeof resultRt);
        memcpy(t, tcurr, sizeof tcurr);
        memcpy(R, Rcurr, sizeof Rcurr);
        return;
    }

    /* :579-583 */
    int reverted = 0;
    if (rgb) {
        const float d0 = tcurr[0] - tprev[0], d1 = tcurr[1] - tprev[1], d2 = tcurr[2] - tprev[2];
        if (sqrtf(d0 * d0 + d1 * d1 + d2 * d2) > 0.3f) {
            memcpy(Rcurr, Rprev, sizeof Rcurr);
            memcpy(tcurr, tprev, sizeof tcurr);
            reverted = 1;
        }
    }
    if (resultRt_out) {
        for (int a = 0; a < 16; ++a) resultRt_out[a] = reverted ? ((a % 5 == 0) ? 1.0 : 0.0) : resultRt[a];
    }
    /* :585-591 */
    if (so3)
        for (int l = 0; l < NPYR; ++l) {
            uint8_t *tmp = o->lastNextImage[l];
            o->lastNextImage[l] = o->nextImage[l];
            o->nextImage[l] = tmp;
        }
    memcpy(t, tcurr, sizeof tcurr);
    memcpy(R, Rcurr, sizeof Rcurr);
}

void orc_get_stats(const orc_odom *o, float s[7])
{
    s[0] = o->lastICPError; s[1] = o->lastICPCount; s[2] = o->lastRGBError; s[3] = o->lastRGBCount;
    s[4] = o->lastSO3Error; s[5] = o->lastSO3Count; s[6] = (float)o->so3_iters;
}

void orc_get_so3_rotation(const orc_odom *o, double R[9])
{
    memcpy(R, o->so3R, sizeof o->so3R);
}

void orc_get_last_system(const orc_odom *o, double A[36], double b[6])
{
    memcpy(A, o->lastA, sizeof o->lastA);
    memcpy(b, o->lastb, sizeof o->lastb);
}

/* RGBDOdometryef.cpp:597-600: LU inverse (Gauss-Jordan, partial pivoting) */
int orc_get_covariance(const orc_odom *o, double cov[36])
{
    double a[36], inv[36];
    memcpy(a, o->lastA, sizeof a);
    for (int i = 0; i < 36; ++i) inv[i] = (i % 7 == 0) ? 1.0 : 0.0;
    for (int c = 0; c < 6; ++c) {
        int p = c;
        for (int r = c + 1; r < 6; ++r)
            if (fabs(a[r * 6 + c]) > fabs(a[p * 6 + c])) p = r;
        if (a[p * 6 + c] == 0.0) return -1;
        if (p != c)
            for (int k = 0; k < 6; ++k) {
                double t = a[c * 6 + k]; a[c * 6 + k] = a[p * 6 + k]; a[p * 6 + k] = t;
                t = inv[c * 6 + k]; inv[c * 6 + k] = inv[p * 6 + k]; inv[p * 6 + k] = t;
            }
        const double d = a[c * 6 + c];
        for (int k = 0; k < 6; ++k) { a[c * 6 + k] /= d; inv[c * 6 + k] /= d; }
        for (int r = 0; r < 6; ++r) {
            if (r == c) continue;
            const double f = a[r * 6 + c];
            for (int k = 0; k < 6; ++k) { a[r * 6 + k] -= f * a[c * 6 + k]; inv[r * 6 + k] -= f * inv[c * 6 + k]; }
        }
    }
    memcpy(cov, inv, sizeof inv);
    return 0;
}

const float *orc_buf_f32(const orc_odom *o, const char *name, int l)
{
    if (l < 0 || l >= NPYR) return NULL;
    if (!strcmp(name, "vmap_curr")) return o->vmap_curr[l];
    if (!strcmp(name, "nmap_curr")) return o->nmap_curr[l];
    if (!strcmp(name, "vmap_g_prev")) return o->vmap_g_prev[l];
    if (!strcmp(name, "nmap_g_prev")) return o->nmap_g_prev[l];
    if (!strcmp(name, "lastDepth")) return o->lastDepth[l];
    if (!strcmp(name, "nextDepth")) return o->nextDepth[l];
    return NULL;
}

const uint8_t *orc_buf_u8(const orc_odom *o, const char *name, int l)
{
    if (l < 0 || l >= NPYR) return NULL;
    if (!strcmp(name, "lastImage")) return o->lastImage[l];
    if (!strcmp(name, "nextImage")) return o->nextImage[l];
    if (!strcmp(name, "lastNextImage")) return o->lastNextImage[l];
    return NULL;
}

const int16_t *orc_buf_s16(const orc_odom *o, const char *name, int l)
{
    if (l < 0 || l >= NPYR) return NULL;
    if (!strcmp(name, "dIdx")) return o->dIdx[l];
    if (!strcmp(name, "dIdy")) return o->dIdy[l];
    return NULL;
}

const uint16_t *orc_buf_u16(const orc_odom *o, const char *name, int l)
{
    if (l < 0 || l >= NPYR) return NULL;
    if (!strcmp(name, "depth_tmp")) return o->depth_tmp[l];
    return NULL;
}
