/*
 * splat_oracle.c -- CPU oracle of the surfel splat prediction and fill-in.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * A restatement in C of the reference's GL prediction:
 *   IndexMap::combinedPredict   src/model/IndexMap.cpp:243-345
 *   vertex stage                src/model/shaders/splat.vert:38-87
 *   fragment stage              src/model/shaders/combo_splat.frag:35-61,
 *                               color.glsl decodeColor
 *   depth test                  GL_LESS with the depth test enabled
 *                               (src/ui/GUI.h:46-48), 24-bit depth
 *                               renderbuffer cleared to 1.0
 *   FillIn::vertex/normal/image src/gl/FillIn.cpp:56-198,
 *                               shaders fill_vertex.frag:32-52,
 *                               fill_normal.frag:34-49, geometry.glsl:44-61,
 *                               fill_rgb.frag:29-37
 * The draw is a loop over surfels in model order; a fragment replaces the
 * pixel's contents when its quantized depth is strictly less than the
 * buffer's, which is what GL_LESS does with primitives processed in order.
 *
 * Choices where GL leaves the result to the implementation:
 * - the inverse pose is the rigid inverse [R^T | -R^T t] (Eigen's general
 *   4x4 inverse differs by rounding only);
 * - the window position of a point is the viewport transform of its
 *   normalized device coordinates without sub-pixel snapping; a point is
 *   dropped when its centre lies outside the view volume (GL clips points
 *   whole);
 * - the sprite covers the pixel centres c with xw - S/2 <= c < xw + S/2 (and
 *   likewise in y); S is clamped to [1, 256]; a NaN size draws nothing;
 * - pow(r, 2) is r * r; normalize(v) is v * (1 / sqrt(v.v)); GLSL min/max
 *   as (b < a ? b : a) / (a < b ? b : a);
 * - a fragment whose disk test is NaN is discarded; the fragment depth is
 *   clamped to [0, 1] and quantized to round-to-nearest-even of
 *   d * (2^24 - 1);
 * - the time output is init_time saturated to [0, 65535]; the colour is
 *   decoded from the float clamped to [0, 2^24 - 1];
 * - textures sampled outside the image clamp to the edge;
 * - every operation is one IEEE FP32 rounding (-ffp-contract=off).
 */
#include "splat_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define DEPTH_CLEAR 0xFFFFFFu

typedef struct {
    float hx, hy, hz, conf, nx, ny, nz, rad, col, tinit;
} surf_t;

static float glsl_min(float a, float b) { return (b < a) ? b : a; }
static float glsl_max(float a, float b) { return (a < b) ? b : a; }

static void rigid_inverse(const float pose[16], float Ti[12])
{
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Ti[r * 4 + c] = pose[c * 4 + r];
        Ti[r * 4 + 3] = -((Ti[r * 4 + 0] * pose[3] + Ti[r * 4 + 1] * pose[7]) + Ti[r * 4 + 2] * pose[11]);
    }
}

/* splat.vert:55-66: t_inv * (p, 1), the discard tests, the normal. */
static int vertex_stage(const float* m, const float Ti[12], const orc_splat_params* P, surf_t* s)
{
    const float px = m[0], py = m[1], pz = m[2], conf = m[3];
    const float col = m[4], tinit = m[6], tlast = m[7];
    const float nx = m[8], ny = m[9], nz = m[10], rad = m[11];
    s->hx = ((Ti[0] * px + Ti[1] * py) + Ti[2] * pz) + Ti[3];
    s->hy = ((Ti[4] * px + Ti[5] * py) + Ti[6] * pz) + Ti[7];
    s->hz = ((Ti[8] * px + Ti[9] * py) + Ti[10] * pz) + Ti[11];
    /* time - vColor.w > timeDelta || vColor.w > maxTime: int operands
     * converted to float */
    if (s->hz > P->depth_cutoff || s->hz < 0.0f || conf < P->conf_threshold ||
        (float)P->time - tlast > (float)P->time_delta || tlast > (float)P->max_time)
        return 0;
    const float ax = (Ti[0] * nx + Ti[1] * ny) + Ti[2] * nz;
    const float ay = (Ti[4] * nx + Ti[5] * ny) + Ti[6] * nz;
    const float az = (Ti[8] * nx + Ti[9] * ny) + Ti[10] * nz;
    const float inv = 1.0f / sqrtf((ax * ax + ay * ay) + az * az);
    s->nx = ax * inv;
    s->ny = ay * inv;
    s->nz = az * inv;
    s->rad = rad;
    s->conf = conf;
    s->col = col;
    s->tinit = tinit;
    return 1;
}

/* splat.vert:38-50 projectPoint / projectPointImage, :68-85 point size. */
static int point_sprite(const surf_t* s, const orc_splat_params* P, float* xw, float* yw, float* size)
{
    const float cols = (float)P->width, rows = (float)P->height;
    const float ndc_x = ((((P->fx * s->hx) / s->hz) + P->cx) - (cols * 0.5f)) / (cols * 0.5f);
    const float ndc_y = ((((P->fy * s->hy) / s->hz) + P->cy) - (rows * 0.5f)) / (rows * 0.5f);
    if (!(ndc_x >= -1.0f && ndc_x <= 1.0f)) return 0;
    if (!(ndc_y >= -1.0f && ndc_y <= 1.0f)) return 0;
    /* viewport transform, viewport (0, 0, cols, rows) */
    *xw = ndc_x * (cols * 0.5f) + (cols * 0.5f);
    *yw = ndc_y * (rows * 0.5f) + (rows * 0.5f);
    /* x1 = normalize(vec3(n.y - n.z, -n.x, n.x)) * r * 1.41421356 */
    const float u[3] = {s->ny - s->nz, -s->nx, s->nx};
    const float iu = 1.0f / sqrtf((u[0] * u[0] + u[1] * u[1]) + u[2] * u[2]);
    float x1[3], y1[3];
    for (int k = 0; k < 3; ++k) x1[k] = ((u[k] * iu) * s->rad) * 1.41421356f;
    /* y1 = cross(n, x1) */
    y1[0] = s->ny * x1[2] - s->nz * x1[1];
    y1[1] = s->nz * x1[0] - s->nx * x1[2];
    y1[2] = s->nx * x1[1] - s->ny * x1[0];
    const float h[3] = {s->hx, s->hy, s->hz};
    float pxs[4], pys[4];
    for (int q = 0; q < 4; ++q) {
        /* h + x1, h + y1, h - y1, h - x1 */
        const float* o = (q == 0 || q == 3) ? x1 : y1;
        const float sg = (q < 2) ? 1.0f : -1.0f;
        float v[3];
        for (int k = 0; k < 3; ++k) v[k] = sg > 0 ? h[k] + o[k] : h[k] - o[k];
        pxs[q] = ((P->fx * v[0]) / v[2]) + P->cx;
        pys[q] = ((P->fy * v[1]) / v[2]) + P->cy;
    }
    const float xmin = glsl_min(pxs[0], glsl_min(pxs[1], glsl_min(pxs[2], pxs[3])));
    const float xmax = glsl_max(pxs[0], glsl_max(pxs[1], glsl_max(pxs[2], pxs[3])));
    const float ymin = glsl_min(pys[0], glsl_min(pys[1], glsl_min(pys[2], pys[3])));
    const float ymax = glsl_max(pys[0], glsl_max(pys[1], glsl_max(pys[2], pys[3])));
    float sz = glsl_max(0.0f, glsl_max(fabsf(xmax - xmin), fabsf(ymax - ymin)));
    if (isnan(sz)) return 0;
    if (sz < 1.0f) sz = 1.0f;
    if (sz > (float)ORC_SPLAT_MAX_POINT_SIZE) sz = (float)ORC_SPLAT_MAX_POINT_SIZE;
    *size = sz;
    return 1;
}

/* combo_splat.frag:37-48, 59: ray-plane intersection, disk test, depth. */
static int fragment_stage(const surf_t* s, const orc_splat_params* P, float fx_, float fy_, float* z,
                          unsigned* qdepth)
{
    const float lx = (fx_ - P->cx) / P->fx;
    const float ly = (fy_ - P->cy) / P->fy;
    const float il = 1.0f / sqrtf((lx * lx + ly * ly) + 1.0f * 1.0f);
    const float l[3] = {lx * il, ly * il, 1.0f * il};
    const float pn = (s->hx * s->nx + s->hy * s->ny) + s->hz * s->nz;
    const float ln = (l[0] * s->nx + l[1] * s->ny) + l[2] * s->nz;
    const float f = pn / ln;
    const float c[3] = {f * l[0], f * l[1], f * l[2]};
    const float d[3] = {c[0] - s->hx, c[1] - s->hy, c[2] - s->hz};
    const float dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    const float sqr = s->rad * s->rad;
    if (dd > sqr || isnan(dd)) return 0;
    float depth = (c[2] / (2.0f * P->depth_cutoff)) + 0.5f;
    if (depth < 0.0f) depth = 0.0f;
    if (depth > 1.0f) depth = 1.0f;
    const unsigned q = (unsigned)rintf(depth * 16777215.0f);
    if (q >= DEPTH_CLEAR) return 0; /* GL_LESS against the cleared 1.0 */
    *z = c[2];
    *qdepth = q;
    return 1;
}

void orc_splat_predict(const orc_splat_params* P, const float pose[16], const float* model, int n, uint8_t* image,
                       float* vertex, float* normal, uint16_t* time_out)
{
    const int W = P->width, H = P->height;
    const size_t npix = (size_t)W * H;
    unsigned* depthbuf = (unsigned*)malloc(npix * sizeof(unsigned));
    for (size_t i = 0; i < npix; ++i) depthbuf[i] = DEPTH_CLEAR;
    memset(image, 0, npix * 4);
    memset(vertex, 0, npix * 16);
    memset(normal, 0, npix * 16);
    memset(time_out, 0, npix * 2);
    float Ti[12];
    rigid_inverse(pose, Ti);
    const float ifx = 1.0f / P->fx, ify = 1.0f / P->fy;
    for (int k = 0; k < n; ++k) {
        surf_t s;
        float xw, yw, size;
        if (!vertex_stage(model + (size_t)k * ORC_SPLAT_FLOATS, Ti, P, &s)) continue;
        if (!point_sprite(&s, P, &xw, &yw, &size)) continue;
        const float half = size * 0.5f;
        /* a few pixels beyond the square each way; the tests below decide */
        const int y0 = (int)(yw - half) - 2 > 0 ? (int)(yw - half) - 2 : 0;
        const int y1 = (int)(yw + half) + 2 < H - 1 ? (int)(yw + half) + 2 : H - 1;
        const int x0 = (int)(xw - half) - 2 > 0 ? (int)(xw - half) - 2 : 0;
        const int x1 = (int)(xw + half) + 2 < W - 1 ? (int)(xw + half) + 2 : W - 1;
        for (int y = y0; y <= y1; ++y) {
            const float cy = (float)y + 0.5f;
            if (!(yw - half <= cy && cy < yw + half)) continue;
            for (int x = x0; x <= x1; ++x) {
                const float cx = (float)x + 0.5f;
                if (!(xw - half <= cx && cx < xw + half)) continue;
                float z;
                unsigned q;
                if (!fragment_stage(&s, P, cx, cy, &z, &q)) continue;
                const size_t p = (size_t)y * W + x;
                if (!(q < depthbuf[p])) continue;
                depthbuf[p] = q;
                /* decodeColor: int(c) >> 16 & 0xFF, ... / 255, stored as unorm8 */
                float cc = s.col;
                if (!(cc >= 0.0f)) cc = 0.0f;
                if (cc > 16777215.0f) cc = 16777215.0f;
                const int ci = (int)cc;
                image[4 * p + 0] = (uint8_t)((ci >> 16) & 0xFF);
                image[4 * p + 1] = (uint8_t)((ci >> 8) & 0xFF);
                image[4 * p + 2] = (uint8_t)(ci & 0xFF);
                image[4 * p + 3] = 255;
                vertex[4 * p + 0] = ((cx - P->cx) * z) * ifx;
                vertex[4 * p + 1] = ((cy - P->cy) * z) * ify;
                vertex[4 * p + 2] = z;
                vertex[4 * p + 3] = s.conf;
                normal[4 * p + 0] = s.nx;
                normal[4 * p + 1] = s.ny;
                normal[4 * p + 2] = s.nz;
                normal[4 * p + 3] = s.rad;
                const float t = s.tinit;
                time_out[p] = t >= 65535.0f ? 65535 : (t >= 0.0f ? (uint16_t)t : 0);
            }
        }
    }
    free(depthbuf);
}

/* geometry.glsl:44-48: z = float(depth) / 1000, integer pixel coordinates. */
static void depth_vertex(const orc_splat_params* P, const uint16_t* depth, int x, int y, float out[3])
{
    const int sx = x < P->width ? x : P->width - 1; /* clamp to edge */
    const int sy = y < P->height ? y : P->height - 1;
    const float z = (float)depth[(size_t)sy * P->width + sx] / 1000.0f;
    const float ifx = 1.0f / P->fx, ify = 1.0f / P->fy;
    out[0] = (((float)x - P->cx) * z) * ifx;
    out[1] = (((float)y - P->cy) * z) * ify;
    out[2] = z;
}

void orc_splat_fill_in(const orc_splat_params* P, const uint8_t* image, const float* vertex, const float* normal,
                       const uint16_t* depth, const uint8_t* rgb, int rgb_channels, int passthrough_geometry,
                       int passthrough_image, uint8_t* fimage, float* fvertex, float* fnormal)
{
    const int W = P->width, H = P->height;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            const size_t p = (size_t)y * W + x;
            /* fill_vertex.frag:43-51 */
            if (vertex[4 * p + 2] == 0.0f || passthrough_geometry) {
                float v[3];
                depth_vertex(P, depth, x, y, v);
                fvertex[4 * p + 0] = v[0];
                fvertex[4 * p + 1] = v[1];
                fvertex[4 * p + 2] = v[2];
                fvertex[4 * p + 3] = 1.0f;
            } else {
                memcpy(fvertex + 4 * p, vertex + 4 * p, 16);
            }
            /* fill_normal.frag:41-48, geometry.glsl:51-61 */
            if (normal[4 * p + 2] == 0.0f || passthrough_geometry) {
                float v0[3], vx[3], vy[3];
                depth_vertex(P, depth, x, y, v0);
                depth_vertex(P, depth, x + 1, y, vx);
                depth_vertex(P, depth, x, y + 1, vy);
                const float a[3] = {vx[0] - v0[0], vx[1] - v0[1], vx[2] - v0[2]};
                const float b[3] = {vy[0] - v0[0], vy[1] - v0[1], vy[2] - v0[2]};
                const float c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
                const float inv = 1.0f / sqrtf((c[0] * c[0] + c[1] * c[1]) + c[2] * c[2]);
                fnormal[4 * p + 0] = c[0] * inv;
                fnormal[4 * p + 1] = c[1] * inv;
                fnormal[4 * p + 2] = c[2] * inv;
                fnormal[4 * p + 3] = 1.0f;
            } else {
                memcpy(fnormal + 4 * p, normal + 4 * p, 16);
            }
            /* fill_rgb.frag:33-36: x + y + z of unorm values */
            const uint8_t* e = image + 4 * p;
            if (e[0] + e[1] + e[2] == 0 || passthrough_image) {
                const uint8_t* r = rgb + p * rgb_channels;
                fimage[4 * p + 0] = r[0];
                fimage[4 * p + 1] = r[1];
                fimage[4 * p + 2] = r[2];
                fimage[4 * p + 3] = rgb_channels == 4 ? r[3] : 255;
            } else {
                memcpy(fimage + 4 * p, e, 4);
            }
        }
}
