/*
 * model_oracle.c -- CPU oracle of the surfel model's fusion: initialisation,
 * index map, data association and update, and the unstable-point cleaning.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * A restatement in C of the reference's GL passes:
 *   depth to metres            src/gl/shaders/depth_metric.frag:29-37
 *   Model::initialise          src/model/Model.cpp:208-256 over
 *                              FeedbackBuffer::compute (FeedbackBuffer.cpp:86-130,
 *                              vertex_feedback.vert:38-68, .geom:33-42) and
 *                              init_unstable.vert:28-36
 *   IndexMap::predictIndices   index_map.vert:37-62, index_map.frag:29-36
 *                              (IndexMap::FACTOR = 1, IndexMap.cpp:21)
 *   Model::fuse                Model.cpp:336-487: data.vert:67-190,
 *                              data.geom:33-49, data.frag:29-38,
 *                              update.vert:37-90
 *   Model::clean               Model.cpp:489-599: copy_unstable.vert:48-142
 *                              (without deformation nodes), .geom:36-43
 *   shared                     surfels.glsl:19-46, color.glsl:20-34,
 *                              geometry.glsl:20-42
 * Draw order is the reference's: pixels column by column (the uv buffers,
 * Model.cpp:82-92, FeedbackBuffer.cpp:41-48), surfels by index.  The update
 * maps keep the first pixel drawn for a surfel (the depth test is on, every
 * point at depth 0.5, GL_LESS), transform feedback appends in draw order.
 *
 * Choices where GL leaves the result to the implementation:
 * - nearest texture sampling is texel floor(u * size) clamped to the edge,
 *   u * size in FP32;
 * - the index map's depth is the window depth z / maxDepth * 0.5 + 0.5
 *   rounded to 24 bits; a point of size 1 covers the pixel centre in
 *   [xw - 0.5, xw + 0.5); points whose centre leaves the view are clipped;
 * - exp is orc_bexpf, GLSL round() rounds half away from zero,
 *   acos(c) < 0.5 is c > cos(0.5), normalize(v) is v * (1 / sqrt(v.v));
 * - Model::initialise pairs the k-th raw vertex with the k-th filtered one,
 *   as the two feedback buffers are drawn together; past the filtered
 *   count the normal and radius read as zero;
 * - the inverse pose is the rigid inverse [R^T | -R^T t];
 * - every operation is one IEEE FP32 rounding (-ffp-contract=off).
 */
#include "model_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "odom_oracle.h"

#define DEPTH_CLEAR 0xFFFFFFu
#define F(k) (12 * (size_t)(k))

struct orc_model {
    int w, h;
    float cx, cy, fx, fy;
    int max_surfels, count;
    float* model;
    float* scratch;
    /* index map */
    uint32_t* index;
    float *vert_conf, *color_time, *norm_rad;
    unsigned* depth;
    /* new unstable points of the last fuse */
    float* fresh;
    int fresh_count;
};

orc_model* orc_model_create(int w, int h, float cx, float cy, float fx, float fy, int max_surfels)
{
    orc_model* m = (orc_model*)calloc(1, sizeof(orc_model));
    m->w = w;
    m->h = h;
    m->cx = cx;
    m->cy = cy;
    m->fx = fx;
    m->fy = fy;
    m->max_surfels = max_surfels;
    m->model = (float*)calloc((size_t)max_surfels * 12, sizeof(float));
    m->scratch = (float*)calloc((size_t)max_surfels * 12, sizeof(float));
    const size_t n = (size_t)w * h;
    m->index = (uint32_t*)calloc(n, 4);
    m->vert_conf = (float*)calloc(n * 4, 4);
    m->color_time = (float*)calloc(n * 4, 4);
    m->norm_rad = (float*)calloc(n * 4, 4);
    m->depth = (unsigned*)calloc(n, 4);
    m->fresh = (float*)calloc(n * 12, 4);
    return m;
}

void orc_model_destroy(orc_model* m)
{
    if (!m) return;
    free(m->model);
    free(m->scratch);
    free(m->index);
    free(m->vert_conf);
    free(m->color_time);
    free(m->norm_rad);
    free(m->depth);
    free(m->fresh);
    free(m);
}

int orc_model_count(const orc_model* m) { return m->count; }
const float* orc_model_data(const orc_model* m) { return m->model; }
int orc_model_fresh_count(const orc_model* m) { return m->fresh_count; }
const float* orc_model_fresh(const orc_model* m) { return m->fresh; }

void orc_model_index_map(const orc_model* m, uint32_t* index, float* vert_conf, float* color_time, float* norm_rad)
{
    const size_t n = (size_t)m->w * m->h;
    memcpy(index, m->index, n * 4);
    memcpy(vert_conf, m->vert_conf, n * 16);
    memcpy(color_time, m->color_time, n * 16);
    memcpy(norm_rad, m->norm_rad, n * 16);
}

int orc_model_set(orc_model* m, const float* model, int n)
{
    if (n < 0 || n > m->max_surfels) return -1;
    memcpy(m->model, model, (size_t)n * 48);
    m->count = n;
    return 0;
}

/* depth_metric.frag:29-37 */
void orc_depth_metric(const uint16_t* d, int n, float maxD, float* out)
{
    const unsigned lim = (unsigned)(maxD * 1000.0f);
    for (int i = 0; i < n; ++i) {
        const unsigned v = d[i];
        out[i] = (v > lim || v < 300u) ? 0.0f : (float)v / 1000.0f;
    }
}

/* ------------------------------------------------------------- helpers */

/* The uv buffers' texture coordinate of pixel i of n: (float)i / n in float,
 * plus 1.0 / (2 * (float)n) in double, stored as float. */
static float pixel_tc(int i, int n)
{
    return (float)((double)((float)i / (float)n) + 1.0 / (double)(2.0f * (float)n));
}

static int texel(float u, int n)
{
    const float t = floorf(u * (float)n);
    if (!(t >= 0.0f)) return 0;
    if (t > (float)(n - 1)) return n - 1;
    return (int)t;
}

static float sample_f(const float* img, int w, int h, float u, float v)
{
    return img[(size_t)texel(v, h) * w + texel(u, w)];
}

/* geometry.glsl:21-25: float-depth vertex at pixel coordinate (x, y) */
static void vertex_at(const float* depth, int w, int h, float u, float v, float x, float y, const float cam[4],
                      float out[3])
{
    const float z = sample_f(depth, w, h, u, v);
    out[0] = ((x - cam[0]) * z) * cam[2];
    out[1] = ((y - cam[1]) * z) * cam[3];
    out[2] = z;
}

static void normalize3(float v[3])
{
    const float inv = 1.0f / sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
}

static void cross3(const float a[3], const float b[3], float c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* geometry.glsl:28-42: central differences */
static void normal_at(const float* depth, int w, int h, const float vp[3], float u, float v, float x, float y,
                      const float cam[4], float n[3])
{
    const float cols = (float)w, rows = (float)h;
    float xf[3], xb[3], yf[3], yb[3];
    vertex_at(depth, w, h, u + (1.0f / cols), v, x + 1.0f, y, cam, xf);
    vertex_at(depth, w, h, u - (1.0f / cols), v, x - 1.0f, y, cam, xb);
    vertex_at(depth, w, h, u, v + (1.0f / rows), x, y + 1.0f, cam, yf);
    vertex_at(depth, w, h, u, v - (1.0f / rows), x, y - 1.0f, cam, yb);
    float dx[3], dy[3];
    for (int k = 0; k < 3; ++k) {
        dx[k] = ((xb[k] + vp[k]) / 2.0f) - ((xf[k] + vp[k]) / 2.0f);
        dy[k] = ((yb[k] + vp[k]) / 2.0f) - ((yf[k] + vp[k]) / 2.0f);
    }
    cross3(dx, dy, n);
    normalize3(n);
}

/* surfels.glsl:19-34 */
static float radius_of(float depth, float nz, const float cam[4])
{
    const float mean_focal = ((1.0f / fabsf(cam[2])) + (1.0f / fabsf(cam[3]))) / 2.0f;
    const float radius = (depth / mean_focal) * 1.41421356237f;
    const float rn = radius / fabsf(nz);
    const float r2 = 2.0f * radius;
    return (rn < r2) ? rn : r2; /* GLSL min(2r, rn) = (rn < 2r ? rn : 2r) */
}

/* surfels.glsl:36-46 */
static float confidence_of(float x, float y, float weighting, const float cam[4])
{
    const float px = x - cam[0], py = y - cam[1];
    const float rd = sqrtf(px * px + py * py) / 400.0f;
    return orc_bexpf((-(rd * rd)) / 0.72f) * weighting;
}

static float encode_bytes(const uint8_t* c)
{
    /* encodeColor(round(c / 255 * 255)): exact for 8-bit colour */
    return (float)(((int)c[0] << 16) | ((int)c[1] << 8) | (int)c[2]);
}

static void decode_color(float c, float out[3])
{
    const int i = (int)c;
    out[0] = (float)((i >> 16) & 0xFF) / 255.0f;
    out[1] = (float)((i >> 8) & 0xFF) / 255.0f;
    out[2] = (float)(i & 0xFF) / 255.0f;
}

static float encode_color(const float c[3])
{
    int rgb = (int)roundf(c[0] * 255.0f);
    rgb = (rgb << 8) + (int)roundf(c[1] * 255.0f);
    rgb = (rgb << 8) + (int)roundf(c[2] * 255.0f);
    return (float)rgb;
}

static void rigid_inverse(const float pose[16], float Ti[16])
{
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Ti[r * 4 + c] = pose[c * 4 + r];
        Ti[r * 4 + 3] = -((Ti[r * 4 + 0] * pose[3] + Ti[r * 4 + 1] * pose[7]) + Ti[r * 4 + 2] * pose[11]);
    }
    Ti[12] = Ti[13] = Ti[14] = 0.0f;
    Ti[15] = 1.0f;
}

/* mat4 * vec4(p, 1) and mat3 * n, summed left to right */
static void xform_point(const float M[16], const float p[3], float o[3])
{
    for (int r = 0; r < 3; ++r) o[r] = ((M[r * 4] * p[0] + M[r * 4 + 1] * p[1]) + M[r * 4 + 2] * p[2]) + M[r * 4 + 3];
}

static void xform_dir(const float M[16], const float p[3], float o[3])
{
    for (int r = 0; r < 3; ++r) o[r] = (M[r * 4] * p[0] + M[r * 4 + 1] * p[1]) + M[r * 4 + 2] * p[2];
}

/* --------------------------------------------------------- initialise */

/* vertex_feedback.vert:38-68 for one pixel; returns zVal > 0 */
static int feedback_pixel(const float* depth, const uint8_t* rgb, int w, int h, int i, int j, const float cam[4],
                          int time, float maxDepth, float out[12])
{
    const float u = pixel_tc(i, w), v = pixel_tc(j, h);
    const float x = u * (float)w, y = v * (float)h;
    float p[3], n[3];
    vertex_at(depth, w, h, u, v, x, y, cam, p);
    const uint8_t* c = rgb + 3 * ((size_t)texel(v, h) * w + texel(u, w));
    normal_at(depth, w, h, p, u, v, x, y, cam, n);
    out[0] = p[0];
    out[1] = p[1];
    out[2] = p[2];
    out[3] = confidence_of(x, y, 1.0f, cam);
    out[4] = encode_bytes(c);
    out[5] = 0.0f;
    out[6] = (float)c[2] / 255.0f; /* vColor.z keeps the sampled blue */
    out[7] = (float)time;
    out[8] = n[0];
    out[9] = n[1];
    out[10] = n[2];
    out[11] = radius_of(p[2], n[2], cam);
    return !(p[2] <= 0.0f || p[2] > maxDepth);
}

int orc_model_initialise(orc_model* m, const uint8_t* rgb, const uint16_t* raw, const uint16_t* filtered, float maxD,
                         int time, float maxDepth)
{
    const int w = m->w, h = m->h;
    const size_t n = (size_t)w * h;
    float* dr = (float*)malloc(n * 4);
    float* df = (float*)malloc(n * 4);
    orc_depth_metric(raw, (int)n, maxD, dr);
    orc_depth_metric(filtered, (int)n, maxD, df);
    /* FeedbackBuffer::compute uses 1.0f / f (FeedbackBuffer.cpp:92-95) */
    const float cam[4] = {m->cx, m->cy, 1.0f / m->fx, 1.0f / m->fy};
    float* fr = (float*)calloc(n * 12, 4);
    int nr = 0, nf = 0;
    float tmp[12];
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < h; ++j) {
            if (feedback_pixel(dr, rgb, w, h, i, j, cam, time, maxDepth, tmp)) {
                if (nr < m->max_surfels) memcpy(m->scratch + F(nr), tmp, 48); /* a full buffer drops the rest */
                ++nr;
            }
            if (feedback_pixel(df, rgb, w, h, i, j, cam, time, maxDepth, tmp)) memcpy(fr + F(nf++), tmp, 48);
        }
    if (nr > m->max_surfels) nr = m->max_surfels;
    for (int k = 0; k < nr; ++k) {
        float* o = m->model + F(k);
        const float* a = m->scratch + F(k);
        memcpy(o, a, 32); /* init_unstable.vert: position and colour from the raw buffer */
        o[5] = 0.0f;
        o[6] = 1.0f;
        if (k < nf)
            memcpy(o + 8, fr + F(k) + 8, 16); /* normal and radius from the filtered buffer */
        else
            memset(o + 8, 0, 16);
    }
    m->count = nr;
    free(fr);
    free(dr);
    free(df);
    return nr;
}

/* ------------------------------------------------------ predictIndices */

void orc_model_predict_indices(orc_model* m, const float pose[16], int time, float maxDepth, int timeDelta)
{
    const int w = m->w, h = m->h;
    const size_t n = (size_t)w * h;
    memset(m->index, 0, n * 4);
    memset(m->vert_conf, 0, n * 16);
    memset(m->color_time, 0, n * 16);
    memset(m->norm_rad, 0, n * 16);
    for (size_t p = 0; p < n; ++p) m->depth[p] = DEPTH_CLEAR;
    float Ti[16];
    rigid_inverse(pose, Ti);
    const float cols = (float)w, rows = (float)h;
    for (int k = 0; k < m->count; ++k) {
        const float* s = m->model + F(k);
        float hp[3];
        xform_point(Ti, s, hp);
        if (hp[2] > maxDepth || hp[2] < 0.0f || (float)time - s[7] > (float)timeDelta) continue;
        const float ndx = ((((m->fx * hp[0]) / hp[2]) + m->cx) - (cols * 0.5f)) / (cols * 0.5f);
        const float ndy = ((((m->fy * hp[1]) / hp[2]) + m->cy) - (rows * 0.5f)) / (rows * 0.5f);
        if (!(ndx >= -1.0f && ndx <= 1.0f && ndy >= -1.0f && ndy <= 1.0f)) continue;
        const float xw = ndx * (cols * 0.5f) + (cols * 0.5f);
        const float yw = ndy * (rows * 0.5f) + (rows * 0.5f);
        /* the pixel whose centre c satisfies xw - 0.5 <= c < xw + 0.5 */
        int px = -1, py = -1;
        for (int i = (int)xw - 1; i <= (int)xw + 1; ++i) {
            const float c = (float)i + 0.5f;
            if (i >= 0 && i < w && xw - 0.5f <= c && c < xw + 0.5f) px = i;
        }
        for (int j = (int)yw - 1; j <= (int)yw + 1; ++j) {
            const float c = (float)j + 0.5f;
            if (j >= 0 && j < h && yw - 0.5f <= c && c < yw + 0.5f) py = j;
        }
        if (px < 0 || py < 0) continue;
        float d = (hp[2] / maxDepth) * 0.5f + 0.5f;
        if (d < 0.0f) d = 0.0f;
        if (d > 1.0f) d = 1.0f;
        const unsigned q = (unsigned)rintf(d * 16777215.0f);
        const size_t p = (size_t)py * w + px;
        if (!(q < m->depth[p])) continue;
        m->depth[p] = q;
        m->index[p] = (uint32_t)k;
        float nn[3];
        xform_dir(Ti, s + 8, nn);
        normalize3(nn);
        float* vc = m->vert_conf + 4 * p;
        vc[0] = hp[0];
        vc[1] = hp[1];
        vc[2] = hp[2];
        vc[3] = s[3];
        memcpy(m->color_time + 4 * p, s + 4, 16);
        float* nr = m->norm_rad + 4 * p;
        nr[0] = nn[0];
        nr[1] = nn[1];
        nr[2] = nn[2];
        nr[3] = s[11];
    }
}

/* ---------------------------------------------------------------- fuse */

void orc_model_fuse(orc_model* m, const float pose[16], int time, const uint8_t* rgb, const uint16_t* raw,
                    const uint16_t* filtered, float maxD, float maxDepth, float weighting)
{
    const int w = m->w, h = m->h;
    const size_t n = (size_t)w * h;
    float* dr = (float*)malloc(n * 4);
    float* df = (float*)malloc(n * 4);
    orc_depth_metric(raw, (int)n, maxD, dr);
    orc_depth_metric(filtered, (int)n, maxD, df);
    /* Model::fuse's uniform: 1.0 / f in double, stored as float */
    const float cam[4] = {m->cx, m->cy, (float)(1.0 / (double)m->fx), (float)(1.0 / (double)m->fy)};
    const float cols = (float)w, rows = (float)h, scale = 1.0f, tm = (float)time;
    /* update maps: the first pixel drawn for a surfel */
    int* claim = (int*)malloc((size_t)m->max_surfels * sizeof(int));
    for (int k = 0; k < m->max_surfels; ++k) claim[k] = -1;
    float* pix = (float*)malloc(n * 48);
    m->fresh_count = 0;
    for (int i = 0; i < w; ++i)
        for (int j = 0; j < h; ++j) {
            const size_t d = (size_t)i * h + j; /* draw index */
            float* o = pix + F(d);
            const float u = pixel_tc(i, w), v = pixel_tc(j, h);
            const float x = u * cols, y = v * rows;
            float pl[3], pf[3], nl[3], pg[3], ng[3];
            vertex_at(dr, w, h, u, v, x, y, cam, pl);
            xform_point(pose, pl, pg);
            vertex_at(df, w, h, u, v, x, y, cam, pf);
            const uint8_t* c = rgb + 3 * ((size_t)texel(v, h) * w + texel(u, w));
            normal_at(df, w, h, pf, u, v, x, y, cam, nl);
            xform_dir(pose, nl, ng);
            o[0] = pg[0];
            o[1] = pg[1];
            o[2] = pg[2];
            o[3] = confidence_of(x, y, weighting, cam);
            o[4] = encode_bytes(c);
            o[5] = 0.0f;
            o[6] = tm;
            o[7] = 0.0f;
            o[8] = ng[0];
            o[9] = ng[1];
            o[10] = ng[2];
            o[11] = radius_of(pf[2], nl[2], cam);
            int update = 0;
            unsigned best = 0;
            const int tpar = (int)tm % 2;
            /* checkNeighbours (data.vert:49-67) */
            int nb = sample_f(dr, w, h, u - (1.0f / cols), v) != 0.0f && sample_f(dr, w, h, u, v - (1.0f / rows)) != 0.0f &&
                     sample_f(dr, w, h, u + (1.0f / cols), v) != 0.0f && sample_f(dr, w, h, u, v + (1.0f / rows)) != 0.0f;
            if ((int)x % 2 == tpar && (int)y % 2 == tpar && nb && pl[2] > 0.0f && pl[2] <= maxDepth) {
                int counter = 0;
                const float sx = (1.0f / (cols * scale)) * 0.5f;
                const float sy = (1.0f / (rows * scale)) * 0.5f;
                float best_dist = 1000.0f;
                const float wm = 2.0f;
                const float xl = (x - cam[0]) * cam[2];
                const float yl = (y - cam[1]) * cam[3];
                const float lambda = sqrtf(xl * xl + yl * yl + 1.0f);
                const float ray[3] = {xl, yl, 1.0f};
                const float ray_len = sqrtf(dot3(ray, ray));
                for (float ii = u - ((scale * sx) * wm); ii < u + ((scale * sx) * wm); ii += sx)
                    for (float jj = v - ((scale * sy) * wm); jj < v + ((scale * sy) * wm); jj += sy) {
                        const size_t t = (size_t)texel(jj, h) * w + texel(ii, w);
                        const uint32_t cur = m->index[t];
                        if (cur == 0) continue;
                        const float* vc = m->vert_conf + 4 * t;
                        if (!(fabsf((vc[2] * lambda) - (pl[2] * lambda)) < 0.05f)) continue;
                        float cr[3];
                        cross3(ray, vc, cr);
                        const float dist = sqrtf(dot3(cr, cr)) / ray_len;
                        const float* nr = m->norm_rad + 4 * t;
                        const float cosang = dot3(nr, nl) / (sqrtf(dot3(nr, nr)) * sqrtf(dot3(nl, nl)));
                        if (dist < best_dist && (fabsf(nr[2]) < 0.75f || cosang > 0.87758255f)) {
                            counter++;
                            best_dist = dist;
                            best = cur;
                        }
                    }
                update = counter > 0 ? 1 : 2;
                o[7] = counter > 0 ? -1.0f : -2.0f;
            }
            /* an index map from a larger model can name surfels that are gone */
            if (update == 1 && (int)best < m->count && claim[best] < 0) claim[best] = (int)d;
            if (update > 0) memcpy(m->fresh + F(m->fresh_count++), o, 48);
        }
    /* update.vert:37-90 */
    for (int k = 0; k < m->count; ++k) {
        if (claim[k] < 0) continue;
        float* s = m->model + F(k);
        const float* g = pix + F(claim[k]);
        const float ck = s[3], a = g[3];
        if (g[11] < (1.0f + 0.5f) * s[11]) {
            float np[3];
            for (int r = 0; r < 3; ++r) np[r] = ((ck * s[r]) + (a * g[r])) / (ck + a);
            float oc[3], nc[3], avg[3];
            decode_color(s[4], oc);
            decode_color(g[4], nc);
            for (int r = 0; r < 3; ++r) avg[r] = ((ck * oc[r]) + (a * nc[r])) / (ck + a);
            float nn[4];
            for (int r = 0; r < 4; ++r) nn[r] = ((ck * s[8 + r]) + (a * g[8 + r])) / (ck + a);
            normalize3(nn);
            s[0] = np[0];
            s[1] = np[1];
            s[2] = np[2];
            s[3] = ck + a;
            s[4] = encode_color(avg);
            s[7] = tm;
            memcpy(s + 8, nn, 16);
        } else {
            s[3] = ck + a;
            s[7] = tm;
        }
    }
    free(pix);
    free(claim);
    free(dr);
    free(df);
}

/* --------------------------------------------------------------- clean */

/* copy_unstable.vert:48-102 (no deformation nodes); 1 = keep */
static int keep_surfel(const orc_model* m, float* s, const float Ti[16], int time, float conf_threshold,
                       int time_delta)
{
    const int w = m->w, h = m->h;
    const float cols = (float)w, rows = (float)h, scale = 1.0f, tm = (float)time;
    int test = 1;
    float lp[3], ln[3];
    xform_point(Ti, s, lp);
    const float x = ((m->fx * lp[0]) / lp[2]) + m->cx;
    const float y = ((m->fy * lp[1]) / lp[2]) + m->cy;
    xform_dir(Ti, s + 8, ln);
    normalize3(ln);
    const float sx = (1.0f / (cols * scale)) * 0.5f;
    const float sy = (1.0f / (rows * scale)) * 0.5f;
    const float wm = 2.0f;
    int count = 0, zcount = 0;
    if (tm - s[7] < (float)time_delta && lp[2] > 0.0f && x > 0.0f && y > 0.0f && x < cols && y < rows) {
        for (float ii = x / cols - ((scale * sx) * wm); ii < x / cols + ((scale * sx) * wm); ii += sx)
            for (float jj = y / rows - ((scale * sy) * wm); jj < y / rows + ((scale * sy) * wm); jj += sy) {
                const size_t t = (size_t)texel(jj, h) * w + texel(ii, w);
                if (m->index[t] == 0) continue;
                const float* vc = m->vert_conf + 4 * t;
                const float* ct = m->color_time + 4 * t;
                const float dx = vc[0] - lp[0], dy = vc[1] - lp[1];
                if (ct[2] < s[6] && vc[3] > conf_threshold && vc[2] > lp[2] && vc[2] - lp[2] < 0.01f &&
                    sqrtf(dx * dx + dy * dy) < s[11] * 1.4f)
                    count++;
                if (ct[3] == tm && vc[3] > conf_threshold && vc[2] > lp[2] && vc[2] - lp[2] > 0.01f &&
                    fabsf(ln[2]) > 0.85f)
                    zcount++;
            }
    }
    if (count > 8 || zcount > 4) test = 0;
    if (s[7] == -2.0f) s[7] = tm; /* new unstable point */
    if (s[7] == -1.0f || ((tm - s[7]) > 20.0f && s[3] < conf_threshold)) test = 0;
    if (s[7] > 0.0f && tm - s[7] > (float)time_delta) test = 1;
    return test;
}

int orc_model_clean(orc_model* m, const float pose[16], int time, float conf_threshold, int time_delta)
{
    float Ti[16];
    rigid_inverse(pose, Ti);
    int out = 0;
    for (int pass = 0; pass < 2; ++pass) {
        const int n = pass == 0 ? m->count : m->fresh_count;
        for (int k = 0; k < n; ++k) {
            float s[12];
            memcpy(s, (pass == 0 ? m->model : m->fresh) + F(k), 48);
            if (keep_surfel(m, s, Ti, time, conf_threshold, time_delta) && out < m->max_surfels)
                memcpy(m->scratch + F(out++), s, 48);
        }
    }
    float* t = m->model;
    m->model = m->scratch;
    m->scratch = t;
    m->count = out;
    m->fresh_count = 0;
    return out;
}
