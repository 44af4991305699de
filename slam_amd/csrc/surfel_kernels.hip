// surfel_kernels.hip -- the surfel model's fusion on gfx950, behind
// include/slam_surfel.h.
//
// The reference keeps the model in GL vertex buffers and updates it with
// transform feedback and rasterized points (Model.cpp, IndexMap.cpp and their
// shaders).  Three GL mechanisms carry the semantics, and each becomes plain
// device code here:
//   - points drawn with the depth test on (the index map): a 64-bit atomic
//     minimum of (24-bit depth << 32 | surfel index) per pixel, then a
//     per-pixel resolve that recomputes the winner's outputs;
//   - points drawn at a surfel's texel of the update maps, where the depth
//     test at equal depth keeps the first pixel drawn: an atomic minimum of
//     the pixel's draw index per surfel;
//   - transform feedback, which appends in draw order: flags, an exclusive
//     scan in draw order, and a scatter.
// Draw order is the reference's: pixels column by column (its uv buffers),
// surfels by index.  The per-pixel maps of this file (metric depth, index map,
// per-pixel fusion data) are stored in that order too, at d = x * height + y:
// per-pixel kernels run over d, and a wave of surfels, which the model holds
// mostly in draw order, reads its neighbourhoods from a few contiguous runs
// instead of one cache line per lane.
//
// Arithmetic is FP32 with -ffp-contract=off in the shaders' operation order.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <vector>

#include "host_wait.h"
#include "odom_math.h"
#include "slam_surfel.h"
#include "timing_events.h"

using slam::bexpf;

namespace {

constexpr unsigned kDepthClear = 0xFFFFFFu;
constexpr unsigned long long kKeyClear = ((unsigned long long)kDepthClear << 32) | 0xFFFFFFFFull;
constexpr unsigned kNoClaim = 0xFFFFFFFFu;
constexpr int kScanItems = 2048;  // per scan block: 256 threads x 8
constexpr int kMaxItems = 1 << 23;

struct Img {
    int w, h;
    float cols, rows;
};

struct Cam4 {
    float cx, cy, a, b;  // the shaders' cam uniform: (cx, cy, 1/fx, 1/fy) or (cx, cy, fx, fy)
};

struct Mat {
    float m[12];  // rows of [R | t]
};

__device__ __forceinline__ void xpoint(const Mat& M, float x, float y, float z, float* o)
{
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = ((M.m[4 * r] * x + M.m[4 * r + 1] * y) + M.m[4 * r + 2] * z) + M.m[4 * r + 3];
}

__device__ __forceinline__ void xdir(const Mat& M, float x, float y, float z, float* o)
{
#pragma unroll
    for (int r = 0; r < 3; ++r) o[r] = (M.m[4 * r] * x + M.m[4 * r + 1] * y) + M.m[4 * r + 2] * z;
}

__device__ __forceinline__ void normalize3(float* v)
{
    const float inv = 1.0f / sqrtf((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
    v[0] *= inv;
    v[1] *= inv;
    v[2] *= inv;
}

__device__ __forceinline__ float dot3(const float* a, const float* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

// Texture coordinate of pixel i of n as the reference's uv buffers store it:
// (float)i / n in FP32 plus 1.0 / (2 * (float)n) in FP64, rounded to FP32.
__device__ __forceinline__ float pixel_tc(int i, int n)
{
    return (float)((double)((float)i / (float)n) + 1.0 / (double)(2.0f * (float)n));
}

// Nearest sampling: texel floor(u * n), clamped to the edge.
__device__ __forceinline__ int texel(float u, int n)
{
    const float t = floorf(u * (float)n);
    return !(t >= 0.0f) ? 0 : (t > (float)(n - 1) ? n - 1 : (int)t);
}

// Address of pixel (x, y) in a map stored in draw order.
__device__ __forceinline__ size_t cm(const Img& I, int x, int y) { return (size_t)x * I.h + y; }

__device__ __forceinline__ float fetch(const float* img, const Img& I, float u, float v)
{
    return img[cm(I, texel(u, I.w), texel(v, I.h))];
}

// The texels the shaders' window loops, for (float s = lo; s < hi; s += step),
// sample along one axis: a base texel and the repetitions of it and the three
// after it (samples are half a texel apart, so the texels they hit rise by at
// most one per step).  Visiting these texels once each, in ascending order,
// with their repetitions as weights, gives what the loop gives: a repeated
// sample sees the same texel.  false (walk the loop itself) if the loop
// would run more than 64 times or reach a fifth texel.
struct Runs {
    int t0;
    int m[4];
};

__device__ __forceinline__ bool window_runs(float lo, float hi, float step, int size, Runs& R)
{
    R.t0 = texel(lo, size);
    R.m[0] = R.m[1] = R.m[2] = R.m[3] = 0;
    int k = 0;
    for (float s = lo; s < hi; s += step) {
        if (++k > 64) return false;
        const int r = texel(s, size) - R.t0;
        if (r < 0 || r > 3) return false;
        R.m[0] += r == 0;
        R.m[1] += r == 1;
        R.m[2] += r == 2;
        R.m[3] += r == 3;
    }
    return true;
}

// geometry.glsl:21-25
__device__ __forceinline__ void vertex_at(const float* depth, const Img& I, float u, float v, float x, float y,
                                          const Cam4& c, float* o)
{
    const float z = fetch(depth, I, u, v);
    o[0] = ((x - c.cx) * z) * c.a;
    o[1] = ((y - c.cy) * z) * c.b;
    o[2] = z;
}

// geometry.glsl:28-42: central differences, the neighbours averaged with
// the centre first
__device__ __forceinline__ void normal_at(const float* depth, const Img& I, const float* vp, float u, float v,
                                          float x, float y, const Cam4& c, float* n)
{
    float xf[3], xb[3], yf[3], yb[3];
    vertex_at(depth, I, u + (1.0f / I.cols), v, x + 1.0f, y, c, xf);
    vertex_at(depth, I, u - (1.0f / I.cols), v, x - 1.0f, y, c, xb);
    vertex_at(depth, I, u, v + (1.0f / I.rows), x, y + 1.0f, c, yf);
    vertex_at(depth, I, u, v - (1.0f / I.rows), x, y - 1.0f, c, yb);
    float dx[3], dy[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dx[k] = ((xb[k] + vp[k]) / 2.0f) - ((xf[k] + vp[k]) / 2.0f);
        dy[k] = ((yb[k] + vp[k]) / 2.0f) - ((yf[k] + vp[k]) / 2.0f);
    }
    n[0] = dx[1] * dy[2] - dx[2] * dy[1];
    n[1] = dx[2] * dy[0] - dx[0] * dy[2];
    n[2] = dx[0] * dy[1] - dx[1] * dy[0];
    normalize3(n);
}

// surfels.glsl:19-34
__device__ __forceinline__ float radius_of(float depth, float nz, const Cam4& c)
{
    const float mean_focal = ((1.0f / fabsf(c.a)) + (1.0f / fabsf(c.b))) / 2.0f;
    const float r = (depth / mean_focal) * 1.41421356237f;
    const float rn = r / fabsf(nz);
    const float r2 = 2.0f * r;
    return rn < r2 ? rn : r2;
}

// surfels.glsl:36-46
__device__ __forceinline__ float confidence_of(float x, float y, float weighting, const Cam4& c)
{
    const float px = x - c.cx, py = y - c.cy;
    const float rd = sqrtf(px * px + py * py) / 400.0f;
    return bexpf((-(rd * rd)) / 0.72f) * weighting;
}

__device__ __forceinline__ float encode_rgb8(const uint8_t* c)
{
    return (float)(((int)c[0] << 16) | ((int)c[1] << 8) | (int)c[2]);
}

__device__ __forceinline__ void decode_color(float c, float* o)
{
    const int i = (int)c;
    o[0] = (float)((i >> 16) & 0xFF) / 255.0f;
    o[1] = (float)((i >> 8) & 0xFF) / 255.0f;
    o[2] = (float)(i & 0xFF) / 255.0f;
}

__device__ __forceinline__ float encode_color(const float* c)
{
    int rgb = (int)roundf(c[0] * 255.0f);
    rgb = (rgb << 8) + (int)roundf(c[1] * 255.0f);
    rgb = (rgb << 8) + (int)roundf(c[2] * 255.0f);
    return (float)rgb;
}

__device__ __forceinline__ void load12(const float* s, float* o)
{
    const float4* p = reinterpret_cast<const float4*>(s);
    const float4 a = p[0], b = p[1], c = p[2];
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    o[4] = b.x; o[5] = b.y; o[6] = b.z; o[7] = b.w;
    o[8] = c.x; o[9] = c.y; o[10] = c.z; o[11] = c.w;
}

__device__ __forceinline__ void store12(float* s, const float* o)
{
    float4* p = reinterpret_cast<float4*>(s);
    p[0] = make_float4(o[0], o[1], o[2], o[3]);
    p[1] = make_float4(o[4], o[5], o[6], o[7]);
    p[2] = make_float4(o[8], o[9], o[10], o[11]);
}

// ----------------------------------------------------------------- scan

// Block-exclusive scan of 2048 flags (8 per thread); per-item prefixes and
// the block total.
__global__ __launch_bounds__(256) void k_scan_local(const int* __restrict__ flags, int n, int* __restrict__ prefix,
                                                    int* __restrict__ block_sums)
{
    __shared__ int wsum[4];
    const int base = blockIdx.x * kScanItems + threadIdx.x * 8;
    int v[8], s = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        v[k] = (base + k < n) ? flags[base + k] : 0;
        s += v[k];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int off = 0;
    for (int q = 0; q < wv; ++q) off += wsum[q];
    int run = off + incl - s;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        if (base + k < n) prefix[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 255) block_sums[blockIdx.x] = off + incl;
}

// Exclusive scan of up to 4096 block totals in one block; the grand total.
__global__ __launch_bounds__(1024) void k_scan_top(int* block_sums, int nb, int* total)
{
    __shared__ int wsum[16];
    const int base = threadIdx.x * 4;
    int v[4], s = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[k] = (base + k < nb) ? block_sums[base + k] : 0;
        s += v[k];
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int incl = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int t = __shfl_up(incl, d, 64);
        if (lane >= d) incl += t;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int off = 0;
    for (int q = 0; q < wv; ++q) off += wsum[q];
    int run = off + incl - s;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (base + k < nb) block_sums[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 1023) *total = off + incl;
}

__device__ __forceinline__ int scan_pos(const int* prefix, const int* block_sums, int i)
{
    return prefix[i] + block_sums[i / kScanItems];
}

// ------------------------------------------------------------ metric depth

// depth_metric.frag:29-37 for the raw and the filtered frame.
__global__ __launch_bounds__(256) void k_metric2(const uint8_t* raw, int rpitch, const uint8_t* filt, int fpitch,
                                                 float maxD, Img I, float* dr, float* df)
{
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= I.w * I.h) return;
    const int x = d / I.h, y = d - x * I.h;
    const unsigned lim = (unsigned)(maxD * 1000.0f);
    const unsigned a = *reinterpret_cast<const uint16_t*>(raw + (size_t)y * rpitch + 2 * (size_t)x);
    const unsigned b = *reinterpret_cast<const uint16_t*>(filt + (size_t)y * fpitch + 2 * (size_t)x);
    dr[d] = (a > lim || a < 300u) ? 0.0f : (float)a / 1000.0f;
    df[d] = (b > lim || b < 300u) ? 0.0f : (float)b / 1000.0f;
}

// ---------------------------------------------------------- initialise

struct RGBView {
    const uint8_t* p;
    int pitch;
};

__device__ __forceinline__ const uint8_t* rgb_at(const RGBView& c, const Img& I, float u, float v)
{
    return c.p + (size_t)texel(v, I.h) * c.pitch + 3 * (size_t)texel(u, I.w);
}

// vertex_feedback.vert:38-68 on the raw and the filtered frame; flags and
// items keyed by draw index.
__global__ __launch_bounds__(256) void k_feedback(const float* dr, const float* df, RGBView rgb, Img I, Cam4 c,
                                                  int time, float maxDepth, float* raw_items, float* filt_norm,
                                                  int* flag_raw, int* flag_filt)
{
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= I.w * I.h) return;
    const int i = d / I.h, j = d - i * I.h;
    const float u = pixel_tc(i, I.w), v = pixel_tc(j, I.h);
    const float x = u * I.cols, y = v * I.rows;
    const uint8_t* col = rgb_at(rgb, I, u, v);
    for (int pass = 0; pass < 2; ++pass) {
        const float* depth = pass == 0 ? dr : df;
        float vp[3], n[3];
        vertex_at(depth, I, u, v, x, y, c, vp);
        normal_at(depth, I, vp, u, v, x, y, c, n);
        const bool ok = !(vp[2] <= 0.0f || vp[2] > maxDepth);
        const float rad = radius_of(vp[2], n[2], c);
        if (pass == 0) {
            flag_raw[d] = ok ? 1 : 0;
            // init_unstable.vert: colour y = 0, init time z = 1
            const float o[12] = {vp[0], vp[1], vp[2], confidence_of(x, y, 1.0f, c), encode_rgb8(col), 0.0f, 1.0f,
                                 (float)time, 0.0f, 0.0f, 0.0f, 0.0f};
            store12(raw_items + 12 * (size_t)d, o);
        } else {
            flag_filt[d] = ok ? 1 : 0;
            reinterpret_cast<float4*>(filt_norm)[d] = make_float4(n[0], n[1], n[2], rad);
        }
    }
}

__global__ __launch_bounds__(256) void k_init_scatter(const float* raw_items, const float* filt_norm,
                                                      const int* flag_raw, const int* flag_filt, const int* pre_raw,
                                                      const int* bs_raw, const int* pre_filt, const int* bs_filt,
                                                      const int* total_raw, Img I, int max_surfels, float* model)
{
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= I.w * I.h) return;
    const int nr = min(*total_raw, max_surfels);
    if (flag_raw[d]) {
        const int k = scan_pos(pre_raw, bs_raw, d);
        if (k < nr) {
            float o[12];
            load12(raw_items + 12 * (size_t)d, o);
            // normal and radius: the k-th filtered vertex's, zero if there is none
            o[8] = o[9] = o[10] = o[11] = 0.0f;
            store12(model + 12 * (size_t)k, o);
        }
    }
}

// The filtered buffer's k-th normal goes to the k-th raw surfel (the two
// feedback buffers are drawn side by side); a separate launch orders it
// after the raw scatter.
__global__ __launch_bounds__(256) void k_init_normals(const float* filt_norm, const int* flag_filt,
                                                      const int* pre_filt, const int* bs_filt, const int* total_raw,
                                                      Img I, int max_surfels, float* model)
{
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= I.w * I.h || !flag_filt[d]) return;
    const int nr = min(*total_raw, max_surfels);
    const int k = scan_pos(pre_filt, bs_filt, d);
    if (k < nr) reinterpret_cast<float4*>(model + 12 * (size_t)k)[2] = reinterpret_cast<const float4*>(filt_norm)[d];
}

// ------------------------------------------------------------ index map

struct IndexArgs {
    const float* model;
    int count;
    Mat Ti;
    float fx, fy, cx, cy;
    float time, maxDepth, timeDelta;
    Img I;
};

// index_map.vert:37-52 for one surfel: camera-frame position; false if
// dropped or clipped.
__device__ __forceinline__ bool index_stage(const IndexArgs& A, const float* s, float* hp, float& xw, float& yw)
{
    xpoint(A.Ti, s[0], s[1], s[2], hp);
    if (hp[2] > A.maxDepth || hp[2] < 0.0f || A.time - s[7] > A.timeDelta) return false;
    const float hc = A.I.cols * 0.5f, hr = A.I.rows * 0.5f;
    const float ndx = ((((A.fx * hp[0]) / hp[2]) + A.cx) - hc) / hc;
    const float ndy = ((((A.fy * hp[1]) / hp[2]) + A.cy) - hr) / hr;
    if (!(ndx >= -1.0f && ndx <= 1.0f && ndy >= -1.0f && ndy <= 1.0f)) return false;
    xw = ndx * hc + hc;
    yw = ndy * hr + hr;
    return true;
}

// The pixel whose centre lies in [c - 0.5, c + 0.5), or -1.
__device__ __forceinline__ int point_pixel(float c, int n)
{
    const int i = (int)ceilf(c - 1.0f);   // smallest i with i + 0.5 >= c - 0.5
    int r = -1;
#pragma unroll
    for (int k = i - 1; k <= i + 1; ++k) {
        const float ctr = (float)k + 0.5f;
        if (k >= 0 && k < n && c - 0.5f <= ctr && ctr < c + 0.5f) r = k;
    }
    return r;
}

__device__ __forceinline__ unsigned index_depth(const IndexArgs& A, float z)
{
    float d = (z / A.maxDepth) * 0.5f + 0.5f;
    d = d < 0.0f ? 0.0f : (d > 1.0f ? 1.0f : d);
    return (unsigned)rintf(d * 16777215.0f);
}

__global__ __launch_bounds__(256) void k_index_raster(IndexArgs A, unsigned long long* keys)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= A.count) return;
    float s[12], hp[3], xw, yw;
    load12(A.model + 12 * (size_t)k, s);
    if (!index_stage(A, s, hp, xw, yw)) return;
    const int px = point_pixel(xw, A.I.w), py = point_pixel(yw, A.I.h);
    if (px < 0 || py < 0) return;
    const unsigned q = index_depth(A, hp[2]);
    if (q >= kDepthClear) return;
    atomicMin(keys + cm(A.I, px, py), ((unsigned long long)q << 32) | (unsigned)k);
}

__global__ __launch_bounds__(256) void k_index_resolve(IndexArgs A, unsigned long long* keys, uint32_t* index,
                                                       float4* vert_conf, float4* color_time, float4* norm_rad)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= A.I.w * A.I.h) return;
    const unsigned long long key = keys[p];
    keys[p] = kKeyClear;
    uint32_t id = 0;
    float4 vc = make_float4(0.f, 0.f, 0.f, 0.f), ct = vc, nr = vc;
    if ((unsigned)(key >> 32) < kDepthClear) {
        const unsigned k = (unsigned)key;
        float s[12], hp[3], nn[3];
        load12(A.model + 12 * (size_t)k, s);
        xpoint(A.Ti, s[0], s[1], s[2], hp);
        xdir(A.Ti, s[8], s[9], s[10], nn);
        normalize3(nn);
        id = k;
        vc = make_float4(hp[0], hp[1], hp[2], s[3]);
        ct = make_float4(s[4], s[5], s[6], s[7]);
        nr = make_float4(nn[0], nn[1], nn[2], s[11]);
    }
    index[p] = id;
    vert_conf[p] = vc;
    color_time[p] = ct;
    norm_rad[p] = nr;
}

// ---------------------------------------------------------------- fuse

struct DataArgs {
    const float* dr;
    const float* df;
    RGBView rgb;
    Img I;
    Cam4 c;                 // (cx, cy, 1.0 / fx, 1.0 / fy): 1.0 / f in FP64, stored as FP32
    Mat pose;
    float time, maxDepth, weighting;
    const uint32_t* index;
    const float4* vert_conf;
    const float4* norm_rad;
    float* pix;             // 12 floats per pixel (draw order)
    int* flag_fresh;        // by draw index
    unsigned* claim;        // per surfel: first draw index that matched it
    int count;              // surfels in the model the claims refer to
};

// data.vert:67-190 for one pixel.
__global__ __launch_bounds__(256) void k_data(DataArgs A)
{
    const Img& I = A.I;
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= I.w * I.h) return;
    const int i = d / I.h, j = d - i * I.h;
    const float u = pixel_tc(i, I.w), v = pixel_tc(j, I.h);
    const float x = u * I.cols, y = v * I.rows;
    const Cam4& c = A.c;
    float pl[3], pf[3], nl[3], o[12];
    vertex_at(A.dr, I, u, v, x, y, c, pl);
    xpoint(A.pose, pl[0], pl[1], pl[2], o);
    vertex_at(A.df, I, u, v, x, y, c, pf);
    normal_at(A.df, I, pf, u, v, x, y, c, nl);
    xdir(A.pose, nl[0], nl[1], nl[2], o + 8);
    o[3] = confidence_of(x, y, A.weighting, c);
    o[4] = encode_rgb8(rgb_at(A.rgb, I, u, v));
    o[5] = 0.0f;
    o[6] = A.time;
    o[7] = 0.0f;
    o[11] = radius_of(pf[2], nl[2], c);
    int update = 0;
    unsigned best = 0;
    const int tpar = (int)A.time % 2;
    const bool nb = fetch(A.dr, I, u - (1.0f / I.cols), v) != 0.0f && fetch(A.dr, I, u, v - (1.0f / I.rows)) != 0.0f &&
                    fetch(A.dr, I, u + (1.0f / I.cols), v) != 0.0f && fetch(A.dr, I, u, v + (1.0f / I.rows)) != 0.0f;
    if ((int)x % 2 == tpar && (int)y % 2 == tpar && nb && pl[2] > 0.0f && pl[2] <= A.maxDepth) {
        int counter = 0;
        const float scale = 1.0f, wm = 2.0f;
        const float sx = (1.0f / (I.cols * scale)) * 0.5f;
        const float sy = (1.0f / (I.rows * scale)) * 0.5f;
        float best_dist = 1000.0f;
        const float xl = (x - c.cx) * c.a;
        const float yl = (y - c.cy) * c.b;
        const float lambda = sqrtf(xl * xl + yl * yl + 1.0f);
        const float ray[3] = {xl, yl, 1.0f};
        const float ray_len = sqrtf(dot3(ray, ray));
        const float u0 = u - ((scale * sx) * wm), u1 = u + ((scale * sx) * wm);
        const float v0 = v - ((scale * sy) * wm), v1 = v + ((scale * sy) * wm);
        // one candidate texel: the nearest ray wins, the first visited on ties
        auto visit = [&](size_t t) {
            const uint32_t cur = A.index[t];
            if (cur == 0) return;
            const float4 vc = A.vert_conf[t];
            if (!(fabsf((vc.z * lambda) - (pl[2] * lambda)) < 0.05f)) return;
            const float cr[3] = {ray[1] * vc.z - ray[2] * vc.y, ray[2] * vc.x - ray[0] * vc.z,
                                 ray[0] * vc.y - ray[1] * vc.x};
            const float dist = sqrtf(dot3(cr, cr)) / ray_len;
            const float4 nr4 = A.norm_rad[t];
            const float nr[3] = {nr4.x, nr4.y, nr4.z};
            // acos(c) < 0.5 as c > cos(0.5)
            const float cosang = dot3(nr, nl) / (sqrtf(dot3(nr, nr)) * sqrtf(dot3(nl, nl)));
            if (dist < best_dist && (fabsf(nr[2]) < 0.75f || cosang > 0.87758255f)) {
                counter++;
                best_dist = dist;
                best = cur;
            }
        };
        Runs rx, ry;
        const bool runs = window_runs(u0, u1, sx, I.w, rx) && window_runs(v0, v1, sy, I.h, ry);
        // a repeated sample revisits a texel with the same distance, which
        // never wins again: each texel once, in the loop's order
        if (runs && !rx.m[3] && !ry.m[3]) {
            // the usual 3 x 3 texels: every load issued before the first use
            // (unused slots read a clamped, valid texel and are skipped)
            uint32_t id[9];
            float4 vcs[9], nrs[9];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const size_t q = cm(I, min(rx.t0 + a, I.w - 1), min(ry.t0 + b, I.h - 1));
                    id[3 * a + b] = A.index[q];
                    vcs[3 * a + b] = A.vert_conf[q];
                    nrs[3 * a + b] = A.norm_rad[q];
                }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const uint32_t cur = id[3 * a + b];
                    if (!rx.m[a] || !ry.m[b] || cur == 0) continue;
                    const float4 vc = vcs[3 * a + b];
                    if (!(fabsf((vc.z * lambda) - (pl[2] * lambda)) < 0.05f)) continue;
                    const float cr[3] = {ray[1] * vc.z - ray[2] * vc.y, ray[2] * vc.x - ray[0] * vc.z,
                                         ray[0] * vc.y - ray[1] * vc.x};
                    const float dist = sqrtf(dot3(cr, cr)) / ray_len;
                    const float nr[3] = {nrs[3 * a + b].x, nrs[3 * a + b].y, nrs[3 * a + b].z};
                    const float cosang = dot3(nr, nl) / (sqrtf(dot3(nr, nr)) * sqrtf(dot3(nl, nl)));
                    if (dist < best_dist && (fabsf(nr[2]) < 0.75f || cosang > 0.87758255f)) {
                        counter++;
                        best_dist = dist;
                        best = cur;
                    }
                }
        } else if (runs) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (rx.m[a] && ry.m[b]) visit(cm(I, rx.t0 + a, ry.t0 + b));
        } else {
            for (float ii = u0; ii < u1; ii += sx)
                for (float jj = v0; jj < v1; jj += sy) visit(cm(I, texel(ii, I.w), texel(jj, I.h)));
        }
        update = counter > 0 ? 1 : 2;
        o[7] = counter > 0 ? -1.0f : -2.0f;
    }
    store12(A.pix + 12 * (size_t)d, o);
    A.flag_fresh[d] = update > 0 ? 1 : 0;
    // an index map predicted from a larger model can name surfels that are gone
    if (update == 1 && best < (unsigned)A.count) atomicMin(A.claim + best, (unsigned)d);
}

// update.vert:37-90 for the surfels a pixel claimed; clears the claims.
__global__ __launch_bounds__(256) void k_update(float* model, int count, const float* pix, unsigned* claim, Img I,
                                                float time)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    if (k >= count) return;
    const unsigned d = claim[k];
    if (d == kNoClaim) return;
    claim[k] = kNoClaim;
    float s[12], g[12];
    load12(model + 12 * (size_t)k, s);
    load12(pix + 12 * (size_t)d, g);
    const float ck = s[3], a = g[3];
    if (g[11] < (1.0f + 0.5f) * s[11]) {
        float np[3], oc[3], nc[3], avg[3], nn[4];
#pragma unroll
        for (int r = 0; r < 3; ++r) np[r] = ((ck * s[r]) + (a * g[r])) / (ck + a);
        decode_color(s[4], oc);
        decode_color(g[4], nc);
#pragma unroll
        for (int r = 0; r < 3; ++r) avg[r] = ((ck * oc[r]) + (a * nc[r])) / (ck + a);
#pragma unroll
        for (int r = 0; r < 4; ++r) nn[r] = ((ck * s[8 + r]) + (a * g[8 + r])) / (ck + a);
        normalize3(nn);
        s[0] = np[0];
        s[1] = np[1];
        s[2] = np[2];
        s[3] = ck + a;
        s[4] = encode_color(avg);
        s[7] = time;
        s[8] = nn[0];
        s[9] = nn[1];
        s[10] = nn[2];
        s[11] = nn[3];
    } else {
        s[3] = ck + a;
        s[7] = time;
    }
    store12(model + 12 * (size_t)k, s);
}

// data.geom's transform feedback: every pixel with an update or a new point,
// in draw order.
__global__ __launch_bounds__(256) void k_fresh_scatter(const float* pix, const int* flags, const int* prefix,
                                                       const int* block_sums, Img I, float* fresh)
{
    const int d = blockIdx.x * 256 + threadIdx.x;
    if (d >= I.w * I.h || !flags[d]) return;
    float o[12];
    load12(pix + 12 * (size_t)d, o);
    store12(fresh + 12 * (size_t)scan_pos(prefix, block_sums, d), o);
}

// --------------------------------------------------------------- clean

struct CleanArgs {
    const float* model;
    int count;
    const float* fresh;
    const int* nfresh;
    Mat Ti;
    float fx, fy, cx, cy;
    Img I;
    float time, confThreshold, timeDelta;
    const uint32_t* index;
    const float4* vert_conf;
    const float4* color_time;
    float* items;           // the surfel as the cleaning leaves it
    int* flags;
};

// copy_unstable.vert:48-102 (no deformation nodes).
__global__ __launch_bounds__(256) void k_clean(CleanArgs A)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int nf = *A.nfresh;
    if (t >= A.count + A.I.w * A.I.h) return;
    if (t >= A.count + nf) {
        A.flags[t] = 0;
        return;
    }
    float s[12];
    load12(t < A.count ? A.model + 12 * (size_t)t : A.fresh + 12 * (size_t)(t - A.count), s);
    const Img& I = A.I;
    float lp[3], ln[3];
    xpoint(A.Ti, s[0], s[1], s[2], lp);
    const float x = ((A.fx * lp[0]) / lp[2]) + A.cx;
    const float y = ((A.fy * lp[1]) / lp[2]) + A.cy;
    xdir(A.Ti, s[8], s[9], s[10], ln);
    normalize3(ln);
    const float scale = 1.0f, wm = 2.0f;
    const float sx = (1.0f / (I.cols * scale)) * 0.5f;
    const float sy = (1.0f / (I.rows * scale)) * 0.5f;
    int count = 0, zcount = 0;
    const float tm = A.time;
    if (tm - s[7] < A.timeDelta && lp[2] > 0.0f && x > 0.0f && y > 0.0f && x < I.cols && y < I.rows) {
        const float u0 = x / I.cols - ((scale * sx) * wm), u1 = x / I.cols + ((scale * sx) * wm);
        const float v0 = y / I.rows - ((scale * sy) * wm), v1 = y / I.rows + ((scale * sy) * wm);
        // one texel, counted as often as the loop samples it
        auto visit = [&](size_t q, int times) {
            if (A.index[q] == 0) return;
            const float4 vc = A.vert_conf[q];
            const float4 ct = A.color_time[q];
            const float dx = vc.x - lp[0], dy = vc.y - lp[1];
            if (ct.z < s[6] && vc.w > A.confThreshold && vc.z > lp[2] && vc.z - lp[2] < 0.01f &&
                sqrtf(dx * dx + dy * dy) < s[11] * 1.4f)
                count += times;
            if (ct.w == tm && vc.w > A.confThreshold && vc.z > lp[2] && vc.z - lp[2] > 0.01f &&
                fabsf(ln[2]) > 0.85f)
                zcount += times;
        };
        Runs rx, ry;
        const bool runs = window_runs(u0, u1, sx, I.w, rx) && window_runs(v0, v1, sy, I.h, ry);
        if (runs && !rx.m[3] && !ry.m[3]) {
            // the usual 3 x 3 texels: every load issued before the first use
            // (unused slots read a clamped, valid texel and weigh 0)
            uint32_t id[9];
            float4 vcs[9], cts[9];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const size_t q = cm(I, min(rx.t0 + a, I.w - 1), min(ry.t0 + b, I.h - 1));
                    id[3 * a + b] = A.index[q];
                    vcs[3 * a + b] = A.vert_conf[q];
                    cts[3 * a + b] = A.color_time[q];
                }
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    const int times = rx.m[a] * ry.m[b];
                    if (!times || id[3 * a + b] == 0) continue;
                    const float4 vc = vcs[3 * a + b], ct = cts[3 * a + b];
                    const float dx = vc.x - lp[0], dy = vc.y - lp[1];
                    if (ct.z < s[6] && vc.w > A.confThreshold && vc.z > lp[2] && vc.z - lp[2] < 0.01f &&
                        sqrtf(dx * dx + dy * dy) < s[11] * 1.4f)
                        count += times;
                    if (ct.w == tm && vc.w > A.confThreshold && vc.z > lp[2] && vc.z - lp[2] > 0.01f &&
                        fabsf(ln[2]) > 0.85f)
                        zcount += times;
                }
        } else if (runs) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (rx.m[a] && ry.m[b]) visit(cm(I, rx.t0 + a, ry.t0 + b), rx.m[a] * ry.m[b]);
        } else {
            for (float ii = u0; ii < u1; ii += sx)
                for (float jj = v0; jj < v1; jj += sy) visit(cm(I, texel(ii, I.w), texel(jj, I.h)), 1);
        }
    }
    int test = 1;
    if (count > 8 || zcount > 4) test = 0;
    if (s[7] == -2.0f) s[7] = tm;
    if (s[7] == -1.0f || ((tm - s[7]) > 20.0f && s[3] < A.confThreshold)) test = 0;
    if (s[7] > 0.0f && tm - s[7] > A.timeDelta) test = 1;
    store12(A.items + 12 * (size_t)t, s);
    A.flags[t] = test;
}

__global__ __launch_bounds__(256) void k_clean_scatter(const float* items, const int* flags, const int* prefix,
                                                       const int* block_sums, int n, int max_surfels, float* out)
{
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= n || !flags[t]) return;
    const int k = scan_pos(prefix, block_sums, t);
    if (k >= max_surfels) return;
    float s[12];
    load12(items + 12 * (size_t)t, s);
    store12(out + 12 * (size_t)k, s);
}

__global__ __launch_bounds__(256) void k_fill_u64(unsigned long long* p, int n, unsigned long long v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_u32(unsigned* p, int n, unsigned v)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }

}  // namespace

struct surfel_ctx {
    Img I{};
    float cx = 0, cy = 0, fx = 0, fy = 0;
    int max_surfels = 0, count = 0;
    hipStream_t own = nullptr, stream = nullptr;
    float* model[2] = {nullptr, nullptr};
    int cur = 0;
    float *dr = nullptr, *df = nullptr;
    void *raw_stage = nullptr, *filt_stage = nullptr, *rgb_stage = nullptr;
    // index map
    unsigned long long* keys = nullptr;
    uint32_t* index = nullptr;
    float4 *vert_conf = nullptr, *color_time = nullptr, *norm_rad = nullptr;
    // fusion
    float* pix = nullptr;
    float* fresh = nullptr;
    unsigned* claim = nullptr;
    float* items = nullptr;           // max(max_surfels + npix) surfels: cleaning output before compaction
    float* feed_norm = nullptr;       // npix float4
    int *flags = nullptr, *flags2 = nullptr;
    int *prefix = nullptr, *prefix2 = nullptr;
    int *bsums = nullptr, *bsums2 = nullptr;
    int* totals = nullptr;            // [0] scan total, [1] second scan total
    int* h_count = nullptr;           // pinned landing buffer of the count readback
};

#define SCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) {                                                                      \
            fprintf(stderr, "slam_surfel: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                       \
            return ODOM_ERR_HIP;                                                                     \
        }                                                                                            \
    } while (0)

namespace {

void free_ctx(surfel_ctx* c)
{
    void* b[] = {c->model[0], c->model[1], c->dr, c->df, c->raw_stage, c->filt_stage, c->rgb_stage, c->keys,
                 c->index, c->vert_conf, c->color_time, c->norm_rad, c->pix, c->fresh, c->claim, c->items,
                 c->feed_norm, c->flags, c->flags2, c->prefix, c->prefix2, c->bsums, c->bsums2, c->totals};
    for (void* p : b)
        if (p) (void)hipFree(p);
    if (c->h_count) (void)hipHostFree(c->h_count);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

Mat rigid(const float pose[16])
{
    Mat M;
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 4; ++k) M.m[4 * r + k] = pose[4 * r + k];
    return M;
}

Mat rigid_inverse(const float pose[16])
{
    Mat M;
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) M.m[4 * r + k] = pose[4 * k + r];
        M.m[4 * r + 3] = -((M.m[4 * r] * pose[3] + M.m[4 * r + 1] * pose[7]) + M.m[4 * r + 2] * pose[11]);
    }
    return M;
}

int view_dev(surfel_ctx* c, const odom_frame_view* v, int fmt, void* stage, const uint8_t** d, int* pitch)
{
    if (!v || !v->ptr || v->format != fmt || v->width != c->I.w || v->height != c->I.h) return ODOM_ERR_ARG;
    const int bpp = fmt == ODOM_FMT_DEPTH_U16 ? 2 : 3;
    const int row = v->width * bpp;
    const int p = v->pitch_bytes ? v->pitch_bytes : row;
    if (p < row || (bpp == 2 && (p % 2) != 0)) return ODOM_ERR_ARG;
    if (v->mem == ODOM_MEM_HOST) {
        SCHK(hipMemcpy2DAsync(stage, row, v->ptr, p, row, v->height, hipMemcpyHostToDevice, c->stream));
        SCHK(hipStreamSynchronize(c->stream));
        *d = (const uint8_t*)stage;
        *pitch = row;
    } else if (v->mem == ODOM_MEM_DEVICE) {
        *d = (const uint8_t*)v->ptr;
        *pitch = p;
    } else {
        return ODOM_ERR_ARG;
    }
    return ODOM_OK;
}

int frames(surfel_ctx* c, const odom_frame_view* rgb, const odom_frame_view* raw, const odom_frame_view* filt,
           float depth_cutoff, RGBView* rv)
{
    const uint8_t *r, *f;
    int rp, fp;
    int rc = view_dev(c, raw, ODOM_FMT_DEPTH_U16, c->raw_stage, &r, &rp);
    if (!rc) rc = view_dev(c, filt, ODOM_FMT_DEPTH_U16, c->filt_stage, &f, &fp);
    if (!rc) rc = view_dev(c, rgb, ODOM_FMT_RGB8, c->rgb_stage, &rv->p, &rv->pitch);
    if (rc) return rc;
    const int npix = c->I.w * c->I.h;
    hipLaunchKernelGGL(k_metric2, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, r, rp, f, fp, depth_cutoff, c->I,
                       c->dr, c->df);
    return ODOM_OK;
}

// Exclusive scan of n flags into prefix/bsums; total to *total (device).
void scan(surfel_ctx* c, const int* flags, int n, int* prefix, int* bsums, int* total)
{
    const int nb = cdiv(n, kScanItems);
    hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, c->stream, flags, n, prefix, bsums);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, c->stream, bsums, nb, total);
}

int read_count(surfel_ctx* c, const int* dev, int* out)
{
    SCHK(hipMemcpyAsync(c->h_count, dev, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    SCHK(slam::poll_stream(c->stream));
    *out = *c->h_count;
    return ODOM_OK;
}

IndexArgs index_args(const surfel_ctx* c, const float pose[16], int time, float max_depth, int time_delta)
{
    IndexArgs A;
    A.model = c->model[c->cur];
    A.count = c->count;
    A.Ti = rigid_inverse(pose);
    A.fx = c->fx;
    A.fy = c->fy;
    A.cx = c->cx;
    A.cy = c->cy;
    A.time = (float)time;
    A.maxDepth = max_depth;
    A.timeDelta = (float)time_delta;
    A.I = c->I;
    return A;
}

void enqueue_index(surfel_ctx* c, const IndexArgs& A)
{
    const int npix = c->I.w * c->I.h;
    if (A.count > 0)
        hipLaunchKernelGGL(k_index_raster, dim3(cdiv(A.count, 256)), dim3(256), 0, c->stream, A, c->keys);
    hipLaunchKernelGGL(k_index_resolve, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, A, c->keys, c->index,
                       c->vert_conf, c->color_time, c->norm_rad);
}

void enqueue_fuse(surfel_ctx* c, const float pose[16], int time, const RGBView& rv, float max_depth,
                  float weighting)
{
    const int npix = c->I.w * c->I.h;
    DataArgs A;
    A.dr = c->dr;
    A.df = c->df;
    A.rgb = rv;
    A.I = c->I;
    A.c = Cam4{c->cx, c->cy, (float)(1.0 / (double)c->fx), (float)(1.0 / (double)c->fy)};
    A.pose = rigid(pose);
    A.time = (float)time;
    A.maxDepth = max_depth;
    A.weighting = weighting;
    A.index = c->index;
    A.vert_conf = c->vert_conf;
    A.norm_rad = c->norm_rad;
    A.pix = c->pix;
    A.flag_fresh = c->flags2;
    A.claim = c->claim;
    A.count = c->count;
    hipLaunchKernelGGL(k_data, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, A);
    if (c->count > 0)
        hipLaunchKernelGGL(k_update, dim3(cdiv(c->count, 256)), dim3(256), 0, c->stream, c->model[c->cur], c->count,
                           c->pix, c->claim, c->I, (float)time);
    scan(c, c->flags2, npix, c->prefix2, c->bsums2, c->totals + 1);
    hipLaunchKernelGGL(k_fresh_scatter, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, c->pix, c->flags2, c->prefix2,
                       c->bsums2, c->I, c->fresh);
}

void enqueue_clean(surfel_ctx* c, const float pose[16], int time, float conf_threshold, int time_delta)
{
    const int npix = c->I.w * c->I.h;
    const int n = c->count + npix;
    CleanArgs A;
    A.model = c->model[c->cur];
    A.count = c->count;
    A.fresh = c->fresh;
    A.nfresh = c->totals + 1;
    A.Ti = rigid_inverse(pose);
    A.fx = c->fx;
    A.fy = c->fy;
    A.cx = c->cx;
    A.cy = c->cy;
    A.I = c->I;
    A.time = (float)time;
    A.confThreshold = conf_threshold;
    A.timeDelta = (float)time_delta;
    A.index = c->index;
    A.vert_conf = c->vert_conf;
    A.color_time = c->color_time;
    A.items = c->items;
    A.flags = c->flags;
    hipLaunchKernelGGL(k_clean, dim3(cdiv(n, 256)), dim3(256), 0, c->stream, A);
    scan(c, c->flags, n, c->prefix, c->bsums, c->totals);
    hipLaunchKernelGGL(k_clean_scatter, dim3(cdiv(n, 256)), dim3(256), 0, c->stream, c->items, c->flags, c->prefix,
                       c->bsums, n, c->max_surfels, c->model[c->cur ^ 1]);
}

}  // namespace

extern "C" {

int surfel_create(int width, int height, float cx, float cy, float fx, float fy, int max_surfels, surfel_handle* out)
{
    if (!out) return ODOM_ERR_ARG;
    *out = nullptr;
    if (width < 2 || height < 2 || width > 8192 || height > 8192 || max_surfels < 1 || !(fx > 0.0f) || !(fy > 0.0f))
        return ODOM_ERR_ARG;
    const long long npix = (long long)width * height;
    if ((long long)max_surfels + npix >= kMaxItems) return ODOM_ERR_ARG;
    char arch[128] = {0};
    if (odom_device_info(nullptr, 0, arch, sizeof arch) != ODOM_OK) return ODOM_ERR_NODEV;
    if (strncmp(arch, "gfx950", 6) != 0) {
        fprintf(stderr, "slam_surfel: device is %s; this library is built for gfx950 only\n", arch);
        return ODOM_ERR_NODEV;
    }
    surfel_ctx* c = new surfel_ctx();
    c->I = Img{width, height, (float)width, (float)height};
    c->cx = cx;
    c->cy = cy;
    c->fx = fx;
    c->fy = fy;
    c->max_surfels = max_surfels;
    const size_t ms = (size_t)max_surfels, np = (size_t)npix, ni = ms + np;
    bool ok = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) == hipSuccess;
    c->stream = c->own;
    auto A = [&](void** p, size_t bytes) { ok = ok && hipMalloc(p, bytes) == hipSuccess; };
    A((void**)&c->model[0], ms * 48);
    A((void**)&c->model[1], ms * 48);
    A((void**)&c->dr, np * 4);
    A((void**)&c->df, np * 4);
    A(&c->raw_stage, np * 2);
    A(&c->filt_stage, np * 2);
    A(&c->rgb_stage, np * 3);
    A((void**)&c->keys, np * 8);
    A((void**)&c->index, np * 4);
    A((void**)&c->vert_conf, np * 16);
    A((void**)&c->color_time, np * 16);
    A((void**)&c->norm_rad, np * 16);
    A((void**)&c->pix, np * 48);
    A((void**)&c->fresh, np * 48);
    A((void**)&c->claim, ms * 4);
    A((void**)&c->items, ni * 48);
    A((void**)&c->feed_norm, np * 16);
    A((void**)&c->flags, ni * 4);
    A((void**)&c->flags2, ni * 4);
    A((void**)&c->prefix, ni * 4);
    A((void**)&c->prefix2, ni * 4);
    A((void**)&c->bsums, 4096 * 4);
    A((void**)&c->bsums2, 4096 * 4);
    A((void**)&c->totals, 16);
    ok = ok && hipHostMalloc((void**)&c->h_count, sizeof(int), 0) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(k_fill_u64, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, c->keys, (int)npix, kKeyClear);
        hipLaunchKernelGGL(k_fill_u32, dim3(cdiv(max_surfels, 256)), dim3(256), 0, c->stream, c->claim, max_surfels,
                           kNoClaim);
        ok = hipGetLastError() == hipSuccess;
        ok = ok && hipMemsetAsync(c->totals, 0, 16, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->index, 0, np * 4, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->vert_conf, 0, np * 16, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->color_time, 0, np * 16, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->norm_rad, 0, np * 16, c->stream) == hipSuccess;
        ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    }
    if (!ok) {
        fprintf(stderr, "slam_surfel: device allocation failed\n");
        free_ctx(c);
        return ODOM_ERR_HIP;
    }
    *out = c;
    return ODOM_OK;
}

int surfel_destroy(surfel_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return ODOM_OK;
}

int surfel_set_stream(surfel_handle c, void* stream)
{
    if (!c) return ODOM_ERR_ARG;
    SCHK(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own;
    return ODOM_OK;
}

int surfel_initialise(surfel_handle c, const odom_frame_view* rgb, const odom_frame_view* raw_depth,
                      const odom_frame_view* filtered_depth, float depth_cutoff, int time, float max_depth)
{
    if (!c) return ODOM_ERR_ARG;
    RGBView rv;
    int rc = frames(c, rgb, raw_depth, filtered_depth, depth_cutoff, &rv);
    if (rc) return rc;
    const int npix = c->I.w * c->I.h;
    const Cam4 cam{c->cx, c->cy, 1.0f / c->fx, 1.0f / c->fy};  // FeedbackBuffer::compute's 1.0f / f
    c->cur = 0;
    hipLaunchKernelGGL(k_feedback, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, c->dr, c->df, rv, c->I, cam, time,
                       max_depth, c->pix, c->feed_norm, c->flags, c->flags2);
    scan(c, c->flags, npix, c->prefix, c->bsums, c->totals);
    scan(c, c->flags2, npix, c->prefix2, c->bsums2, c->totals + 1);
    hipLaunchKernelGGL(k_init_scatter, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, c->pix, c->feed_norm, c->flags,
                       c->flags2, c->prefix, c->bsums, c->prefix2, c->bsums2, c->totals, c->I, c->max_surfels,
                       c->model[0]);
    hipLaunchKernelGGL(k_init_normals, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, c->feed_norm, c->flags2,
                       c->prefix2, c->bsums2, c->totals, c->I, c->max_surfels, c->model[0]);
    SCHK(hipMemsetAsync(c->totals + 1, 0, 4, c->stream));  // no new points pending
    SCHK(hipGetLastError());
    int n = 0;
    rc = read_count(c, c->totals, &n);
    if (rc) return rc;
    c->count = n < c->max_surfels ? n : c->max_surfels;
    return ODOM_OK;
}

int surfel_predict_indices(surfel_handle c, const float pose[16], int time, float max_depth, int time_delta)
{
    if (!c || !pose) return ODOM_ERR_ARG;
    enqueue_index(c, index_args(c, pose, time, max_depth, time_delta));
    SCHK(hipGetLastError());
    return ODOM_OK;
}

int surfel_fuse(surfel_handle c, const float pose[16], int time, const odom_frame_view* rgb,
                const odom_frame_view* raw_depth, const odom_frame_view* filtered_depth, float depth_cutoff,
                float max_depth, float weighting)
{
    if (!c || !pose) return ODOM_ERR_ARG;
    RGBView rv;
    const int rc = frames(c, rgb, raw_depth, filtered_depth, depth_cutoff, &rv);
    if (rc) return rc;
    enqueue_fuse(c, pose, time, rv, max_depth, weighting);
    SCHK(hipGetLastError());
    return ODOM_OK;
}

int surfel_clean(surfel_handle c, const float pose[16], int time, float conf_threshold, int time_delta)
{
    if (!c || !pose) return ODOM_ERR_ARG;
    enqueue_clean(c, pose, time, conf_threshold, time_delta);
    SCHK(hipMemsetAsync(c->totals + 1, 0, 4, c->stream));  // the new points are appended
    SCHK(hipGetLastError());
    int n = 0;
    const int rc = read_count(c, c->totals, &n);
    if (rc) return rc;
    c->cur ^= 1;
    c->count = n < c->max_surfels ? n : c->max_surfels;
    return ODOM_OK;
}

int surfel_count(surfel_handle c, int* count)
{
    if (!c || !count) return ODOM_ERR_ARG;
    *count = c->count;
    return ODOM_OK;
}

int surfel_get_model(surfel_handle c, const float** dev_ptr)
{
    if (!c || !dev_ptr) return ODOM_ERR_ARG;
    *dev_ptr = c->model[c->cur];
    return ODOM_OK;
}

int surfel_read_model(surfel_handle c, float* host_dst, int max_surfels, int* count)
{
    if (!c || !host_dst || !count || max_surfels < 0) return ODOM_ERR_ARG;
    const int n = c->count < max_surfels ? c->count : max_surfels;
    if (n > 0)
        SCHK(hipMemcpyAsync(host_dst, c->model[c->cur], (size_t)n * 48, hipMemcpyDeviceToHost, c->stream));
    SCHK(hipStreamSynchronize(c->stream));
    *count = c->count;
    return ODOM_OK;
}

int surfel_set_model(surfel_handle c, const float* model, int n, int mem)
{
    if (!c || n < 0 || n > c->max_surfels || (n > 0 && !model)) return ODOM_ERR_ARG;
    if (mem != ODOM_MEM_HOST && mem != ODOM_MEM_DEVICE) return ODOM_ERR_ARG;
    if (n > 0)
        SCHK(hipMemcpyAsync(c->model[c->cur], model, (size_t)n * 48,
                            mem == ODOM_MEM_HOST ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, c->stream));
    SCHK(hipMemsetAsync(c->totals + 1, 0, 4, c->stream));
    SCHK(hipStreamSynchronize(c->stream));
    c->count = n;
    return ODOM_OK;
}

int surfel_read_index_map(surfel_handle c, uint32_t* index, float* vert_conf, float* color_time, float* norm_rad)
{
    if (!c || !index || !vert_conf || !color_time || !norm_rad) return ODOM_ERR_ARG;
    const int w = c->I.w, h = c->I.h;
    const size_t np = (size_t)w * h;
    // the device maps are in draw order (column by column); the host gets rows
    std::vector<uint32_t> ti(np);
    std::vector<float> tv(np * 4), tc(np * 4), tn(np * 4);
    SCHK(hipMemcpyAsync(ti.data(), c->index, np * 4, hipMemcpyDeviceToHost, c->stream));
    SCHK(hipMemcpyAsync(tv.data(), c->vert_conf, np * 16, hipMemcpyDeviceToHost, c->stream));
    SCHK(hipMemcpyAsync(tc.data(), c->color_time, np * 16, hipMemcpyDeviceToHost, c->stream));
    SCHK(hipMemcpyAsync(tn.data(), c->norm_rad, np * 16, hipMemcpyDeviceToHost, c->stream));
    SCHK(hipStreamSynchronize(c->stream));
    for (int x = 0; x < w; ++x)
        for (int y = 0; y < h; ++y) {
            const size_t d = (size_t)x * h + y, r = (size_t)y * w + x;
            index[r] = ti[d];
            memcpy(vert_conf + 4 * r, &tv[4 * d], 16);
            memcpy(color_time + 4 * r, &tc[4 * d], 16);
            memcpy(norm_rad + 4 * r, &tn[4 * d], 16);
        }
    return ODOM_OK;
}

int surfel_read_fresh(surfel_handle c, float* host_dst, int max_points, int* count)
{
    if (!c || !host_dst || !count || max_points < 0) return ODOM_ERR_ARG;
    int n = 0;
    const int rc = read_count(c, c->totals + 1, &n);
    if (rc) return rc;
    const int m = n < max_points ? n : max_points;
    if (m > 0) SCHK(hipMemcpyAsync(host_dst, c->fresh, (size_t)m * 48, hipMemcpyDeviceToHost, c->stream));
    SCHK(hipStreamSynchronize(c->stream));
    *count = n;
    return ODOM_OK;
}

int surfel_profile(surfel_handle c, const float pose[16], int time, const odom_frame_view* rgb,
                   const odom_frame_view* raw_depth, const odom_frame_view* filtered_depth, float depth_cutoff,
                   float max_depth, float weighting, float conf_threshold, int time_delta, int reps, float us[4])
{
    if (!c || !pose || !us || reps < 1) return ODOM_ERR_ARG;
    RGBView rv;
    int rc = frames(c, rgb, raw_depth, filtered_depth, depth_cutoff, &rv);
    if (rc) return rc;
    // keep the starting model: every rep fuses the same frame into it
    const int n0 = c->count, cur0 = c->cur;
    DeviceScratch start_buf;  // a copy of the starting model (the one allocation outside create)
    SCHK(hipMalloc(&start_buf.p, (size_t)(n0 > 0 ? n0 : 1) * 48));
    float* start = (float*)start_buf.p;
    if (n0 > 0)
        SCHK(hipMemcpyAsync(start, c->model[cur0], (size_t)n0 * 48, hipMemcpyDeviceToDevice, c->stream));
    TimingEvents ev(5);
    SCHK(ev.create());
    double acc[4] = {0, 0, 0, 0};
    int n = 0;
    for (int r = 0; r < reps; ++r) {
        c->cur = cur0;
        c->count = n0;
        if (n0 > 0)
            SCHK(hipMemcpyAsync(c->model[cur0], start, (size_t)n0 * 48, hipMemcpyDeviceToDevice, c->stream));
        SCHK(hipEventRecord(ev[0], c->stream));
        enqueue_index(c, index_args(c, pose, time, max_depth, time_delta));
        SCHK(hipEventRecord(ev[1], c->stream));
        enqueue_fuse(c, pose, time, rv, max_depth, weighting);
        SCHK(hipEventRecord(ev[2], c->stream));
        enqueue_index(c, index_args(c, pose, time, max_depth, time_delta));
        SCHK(hipEventRecord(ev[3], c->stream));
        enqueue_clean(c, pose, time, conf_threshold, time_delta);
        SCHK(hipEventRecord(ev[4], c->stream));
        SCHK(hipGetLastError());
        SCHK(hipEventSynchronize(ev[4]));
        float ms;
        for (int k = 0; k < 4; ++k) {
            SCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
        rc = read_count(c, c->totals, &n);
        if (rc) return rc;
    }
    c->cur = cur0 ^ 1;
    c->count = n < c->max_surfels ? n : c->max_surfels;
    SCHK(hipMemsetAsync(c->totals + 1, 0, 4, c->stream));
    SCHK(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) us[k] = (float)(acc[k] * 1000.0 / reps);
    return ODOM_OK;
}

}  // extern "C"
