// splat_kernels.hip -- surfel splat prediction and fill-in on gfx950, behind
// include/slam_splat.h.
//
// The reference draws the surfel model as GL point sprites
// (IndexMap::combinedPredict, IndexMap.cpp:243-345): the vertex shader
// (splat.vert:53-87) transforms each surfel into the camera, drops it when it
// is out of range, unconfident or outside the time window, and sizes a square
// sprite from the projections of four points on the surfel's disk; the
// fragment shader (combo_splat.frag:35-61) intersects each pixel's ray with
// the disk plane, discards it outside the radius, writes colour, vertex,
// normal and time, and sets the depth to the intersection's z, which GL_LESS
// resolves to the nearest surfel, the first drawn on ties.
//
// Here that is two launches and no rasterizer:
//   k_splat_raster   one thread per surfel: the vertex stage, then every
//                    pixel centre of the sprite square (exact bounds) that
//                    the disk's window-space extent can reach (disk_bounds);
//                    a fragment that passes the disk test offers
//                    ((24-bit depth) << 32 | surfel index) to the pixel with a
//                    64-bit atomic minimum, which is GL_LESS in draw order: the
//                    nearest depth wins and, among equal depths, the lowest
//                    index (the first drawn); key_min skips the atomics that
//                    cannot win.  The ray through each pixel centre is read
//                    from a per-handle table (k_rays);
//   k_splat_resolve  one thread per pixel: reads the winning key, clears it
//                    for the next prediction, and recomputes that surfel's
//                    fragment at this pixel to write the four outputs.
// Recomputing the winner instead of storing every fragment's outputs keeps
// the contended traffic at 8 bytes per fragment.  FillIn's three passes
// (FillIn.cpp:56-198, fill_*.frag) are one per-pixel launch, k_fill_in.
//
// Arithmetic is FP32 with -ffp-contract=off and the shaders' operation order,
// so every output is a function of the inputs alone (bit-identical to a CPU
// evaluation of the same expressions).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "slam_splat.h"
#include "timing_events.h"

namespace {

constexpr unsigned kDepthClear = 0xFFFFFFu;  // the depth buffer's clear value 1.0 in 24-bit units
constexpr unsigned long long kKeyClear = ((unsigned long long)kDepthClear << 32) | 0xFFFFFFFFull;
constexpr float kMaxPointSize = (float)SPLAT_MAX_POINT_SIZE;

struct SplatCam {
    float cx, cy, fx, fy;
    float invfx, invfy;              // 1.f / f, as the shaders' uniforms
    float cols, rows;
    int w, h;
};

struct SplatUniforms {
    float Ti[12];                    // rows of the inverse pose [R^T | -R^T t]
    float maxDepth, confThreshold;
    float time, maxTime, timeDelta;  // the shader compares them with float attributes
    SplatCam cam;
};

// One surfel in the camera frame (the vertex shader's outputs).
struct Surf {
    float hx, hy, hz, conf;          // position (vPosHome) and confidence
    float nx, ny, nz, rad;           // normRad
    float col, tinit;                // colTime.x, colTime.z
};

__device__ __forceinline__ float gmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return (a < b) ? b : a; }

// splat.vert:53-66: transform, range/confidence/time tests, normal.
__device__ __forceinline__ bool surfel_stage(const float4 a, const float4 b, const float4 c, const SplatUniforms& U,
                                             Surf& s)
{
    const float* T = U.Ti;
    s.hx = ((T[0] * a.x + T[1] * a.y) + T[2] * a.z) + T[3];
    s.hy = ((T[4] * a.x + T[5] * a.y) + T[6] * a.z) + T[7];
    s.hz = ((T[8] * a.x + T[9] * a.y) + T[10] * a.z) + T[11];
    if (s.hz > U.maxDepth || s.hz < 0.0f || a.w < U.confThreshold || U.time - b.w > U.timeDelta ||
        b.w > U.maxTime)
        return false;
    const float mx = (T[0] * c.x + T[1] * c.y) + T[2] * c.z;
    const float my = (T[4] * c.x + T[5] * c.y) + T[6] * c.z;
    const float mz = (T[8] * c.x + T[9] * c.y) + T[10] * c.z;
    const float rn = 1.0f / sqrtf((mx * mx + my * my) + mz * mz);
    s.nx = mx * rn;
    s.ny = my * rn;
    s.nz = mz * rn;
    s.rad = c.w;
    s.conf = a.w;
    s.col = b.x;
    s.tinit = b.z;
    return true;
}

// splat.vert:38-50, 68-85: window position of the sprite centre and its
// size; false when GL would clip the point (centre outside the view volume)
// or the size is not a number.
__device__ __forceinline__ bool sprite(const Surf& s, const SplatCam& K, float& xw, float& yw, float& size)
{
    const float hc = K.cols * 0.5f, hr = K.rows * 0.5f;
    const float ndx = (((K.fx * s.hx) / s.hz + K.cx) - hc) / hc;
    const float ndy = (((K.fy * s.hy) / s.hz + K.cy) - hr) / hr;
    if (!(ndx >= -1.0f && ndx <= 1.0f && ndy >= -1.0f && ndy <= 1.0f)) return false;
    xw = ndx * hc + hc;
    yw = ndy * hr + hr;
    const float ux = s.ny - s.nz, uy = -s.nx, uz = s.nx;
    const float ru = 1.0f / sqrtf((ux * ux + uy * uy) + uz * uz);
    const float x1x = ((ux * ru) * s.rad) * 1.41421356f;
    const float x1y = ((uy * ru) * s.rad) * 1.41421356f;
    const float x1z = ((uz * ru) * s.rad) * 1.41421356f;
    const float y1x = s.ny * x1z - s.nz * x1y;
    const float y1y = s.nz * x1x - s.nx * x1z;
    const float y1z = s.nx * x1y - s.ny * x1x;
    float px[4], py[4];
    const float ox[4] = {x1x, y1x, -y1x, -x1x};
    const float oy[4] = {x1y, y1y, -y1y, -x1y};
    const float oz[4] = {x1z, y1z, -y1z, -x1z};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float qx = s.hx + ox[q], qy = s.hy + oy[q], qz = s.hz + oz[q];
        px[q] = (K.fx * qx) / qz + K.cx;
        py[q] = (K.fy * qy) / qz + K.cy;
    }
    const float x0 = gmin(px[0], gmin(px[1], gmin(px[2], px[3])));
    const float x1 = gmax(px[0], gmax(px[1], gmax(px[2], px[3])));
    const float y0 = gmin(py[0], gmin(py[1], gmin(py[2], py[3])));
    const float y1 = gmax(py[0], gmax(py[1], gmax(py[2], py[3])));
    const float sz = gmax(0.0f, gmax(fabsf(x1 - x0), fabsf(y1 - y0)));
    if (isnan(sz)) return false;
    size = sz < 1.0f ? 1.0f : (sz > kMaxPointSize ? kMaxPointSize : sz);
    return true;
}

// combo_splat.frag:37: the normalized ray through the pixel centre (fcx,
// fcy).  It depends on the pixel only, so a handle computes it once per
// pixel (k_rays) and every fragment reads it.
__device__ __forceinline__ float4 pixel_ray(const SplatCam& K, float fcx, float fcy)
{
    const float lx = (fcx - K.cx) / K.fx;
    const float ly = (fcy - K.cy) / K.fy;
    const float rl = 1.0f / sqrtf((lx * lx + ly * ly) + 1.0f);
    return make_float4(lx * rl, ly * rl, 1.0f * rl, 0.0f);
}

// combo_splat.frag:38-48, 59: the ray l meets the surfel plane at z (num is
// dot(position, normal)); false when the point lies outside the disk or the
// depth does not pass against the cleared buffer.
__device__ __forceinline__ bool fragment(const Surf& s, float num, float maxDepth, const float4 l, float& z,
                                         unsigned& q)
{
    const float den = (l.x * s.nx + l.y * s.ny) + l.z * s.nz;
    const float t = num / den;
    const float c0 = t * l.x, c1 = t * l.y, c2 = t * l.z;
    const float d0 = c0 - s.hx, d1 = c1 - s.hy, d2 = c2 - s.hz;
    const float dd = (d0 * d0 + d1 * d1) + d2 * d2;
    if (!(dd <= s.rad * s.rad)) return false;
    float depth = c2 / (2.0f * maxDepth) + 0.5f;
    depth = depth < 0.0f ? 0.0f : (depth > 1.0f ? 1.0f : depth);
    q = (unsigned)rintf(depth * 16777215.0f);
    if (q >= kDepthClear) return false;
    z = c2;
    return true;
}

__device__ __forceinline__ float plane_num(const Surf& s) { return (s.hx * s.nx + s.hy * s.ny) + s.hz * s.nz; }

// A bound on the pixel centres whose fragment can pass: the window-space
// extent of the surfel's disk (centre h, normal n, radius rad).  The disk is
// h + rad (cos p u + sin p v) for an orthonormal pair u, v of its plane;
// x = fx Cx / Cz + cx is extremal where A sin p + B cos p + K = 0 (A, B, K
// below, the same for y), which has the two solutions taken here.  A fragment
// passes only when its ray meets the disk up to the fragment test's FP32
// rounding, about 1e-4 pixel in window space, and this FP32 evaluation (with
// hardware reciprocals and square roots) is within about 3e-4 pixel of the
// exact extent, so the bound widened by kDiskMargin holds every passing
// centre.  False (no bound: the sprite square alone is used) for views
// within about 3 degrees of the plane and for radii above 2 % of the distance
// (rays that meet such disks may graze the plane, where the test's rounding
// grows), for disks that reach z <= 1 % of the centre's depth (their
// projections are not bounded ellipses), and where the extremes are poorly
// conditioned (D < R2 / 2; with the radius bound D is above 0.999 R2).
constexpr float kDiskMargin = 0.05f;

__device__ bool disk_bounds(const Surf& s, const SplatCam& K, float& x0, float& x1, float& y0, float& y1)
{
    const float h[3] = {s.hx, s.hy, s.hz};
    const float inl = __builtin_amdgcn_rsqf(s.nx * s.nx + s.ny * s.ny + s.nz * s.nz);
    const float n[3] = {s.nx * inl, s.ny * inl, s.nz * inl};
    const float r = s.rad;
    const float hl = __builtin_amdgcn_sqrtf(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    if (!(fabsf(h[0] * n[0] + h[1] * n[1] + h[2] * n[2]) >= 0.05f * hl) || !(r <= 0.02f * hl)) return false;
    if (!(h[2] - r * __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - n[2] * n[2])) > 0.01f * h[2])) return false;
    // u = n x e (e the axis least aligned with n), v = n x u
    const bool ax = fabsf(n[0]) < 0.9f;
    float u[3] = {ax ? 0.0f : -n[2], ax ? n[2] : 0.0f, ax ? -n[1] : n[0]};
    const float iul = __builtin_amdgcn_rsqf(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] *= iul, u[1] *= iul, u[2] *= iul;
    const float v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
    float lo[2], hi[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const float A = h[c] * u[2] - h[2] * u[c];
        const float B = h[2] * v[c] - h[c] * v[2];
        const float Kc = r * (v[c] * u[2] - u[c] * v[2]);
        const float R2 = A * A + B * B;
        const float D = R2 - Kc * Kc;
        if (!(D > 0.5f * R2)) return false;
        const float sd = __builtin_amdgcn_sqrtf(D), iR2 = __builtin_amdgcn_rcpf(R2);
        const float f = c == 0 ? K.fx : K.fy, ctr = c == 0 ? K.cx : K.cy;
        float e[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const float sg = q == 0 ? 1.0f : -1.0f;
            const float sn = (-Kc * A + sg * B * sd) * iR2, cs = (-Kc * B - sg * A * sd) * iR2;
            const float Cc = h[c] + r * (cs * u[c] + sn * v[c]);
            const float Cz = h[2] + r * (cs * u[2] + sn * v[2]);
            e[q] = f * Cc * __builtin_amdgcn_rcpf(Cz) + ctr;
        }
        if (!(fabsf(e[0]) < 1e6f && fabsf(e[1]) < 1e6f)) return false;   // also NaN
        lo[c] = fminf(e[0], e[1]) - kDiskMargin;
        hi[c] = fmaxf(e[0], e[1]) + kDiskMargin;
    }
    x0 = lo[0], x1 = hi[0], y0 = lo[1], y1 = hi[1];
    return true;
}

// The pixel centres c = i + 0.5 of the sprite: lo <= c < hi.  Window
// coordinates stay below 2^13 and the size at most 256, so lo - 0.5 is exact
// and the ceilings give the bounds; the checks only guard that reasoning.
__device__ __forceinline__ int first_centre(float lo)
{
    int i = (int)ceilf(lo - 0.5f);
    if ((float)(i - 1) + 0.5f >= lo) i -= 1;
    else if ((float)i + 0.5f < lo) i += 1;
    return i;
}

__device__ __forceinline__ int last_centre(float hi)
{
    int i = (int)ceilf(hi - 0.5f) - 1;
    if ((float)(i + 1) + 0.5f < hi) i += 1;
    else if ((float)i + 0.5f >= hi) i -= 1;
    return i;
}

// GL_LESS as a 64-bit minimum.  Keys only decrease while a raster runs, so
// when a coherent (device-scope) read of the key, which returns one of its
// values in update order, is already at or below v, the minimum would change
// nothing and the atomic is skipped.  Most fragments lose to one drawn
// earlier, so most of the atomics, which device scope performs beyond the
// compute die's cache, are saved; the result is the same.
__device__ __forceinline__ void key_min(unsigned long long* key, unsigned long long v)
{
    if (__hip_atomic_load(key, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > v) atomicMin(key, v);
}

__global__ __launch_bounds__(256) void k_rays(SplatCam K, float4* rays)
{
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= K.w * K.h) return;
    rays[p] = pixel_ray(K, (float)(p % K.w) + 0.5f, (float)(p / K.w) + 0.5f);
}

__global__ __launch_bounds__(256) void k_fill_keys(unsigned long long* keys, int n)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) keys[i] = kKeyClear;
}

// One thread per surfel: the vertex stage, then the pixel centres of its
// sprite (exact bounds, so every iteration is a candidate fragment).  A
// sprite of more than kBigSprite candidates would hold its whole wave for its
// length, so the wave works through such sprites together, 64 candidates at
// a time, after the lanes' own small ones.
constexpr int kBigSprite = 100;

__global__ __launch_bounds__(256) void k_splat_raster(const float4* __restrict__ model, int n, SplatUniforms U,
                                                      const float4* __restrict__ rays, unsigned long long* keys,
                                                      unsigned long long* frag_count)
{
    const int k = blockIdx.x * 256 + threadIdx.x;
    Surf s{};
    int i0 = 0, j0 = 0, ncols = 1, cnt = 0;
    if (k < n) {
        const size_t b = (size_t)k * 3;
        float xw, yw, size;
        if (surfel_stage(model[b], model[b + 1], model[b + 2], U, s) && sprite(s, U.cam, xw, yw, size)) {
            const float half = size * 0.5f;
            i0 = max(0, first_centre(xw - half));
            int i1 = min(U.cam.w - 1, last_centre(xw + half));
            j0 = max(0, first_centre(yw - half));
            int j1 = min(U.cam.h - 1, last_centre(yw + half));
            float bx0, bx1, by0, by1;
            if (disk_bounds(s, U.cam, bx0, bx1, by0, by1)) {
                // centres c = i + 0.5 inside [b0, b1] (|b| < 1e6, so the conversions are exact)
                i0 = max(i0, (int)ceilf(bx0 - 0.5f));
                i1 = min(i1, (int)floorf(bx1 - 0.5f));
                j0 = max(j0, (int)ceilf(by0 - 0.5f));
                j1 = min(j1, (int)floorf(by1 - 0.5f));
            }
            if (i1 >= i0 && j1 >= j0) {
                ncols = i1 - i0 + 1;
                cnt = ncols * (j1 - j0 + 1);
            }
        }
    }
    const float num = plane_num(s);
    const int w = U.cam.w;
    unsigned drawn = 0;
    auto candidate = [&](const Surf& f, float fnum, int i, int j, unsigned idx) {
        const size_t p = (size_t)j * w + i;
        float z;
        unsigned q;
        if (!fragment(f, fnum, U.maxDepth, rays[p], z, q)) return;
        ++drawn;
        key_min(keys + p, ((unsigned long long)q << 32) | idx);
    };
    if (cnt > 0 && cnt <= kBigSprite) {
        const int j1 = j0 + cnt / ncols - 1, i1 = i0 + ncols - 1;
        for (int j = j0; j <= j1; ++j)
            for (int i = i0; i <= i1; ++i) candidate(s, num, i, j, (unsigned)k);
    }
    // the wave's big sprites, one after another, each spread over the lanes
    const int lane = threadIdx.x & 63;
    unsigned long long big = __ballot(cnt > kBigSprite);
    while (big) {
        const int src = __ffsll((long long)big) - 1;
        big &= big - 1;
        Surf f;
        f.hx = __shfl(s.hx, src, 64);
        f.hy = __shfl(s.hy, src, 64);
        f.hz = __shfl(s.hz, src, 64);
        f.nx = __shfl(s.nx, src, 64);
        f.ny = __shfl(s.ny, src, 64);
        f.nz = __shfl(s.nz, src, 64);
        f.rad = __shfl(s.rad, src, 64);
        const float fnum = __shfl(num, src, 64);
        const int bi0 = __shfl(i0, src, 64), bj0 = __shfl(j0, src, 64);
        const int bcols = __shfl(ncols, src, 64), bcnt = __shfl(cnt, src, 64);
        const unsigned bk = (unsigned)__shfl(k, src, 64);
        for (int t = lane; t < bcnt; t += 64) {
            const int r = t / bcols;
            candidate(f, fnum, bi0 + (t - r * bcols), bj0 + r, bk);
        }
    }
    if (frag_count && drawn) atomicAdd(frag_count, (unsigned long long)drawn);
}

__global__ __launch_bounds__(256) void k_splat_resolve(const float4* __restrict__ model, SplatUniforms U,
                                                       const float4* __restrict__ rays, unsigned long long* keys,
                                                       uchar4* image, float4* vertex, float4* normal, uint16_t* time)
{
    const int w = U.cam.w;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= w * U.cam.h) return;
    const unsigned long long key = keys[p];
    keys[p] = kKeyClear;
    uchar4 im = make_uchar4(0, 0, 0, 0);
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f), nn = make_float4(0.f, 0.f, 0.f, 0.f);
    uint16_t tm = 0;
    if ((unsigned)(key >> 32) < kDepthClear) {
        const size_t b = (size_t)(unsigned)key * 3;
        Surf s;
        float z;
        unsigned q;
        const int i = p % w, j = p / w;
        const float fcx = (float)i + 0.5f, fcy = (float)j + 0.5f;
        // the winner passed both stages in the raster pass; the same
        // expressions give the same values here
        if (surfel_stage(model[b], model[b + 1], model[b + 2], U, s) &&
            fragment(s, plane_num(s), U.maxDepth, rays[p], z, q)) {
            // color.glsl decodeColor, stored to an 8-bit unorm target
            const int ci = (int)fminf(fmaxf(s.col, 0.0f), 16777215.0f);
            im = make_uchar4((ci >> 16) & 0xFF, (ci >> 8) & 0xFF, ci & 0xFF, 255);
            v = make_float4(((fcx - U.cam.cx) * z) * U.cam.invfx, ((fcy - U.cam.cy) * z) * U.cam.invfy, z, s.conf);
            nn = make_float4(s.nx, s.ny, s.nz, s.rad);
            const float t = s.tinit;
            tm = t >= 65535.0f ? (uint16_t)65535 : (t >= 0.0f ? (uint16_t)t : (uint16_t)0);
        }
    }
    image[p] = im;
    vertex[p] = v;
    normal[p] = nn;
    time[p] = tm;
}

struct FillArgs {
    const uchar4* image;
    const float4* vertex;
    const float4* normal;
    const uint8_t* depth;            // uint16 rows of dpitch bytes
    const uint8_t* rgb;              // rows of rpitch bytes, rch bytes per pixel
    int dpitch, rpitch, rch;
    int pass_geom, pass_img;
    uchar4* fimage;
    float4* fvertex;
    float4* fnormal;
    SplatCam cam;
};

__device__ __forceinline__ float raw_z(const FillArgs& A, int x, int y)
{
    // textures clamp to the edge: the neighbour of the last column or row is itself
    x = min(x, A.cam.w - 1);
    y = min(y, A.cam.h - 1);
    const uint16_t d = *reinterpret_cast<const uint16_t*>(A.depth + (size_t)y * A.dpitch + 2 * (size_t)x);
    return (float)d / 1000.0f;
}

// fill_vertex.frag:32-35, geometry.glsl:44-48: vertex of integer pixel
// (x, y) at depth z (no half-pixel offset, unlike the splat).
__device__ __forceinline__ float3 raw_vertex(const SplatCam& K, int x, int y, float z)
{
    return make_float3((((float)x - K.cx) * z) * K.invfx, (((float)y - K.cy) * z) * K.invfy, z);
}

__global__ __launch_bounds__(256) void k_fill_in(FillArgs A)
{
    const int w = A.cam.w;
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= w * A.cam.h) return;
    const int i = p % w, j = p / w;
    // fill_vertex.frag:37-52
    const float4 v = A.vertex[p];
    float4 fv = v;
    if (v.z == 0.0f || A.pass_geom) {
        const float3 r = raw_vertex(A.cam, i, j, raw_z(A, i, j));
        fv = make_float4(r.x, r.y, r.z, 1.0f);
    }
    A.fvertex[p] = fv;
    // fill_normal.frag:34-49, geometry.glsl:51-61 (forward differences)
    const float4 n = A.normal[p];
    float4 fn = n;
    if (n.z == 0.0f || A.pass_geom) {
        const float3 v0 = raw_vertex(A.cam, i, j, raw_z(A, i, j));
        const float3 vx = raw_vertex(A.cam, i + 1, j, raw_z(A, i + 1, j));
        const float3 vy = raw_vertex(A.cam, i, j + 1, raw_z(A, i, j + 1));
        const float ax = vx.x - v0.x, ay = vx.y - v0.y, az = vx.z - v0.z;
        const float bx = vy.x - v0.x, by = vy.y - v0.y, bz = vy.z - v0.z;
        const float cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
        const float rn = 1.0f / sqrtf((cx * cx + cy * cy) + cz * cz);
        fn = make_float4(cx * rn, cy * rn, cz * rn, 1.0f);
    }
    A.fnormal[p] = fn;
    // fill_rgb.frag:29-37: r + g + b of unorm values is 0 only if all are 0
    const uchar4 im = A.image[p];
    uchar4 fi = im;
    if ((im.x | im.y | im.z) == 0 || A.pass_img) {
        const uint8_t* px = A.rgb + (size_t)j * A.rpitch + (size_t)i * A.rch;
        fi = make_uchar4(px[0], px[1], px[2], A.rch == 4 ? px[3] : 255);
    }
    A.fimage[p] = fi;
}

int cdiv(long long a, int b) { return (int)((a + b - 1) / b); }

}  // namespace

struct splat_ctx {
    SplatCam cam{};
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    unsigned long long* keys = nullptr;
    float4* rays = nullptr;          // per-pixel normalized ray through the pixel centre
    unsigned long long* frag_count = nullptr;
    uchar4 *image = nullptr, *fimage = nullptr;
    float4 *vertex = nullptr, *normal = nullptr, *fvertex = nullptr, *fnormal = nullptr;
    uint16_t* time = nullptr;
    float* staging = nullptr;        // host models
    int staging_cap = 0;
    void* depth_staging = nullptr;
    void* rgb_staging = nullptr;
};

#define SCHK(x)                                                                                      \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess) {                                                                      \
            fprintf(stderr, "slam_splat: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, \
                    __LINE__);                                                                       \
            return ODOM_ERR_HIP;                                                                     \
        }                                                                                            \
    } while (0)

namespace {

void free_ctx(splat_ctx* c)
{
    void* bufs[] = {c->keys, c->rays, c->frag_count, c->image, c->fimage, c->vertex, c->normal, c->fvertex,
                    c->fnormal, c->time, c->staging, c->depth_staging, c->rgb_staging};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

// Inverse of a rigid camera-to-global pose: [R^T | -R^T t].
void make_uniforms(const splat_ctx* c, const float pose[16], float depth_cutoff, float conf_threshold, int time,
                   int max_time, int time_delta, SplatUniforms* U)
{
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k) U->Ti[r * 4 + k] = pose[k * 4 + r];
        U->Ti[r * 4 + 3] = -((U->Ti[r * 4] * pose[3] + U->Ti[r * 4 + 1] * pose[7]) + U->Ti[r * 4 + 2] * pose[11]);
    }
    U->maxDepth = depth_cutoff;
    U->confThreshold = conf_threshold;
    U->time = (float)time;
    U->maxTime = (float)max_time;
    U->timeDelta = (float)time_delta;
    U->cam = c->cam;
}

int view_to_device(splat_ctx* c, const odom_frame_view* v, int fmt_a, int fmt_b, void* staging, const uint8_t** d,
                   int* pitch)
{
    if (!v || !v->ptr) return ODOM_ERR_ARG;
    if (v->format != fmt_a && v->format != fmt_b) return ODOM_ERR_ARG;
    if (v->width != c->cam.w || v->height != c->cam.h) return ODOM_ERR_ARG;
    const int bpp = v->format == ODOM_FMT_DEPTH_U16 ? 2 : (v->format == ODOM_FMT_RGB8 ? 3 : 4);
    const int row = v->width * bpp;
    const int p = v->pitch_bytes ? v->pitch_bytes : row;
    if (p < row || (bpp == 2 && (p % 2) != 0)) return ODOM_ERR_ARG;
    if (v->mem == ODOM_MEM_HOST) {
        SCHK(hipMemcpy2DAsync(staging, row, v->ptr, p, row, v->height, hipMemcpyHostToDevice, c->stream));
        SCHK(hipStreamSynchronize(c->stream));
        *d = (const uint8_t*)staging;
        *pitch = row;
    } else if (v->mem == ODOM_MEM_DEVICE) {
        *d = (const uint8_t*)v->ptr;
        *pitch = p;
    } else {
        return ODOM_ERR_ARG;
    }
    return ODOM_OK;
}

int enqueue_predict(splat_ctx* c, const SplatUniforms& U, const float* model, int n, bool count)
{
    if (n > 0)
        hipLaunchKernelGGL(k_splat_raster, dim3(cdiv(n, 256)), dim3(256), 0, c->stream, (const float4*)model, n, U,
                           c->rays, c->keys, count ? c->frag_count : nullptr);
    const int npix = c->cam.w * c->cam.h;
    hipLaunchKernelGGL(k_splat_resolve, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, (const float4*)model, U,
                       c->rays, c->keys, c->image, c->vertex, c->normal, c->time);
    SCHK(hipGetLastError());
    return ODOM_OK;
}

int make_fill_args(splat_ctx* c, const odom_frame_view* depth, const odom_frame_view* rgb, int pass_geom,
                   int pass_img, FillArgs* A)
{
    int rc = view_to_device(c, depth, ODOM_FMT_DEPTH_U16, ODOM_FMT_DEPTH_U16, c->depth_staging, &A->depth, &A->dpitch);
    if (rc) return rc;
    rc = view_to_device(c, rgb, ODOM_FMT_RGB8, ODOM_FMT_RGBA8, c->rgb_staging, &A->rgb, &A->rpitch);
    if (rc) return rc;
    A->rch = rgb->format == ODOM_FMT_RGB8 ? 3 : 4;
    A->image = c->image;
    A->vertex = c->vertex;
    A->normal = c->normal;
    A->pass_geom = pass_geom != 0;
    A->pass_img = pass_img != 0;
    A->fimage = c->fimage;
    A->fvertex = c->fvertex;
    A->fnormal = c->fnormal;
    A->cam = c->cam;
    return ODOM_OK;
}

}  // namespace

extern "C" {

int splat_create(int width, int height, float cx, float cy, float fx, float fy, int max_host_surfels,
                 splat_handle* out)
{
    if (!out) return ODOM_ERR_ARG;
    *out = nullptr;
    if (width < 1 || height < 1 || width > 8192 || height > 8192 || max_host_surfels < 0 || !(fx > 0.0f) ||
        !(fy > 0.0f))
        return ODOM_ERR_ARG;
    char arch[128] = {0};
    if (odom_device_info(nullptr, 0, arch, sizeof arch) != ODOM_OK) return ODOM_ERR_NODEV;
    if (strncmp(arch, "gfx950", 6) != 0) {
        fprintf(stderr, "slam_splat: device is %s; this library is built for gfx950 only\n", arch);
        return ODOM_ERR_NODEV;
    }
    splat_ctx* c = new splat_ctx();
    c->cam.cx = cx;
    c->cam.cy = cy;
    c->cam.fx = fx;
    c->cam.fy = fy;
    c->cam.invfx = 1.0f / fx;
    c->cam.invfy = 1.0f / fy;
    c->cam.cols = (float)width;
    c->cam.rows = (float)height;
    c->cam.w = width;
    c->cam.h = height;
    const size_t npix = (size_t)width * height;
    bool ok = hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) == hipSuccess;
    c->stream = c->own;
    ok = ok && hipMalloc((void**)&c->keys, npix * 8) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->rays, npix * 16) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->frag_count, 8) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->image, npix * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->fimage, npix * 4) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->vertex, npix * 16) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->normal, npix * 16) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->fvertex, npix * 16) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->fnormal, npix * 16) == hipSuccess;
    ok = ok && hipMalloc((void**)&c->time, npix * 2) == hipSuccess;
    ok = ok && hipMalloc(&c->depth_staging, npix * 2) == hipSuccess;
    ok = ok && hipMalloc(&c->rgb_staging, npix * 4) == hipSuccess;
    if (ok && max_host_surfels > 0) {
        ok = hipMalloc((void**)&c->staging, (size_t)max_host_surfels * SPLAT_FLOATS * sizeof(float)) == hipSuccess;
        c->staging_cap = max_host_surfels;
    }
    if (ok) {
        hipLaunchKernelGGL(k_fill_keys, dim3(cdiv((long long)npix, 256)), dim3(256), 0, c->stream, c->keys,
                           (int)npix);
        hipLaunchKernelGGL(k_rays, dim3(cdiv((long long)npix, 256)), dim3(256), 0, c->stream, c->cam, c->rays);
        ok = hipGetLastError() == hipSuccess;
        // outputs start as a prediction of an empty model
        ok = ok && hipMemsetAsync(c->image, 0, npix * 4, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->vertex, 0, npix * 16, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->normal, 0, npix * 16, c->stream) == hipSuccess;
        ok = ok && hipMemsetAsync(c->time, 0, npix * 2, c->stream) == hipSuccess;
        ok = ok && hipStreamSynchronize(c->stream) == hipSuccess;
    }
    if (!ok) {
        fprintf(stderr, "slam_splat: device allocation failed\n");
        free_ctx(c);
        return ODOM_ERR_HIP;
    }
    *out = c;
    return ODOM_OK;
}

int splat_destroy(splat_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    (void)hipStreamSynchronize(c->stream);
    free_ctx(c);
    return ODOM_OK;
}

int splat_set_stream(splat_handle c, void* stream)
{
    if (!c) return ODOM_ERR_ARG;
    SCHK(hipStreamSynchronize(c->stream));
    c->stream = stream ? (hipStream_t)stream : c->own;
    return ODOM_OK;
}

int splat_combined_predict(splat_handle c, const float pose[16], const float* model, int n, int mem,
                           float depth_cutoff, float conf_threshold, int time, int max_time, int time_delta)
{
    if (!c || !pose || n < 0 || (n > 0 && !model)) return ODOM_ERR_ARG;
    if (model && ((uintptr_t)model % 16) != 0) return ODOM_ERR_ARG;
    const float* dmodel = model;
    if (mem == ODOM_MEM_HOST) {
        if (n > c->staging_cap) return ODOM_ERR_ARG;
        if (n > 0) {
            SCHK(hipMemcpyAsync(c->staging, model, (size_t)n * SPLAT_FLOATS * sizeof(float), hipMemcpyHostToDevice,
                                c->stream));
            SCHK(hipStreamSynchronize(c->stream));
        }
        dmodel = c->staging;
    } else if (mem != ODOM_MEM_DEVICE) {
        return ODOM_ERR_ARG;
    }
    SplatUniforms U;
    make_uniforms(c, pose, depth_cutoff, conf_threshold, time, max_time, time_delta, &U);
    return enqueue_predict(c, U, dmodel, n, false);
}

int splat_fill_in(splat_handle c, const odom_frame_view* depth, const odom_frame_view* rgb, int passthrough_geometry,
                  int passthrough_image)
{
    if (!c || !depth || !rgb) return ODOM_ERR_ARG;
    FillArgs A;
    const int rc = make_fill_args(c, depth, rgb, passthrough_geometry, passthrough_image, &A);
    if (rc) return rc;
    hipLaunchKernelGGL(k_fill_in, dim3(cdiv((long long)c->cam.w * c->cam.h, 256)), dim3(256), 0, c->stream, A);
    SCHK(hipGetLastError());
    return ODOM_OK;
}

int splat_get_buffer(splat_handle c, int which, void** dev_ptr)
{
    if (!c || !dev_ptr) return ODOM_ERR_ARG;
    void* p[] = {c->image, c->vertex, c->normal, c->time, c->fimage, c->fvertex, c->fnormal};
    if (which < 0 || which > SPLAT_FILL_NORMAL) return ODOM_ERR_ARG;
    *dev_ptr = p[which];
    return ODOM_OK;
}

int splat_read(splat_handle c, int which, void* host_dst)
{
    if (!c || !host_dst || which < 0 || which > SPLAT_FILL_NORMAL) return ODOM_ERR_ARG;
    static const int bpp[] = {4, 16, 16, 2, 4, 16, 16};
    void* src;
    splat_get_buffer(c, which, &src);
    SCHK(hipMemcpyAsync(host_dst, src, (size_t)c->cam.w * c->cam.h * bpp[which], hipMemcpyDeviceToHost, c->stream));
    SCHK(hipStreamSynchronize(c->stream));
    return ODOM_OK;
}

int splat_synchronize(splat_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    SCHK(hipStreamSynchronize(c->stream));
    return ODOM_OK;
}

int splat_profile(splat_handle c, const float pose[16], const float* model_dev, int n, float depth_cutoff,
                  float conf_threshold, int time, int max_time, int time_delta, const odom_frame_view* fill_depth,
                  const odom_frame_view* fill_rgb, int reps, float us[3], long long* fragments)
{
    if (!c || !pose || !us || reps < 1 || n < 1 || !model_dev || ((uintptr_t)model_dev % 16) != 0)
        return ODOM_ERR_ARG;
    if ((fill_depth == nullptr) != (fill_rgb == nullptr)) return ODOM_ERR_ARG;
    SplatUniforms U;
    make_uniforms(c, pose, depth_cutoff, conf_threshold, time, max_time, time_delta, &U);
    FillArgs A;
    if (fill_depth) {
        const int rc = make_fill_args(c, fill_depth, fill_rgb, 0, 0, &A);
        if (rc) return rc;
    }
    const int npix = c->cam.w * c->cam.h;
    TimingEvents ev(4);
    SCHK(ev.create());
    double acc[3] = {0, 0, 0};
    for (int r = 0; r < reps; ++r) {
        const bool last = r == reps - 1;
        if (last) SCHK(hipMemsetAsync(c->frag_count, 0, 8, c->stream));
        SCHK(hipEventRecord(ev[0], c->stream));
        hipLaunchKernelGGL(k_splat_raster, dim3(cdiv(n, 256)), dim3(256), 0, c->stream, (const float4*)model_dev, n,
                           U, c->rays, c->keys, last ? c->frag_count : nullptr);
        SCHK(hipEventRecord(ev[1], c->stream));
        hipLaunchKernelGGL(k_splat_resolve, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream,
                           (const float4*)model_dev, U, c->rays, c->keys, c->image, c->vertex, c->normal, c->time);
        SCHK(hipEventRecord(ev[2], c->stream));
        if (fill_depth)
            hipLaunchKernelGGL(k_fill_in, dim3(cdiv(npix, 256)), dim3(256), 0, c->stream, A);
        SCHK(hipEventRecord(ev[3], c->stream));
        SCHK(hipGetLastError());
        SCHK(hipEventSynchronize(ev[3]));
        float ms;
        for (int k = 0; k < 3; ++k) {
            SCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            acc[k] += ms;
        }
    }
    for (int k = 0; k < 3; ++k) us[k] = (float)(acc[k] * 1000.0 / reps);
    if (!fill_depth) us[2] = 0.0f;
    if (fragments) {
        unsigned long long f = 0;
        SCHK(hipMemcpy(&f, c->frag_count, 8, hipMemcpyDeviceToHost));
        *fragments = (long long)f;
    }
    return ODOM_OK;
}

}  // extern "C"
