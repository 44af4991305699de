// odom_kernels.hip -- gfx950 kernels of the dense RGB-D odometry path.
//
// Preprocessing (bilateral, pyramids, vertex/normal maps, model-map unpack,
// intensity, derivatives) and the Gauss-Newton steps: the fused ICP +
// photometric-association pass (A), the photometric pass (B) and the SO(3)
// pre-alignment.  Each step folds its per-pixel products through wave64
// shuffles and LDS into one block partial; the last workgroup to finish (an
// atomic ticket) reduces the partials of its pair in FP64 in a fixed order
// and runs the 6x6 solve and SE(3) update, so a tracking iteration never
// returns to the host.
//
// Every expression that decides a correspondence (projection, rounding,
// thresholds, NaN tests) is written in the reference's operation order and
// the file is compiled with -ffp-contract=off, so masks are bit-identical to
// the CPU restatement of the same algorithm.
#include <float.h>

#include "odom_internal.h"
#include "odom_math.h"

namespace slam {

#define QNAN __builtin_nanf("")

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

float bexpf_host(float x) { return bexpf(x); }

// =============================================================== a1 bilateral
// depth_bilateral.frag:30-76.  Output tile 32x32 per 256-thread block (4
// pixels per thread), input tile with a 6-pixel halo in LDS.  The weight is
// the product of a 13x13 spatial table and a range table over |value - tmp|
// (factors of exp(-(space2*0.024691358 + color2*0.000555556))), both built
// per block with bexpf.  Range factors for |value - tmp| >= 380 are 0.
constexpr int kBilatRange = 380;
constexpr int kBT = 32;                 // output tile edge
constexpr int kBL = kBT + 12;           // LDS tile edge

__global__ __launch_bounds__(256) void k_bilateral(const uint16_t* __restrict__ src, int pitch,
                                                   uint16_t* __restrict__ dst, int w, int h,
                                                   unsigned int maxv)
{
    __shared__ uint16_t tile[kBL * kBL];
    __shared__ float sES[169];
    __shared__ float sEC[kBilatRange + 1];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kBT, y0 = blockIdx.y * kBT;
    for (int i = tid; i < kBL * kBL; i += 256) {
        const int gx = x0 - 6 + i % kBL, gy = y0 - 6 + i / kBL;
        tile[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? src[(size_t)gy * pitch + gx] : (uint16_t)0;
    }
    for (int i = tid; i < 169; i += 256) {
        const int dy = i / 13 - 6, dx = i % 13 - 6;
        sES[i] = bexpf(-((float)(dx * dx + dy * dy) * 0.024691358f));
    }
    for (int i = tid; i <= kBilatRange; i += 256)
        sEC[i] = (i < kBilatRange) ? bexpf(-((float)(i * i) * 0.000555556f)) : 0.0f;
    __syncthreads();
    const int lx = tid & 31;
    for (int r = 0; r < 4; ++r) {
        const int ly = (tid >> 5) + 8 * r;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= w || y >= h) continue;
        const int value = tile[(ly + 6) * kBL + lx + 6];
        if ((unsigned int)value > maxv || value < 300) {
            dst[(size_t)y * w + x] = 0;
            continue;
        }
        const int cy0 = max(y - 6, 0), cy1 = min(y + 7, h);
        const int cx0 = max(x - 6, 0), cx1 = min(x + 7, w);
        float sum1 = 0.0f, sum2 = 0.0f;
        for (int cy = cy0; cy < cy1; ++cy) {
            const int trow = (cy - y0 + 6) * kBL - x0 + 6;
            const int erow = (cy - y + 6) * 13 + 6 - x;
            for (int cx = cx0; cx < cx1; ++cx) {
                const int t = tile[trow + cx];
                const int a = min(abs(value - t), kBilatRange);
                const float wgt = sES[erow + cx] * sEC[a];
                sum1 = sum1 + (float)t * wgt;
                sum2 = sum2 + wgt;
            }
        }
        dst[(size_t)y * w + x] = (uint16_t)(unsigned int)roundf(sum1 / sum2);
    }
}

void launch_bilateral(const uint16_t* src, int pitch, uint16_t* dst, int w, int h, float maxD, hipStream_t s)
{
    const unsigned int maxv = (unsigned int)(maxD * 1000.0f);
    dim3 grid(cdiv(w, kBT), cdiv(h, kBT));
    hipLaunchKernelGGL(k_bilateral, grid, dim3(256), 0, s, src, pitch, dst, w, h, maxv);
}

// ================================================================ a2 pyrDown
// utils.cu:57-94, sigma_color 30
__global__ __launch_bounds__(256) void k_pyrdown_u16(const uint16_t* __restrict__ src, int sw, int sh,
                                                     uint16_t* __restrict__ dst, int dw, int dh)
{
    const int x = blockIdx.x * 32 + (threadIdx.x & 31);
    const int y = blockIdx.y * 8 + (threadIdx.x >> 5);
    if (x >= dw || y >= dh) return;
    const float wts[3] = {0.375f, 0.25f, 0.0625f};
    const int center = src[(size_t)(2 * y) * sw + 2 * x];
    const int x_mi = max(0, 2 * x - 2) - 2 * x, y_mi = max(0, 2 * y - 2) - 2 * y;
    const int x_ma = min(sw, 2 * x + 3) - 2 * x, y_ma = min(sh, 2 * y + 3) - 2 * y;
    float sum = 0.0f, wall = 0.0f;
    for (int yi = y_mi; yi < y_ma; ++yi)
        for (int xi = x_mi; xi < x_ma; ++xi) {
            const int val = src[(size_t)(2 * y + yi) * sw + 2 * x + xi];
            if ((float)abs(val - center) < 90.0f) {
                sum = sum + (float)val * wts[abs(xi)] * wts[abs(yi)];
                wall = wall + wts[abs(xi)] * wts[abs(yi)];
            }
        }
    dst[(size_t)y * dw + x] = (uint16_t)(int)(sum / wall);
}

void launch_pyrdown_u16(const uint16_t* src, int w, int h, uint16_t* dst, hipStream_t s)
{
    const int dw = w / 2, dh = h / 2;
    hipLaunchKernelGGL(k_pyrdown_u16, dim3(cdiv(dw, 32), cdiv(dh, 8)), dim3(256), 0, s, src, w, h, dst, dw, dh);
}

// Level segments of a one-dimensional grid: blocks [b0[l], b0[l+1]) cover level l.
struct Seg3 {
    int b0[kLevels + 1];
};

__device__ __forceinline__ int seg_level(const Seg3& g, int b)
{
    return b < g.b0[1] ? 0 : (b < g.b0[2] ? 1 : 2);
}

static Seg3 make_seg(const int* n, int per_block)
{
    Seg3 g;
    g.b0[0] = 0;
    for (int l = 0; l < kLevels; ++l) g.b0[l + 1] = g.b0[l] + cdiv(n[l], per_block);
    return g;
}

// ========================================================= a3 + a4 vmap/nmap
// utils.cu:109-133 and 151-188, fused: the normal's two neighbours are
// recomputed from depth with the same expression as the vertex map.
__device__ __forceinline__ bool vpoint(const uint16_t* __restrict__ d, int w, int u, int v, float fxi,
                                       float fyi, float cx, float cy, float cutoff, float* o)
{
    const float z = (float)d[(size_t)v * w + u] / 1000.f;
    if (z != 0 && z < cutoff) {
        o[0] = z * ((float)u - cx) * fxi;
        o[1] = z * ((float)v - cy) * fyi;
        o[2] = z;
        return true;
    }
    return false;
}

__global__ __launch_bounds__(256) void k_vmap_nmap(VNLevels a, Seg3 g)
{
    const int l = seg_level(g, blockIdx.x);
    const int i = (blockIdx.x - g.b0[l]) * 256 + threadIdx.x;
    const int w = a.w[l], h = a.h[l];
    if (i >= w * h) return;
    const int v = i / w, u = i - v * w;
    const size_t P = (size_t)w * h;
    const uint16_t* d = a.depth[l];
    float* vm = a.vmap[l];
    float* nm = a.nmap[l];
    const float fxi = a.fx_inv[l], fyi = a.fy_inv[l], cx = a.cx[l], cy = a.cy[l];
    float p00[3];
    const bool ok00 = vpoint(d, w, u, v, fxi, fyi, cx, cy, a.cutoff, p00);
    if (ok00) {
        vm[i] = p00[0]; vm[i + P] = p00[1]; vm[i + 2 * P] = p00[2];
    } else {
        vm[i] = QNAN; vm[i + P] = QNAN; vm[i + 2 * P] = QNAN;
    }
    float n[3] = {QNAN, QNAN, QNAN};
    if (u != w - 1 && v != h - 1 && ok00) {
        float p01[3], p10[3];
        const bool ok01 = vpoint(d, w, u + 1, v, fxi, fyi, cx, cy, a.cutoff, p01);
        const bool ok10 = vpoint(d, w, u, v + 1, fxi, fyi, cx, cy, a.cutoff, p10);
        if (ok01 && ok10) {
            const float e0[3] = {p01[0] - p00[0], p01[1] - p00[1], p01[2] - p00[2]};
            const float e1[3] = {p10[0] - p00[0], p10[1] - p00[1], p10[2] - p00[2]};
            float r[3] = {e0[1] * e1[2] - e0[2] * e1[1], e0[2] * e1[0] - e0[0] * e1[2],
                          e0[0] * e1[1] - e0[1] * e1[0]};
            const float rn = 1.0f / sqrtf(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
            n[0] = r[0] * rn; n[1] = r[1] * rn; n[2] = r[2] * rn;
        }
    }
    nm[i] = n[0]; nm[i + P] = n[1]; nm[i + 2 * P] = n[2];
}

void launch_vmap_nmap(const VNLevels& a, hipStream_t s)
{
    int n[kLevels];
    for (int l = 0; l < kLevels; ++l) n[l] = a.w[l] * a.h[l];
    const Seg3 g = make_seg(n, 256);
    hipLaunchKernelGGL(k_vmap_nmap, dim3(g.b0[kLevels]), dim3(256), 0, s, a, g);
}

// ================================================ a5 + a6 + a7 model maps
// copyMaps (utils.cu:270-310), resizeVMap/resizeNMap (utils.cu:365-444) and
// tranformMaps (utils.cu:206-248).  A level-l pixel recomputes its
// camera-frame value from the 2^l x 2^l block of input texels with the
// reference's staged means (NaN if any child's x is NaN; normals
// renormalised at every stage), then applies the rigid transform.
struct Cam6 {
    float v[3], n[3];
};

__device__ __forceinline__ Cam6 model_l0(const ModelLevels& a, int x, int y)
{
    const float4 vv = a.v4[(size_t)y * a.v_pitch + x];
    const float4 nn = a.n4[(size_t)y * a.n_pitch + x];
    Cam6 c;
    if (!(vv.z == 0)) {
        c.v[0] = vv.x; c.v[1] = vv.y; c.v[2] = vv.z;
        c.n[0] = nn.x; c.n[1] = nn.y; c.n[2] = nn.z;
    } else {
        c.v[0] = c.v[1] = c.v[2] = QNAN;
        c.n[0] = c.n[1] = c.n[2] = QNAN;
    }
    return c;
}

__device__ __forceinline__ void mean4(const float* a, const float* b, const float* c, const float* d,
                                      bool normalize, float* o)
{
    if (isnan(a[0]) || isnan(b[0]) || isnan(c[0]) || isnan(d[0])) {
        o[0] = o[1] = o[2] = QNAN;
        return;
    }
    for (int k = 0; k < 3; ++k) o[k] = (a[k] + b[k] + c[k] + d[k]) / 4.0f;
    if (normalize) {
        const float rn = 1.0f / sqrtf(o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
        o[0] = o[0] * rn; o[1] = o[1] * rn; o[2] = o[2] * rn;
    }
}

__device__ __forceinline__ Cam6 model_l1(const ModelLevels& a, int x, int y)
{
    const Cam6 c00 = model_l0(a, 2 * x, 2 * y), c01 = model_l0(a, 2 * x + 1, 2 * y);
    const Cam6 c10 = model_l0(a, 2 * x, 2 * y + 1), c11 = model_l0(a, 2 * x + 1, 2 * y + 1);
    Cam6 o;
    mean4(c00.v, c01.v, c10.v, c11.v, false, o.v);
    mean4(c00.n, c01.n, c10.n, c11.n, true, o.n);
    return o;
}

__device__ __forceinline__ Cam6 model_l2(const ModelLevels& a, int x, int y)
{
    const Cam6 c00 = model_l1(a, 2 * x, 2 * y), c01 = model_l1(a, 2 * x + 1, 2 * y);
    const Cam6 c10 = model_l1(a, 2 * x, 2 * y + 1), c11 = model_l1(a, 2 * x + 1, 2 * y + 1);
    Cam6 o;
    mean4(c00.v, c01.v, c10.v, c11.v, false, o.v);
    mean4(c00.n, c01.n, c10.n, c11.n, true, o.n);
    return o;
}

__global__ __launch_bounds__(256) void k_model_maps(ModelLevels a, Seg3 g)
{
    const int l = seg_level(g, blockIdx.x);
    const int i = (blockIdx.x - g.b0[l]) * 256 + threadIdx.x;
    const int w = a.w[l], h = a.h[l];
    if (i >= w * h) return;
    const int y = i / w, x = i - y * w;
    Cam6 c = (l == 0) ? model_l0(a, x, y) : (l == 1 ? model_l1(a, x, y) : model_l2(a, x, y));
    if (l == 0 && a.vz) a.vz[i] = a.v4[(size_t)y * a.v_pitch + x].z;
    if (a.transform) {
        const float* R = a.R;
        if (!isnan(c.v[0])) {
            const float v0 = c.v[0], v1 = c.v[1], v2 = c.v[2];
            for (int r = 0; r < 3; ++r) c.v[r] = R[3 * r] * v0 + R[3 * r + 1] * v1 + R[3 * r + 2] * v2 + a.t[r];
        } else {
            c.v[0] = c.v[1] = c.v[2] = QNAN;
        }
        if (!isnan(c.n[0])) {
            const float n0 = c.n[0], n1 = c.n[1], n2 = c.n[2];
            for (int r = 0; r < 3; ++r) c.n[r] = R[3 * r] * n0 + R[3 * r + 1] * n1 + R[3 * r + 2] * n2;
        } else {
            c.n[0] = c.n[1] = c.n[2] = QNAN;
        }
    }
    const size_t P = (size_t)w * h;
    float* vm = a.vmap[l];
    float* nm = a.nmap[l];
    vm[i] = c.v[0]; vm[i + P] = c.v[1]; vm[i + 2 * P] = c.v[2];
    nm[i] = c.n[0]; nm[i + P] = c.n[1]; nm[i + 2 * P] = c.n[2];
}

void launch_model_maps(const ModelLevels& a, hipStream_t s)
{
    int n[kLevels];
    for (int l = 0; l < kLevels; ++l) n[l] = a.w[l] * a.h[l];
    const Seg3 g = make_seg(n, 256);
    hipLaunchKernelGGL(k_model_maps, dim3(g.b0[kLevels]), dim3(256), 0, s, a, g);
}

// ================================================== a8 verticesToDepth
// utils.cu:526-537 (input: the z plane of the last vertex map)
__global__ __launch_bounds__(256) void k_vertices_to_depth(const float* __restrict__ vz, int n, float cutoff,
                                                           float* __restrict__ depth)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float z = vz[i];
    depth[i] = (z > cutoff || z <= 0) ? QNAN : z;
}

void launch_vertices_to_depth(const float* vz, int w, int h, float cutoff, float* depth, hipStream_t s)
{
    const int n = w * h;
    hipLaunchKernelGGL(k_vertices_to_depth, dim3(cdiv(n, 256)), dim3(256), 0, s, vz, n, cutoff, depth);
}

// ============================================== a9 + a11 Gaussian pyrDown
// utils.cu:332-363 and 470-500: loop bound min(2x+3, cols-1) and the
// reversed weight index (ty-cy-1)*5+(tx-cx-1) are kept.
__constant__ float c_gauss25[25] = {1, 4, 6, 4, 1, 4, 16, 24, 16, 4, 6, 24, 36, 24, 6,
                                    4, 16, 24, 16, 4, 1, 4, 6, 4, 1};

__global__ __launch_bounds__(256) void k_pyrdown_f32(const float* __restrict__ src, int sw, int sh,
                                                     float* __restrict__ dst, int dw, int dh)
{
    const int x = blockIdx.x * 32 + (threadIdx.x & 31);
    const int y = blockIdx.y * 8 + (threadIdx.x >> 5);
    if (x >= dw || y >= dh) return;
    const int tx = min(2 * x + 3, sw - 1), ty = min(2 * y + 3, sh - 1);
    float sum = 0.0f;
    int count = 0;
    for (int cy = max(0, 2 * y - 2); cy < ty; ++cy)
        for (int cx = max(0, 2 * x - 2); cx < tx; ++cx) {
            const float v = src[(size_t)cy * sw + cx];
            if (!isnan(v)) {
                const float gk = c_gauss25[(ty - cy - 1) * 5 + (tx - cx - 1)];
                sum = sum + v * gk;
                count = (int)((float)count + gk);
            }
        }
    dst[(size_t)y * dw + x] = sum / (float)count;
}

void launch_pyrdown_f32(const float* src, int w, int h, float* dst, hipStream_t s)
{
    const int dw = w / 2, dh = h / 2;
    hipLaunchKernelGGL(k_pyrdown_f32, dim3(cdiv(dw, 32), cdiv(dh, 8)), dim3(256), 0, s, src, w, h, dst, dw, dh);
}

__global__ __launch_bounds__(256) void k_pyrdown_u8(const uint8_t* __restrict__ src, int sw, int sh,
                                                    uint8_t* __restrict__ dst, int dw, int dh)
{
    const int x = blockIdx.x * 32 + (threadIdx.x & 31);
    const int y = blockIdx.y * 8 + (threadIdx.x >> 5);
    if (x >= dw || y >= dh) return;
    const int tx = min(2 * x + 3, sw - 1), ty = min(2 * y + 3, sh - 1);
    float sum = 0.0f;
    int count = 0;
    for (int cy = max(0, 2 * y - 2); cy < ty; ++cy)
        for (int cx = max(0, 2 * x - 2); cx < tx; ++cx) {
            const uint8_t v = src[(size_t)cy * sw + cx];
            if (v > 0) {
                const float gk = c_gauss25[(ty - cy - 1) * 5 + (tx - cx - 1)];
                sum = sum + (float)v * gk;
                count = (int)((float)count + gk);
            }
        }
    const float q = sum / (float)count;
    dst[(size_t)y * dw + x] = (count == 0) ? (uint8_t)0 : (uint8_t)q;
}

void launch_pyrdown_u8(const uint8_t* src, int w, int h, uint8_t* dst, hipStream_t s)
{
    const int dw = w / 2, dh = h / 2;
    hipLaunchKernelGGL(k_pyrdown_u8, dim3(cdiv(dw, 32), cdiv(dh, 8)), dim3(256), 0, s, src, w, h, dst, dw, dh);
}

// ====================================================== a10 intensity
// utils.cu:550-563: 0.114/0.299/0.587 on bytes 0/1/2, truncated
__global__ __launch_bounds__(256) void k_intensity(const uint8_t* __restrict__ rgb, int pitch, int ch, int w,
                                                   int h, uint8_t* __restrict__ dst)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= w || y >= h) return;
    const uint8_t* p = rgb + (size_t)y * pitch + (size_t)x * ch;
    const int value = (int)((float)p[0] * 0.114f + (float)p[1] * 0.299f + (float)p[2] * 0.587f);
    dst[(size_t)y * w + x] = (uint8_t)value;
}

void launch_intensity(const uint8_t* rgb, int pitch, int ch, int w, int h, uint8_t* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_intensity, dim3(cdiv(w, 64), cdiv(h, 4)), dim3(256), 0, s, rgb, pitch, ch, w, h, dst);
}

// ================================ a12 derivatives + photometric source mask
// computeDerivativeImages (utils.cu:582-606: kernelIndex counts down over the
// taps visited, so border pixels shift the kernel) and the pixel tests of
// RGBResidual::getProducts that do not depend on the pose (reduce.cu:780-810:
// j < cols-5, i < rows-1, the 4x4 window of nextImage > 0, gradient
// magnitude >= minScale, nextDepth not NaN).  zsrc = nextDepth where all of
// them hold, NaN elsewhere.
__constant__ float c_gsx[9] = {0.52201f, 0.00000f, -0.52201f, 0.79451f, -0.00000f, -0.79451f,
                               0.52201f, 0.00000f, -0.52201f};
__constant__ float c_gsy[9] = {0.52201f, 0.79451f, 0.52201f, 0.00000f, 0.00000f, 0.00000f,
                               -0.52201f, -0.79451f, -0.52201f};

__global__ __launch_bounds__(256) void k_rgb_prep(PrepLevels a, Seg3 g)
{
    const int l = seg_level(g, blockIdx.x);
    const int k = (blockIdx.x - g.b0[l]) * 256 + threadIdx.x;
    const int w = a.w[l], h = a.h[l];
    if (k >= w * h) return;
    const int y = k / w, x = k - y * w;
    const uint8_t* img = a.img[l];
    float dxv = 0.0f, dyv = 0.0f;
    int ki = 8;
    for (int j = max(y - 1, 0); j <= min(y + 1, h - 1); j++)
        for (int i = max(x - 1, 0); i <= min(x + 1, w - 1); i++) {
            const float sv = (float)img[(size_t)j * w + i];
            dxv = dxv + sv * c_gsx[ki];
            dyv = dyv + sv * c_gsy[ki];
            --ki;
        }
    const short sx = (short)dxv, sy = (short)dyv;
    a.dxy[l][k] = make_short2(sx, sy);
    float z = QNAN;
    if (x < w - 5 && y < h - 1) {
        bool valid = true;
        for (int u = max(y - 2, 0); u < min(y + 2, h); u++)
            for (int v = max(x - 2, 0); v < min(x + 2, w); v++) valid = valid && (img[(size_t)u * w + v] > 0);
        if (valid) {
            const float mTwo = (float)((int)sx * (int)sx + (int)sy * (int)sy);
            if (mTwo >= a.minScale[l]) {
                const float d1 = a.depth[l][k];
                if (!isnan(d1)) z = d1;
            }
        }
    }
    a.zsrc[l][k] = z;
}

void launch_rgb_prep(const PrepLevels& a, hipStream_t s)
{
    int n[kLevels];
    for (int l = 0; l < kLevels; ++l) n[l] = a.w[l] * a.h[l];
    const Seg3 g = make_seg(n, 256);
    hipLaunchKernelGGL(k_rgb_prep, dim3(g.b0[kLevels]), dim3(256), 0, s, a, g);
}

__global__ __launch_bounds__(256) void k_copy_u8(const uint8_t* __restrict__ src, uint8_t* __restrict__ dst,
                                                 size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

void launch_copy_u8(const uint8_t* src, uint8_t* dst, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, src, dst, n);
}

// std::swap(lastNextImage[i], nextImage[i]) (RGBDOdometryef.cpp:585-591) as a
// swap of contents, so buffer addresses never change.
__global__ __launch_bounds__(256) void k_swap_u8(uint8_t* __restrict__ a, uint8_t* __restrict__ b, size_t n)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const uint8_t t = a[i];
        a[i] = b[i];
        b[i] = t;
    }
}

void launch_swap_u8(uint8_t* a, uint8_t* b, size_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_swap_u8, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a, b, n);
}

// ============================================================ reductions
template <int V>
__device__ __forceinline__ void wave_sum(float (&a)[V])
{
#pragma unroll
    for (int k = 0; k < V; ++k) {
        float v = a[k];
        v += __shfl_xor(v, 32);
        v += __shfl_xor(v, 16);
        v += __shfl_xor(v, 8);
        v += __shfl_xor(v, 4);
        v += __shfl_xor(v, 2);
        v += __shfl_xor(v, 1);
        a[k] = v;
    }
}

// Block partial of V floats: wave sums, then waves in order.  Deterministic.
template <int V, int NT>
__device__ __forceinline__ void block_partial(float (&a)[V], float* __restrict__ out, float (*red)[V])
{
    wave_sum<V>(a);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0)
        for (int k = 0; k < V; ++k) red[wid][k] = a[k];
    __syncthreads();
    if (threadIdx.x < V) {
        float s = 0.0f;
        for (int wv = 0; wv < NT / 64; ++wv) s += red[wv][threadIdx.x];
        out[threadIdx.x] = s;
    }
}

// FP64 total over nblk block partials (stride kPartStride), fixed order.
template <int V, int NT>
__device__ __forceinline__ void final_sum(const float* __restrict__ part, int nblk, double* tot,
                                          double (*red)[NT / 32])
{
    constexpr int G = NT / 32;
    const int t = threadIdx.x;
    if (t < V * G) {
        const int k = t / G, gi = t - k * G;
        double s = 0.0;
        for (int b = gi; b < nblk; b += G) s += (double)part[(size_t)b * kPartStride + k];
        red[k][gi] = s;
    }
    __syncthreads();
    if (t < V) {
        double s = 0.0;
        for (int gi = 0; gi < G; ++gi) s += red[t][gi];
        tot[t] = s;
    }
    __syncthreads();
}

// Last-block-done ticket.  Every thread's partial writes are released at
// device scope before thread 0 takes a ticket; the last block acquires.
__device__ __forceinline__ bool take_ticket(unsigned int* ticket, int nblk, int* s_last)
{
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned int prev = atomicAdd(ticket, 1u);
        *s_last = (prev == (unsigned int)(nblk - 1));
    }
    __syncthreads();
    const bool last = *s_last != 0;
    if (last) __threadfence();
    return last;
}

// ======================================================= per-pixel steps
// a14 ICPReduction::search/getProducts, reduce.cu:282-387.  A pixel whose
// vertex x is NaN is rejected up front: in the reference its NaN reaches
// `dist` and fails `dist <= distThres`, so the outcome is the same.
struct IcpC {
    float Rc[9], tc[3], Rpi[9], tp[3];
    float fx, fy, cx, cy, dist, angle;
};

__device__ __forceinline__ bool icp_pixel(const LevelDesc& L, const IcpC& c, int i, size_t P, float* row)
{
    const float vx = L.vcur[i];
    if (isnan(vx)) return false;
    const float vy = L.vcur[i + P], vz = L.vcur[i + 2 * P];
    float vg[3], d3[3], vcp[3];
    for (int r = 0; r < 3; ++r) vg[r] = c.Rc[3 * r] * vx + c.Rc[3 * r + 1] * vy + c.Rc[3 * r + 2] * vz + c.tc[r];
    for (int r = 0; r < 3; ++r) d3[r] = vg[r] - c.tp[r];
    for (int r = 0; r < 3; ++r) vcp[r] = c.Rpi[3 * r] * d3[0] + c.Rpi[3 * r + 1] * d3[1] + c.Rpi[3 * r + 2] * d3[2];
    const int ux = f2i_rn(vcp[0] * c.fx / vcp[2] + c.cx);
    const int uy = f2i_rn(vcp[1] * c.fy / vcp[2] + c.cy);
    if (ux < 0 || uy < 0 || ux >= L.w || uy >= L.h || vcp[2] < 0) return false;
    const size_t j = (size_t)uy * L.w + ux;
    const float vp[3] = {L.vprev[j], L.vprev[j + P], L.vprev[j + 2 * P]};
    const float np[3] = {L.nprev[j], L.nprev[j + P], L.nprev[j + 2 * P]};
    const float nc[3] = {L.ncur[i], L.ncur[i + P], L.ncur[i + 2 * P]};
    float ncg[3];
    for (int r = 0; r < 3; ++r) ncg[r] = c.Rc[3 * r] * nc[0] + c.Rc[3 * r + 1] * nc[1] + c.Rc[3 * r + 2] * nc[2];
    const float df[3] = {vp[0] - vg[0], vp[1] - vg[1], vp[2] - vg[2]};
    const float dist = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
    const float cr[3] = {ncg[1] * np[2] - ncg[2] * np[1], ncg[2] * np[0] - ncg[0] * np[2],
                         ncg[0] * np[1] - ncg[1] * np[0]};
    const float sine = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    if (!(sine < c.angle && dist <= c.dist && !isnan(nc[0]) && !isnan(np[0]))) return false;
    float dd[3], d[3], n[3];
    for (int r = 0; r < 3; ++r) dd[r] = vp[r] - c.tp[r];
    for (int r = 0; r < 3; ++r) d[r] = c.Rpi[3 * r] * dd[0] + c.Rpi[3 * r + 1] * dd[1] + c.Rpi[3 * r + 2] * dd[2];
    for (int r = 0; r < 3; ++r) n[r] = c.Rpi[3 * r] * np[0] + c.Rpi[3 * r + 1] * np[1] + c.Rpi[3 * r + 2] * np[2];
    const float* s = vcp;  // Rprev_inv * (vcurr_g - tprev), bit for bit
    row[0] = n[0]; row[1] = n[1]; row[2] = n[2];
    row[3] = s[1] * n[2] - s[2] * n[1];
    row[4] = s[2] * n[0] - s[0] * n[2];
    row[5] = s[0] * n[1] - s[1] * n[0];
    const float sd[3] = {s[0] - d[0], s[1] - d[1], s[2] - d[2]};
    row[6] = n[0] * sd[0] + n[1] * sd[1] + n[2] * sd[2];
    return true;
}

// a15 RGBResidual::getProducts, reduce.cu:808-830 (the pose-independent
// tests are folded into zsrc by k_rgb_prep).
struct RgbC {
    float K[9], kt[3];
    float maxDepthDelta;
};

__device__ __forceinline__ bool rgb_pixel(const LevelDesc& L, const RgbC& c, int x, int y, int idx,
                                          float& diff, float& Z, int& u0, int& v0)
{
    const float d1 = L.zsrc[idx];
    if (isnan(d1)) return false;
    const float xf = (float)x, yf = (float)y;
    const float tz = d1 * (c.K[6] * xf + c.K[7] * yf + c.K[8]) + c.kt[2];
    u0 = f2i_rn((d1 * (c.K[0] * xf + c.K[1] * yf + c.K[2]) + c.kt[0]) / tz);
    v0 = f2i_rn((d1 * (c.K[3] * xf + c.K[4] * yf + c.K[5]) + c.kt[1]) / tz);
    if (!(u0 >= 0 && v0 >= 0 && u0 < L.w && v0 < L.h)) return false;
    const size_t m = (size_t)v0 * L.w + u0;
    const float d0 = L.zlast[m];
    const uint8_t il = L.ilast[m];
    if (!(d0 > 0 && fabsf(tz - d0) <= c.maxDepthDelta && il != 0)) return false;
    diff = (float)L.inext[idx] - (float)il;
    Z = d0;
    return true;
}

// a16 RGBReduction::getProducts, reduce.cu:512-552; the cloud point is
// projectPointsKernel's (utils.cu:655-657) recomputed from lastDepth.
__device__ __forceinline__ void rgb_row(const LevelDesc& L, float sigma, float sobelScale, int idx, float diff,
                                        float Z, int u0, int v0, float* row)
{
    float wt = sigma + fabsf(diff);
    wt = wt > FLT_EPSILON ? 1.0f / wt : 1.0f;
    if (sigma == -1) wt = 1;
    row[6] = -wt * diff;
    const float X = ((float)u0 - L.cx) * Z * L.invFx;
    const float Y = ((float)v0 - L.cy) * Z * L.invFy;
    const float invz = 1.0f / Z;
    const short2 gr = L.dxy[idx];
    const float gx = wt * sobelScale * (float)gr.x;
    const float gy = wt * sobelScale * (float)gr.y;
    const float a0 = gx * L.fx * invz;
    const float a1 = gy * L.fy * invz;
    const float a2 = -(a0 * X + a1 * Y) * invz;
    row[0] = a0;
    row[1] = a1;
    row[2] = a2;
    row[3] = -Z * a1 + Y * a2;
    row[4] = Z * a0 - X * a2;
    row[5] = -Y * a0 + X * a1;
}

__device__ __forceinline__ void acc_row7(float (&acc)[kSE3], const float* row)
{
    int s = 0;
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 7; ++j) {
            acc[s] = __builtin_fmaf(row[i], row[j], acc[s]);
            ++s;
        }
    acc[27] = __builtin_fmaf(row[6], row[6], acc[27]);
    acc[28] += 1.0f;
}

// ===================================================== single-thread math
// Per-iteration warp constants from resultRt (RGBDOdometryef.cpp:422-432):
// Rt = resultRt^-1 taken as the rigid inverse [R^T | -R^T t].
__device__ void iter_constants(TrackState* st, const LevelDesc& L)
{
    const double* Rt = st->resultRt;
    double Rinv[9], tinv[3], K[9], Ki[9], KR[9], KRKi[9];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Rinv[a * 3 + b] = Rt[b * 4 + a];
    for (int a = 0; a < 3; ++a) tinv[a] = -(Rinv[a * 3] * Rt[3] + Rinv[a * 3 + 1] * Rt[7] + Rinv[a * 3 + 2] * Rt[11]);
    kmat(L.fx, L.fy, L.cx, L.cy, K, Ki);
    mm3d(K, Rinv, KR);
    mm3d(KR, Ki, KRKi);
    for (int a = 0; a < 9; ++a) st->krkinv[a] = (float)KRKi[a];
    for (int a = 0; a < 3; ++a) st->kt[a] = (float)(K[a * 3] * tinv[0] + K[a * 3 + 1] * tinv[1] + K[a * 3 + 2] * tinv[2]);
}

// RGBDOdometryef.cpp:533-575: combine, solve, computeUpdateSE3, pose update.
__device__ void solve_update(TrackState* st, const TrackParams& P)
{
    double Ai[36], bi[6], Ar[36], br[6], A[36], b[6], x[6];
    for (int k = 0; k < 36; ++k) { Ai[k] = 0; Ar[k] = 0; }
    for (int k = 0; k < 6; ++k) { bi[k] = 0; br[k] = 0; }
    if (P.icp) unpack_se3(st->icp_sums, Ai, bi);
    if (P.rgb) unpack_se3(st->rgb_sums, Ar, br);
    if (P.icp && P.rgb) {
        const double wgt = P.icpWeight;
        for (int k = 0; k < 36; ++k) A[k] = Ar[k] + wgt * wgt * Ai[k];
        for (int k = 0; k < 6; ++k) b[k] = br[k] + wgt * bi[k];
    } else if (P.icp) {
        for (int k = 0; k < 36; ++k) A[k] = Ai[k];
        for (int k = 0; k < 6; ++k) b[k] = bi[k];
    } else {
        for (int k = 0; k < 36; ++k) A[k] = Ar[k];
        for (int k = 0; k < 6; ++k) b[k] = br[k];
    }
    for (int k = 0; k < 36; ++k) st->lastA[k] = A[k];
    for (int k = 0; k < 6; ++k) st->lastb[k] = b[k];
    const int zeros = ldlt_solve(6, A, b, x);
    st->degenerate = (zeros == 6);
    for (int k = 0; k < 6; ++k) st->lastx[k] = x[k];
    double rot[9], upd[16], nrt[16];
    rodrigues(x + 3, rot);
    for (int a = 0; a < 16; ++a) upd[a] = 0.0;
    for (int a = 0; a < 3; ++a) {
        for (int c = 0; c < 3; ++c) upd[a * 4 + c] = rot[a * 3 + c];
        upd[a * 4 + 3] = x[a];
    }
    upd[15] = 1.0;
    mm4d(upd, st->resultRt, nrt);
    for (int a = 0; a < 16; ++a) st->resultRt[a] = nrt[a];
    float Rf[9], tf[3], Ri[9], ti[3], Rn[9];
    for (int a = 0; a < 3; ++a) {
        for (int c = 0; c < 3; ++c) Rf[a * 3 + c] = (float)nrt[a * 4 + c];
        tf[a] = (float)nrt[a * 4 + 3];
    }
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) Ri[a * 3 + c] = Rf[c * 3 + a];
    for (int a = 0; a < 3; ++a) ti[a] = -(Ri[a * 3] * tf[0] + Ri[a * 3 + 1] * tf[1] + Ri[a * 3 + 2] * tf[2]);
    mm3f(st->Rprev, Ri, Rn);
    for (int a = 0; a < 9; ++a) st->Rcurr[a] = Rn[a];
    for (int a = 0; a < 3; ++a)
        st->tcurr[a] = st->Rprev[a * 3] * ti[0] + st->Rprev[a * 3 + 1] * ti[1] + st->Rprev[a * 3 + 2] * ti[2] + st->tprev[a];
    st->iters_done++;
}

// RGBDOdometryef.cpp:434-471 and 509-510: photometric bookkeeping after the
// association pass.  Returns false when the rgbOnly early exit fires.
__device__ bool after_association(TrackState* st, const PairDesc& D, const TrackParams& P, int level, int iter,
                                  int break_next, int count, int sigma)
{
    float sigmaVal, rgbError;
    if (P.rgb) {
        st->rgb_count = count;
        st->rgb_sigma = sigma;
        sigmaVal = (float)sqrt((double)((((float)sigma / (float)count) == 0.0f) ? 1 : count));
        rgbError = (float)(sqrt((double)sigma) / (double)(count == 0 ? 1 : count));
    } else {
        sigmaVal = 0.0f;
        rgbError = 0.0f;
        count = 0;
    }
    const float cmp = (iter == 0) ? FLT_MAX : st->lastRGBError;
    if (P.rgbOnly && rgbError > cmp) {
        // the level ends here; the next level starts from the unchanged increment
        st->break_level = level;
        if (break_next >= 0) iter_constants(st, D.lv[break_next]);
        return false;
    }
    st->lastRGBError = rgbError;
    st->lastRGBCount = (float)count;
    if (P.rgbOnly) sigmaVal = -1;
    st->sigmaVal = sigmaVal;
    if (P.icp) {
        st->res0 = (float)st->icp_sums[27];
        st->res1 = (float)st->icp_sums[28];
    }
    st->lastICPError = sqrtf(st->res0) / st->res1;
    st->lastICPCount = st->res1;
    return true;
}

// ======================================================== step kernel A
// ICP products (a14) and the photometric association count (a15) of one
// level of every pair.  The last block per pair reduces in FP64 and either
// records the sums (mode 1) or continues the Gauss-Newton iteration.
template <int NT, int PPT>
__global__ __launch_bounds__(NT) void k_step_a(const PairDesc* __restrict__ pairs, const int* __restrict__ blk_pair,
                                               int level, int iter, int next_level, int break_next, TrackParams P,
                                               int mode, uint8_t* __restrict__ dbg_mask)
{
    __shared__ float s_red[NT / 64][kSE3];
    __shared__ unsigned int s_redi[NT / 64][2];
    __shared__ double s_fin[kSE3][NT / 32];
    __shared__ double s_tot[kSE3];
    __shared__ int s_last;
    const int p = blk_pair ? blk_pair[blockIdx.x] : 0;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    if (mode == 0 && st->break_level == level) return;
    const LevelDesc L = D.lv[level];
    const int lb = blockIdx.x - D.blk0_a[level];
    const int N = L.w * L.h;
    const size_t P3 = (size_t)N;
    IcpC ic;
    RgbC rc;
    for (int k = 0; k < 9; ++k) { ic.Rc[k] = st->Rcurr[k]; ic.Rpi[k] = st->Rprev_inv[k]; rc.K[k] = st->krkinv[k]; }
    for (int k = 0; k < 3; ++k) { ic.tc[k] = st->tcurr[k]; ic.tp[k] = st->tprev[k]; rc.kt[k] = st->kt[k]; }
    ic.fx = L.fx; ic.fy = L.fy; ic.cx = L.cx; ic.cy = L.cy;
    ic.dist = P.distThres; ic.angle = P.angleThres;
    rc.maxDepthDelta = P.maxDepthDelta;

    float acc[kSE3];
#pragma unroll
    for (int k = 0; k < kSE3; ++k) acc[k] = 0.0f;
    unsigned int cnt = 0, sig = 0;
    const int base = lb * NT * PPT + threadIdx.x;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int idx = base + k * NT;
        if (idx < N) {
            unsigned int m = 0;
            if (P.icp) {
                float row[7];
                if (icp_pixel(L, ic, idx, P3, row)) {
                    acc_row7(acc, row);
                    m |= 1u;
                }
            }
            if (P.rgb) {
                const int y = idx / L.w, x = idx - y * L.w;
                float diff, Z;
                int u0, v0;
                if (rgb_pixel(L, rc, x, y, idx, diff, Z, u0, v0)) {
                    cnt += 1u;
                    sig += (unsigned int)(int)(diff * diff);
                    m |= 2u;
                }
            }
            if (dbg_mask) dbg_mask[idx] = (uint8_t)m;
        }
    }
    block_partial<kSE3, NT>(acc, D.part_a + (size_t)lb * kPartStride, s_red);
    {
        // integer partials: count and the wrapping int32 sum of diff^2
        unsigned int c2[2] = {cnt, sig};
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            c2[0] += __shfl_xor(c2[0], off);
            c2[1] += __shfl_xor(c2[1], off);
        }
        if ((threadIdx.x & 63) == 0) { s_redi[threadIdx.x >> 6][0] = c2[0]; s_redi[threadIdx.x >> 6][1] = c2[1]; }
        __syncthreads();
        if (threadIdx.x < 2) {
            unsigned int s = 0;
            for (int wv = 0; wv < NT / 64; ++wv) s += s_redi[wv][threadIdx.x];
            D.part_ai[(size_t)lb * 2 + threadIdx.x] = s;
        }
    }
    const int nblk = D.nblk_a[level];
    if (!take_ticket(&st->ticket_a, nblk, &s_last)) return;
    final_sum<kSE3, NT>(D.part_a, nblk, s_tot, s_fin);
    if (threadIdx.x == 0) {
        unsigned int c = 0, sg = 0;
        for (int b = 0; b < nblk; ++b) { c += D.part_ai[2 * b]; sg += D.part_ai[2 * b + 1]; }
        st->ticket_a = 0;
        if (P.icp)
            for (int k = 0; k < kSE3; ++k) st->icp_sums[k] = s_tot[k];
        if (mode == 1) {
            st->rgb_count = (int)c;
            st->rgb_sigma = (int)sg;
        } else if (after_association(st, D, P, level, iter, break_next, (int)c, (int)sg) && !P.rgb) {
            solve_update(st, P);
            if (next_level >= 0) iter_constants(st, D.lv[next_level]);
        }
    }
}

// ======================================================== step kernel B
// Photometric products (a16) with the robust weight of the global count, then
// (mode 0) the combined solve and SE(3) update by the last block per pair.
template <int NT, int PPT>
__global__ __launch_bounds__(NT) void k_step_b(const PairDesc* __restrict__ pairs, const int* __restrict__ blk_pair,
                                               int level, int iter, int next_level, TrackParams P, int mode)
{
    __shared__ float s_red[NT / 64][kSE3];
    __shared__ double s_fin[kSE3][NT / 32];
    __shared__ double s_tot[kSE3];
    __shared__ int s_last;
    const int p = blk_pair ? blk_pair[blockIdx.x] : 0;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    if (mode == 0 && st->break_level == level) return;
    const LevelDesc L = D.lv[level];
    const int lb = blockIdx.x - D.blk0_b[level];
    const int N = L.w * L.h;
    RgbC rc;
    for (int k = 0; k < 9; ++k) rc.K[k] = st->krkinv[k];
    for (int k = 0; k < 3; ++k) rc.kt[k] = st->kt[k];
    rc.maxDepthDelta = P.maxDepthDelta;
    const float sigma = st->sigmaVal;

    float acc[kSE3];
#pragma unroll
    for (int k = 0; k < kSE3; ++k) acc[k] = 0.0f;
    const int base = lb * NT * PPT + threadIdx.x;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
        const int idx = base + k * NT;
        if (idx < N) {
            const int y = idx / L.w, x = idx - y * L.w;
            float diff, Z;
            int u0, v0;
            if (rgb_pixel(L, rc, x, y, idx, diff, Z, u0, v0)) {
                float row[7];
                rgb_row(L, sigma, P.sobelScale, idx, diff, Z, u0, v0, row);
                acc_row7(acc, row);
            }
        }
    }
    block_partial<kSE3, NT>(acc, D.part_b + (size_t)lb * kPartStride, s_red);
    const int nblk = D.nblk_b[level];
    if (!take_ticket(&st->ticket_b, nblk, &s_last)) return;
    final_sum<kSE3, NT>(D.part_b, nblk, s_tot, s_fin);
    if (threadIdx.x == 0) {
        st->ticket_b = 0;
        for (int k = 0; k < kSE3; ++k) st->rgb_sums[k] = s_tot[k];
        if (mode == 0) {
            solve_update(st, P);
            if (next_level >= 0) iter_constants(st, D.lv[next_level]);
        }
    }
}

int step_shape_supported(StepShape sh)
{
    const bool t = sh.threads == 128 || sh.threads == 256 || sh.threads == 512;
    const bool p = sh.ppt == 1 || sh.ppt == 2 || sh.ppt == 4 || sh.ppt == 8;
    return t && p;
}

#define STEP_DISPATCH(KER, ...)                                                                         \
    do {                                                                                               \
        const dim3 grid(L.total_blocks);                                                               \
        switch (L.shape.threads * 100 + L.shape.ppt) {                                                 \
            case 12801: hipLaunchKernelGGL((KER<128, 1>), grid, dim3(128), 0, s, __VA_ARGS__); break;  \
            case 12802: hipLaunchKernelGGL((KER<128, 2>), grid, dim3(128), 0, s, __VA_ARGS__); break;  \
            case 12804: hipLaunchKernelGGL((KER<128, 4>), grid, dim3(128), 0, s, __VA_ARGS__); break;  \
            case 12808: hipLaunchKernelGGL((KER<128, 8>), grid, dim3(128), 0, s, __VA_ARGS__); break;  \
            case 25601: hipLaunchKernelGGL((KER<256, 1>), grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 25602: hipLaunchKernelGGL((KER<256, 2>), grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 25604: hipLaunchKernelGGL((KER<256, 4>), grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 25608: hipLaunchKernelGGL((KER<256, 8>), grid, dim3(256), 0, s, __VA_ARGS__); break;  \
            case 51201: hipLaunchKernelGGL((KER<512, 1>), grid, dim3(512), 0, s, __VA_ARGS__); break;  \
            case 51202: hipLaunchKernelGGL((KER<512, 2>), grid, dim3(512), 0, s, __VA_ARGS__); break;  \
            case 51204: hipLaunchKernelGGL((KER<512, 4>), grid, dim3(512), 0, s, __VA_ARGS__); break;  \
            case 51208: hipLaunchKernelGGL((KER<512, 8>), grid, dim3(512), 0, s, __VA_ARGS__); break;  \
            default: break;                                                                            \
        }                                                                                              \
    } while (0)

void launch_step_a(const StepLaunch& L, int level, int iter, int next_level, int break_next, const TrackParams& P,
                   int mode, uint8_t* dbg_mask, hipStream_t s)
{
    STEP_DISPATCH(k_step_a, L.pairs, L.blk_pair, level, iter, next_level, break_next, P, mode, dbg_mask);
}

void launch_step_b(const StepLaunch& L, int level, int iter, int next_level, const TrackParams& P, int mode,
                   hipStream_t s)
{
    STEP_DISPATCH(k_step_b, L.pairs, L.blk_pair, level, iter, next_level, P, mode);
}

// ======================================================= begin / end
// RGBDOdometryef.cpp:278-282, 386-401: pose prior, its inverse, the
// increment seeded with the SO3 rotation, constants of the first iteration.
__global__ void k_track_begin(const PairDesc* __restrict__ pairs, int npairs, TrackParams P, int first_level)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) st->Rprev[a * 3 + b] = st->pose_out[a * 4 + b];
        st->tprev[a] = st->pose_out[a * 4 + 3];
    }
    inv3f(st->Rprev, st->Rprev_inv);
    for (int a = 0; a < 9; ++a) st->Rcurr[a] = st->Rprev[a];
    for (int a = 0; a < 3; ++a) st->tcurr[a] = st->tprev[a];
    for (int a = 0; a < 16; ++a) st->resultRt[a] = (a % 5 == 0) ? 1.0 : 0.0;
    if (P.so3)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) st->resultRt[a * 4 + b] = st->so3R[a * 3 + b];
    st->break_level = -1;
    st->iters_done = 0;
    st->degenerate = 0;
    st->reverted = 0;
    st->res0 = 0.0f;
    st->res1 = 0.0f;
    st->ticket_a = 0;
    st->ticket_b = 0;
    if (first_level >= 0) iter_constants(st, D.lv[first_level]);
}

void launch_track_begin(const PairDesc* pairs, int npairs, const TrackParams& P, int first_level, hipStream_t s)
{
    hipLaunchKernelGGL(k_track_begin, dim3(cdiv(npairs, 64)), dim3(64), 0, s, pairs, npairs, P, first_level);
}

// RGBDOdometryef.cpp:579-583, 593-594
__global__ void k_track_end(const PairDesc* __restrict__ pairs, int npairs, TrackParams P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    TrackState* st = pairs[p].st;
    float R[9], t[3];
    for (int a = 0; a < 9; ++a) R[a] = st->Rcurr[a];
    for (int a = 0; a < 3; ++a) t[a] = st->tcurr[a];
    if (P.rgb) {
        const float d0 = t[0] - st->tprev[0], d1 = t[1] - st->tprev[1], d2 = t[2] - st->tprev[2];
        if (sqrtf(d0 * d0 + d1 * d1 + d2 * d2) > 0.3f) {
            for (int a = 0; a < 9; ++a) R[a] = st->Rprev[a];
            for (int a = 0; a < 3; ++a) t[a] = st->tprev[a];
            st->reverted = 1;
        }
    }
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) st->pose_out[a * 4 + b] = R[a * 3 + b];
        st->pose_out[a * 4 + 3] = t[a];
    }
}

void launch_track_end(const PairDesc* pairs, int npairs, const TrackParams& P, hipStream_t s)
{
    hipLaunchKernelGGL(k_track_end, dim3(cdiv(npairs, 64)), dim3(64), 0, s, pairs, npairs, P);
}

// ================================================================= SO3
// a17 SO3Reduction (reduce.cu:937-1075) and the host loop of
// RGBDOdometryef.cpp:294-380, one workgroup per pair on pyramid level 2: all
// (at most 10) iterations, the FP64 3x3 solve and the convergence tests run
// inside one launch.
__device__ __forceinline__ void so3_grad(const uint8_t* img, int w, int x, int y, float* g)
{
    const float actu = (float)img[(size_t)y * w + x];
    float back = (float)img[(size_t)y * w + x - 1];
    float fore = (float)img[(size_t)y * w + x + 1];
    g[0] = ((back + actu) / 2.0f) - ((fore + actu) / 2.0f);
    back = (float)img[(size_t)(y - 1) * w + x];
    fore = (float)img[(size_t)(y + 1) * w + x];
    g[1] = ((back + actu) / 2.0f) - ((fore + actu) / 2.0f);
}

__device__ __forceinline__ bool so3_pixel(const uint8_t* __restrict__ last, const uint8_t* __restrict__ next,
                                          int w, int h, int x, int y, const float* H, const float* Ki,
                                          const float* KR, float* row)
{
    const float xf = (float)x, yf = (float)y;
    const float wp0 = H[0] * xf + H[1] * yf + H[2] * 1.0f;
    const float wp1 = H[3] * xf + H[4] * yf + H[5] * 1.0f;
    const float wp2 = H[6] * xf + H[7] * yf + H[8] * 1.0f;
    const int wx = f2i_rn(wp0 / wp2), wy = f2i_rn(wp1 / wp2);
    if (!(wx >= 1 && wx < w - 1 && wy >= 1 && wy < h - 1 && x >= 1 && x < w - 1 && y >= 1 && y < h - 1))
        return false;
    float gn[2], gl[2];
    so3_grad(next, w, wx, wy, gn);
    so3_grad(last, w, x, y, gl);
    const float gx = (gn[0] + gl[0]) / 2.0f;
    const float gy = (gn[1] + gl[1]) / 2.0f;
    const float pt0 = Ki[0] * xf + Ki[1] * yf + Ki[2] * 1.0f;
    const float pt1 = Ki[3] * xf + Ki[4] * yf + Ki[5] * 1.0f;
    const float pt2 = Ki[6] * xf + Ki[7] * yf + Ki[8] * 1.0f;
    const float z2 = pt2 * pt2;
    const float a = KR[0], b = KR[1], c = KR[2], d = KR[3], e = KR[4], f = KR[5], g = KR[6], hh = KR[7], ii = KR[8];
    const float lp0 = ((pt2 * (d * gy + a * gx)) - (gy * g * yf) - (gx * g * xf)) / z2;
    const float lp1 = ((pt2 * (e * gy + b * gx)) - (gy * hh * yf) - (gx * hh * xf)) / z2;
    const float lp2 = ((pt2 * (f * gy + c * gx)) - (gy * ii * yf) - (gx * ii * xf)) / z2;
    row[0] = lp1 * pt2 - lp2 * pt1;
    row[1] = lp2 * pt0 - lp0 * pt2;
    row[2] = lp0 * pt1 - lp1 * pt0;
    row[3] = -((float)next[(size_t)wy * w + wx] - (float)last[(size_t)y * w + x]);
    return true;
}

__device__ __forceinline__ void acc_row4(float (&acc)[kSO3], const float* row)
{
    int s = 0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
            acc[s] = __builtin_fmaf(row[i], row[j], acc[s]);
            ++s;
        }
    acc[9] = __builtin_fmaf(row[3], row[3], acc[9]);
    acc[10] += 1.0f;
}

// FP64 total of one workgroup's per-thread sums: wave sums, then waves in order.
template <int NT>
__device__ __forceinline__ void so3_block_total(float (&acc)[kSO3], float (*red)[kSO3], double* tot)
{
    wave_sum<kSO3>(acc);
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < kSO3; ++k) red[threadIdx.x >> 6][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < kSO3) {
        double s = 0.0;
        for (int wv = 0; wv < NT / 64; ++wv) s += (double)red[wv][threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
}

constexpr int kSO3Threads = 1024;

__global__ __launch_bounds__(kSO3Threads) void k_so3(const PairDesc* __restrict__ pairs)
{
    __shared__ float s_red[kSO3Threads / 64][kSO3];
    __shared__ double s_tot[kSO3];
    __shared__ float sH[9], sKi[9], sKR[9];
    __shared__ int s_done;
    const PairDesc& D = pairs[blockIdx.x];
    TrackState* st = D.st;
    const LevelDesc& L = D.lv[2];
    const int w = L.w, h = L.h;
    double K[9], Ki[9];
    kmat(L.fx, L.fy, L.cx, L.cy, K, Ki);
    // loop state, meaningful in thread 0
    float R_lr[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double resultR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    double lastResultR[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    float lastError = FLT_MAX / 2, lastCount = FLT_MAX / 2;
    int iters = 0;
    for (int it = 0; it < 10; ++it) {
        if (threadIdx.x == 0) {
            double KR[9], Hd[9];
            mm3d(K, resultR, KR);
            mm3d(KR, Ki, Hd);
            for (int k = 0; k < 9; ++k) { sH[k] = (float)Hd[k]; sKi[k] = (float)Ki[k]; sKR[k] = (float)KR[k]; }
        }
        __syncthreads();
        float acc[kSO3];
#pragma unroll
        for (int k = 0; k < kSO3; ++k) acc[k] = 0.0f;
        for (int idx = threadIdx.x; idx < w * h; idx += kSO3Threads) {
            const int y = idx / w, x = idx - y * w;
            float row[4];
            if (so3_pixel(D.so3_last, D.so3_next, w, h, x, y, sH, sKi, sKR, row)) acc_row4(acc, row);
        }
        so3_block_total<kSO3Threads>(acc, s_red, s_tot);
        if (threadIdx.x == 0) {
            ++iters;
            const float res0 = (float)s_tot[9], res1 = (float)s_tot[10];
            const float err = sqrtf(res0) / res1;
            st->lastSO3Error = err;
            st->lastSO3Count = res1;
            int done = 0;
            if (err < lastError && lastCount == res1) {
                done = 1;
            } else if (err > lastError + 0.001f) {
                st->lastSO3Error = lastError;
                st->lastSO3Count = lastCount;
                for (int k = 0; k < 9; ++k) resultR[k] = lastResultR[k];
                done = 1;
            }
            if (!done) {
                lastError = err;
                lastCount = res1;
                for (int k = 0; k < 9; ++k) lastResultR[k] = resultR[k];
                double A3[9], b3[3], delta[3], rot[9];
                unpack_so3(s_tot, A3, b3);
                ldlt_solve(3, A3, b3, delta);
                rodrigues(delta, rot);
                float rotf[9], nR[9];
                for (int k = 0; k < 9; ++k) rotf[k] = (float)rot[k];
                mm3f(rotf, R_lr, nR);
                for (int k = 0; k < 9; ++k) { R_lr[k] = nR[k]; resultR[k] = (double)nR[k]; }
            }
            s_done = done;
        }
        __syncthreads();
        if (s_done) break;
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 9; ++k) st->so3R[k] = resultR[k];
        st->so3_iters = iters;
    }
}

void launch_so3(const PairDesc* pairs, int npairs, int mode, hipStream_t s)
{
    (void)mode;
    hipLaunchKernelGGL(k_so3, dim3(npairs), dim3(kSO3Threads), 0, s, pairs);
}

// One SO3 step with explicit constants (H, K^-1, K R as 27 floats): sums in
// FP64 and the correspondence mask, for parity checks.
__global__ __launch_bounds__(kSO3Threads) void k_so3_step_dbg(const PairDesc* __restrict__ pairs,
                                                              const float* __restrict__ c27,
                                                              double* __restrict__ out, uint8_t* __restrict__ mask)
{
    __shared__ float s_red[kSO3Threads / 64][kSO3];
    __shared__ double s_tot[kSO3];
    const PairDesc& D = pairs[0];
    const LevelDesc& L = D.lv[2];
    const int w = L.w, h = L.h;
    float H[9], Ki[9], KR[9];
    for (int k = 0; k < 9; ++k) { H[k] = c27[k]; Ki[k] = c27[9 + k]; KR[k] = c27[18 + k]; }
    float acc[kSO3];
#pragma unroll
    for (int k = 0; k < kSO3; ++k) acc[k] = 0.0f;
    for (int idx = threadIdx.x; idx < w * h; idx += kSO3Threads) {
        const int y = idx / w, x = idx - y * w;
        float row[4];
        const bool f = so3_pixel(D.so3_last, D.so3_next, w, h, x, y, H, Ki, KR, row);
        if (f) acc_row4(acc, row);
        if (mask) mask[idx] = (uint8_t)f;
    }
    so3_block_total<kSO3Threads>(acc, s_red, s_tot);
    if (threadIdx.x < kSO3) out[threadIdx.x] = s_tot[threadIdx.x];
}

void launch_so3_step_dbg(const PairDesc* pairs, const float* c27, double* out11, uint8_t* mask, hipStream_t s)
{
    hipLaunchKernelGGL(k_so3_step_dbg, dim3(1), dim3(kSO3Threads), 0, s, pairs, c27, out11, mask);
}

}  // namespace slam
