// odom_kernels.hip -- gfx950 kernels of the dense RGB-D odometry path.
//
// Preprocessing (bilateral, pyramids, vertex/normal maps, model-map unpack,
// intensity, derivatives) and the Gauss-Newton steps: the fused ICP +
// photometric-association pass (A), the photometric pass (B) and the SO(3)
// pre-alignment.  Each step folds its per-pixel products through wave64
// shuffles and LDS into one block partial per workgroup.  The next launch on
// the stream reduces those partials in FP64 in a fixed order in its prologue
// (every block of the single-pair pass A, or one k_solve workgroup per pair
// of a batch) and runs the 6x6 solve and SE(3) update there, so a tracking
// iteration needs no device atomics and never returns to the host.
//
// Every expression that decides a correspondence (projection, rounding,
// thresholds, NaN tests) is written in the reference's operation order and
// the file is compiled with -ffp-contract=off, so masks are bit-identical to
// the CPU restatement of the same algorithm.
#include <float.h>

#include <atomic>
#include <type_traits>

#include "odom_internal.h"
#include "odom_math.h"

namespace slam {

#define QNAN __builtin_nanf("")

// Device accesses through pointers held in descriptors and state records.
// The compiler cannot tell such a pointer's address space and would use flat
// instructions, which also count against the LDS/scalar wait counter (so a
// wait for a descriptor field or an LDS read would also wait for pixel loads
// in flight) and take a 64-bit address per lane.  gp() makes them global-
// memory accesses, with the uniform base in scalar registers.
#ifdef __HIP_DEVICE_COMPILE__
#define GLOBAL_AS __attribute__((address_space(1)))
#else
#define GLOBAL_AS  // the host pass only parses kernels
#endif
template <class T>
__device__ __forceinline__ const GLOBAL_AS T* gp(const T* p)
{
    return (const GLOBAL_AS T*)p;
}
template <class T>
__device__ __forceinline__ GLOBAL_AS T* gp(T* p)
{
    return (GLOBAL_AS T*)p;
}

static inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Staging of an N-texel tile (row length ld, origin (ox, oy) in a w x h
// image) into LDS with all of a thread's loads in flight together: the
// coordinates are clamped into the image, so every address is valid, and the
// stores follow the last load.  An entry whose texel lies outside the image
// holds `fill` (FILL_OUT) or its clamped neighbour.
template <int N, int NT, bool FILL_OUT, class T, class Load>
__device__ __forceinline__ void stage_tile(T* tile, int ld, int ox, int oy, int w, int h, Load load, T fill = T(0))
{
    constexpr int R = (N + NT - 1) / NT;
    T v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = min((int)threadIdx.x + r * NT, N - 1);
        const int gx = ox + i % ld, gy = oy + i / ld;
        v[r] = load(min(max(gx, 0), w - 1), min(max(gy, 0), h - 1));
        if (FILL_OUT && (gx < 0 || gx >= w || gy < 0 || gy >= h)) v[r] = fill;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int i = (int)threadIdx.x + r * NT;
        if (i < N) tile[i] = v[r];
    }
}

float bexpf_host(float x) { return bexpf(x); }

// =============================================================== a1 bilateral
// depth_bilateral.frag:30-76.  Output tile 32x8 per 256-thread block (one
// pixel per thread), input tile with a 6-pixel halo in LDS.  The weight is
// the product of a 13x13 spatial table and a range table over |value - tmp|
// (factors of exp(-(space2*0.024691358 + color2*0.000555556))), both built
// per block with bexpf.  Range factors for |value - tmp| >= 380 are 0.
// Texels outside the image are staged as 0xFFFF, at least 380 from every
// filtered value when the cut-off is below 65155, so they weigh exactly 0 and
// every tile runs the unclipped 13x13 loop: a zero term leaves both sums'
// bits unchanged, which equals the reference's clipped window.  Above that
// cut-off, border tiles clip the window explicitly.
constexpr int kBilatRange = 380;
constexpr int kBT = 32;                 // output tile width
constexpr int kBTY = 8;                 // output tile height (one pixel per thread)
constexpr int kBL = kBT + 12;           // LDS tile width
constexpr int kBLY = kBTY + 12;         // LDS tile height

// The 13x13 spatial factors, set from the host copy of bexpf (the same
// operation sequence) so the unrolled interior path reads them as scalars.
// Set once per device; handles created concurrently on one device may both
// upload the (identical) table.
__constant__ float c_bilat_es[169];
static std::atomic<bool> g_bilat_es_set[64];

void init_device_tables()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64 || g_bilat_es_set[dev].load()) return;
    float es[169];
    for (int i = 0; i < 169; ++i) {
        const int dy = i / 13 - 6, dx = i % 13 - 6;
        es[i] = bexpf(-((float)(dx * dx + dy * dy) * 0.024691358f));
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_bilat_es), es, sizeof es) == hipSuccess) g_bilat_es_set[dev].store(true);
}

__global__ __launch_bounds__(256) void k_bilateral(const uint16_t* __restrict__ src, int pitch,
                                                   uint16_t* __restrict__ dst, int w, int h,
                                                   unsigned int maxv)
{
    // texels as floats (exact for 16-bit values): the tap's difference, its
    // magnitude and the sum's factor then need no conversions of their own
    __shared__ float tile[kBLY * kBL];
    __shared__ float sEC[kBilatRange + 1];
    const int tid = threadIdx.x;
    const int x0 = blockIdx.x * kBT, y0 = blockIdx.y * kBTY;
    const bool unclipped = maxv < 65535u - kBilatRange ||
                           (x0 >= 6 && y0 >= 6 && x0 + kBT + 6 <= w && y0 + kBTY + 6 <= h);
    stage_tile<kBLY * kBL, 256, true>(tile, kBL, x0 - 6, y0 - 6, w, h,
                                      [&](int gx, int gy) { return (float)gp(src)[(size_t)gy * pitch + gx]; },
                                      65535.0f);
    for (int i = tid; i <= kBilatRange; i += 256)
        sEC[i] = (i < kBilatRange) ? bexpf(-((float)(i * i) * 0.000555556f)) : 0.0f;
    __syncthreads();
    const int lx = tid & 31, ly = tid >> 5;
    const int x = x0 + lx, y = y0 + ly;
    if (x >= w || y >= h) return;
    const float vf = tile[(ly + 6) * kBL + lx + 6];
    const int value = (int)vf;
    if ((unsigned int)value > maxv || value < 300) {
        dst[(size_t)y * w + x] = 0;
        return;
    }
    // |value - tmp| clamped to the table, exactly as in integers: both are
    // integers below 2^17, so the float difference is exact
    auto range = [&](float t) { return sEC[(int)fminf(fabsf(vf - t), (float)kBilatRange)]; };
    float sum1 = 0.0f, sum2 = 0.0f;
    if (unclipped) {
        // no window clipping, 13 columns unrolled per row
#pragma unroll 1
        for (int dy = 0; dy < 13; ++dy) {
            const int trow = (ly + dy) * kBL + lx;
            const float* es = c_bilat_es + dy * 13;
#pragma unroll
            for (int dx = 0; dx < 13; ++dx) {
                const float t = tile[trow + dx];
                const float wgt = es[dx] * range(t);
                sum1 = sum1 + t * wgt;
                sum2 = sum2 + wgt;
            }
        }
    } else {
        const int cy0 = max(y - 6, 0), cy1 = min(y + 7, h);
        const int cx0 = max(x - 6, 0), cx1 = min(x + 7, w);
        for (int cy = cy0; cy < cy1; ++cy) {
            const int trow = (cy - y0 + 6) * kBL - x0 + 6;
            const int erow = (cy - y + 6) * 13 + 6 - x;
            for (int cx = cx0; cx < cx1; ++cx) {
                const float t = tile[trow + cx];
                const float wgt = c_bilat_es[erow + cx] * range(t);
                sum1 = sum1 + t * wgt;
                sum2 = sum2 + wgt;
            }
        }
    }
    dst[(size_t)y * w + x] = (uint16_t)(unsigned int)roundf(sum1 / sum2);
}

void launch_bilateral(const uint16_t* src, int pitch, uint16_t* dst, int w, int h, float maxD, hipStream_t s)
{
    const unsigned int maxv = (unsigned int)(maxD * 1000.0f);
    init_device_tables();
    dim3 grid(cdiv(w, kBT), cdiv(h, kBTY));
    hipLaunchKernelGGL(k_bilateral, grid, dim3(256), 0, s, src, pitch, dst, w, h, maxv);
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
__device__ __forceinline__ bool vpoint(uint16_t dv, int u, int v, float fxi, float fyi, float cx, float cy,
                                       float cutoff, float* o)
{
    const float z = (float)dv / 1000.f;
    if (z != 0 && z < cutoff) {
        o[0] = z * ((float)u - cx) * fxi;
        o[1] = z * ((float)v - cy) * fyi;
        o[2] = z;
        return true;
    }
    return false;
}

// Both maps at pixel (u, v) of level l; depth(x, y) gives the level's texels
// (the right and lower neighbours are read only where the normal needs them).
template <class Depth>
__device__ __forceinline__ void vn_pixel(const VNLevels& a, int l, int u, int v, Depth depth)
{
    const int w = a.w[l], h = a.h[l];
    const size_t i = (size_t)v * w + u, P = (size_t)w * h;
    float* vm = a.vmap[l];
    float* nm = a.nmap[l];
    const float fxi = a.fx_inv[l], fyi = a.fy_inv[l], cx = a.cx[l], cy = a.cy[l];
    float p00[3];
    const bool ok00 = vpoint(depth(u, v), u, v, fxi, fyi, cx, cy, a.cutoff, p00);
    if (ok00) {
        vm[i] = p00[0]; vm[i + P] = p00[1]; vm[i + 2 * P] = p00[2];
    } else {
        vm[i] = QNAN; vm[i + P] = QNAN; vm[i + 2 * P] = QNAN;
    }
    float n[3] = {QNAN, QNAN, QNAN};
    if (u != w - 1 && v != h - 1 && ok00) {
        float p01[3], p10[3];
        const bool ok01 = vpoint(depth(u + 1, v), u + 1, v, fxi, fyi, cx, cy, a.cutoff, p01);
        const bool ok10 = vpoint(depth(u, v + 1), u, v + 1, fxi, fyi, cx, cy, a.cutoff, p10);
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

// ================================================ a5 + a6 + a7 model maps
// copyMaps (utils.cu:270-310), resizeVMap/resizeNMap (utils.cu:365-444) and
// tranformMaps (utils.cu:206-248).  A level-l pixel recomputes its
// camera-frame value from the 2^l x 2^l block of input texels with the
// reference's staged means (NaN if any child's x is NaN; normals
// renormalised at every stage), then applies the rigid transform.
struct Cam6 {
    float v[3], n[3];
};

// The texels are loaded whole and the NaN rules applied by selects, so every
// load of an output is issued before the first wait (with branches the
// compiler splits the loads and waits for each texel's z first).
__device__ __forceinline__ Cam6 model_l0(const ModelLevels& a, int x, int y)
{
    const float4 vv = gp(a.v4)[(size_t)y * a.v_pitch + x];
    const float4 nn = gp(a.n4)[(size_t)y * a.n_pitch + x];
    const bool ok = !(vv.z == 0);
    Cam6 c;
    c.v[0] = ok ? vv.x : QNAN;
    c.v[1] = ok ? vv.y : QNAN;
    c.v[2] = ok ? vv.z : QNAN;
    c.n[0] = ok ? nn.x : QNAN;
    c.n[1] = ok ? nn.y : QNAN;
    c.n[2] = ok ? nn.z : QNAN;
    return c;
}

__device__ __forceinline__ void mean4(const float* a, const float* b, const float* c, const float* d,
                                      bool normalize, float* o)
{
    const bool bad = isnan(a[0]) || isnan(b[0]) || isnan(c[0]) || isnan(d[0]);
    float m[3];
    for (int k = 0; k < 3; ++k) m[k] = (a[k] + b[k] + c[k] + d[k]) / 4.0f;
    if (normalize) {
        const float rn = 1.0f / sqrtf(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
        m[0] = m[0] * rn; m[1] = m[1] * rn; m[2] = m[2] * rn;
    }
    for (int k = 0; k < 3; ++k) o[k] = bad ? QNAN : m[k];
}

// 2x2 mean of four texels (x, y), (x+1, y), (x, y+1), (x+1, y+1).
__device__ __forceinline__ Cam6 mean_cam(const Cam6& c00, const Cam6& c01, const Cam6& c10, const Cam6& c11)
{
    Cam6 o;
    mean4(c00.v, c01.v, c10.v, c11.v, false, o.v);
    mean4(c00.n, c01.n, c10.n, c11.n, true, o.n);
    return o;
}

__device__ __forceinline__ Cam6 model_l1(const ModelLevels& a, int x, int y)
{
    const Cam6 c00 = model_l0(a, 2 * x, 2 * y), c01 = model_l0(a, 2 * x + 1, 2 * y);
    const Cam6 c10 = model_l0(a, 2 * x, 2 * y + 1), c11 = model_l0(a, 2 * x + 1, 2 * y + 1);
    return mean_cam(c00, c01, c10, c11);
}


// Transform (when set) and store output i of level l.
__device__ __forceinline__ void model_store(const ModelLevels& a, int l, size_t i, Cam6 c, const float* R,
                                            const float* t)
{
    if (a.transform) {
        if (!isnan(c.v[0])) {
            const float v0 = c.v[0], v1 = c.v[1], v2 = c.v[2];
            for (int r = 0; r < 3; ++r) c.v[r] = R[3 * r] * v0 + R[3 * r + 1] * v1 + R[3 * r + 2] * v2 + t[r];
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
    const size_t P = (size_t)a.w[l] * a.h[l];
    GLOBAL_AS float* vm = gp(a.vmap[l]);
    GLOBAL_AS float* nm = gp(a.nmap[l]);
    vm[i] = c.v[0]; vm[i + P] = c.v[1]; vm[i + 2 * P] = c.v[2];
    nm[i] = c.n[0]; nm[i + P] = c.n[1]; nm[i + 2 * P] = c.n[2];
}

// Blocks [0, nb1): a 16x16 tile of level 1 each, whose untransformed values
// also give, from LDS, the 8x8 level-2 outputs the tile covers (2x2 means of
// level-1 values, as level 1 is of level 0); then 256 level-0 pixels per block.
__global__ __launch_bounds__(256) void k_model_maps(ModelLevels a, int nb1, int tiles_x1)
{
    __shared__ float s1[6][16][16];
    float R[9], t[3];
    if (a.transform) {
        if (a.pose_dev) {
            const GLOBAL_AS float* pd = gp(a.pose_dev);
            for (int r = 0; r < 3; ++r) {
                for (int k = 0; k < 3; ++k) R[3 * r + k] = pd[4 * r + k];
                t[r] = pd[4 * r + 3];
            }
        } else {
            for (int k = 0; k < 9; ++k) R[k] = a.R[k];
            for (int k = 0; k < 3; ++k) t[k] = a.t[k];
        }
    }
    if ((int)blockIdx.x < nb1) {
        const int tx = blockIdx.x % tiles_x1, ty = blockIdx.x / tiles_x1;
        const int lx = threadIdx.x & 15, ly = threadIdx.x >> 4;
        const int x = 16 * tx + lx, y = 16 * ty + ly;
        if (x < a.w[1] && y < a.h[1]) {
            const Cam6 c = model_l1(a, x, y);
            for (int m = 0; m < 3; ++m) {
                s1[m][ly][lx] = c.v[m];
                s1[3 + m][ly][lx] = c.n[m];
            }
            model_store(a, 1, (size_t)y * a.w[1] + x, c, R, t);
        }
        __syncthreads();
        if (threadIdx.x < 64) {
            const int qx = threadIdx.x & 7, qy = threadIdx.x >> 3;
            const int X = 8 * tx + qx, Y = 8 * ty + qy;
            if (X < a.w[2] && Y < a.h[2]) {
                Cam6 q[4];   // (2X, 2Y), (2X+1, 2Y), (2X, 2Y+1), (2X+1, 2Y+1) of level 1
                for (int k = 0; k < 4; ++k) {
                    const int sx = 2 * qx + (k & 1), sy = 2 * qy + (k >> 1);
                    for (int m = 0; m < 3; ++m) {
                        q[k].v[m] = s1[m][sy][sx];
                        q[k].n[m] = s1[3 + m][sy][sx];
                    }
                }
                model_store(a, 2, (size_t)Y * a.w[2] + X, mean_cam(q[0], q[1], q[2], q[3]), R, t);
            }
        }
        return;
    }
    const int i = (blockIdx.x - nb1) * 256 + threadIdx.x;
    const int w = a.w[0];
    if (i >= w * a.h[0]) return;
    const int y = i / w, x = i - y * w;
    if (a.vz) gp(a.vz)[i] = gp(a.v4)[(size_t)y * a.v_pitch + x].z;
    model_store(a, 0, (size_t)i, model_l0(a, x, y), R, t);
}

void launch_model_maps(const ModelLevels& a, hipStream_t s)
{
    const int tiles_x1 = cdiv(a.w[1], 16), nb1 = tiles_x1 * cdiv(a.h[1], 16);
    const int nb0 = cdiv(a.w[0] * a.h[0], 256);
    hipLaunchKernelGGL(k_model_maps, dim3(nb1 + nb0), dim3(256), 0, s, a, nb1, tiles_x1);
}

// ============================================== a9 + a11 Gaussian pyrDown
// utils.cu:332-363 and 470-500: loop bound min(2x+3, cols-1) and the
// reversed weight index (ty-cy-1)*5+(tx-cx-1) are kept.
// The 5x5 kernel {1 4 6 4 1} x {1 4 6 4 1} of utils.cu:453-457, entry
// (r, c) computed as an exact product instead of a per-lane table lookup.
__device__ __forceinline__ int gauss5i(int k)
{
    return k == 2 ? 6 : ((k & 1) ? 4 : 1);
}

// ===================================================== pyramid launches
// One pyramid level of up to two pyramids per launch.  Each output reads its
// 5x5 window as 25 independent loads (coordinates clamped into the image, so
// every address is valid) that are all in flight together; taps outside the
// window of the reference's loops are then skipped in the accumulation, which
// visits the rest in the loops' row-major order, so every sum is the value
// the loops produce.
template <class T, class Get>
__device__ __forceinline__ void window5(Get get, int sw, int sh, int x, int y, T (&v)[25])
{
#pragma unroll
    for (int dy = 0; dy < 5; ++dy) {
        const int cy = min(max(2 * y - 2 + dy, 0), sh - 1);
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) v[dy * 5 + dx] = get(min(max(2 * x - 2 + dx, 0), sw - 1), cy);
    }
}

template <class Get>
__device__ __forceinline__ float gauss_f32_at(Get get, int sw, int sh, int x, int y)
{
    const int tx = min(2 * x + 3, sw - 1), ty = min(2 * y + 3, sh - 1);
    float v[25];
    window5(get, sw, sh, x, y, v);
    // the reference adds each float weight to an int count, truncating; the
    // weights are integers, so an integer sum is the same count
    float sum = 0.0f;
    int count = 0;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const int cy = 2 * y - 2 + dy, cx = 2 * x - 2 + dx;
            const float val = v[dy * 5 + dx];
            if (cy >= 0 && cy < ty && cx >= 0 && cx < tx && !isnan(val)) {
                const int gk = gauss5i(ty - cy - 1) * gauss5i(tx - cx - 1);
                sum = sum + val * (float)gk;
                count += gk;
            }
        }
    return sum / (float)count;
}

template <class Get>
__device__ __forceinline__ uint8_t gauss_u8_at(Get get, int sw, int sh, int x, int y)
{
    const int tx = min(2 * x + 3, sw - 1), ty = min(2 * y + 3, sh - 1);
    uint8_t v[25];
    window5(get, sw, sh, x, y, v);
    // every term val * gk is an integer below 2^14 and their sum stays below
    // 2^24, so the reference's float sum is exact in any order: integer sums
    // give the same bits
    int sum = 0, count = 0;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const int cy = 2 * y - 2 + dy, cx = 2 * x - 2 + dx;
            const int val = v[dy * 5 + dx];
            if (cy >= 0 && cy < ty && cx >= 0 && cx < tx && val > 0) {
                const int gk = gauss5i(ty - cy - 1) * gauss5i(tx - cx - 1);
                sum += val * gk;
                count += gk;
            }
        }
    const float q = (float)sum / (float)count;
    return (count == 0) ? (uint8_t)0 : (uint8_t)q;
}

// utils.cu:57-94 (sigma_color 30): taps at offsets -2..2 inside the image,
// weighted {0.0625, 0.25, 0.375, 0.25, 0.0625} per axis, kept when within
// 3 sigma of the centre value.
template <class Get>
__device__ __forceinline__ uint16_t pyr_u16_at(Get get, int sw, int sh, int x, int y)
{
    // In 1/256 units the weights are the integers {1 4 6 4 1} x {1 4 6 4 1}:
    // each float term val * w * w is exact, and the sums stay below 2^24
    // units, so the reference's float sums are exact in any order and equal
    // the integer sums; scaling both by 2^-8 leaves their quotient unchanged.
    int v[25];
    window5(get, sw, sh, x, y, v);
    const int center = v[12];
    int sum = 0, wall = 0;
#pragma unroll
    for (int dy = 0; dy < 5; ++dy)
#pragma unroll
        for (int dx = 0; dx < 5; ++dx) {
            const int cy = 2 * y - 2 + dy, cx = 2 * x - 2 + dx;
            const int val = v[dy * 5 + dx];
            if (cy >= 0 && cy < sh && cx >= 0 && cx < sw && abs(val - center) < 90) {
                const int wk = gauss5i(dx) * gauss5i(dy);
                sum += val * wk;
                wall += wk;
            }
        }
    return (uint16_t)(int)((float)sum / (float)wall);
}

// Levels 0 (when generated), 1 and 2 of up to two pyramids in one launch,
// from LDS tiles.  A level-1 block stages the level-0 texels of its 16x16
// outputs' windows (36x36); a level-2 block stages 43x27 level-0 texels,
// computes the 20x12 level-1 values its 8x4 outputs read (the same
// computation the level-1 blocks store; one value per thread, so the block
// runs two dependent Gaussian sums, not three), then its outputs.  Level 0 is read
// from memory or generated (verticesToDepth, or the intensity conversion of
// populateRGBDData, RGBDOdometryef.cpp:212, 225), in which case level-0
// blocks also store it.  The taps, their clamped coordinates and the
// summation order are those of the per-level functions above.
constexpr int kPT = 16;                 // level-1 output tile edge
constexpr int kT1 = 2 * kPT + 4;        // its level-0 texels
constexpr int kPT2X = 8, kPT2Y = 4;     // level-2 output tile (its level-1 values, below, fit
                                        // one pass of the 256 threads)
constexpr int kT21X = 2 * kPT2X + 4, kT21Y = 2 * kPT2Y + 4;   // its level-1 values
constexpr int kT20X = 2 * kT21X + 3, kT20Y = 2 * kT21Y + 3;   // its level-0 texels
static_assert(kT21X * kT21Y <= 256, "one level-1 value per thread");

struct PyrGrid {
    PyrJob j;
    int nseg;
    int b0[7];
    int seg_q[6], seg_t[6];             // pyramid, and level (0 generation, 1, 2) of a segment
    int tiles_x[3];                     // tiles per row at levels 1, 2 (index by level)
};

template <class T>
__device__ __forceinline__ T pyr_l0(const PyrSrc& q, int w0, int x, int y)
{
    const size_t i = (size_t)y * w0 + x;
    if constexpr (sizeof(T) == 4) {
        if (q.gen == 1) {
            const float z = gp(q.vz)[i];
            return (z > q.cutoff || z <= 0) ? QNAN : z;
        }
    } else if constexpr (sizeof(T) == 1) {
        if (q.gen == 2) {
            const GLOBAL_AS uint8_t* px = gp(q.rgb) + (size_t)y * q.pitch + (size_t)x * q.ch;
            return (uint8_t)(int)((float)px[0] * 0.114f + (float)px[1] * 0.299f + (float)px[2] * 0.587f);
        }
    }
    return gp((const T*)q.src0)[i];
}

// Level-0 texels of a tile in LDS (texels outside the image hold their
// clamped neighbour; the windows clamp their coordinates, so never read them).
// The source kind is uniform, so each kind stages with its own straight-line
// loads.
template <int N, class T>
__device__ __forceinline__ void pyr_stage(const PyrSrc& q, T* tile, int ld, int ox, int oy, int w0, int h0)
{
    if constexpr (sizeof(T) == 4) {
        if (q.gen == 1) {
            stage_tile<N, 256, false>(tile, ld, ox, oy, w0, h0, [&](int x, int y) {
                const float z = gp(q.vz)[(size_t)y * w0 + x];
                return (T)((z > q.cutoff || z <= 0) ? QNAN : z);
            });
            return;
        }
    } else if constexpr (sizeof(T) == 1) {
        if (q.gen == 2) {
            stage_tile<N, 256, false>(tile, ld, ox, oy, w0, h0, [&](int x, int y) {
                const GLOBAL_AS uint8_t* px = gp(q.rgb) + (size_t)y * q.pitch + (size_t)x * q.ch;
                return (T)(uint8_t)(int)((float)px[0] * 0.114f + (float)px[1] * 0.299f + (float)px[2] * 0.587f);
            });
            return;
        }
    }
    stage_tile<N, 256, false>(tile, ld, ox, oy, w0, h0,
                              [&](int x, int y) { return gp((const T*)q.src0)[(size_t)y * w0 + x]; });
}

// One output of level `lvl` from its source level `src` (sw x sh texels)
// held in an LDS tile with origin (ox, oy) and row length `ld`.
template <int KIND, class T>
__device__ __forceinline__ T pyr_at(const T* tile, int ld, int ox, int oy, int sw, int sh, int x, int y)
{
    auto get = [&](int cx, int cy) { return tile[(cy - oy) * ld + (cx - ox)]; };
    if constexpr (KIND == 0) {
        return gauss_f32_at(get, sw, sh, x, y);
    } else if constexpr (KIND == 1) {
        return gauss_u8_at(get, sw, sh, x, y);
    } else {
        auto geti = [&](int cx, int cy) { return (int)get(cx, cy); };
        return pyr_u16_at(geti, sw, sh, x, y);
    }
}

template <int KIND, class T>
__device__ void pyr_block(const PyrGrid& g, int q, int t, int lb, unsigned char* smem)
{
    const PyrSrc& P = g.j.p[q];
    const int w0 = g.j.w0, h0 = g.j.h0, w1 = w0 / 2, h1 = h0 / 2, w2 = w1 / 2, h2 = h1 / 2;
    const int tid = threadIdx.x;
    if (t == 0) {
        const int i = lb * 256 + tid;
        if (i < w0 * h0) {
            const int y = i / w0, x = i - y * w0;
            gp((T*)P.dst[0])[i] = pyr_l0<T>(P, w0, x, y);
        }
        return;
    }
    const int tx = lb % g.tiles_x[t], ty = lb / g.tiles_x[t];
    T* t1 = (T*)smem;                            // texels one level below the outputs
    if (t == 1) {
        const int X0 = tx * kPT, Y0 = ty * kPT;
        const int x = X0 + (tid & (kPT - 1)), y = Y0 + tid / kPT;
        const int ox = 2 * X0 - 2, oy = 2 * Y0 - 2;
        pyr_stage<kT1 * kT1, T>(P, t1, kT1, ox, oy, w0, h0);
        __syncthreads();
        if (x < w1 && y < h1) gp((T*)P.dst[1])[(size_t)y * w1 + x] = pyr_at<KIND>(t1, kT1, ox, oy, w0, h0, x, y);
        return;
    }
    const int X0 = tx * kPT2X, Y0 = ty * kPT2Y;
    T* t0 = t1 + kT21X * kT21Y;                  // level-0 texels
    const int o1x = 2 * X0 - 2, o1y = 2 * Y0 - 2;
    const int o0x = 2 * o1x - 2, o0y = 2 * o1y - 2;
    pyr_stage<kT20X * kT20Y, T>(P, t0, kT20X, o0x, o0y, w0, h0);
    __syncthreads();
    if (tid < kT21X * kT21Y) {
        const int x1 = o1x + tid % kT21X, y1 = o1y + tid / kT21X;
        if (x1 >= 0 && x1 < w1 && y1 >= 0 && y1 < h1) t1[tid] = pyr_at<KIND>(t0, kT20X, o0x, o0y, w0, h0, x1, y1);
    }
    __syncthreads();
    const int x = X0 + (tid & (kPT2X - 1)), y = Y0 + tid / kPT2X;
    if (tid < kPT2X * kPT2Y && x < w2 && y < h2)
        gp((T*)P.dst[2])[(size_t)y * w2 + x] = pyr_at<KIND>(t1, kT21X, o1x, o1y, w1, h1, x, y);
}

__global__ __launch_bounds__(256) void k_pyramid(PyrGrid g)
{
    constexpr int kL2 = kT21X * kT21Y + kT20X * kT20Y;
    __shared__ __align__(16) unsigned char smem[(kT1 * kT1 > kL2 ? kT1 * kT1 : kL2) * 4];
    int sgi = 0;
    while (sgi + 1 < g.nseg && (int)blockIdx.x >= g.b0[sgi + 1]) ++sgi;
    const int q = g.seg_q[sgi], t = g.seg_t[sgi], lb = blockIdx.x - g.b0[sgi];
    const int kind = g.j.p[q].kind;
    if (kind == 0) pyr_block<0, float>(g, q, t, lb, smem);
    else if (kind == 1) pyr_block<1, uint8_t>(g, q, t, lb, smem);
    else pyr_block<2, uint16_t>(g, q, t, lb, smem);
}

void launch_pyramid(const PyrJob& j, hipStream_t s)
{
    PyrGrid g{};
    g.j = j;
    const int w1 = j.w0 / 2, h1 = j.h0 / 2, w2 = w1 / 2, h2 = h1 / 2;
    g.tiles_x[1] = cdiv(w1, kPT);
    g.tiles_x[2] = cdiv(w2, kPT2X);
    const int n2 = g.tiles_x[2] * cdiv(h2, kPT2Y), n1 = g.tiles_x[1] * cdiv(h1, kPT);
    int nb = 0;
    // level-2 tiles first (the longest blocks), then level 1, then level-0 generation
    for (int t = 2; t >= 0; --t)
        for (int q = 0; q < j.count; ++q) {
            const int n = t == 2 ? n2 : (t == 1 ? n1 : (j.p[q].gen ? cdiv(j.w0 * j.h0, 256) : 0));
            if (!n) continue;
            g.b0[g.nseg] = nb;
            g.seg_q[g.nseg] = q;
            g.seg_t[g.nseg] = t;
            g.nseg++;
            nb += n;
        }
    g.b0[g.nseg] = nb;
    hipLaunchKernelGGL(k_pyramid, dim3(nb), dim3(256), 0, s, g);
}

// initICP's depth pyramid (a2) and the vertex and normal maps of its three
// levels (a3, a4) in one launch.  Level-2 tiles come first, then level-1
// tiles, then the level-0 maps one pixel per thread from the input.  A tile
// computes its level's depth one output beyond its right and lower edges
// too, in LDS only (the neighbouring tile stores those), so the normals of
// its edge pixels, which read their right and lower neighbours, come from
// the tile itself.  The depths are those of k_pyramid (the same windows and
// sums) and every level's maps come from vn_pixel, the same expressions
// whichever block computes them.
constexpr int kDM1 = 16;                                      // level-1 tile edge
constexpr int kDM1R = kDM1 + 1;                               // with the one-output margin
constexpr int kDM1T = 2 * kDM1R + 3;                          // its level-0 texels (37)
constexpr int kDM2X = 8, kDM2Y = 4;                           // level-2 tile
constexpr int kDM2RX = kDM2X + 1, kDM2RY = kDM2Y + 1;         // with the margin
constexpr int kDM21X = 2 * kDM2RX + 3, kDM21Y = 2 * kDM2RY + 3;   // its level-1 values (21 x 13)
constexpr int kDM20X = 2 * kDM21X + 3, kDM20Y = 2 * kDM21Y + 3;   // their level-0 texels (45 x 29)
static_assert(kDM1R * kDM1R <= 2 * 256 && kDM21X * kDM21Y <= 2 * 256 && kDM2RX * kDM2RY <= 256, "passes");

struct DepthMapsGrid {
    VNLevels a;                   // a.depth[0] is the input; a.depth[1], a.depth[2] are written
    int n2, n1;                   // level-2 and level-1 tiles
    int tx2, tx1;                 // tiles per row
};

__global__ __launch_bounds__(256) void k_depth_pyramid_maps(DepthMapsGrid g)
{
    constexpr int kS0 = kDM20X * kDM20Y > kDM1T * kDM1T ? kDM20X * kDM20Y : kDM1T * kDM1T;
    constexpr int kS1 = kDM21X * kDM21Y > kDM1R * kDM1R ? kDM21X * kDM21Y : kDM1R * kDM1R;
    __shared__ uint16_t s0[kS0];
    __shared__ uint16_t s1[kS1];
    __shared__ uint16_t s2[kDM2RX * kDM2RY];
    const VNLevels& a = g.a;
    const int tid = threadIdx.x;
    const int w0 = a.w[0], h0 = a.h[0], w1 = a.w[1], h1 = a.h[1], w2 = a.w[2], h2 = a.h[2];
    const uint16_t* d0 = a.depth[0];
    auto load0 = [&](int x, int y) { return d0[(size_t)y * w0 + x]; };
    int b = blockIdx.x;
    if (b < g.n2) {
        const int X0 = (b % g.tx2) * kDM2X, Y0 = (b / g.tx2) * kDM2Y;
        const int o1x = 2 * X0 - 2, o1y = 2 * Y0 - 2;
        const int o0x = 2 * o1x - 2, o0y = 2 * o1y - 2;
        stage_tile<kDM20X * kDM20Y, 256, false>(s0, kDM20X, o0x, o0y, w0, h0, load0);
        __syncthreads();
        for (int e = tid; e < kDM21X * kDM21Y; e += 256) {
            const int x1 = o1x + e % kDM21X, y1 = o1y + e / kDM21X;
            if (x1 >= 0 && x1 < w1 && y1 >= 0 && y1 < h1)
                s1[e] = pyr_at<2>((const uint16_t*)s0, kDM20X, o0x, o0y, w0, h0, x1, y1);
        }
        __syncthreads();
        if (tid < kDM2RX * kDM2RY) {
            const int cx = tid % kDM2RX, cy = tid / kDM2RX;
            const int x = X0 + cx, y = Y0 + cy;
            if (x < w2 && y < h2) {
                const uint16_t v = pyr_at<2>((const uint16_t*)s1, kDM21X, o1x, o1y, w1, h1, x, y);
                s2[tid] = v;
                if (cx < kDM2X && cy < kDM2Y) a.depth[2][(size_t)y * w2 + x] = v;
            }
        }
        __syncthreads();
        if (tid < kDM2X * kDM2Y) {
            const int x = X0 + tid % kDM2X, y = Y0 + tid / kDM2X;
            if (x < w2 && y < h2)
                vn_pixel(a, 2, x, y, [&](int u, int v) { return s2[(v - Y0) * kDM2RX + (u - X0)]; });
        }
        return;
    }
    b -= g.n2;
    if (b < g.n1) {
        const int X0 = (b % g.tx1) * kDM1, Y0 = (b / g.tx1) * kDM1;
        const int ox = 2 * X0 - 2, oy = 2 * Y0 - 2;
        stage_tile<kDM1T * kDM1T, 256, false>(s0, kDM1T, ox, oy, w0, h0, load0);
        __syncthreads();
        for (int e = tid; e < kDM1R * kDM1R; e += 256) {
            const int cx = e % kDM1R, cy = e / kDM1R;
            const int x = X0 + cx, y = Y0 + cy;
            if (x < w1 && y < h1) {
                const uint16_t v = pyr_at<2>((const uint16_t*)s0, kDM1T, ox, oy, w0, h0, x, y);
                s1[e] = v;
                if (cx < kDM1 && cy < kDM1) a.depth[1][(size_t)y * w1 + x] = v;
            }
        }
        __syncthreads();
        const int x = X0 + (tid & (kDM1 - 1)), y = Y0 + tid / kDM1;
        if (x < w1 && y < h1) vn_pixel(a, 1, x, y, [&](int u, int v) { return s1[(v - Y0) * kDM1R + (u - X0)]; });
        return;
    }
    b -= g.n1;
    const int i = b * 256 + tid;
    if (i < w0 * h0) {
        const int y = i / w0, x = i - y * w0;
        vn_pixel(a, 0, x, y, load0);
    }
}

void launch_depth_pyramid_maps(const VNLevels& a, hipStream_t s)
{
    DepthMapsGrid g;
    g.a = a;
    g.tx2 = cdiv(a.w[2], kDM2X);
    g.tx1 = cdiv(a.w[1], kDM1);
    g.n2 = g.tx2 * cdiv(a.h[2], kDM2Y);
    g.n1 = g.tx1 * cdiv(a.h[1], kDM1);
    const int n0 = cdiv(a.w[0] * a.h[0], 256);
    hipLaunchKernelGGL(k_depth_pyramid_maps, dim3(g.n2 + g.n1 + n0), dim3(256), 0, s, g);
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

// Pixel k of level l.
__device__ __forceinline__ void rgb_prep_px(const PrepLevels& a, int l, int k)
{
    const int w = a.w[l], h = a.h[l];
    if (k >= w * h) return;
    const int y = k / w, x = k - y * w;
    const uint8_t* img = a.img[l];
    float dxv = 0.0f, dyv = 0.0f;
    if (x >= 1 && x < w - 1 && y >= 1 && y < h - 1) {
        // interior: the nine taps in the same order with fixed weights
        const float kx[9] = {0.52201f, 0.00000f, -0.52201f, 0.79451f, -0.00000f, -0.79451f,
                             0.52201f, 0.00000f, -0.52201f};
        const float ky[9] = {0.52201f, 0.79451f, 0.52201f, 0.00000f, 0.00000f, 0.00000f,
                             -0.52201f, -0.79451f, -0.52201f};
        int ki = 8;
#pragma unroll
        for (int j = -1; j <= 1; j++)
#pragma unroll
            for (int i = -1; i <= 1; i++) {
                const float sv = (float)img[(size_t)(y + j) * w + x + i];
                dxv = dxv + sv * kx[ki];
                dyv = dyv + sv * ky[ki];
                --ki;
            }
    } else {
        int ki = 8;
        for (int j = max(y - 1, 0); j <= min(y + 1, h - 1); j++)
            for (int i = max(x - 1, 0); i <= min(x + 1, w - 1); i++) {
                const float sv = (float)img[(size_t)j * w + i];
                dxv = dxv + sv * c_gsx[ki];
                dyv = dyv + sv * c_gsy[ki];
                --ki;
            }
    }
    const short sx = (short)dxv, sy = (short)dyv;
    a.dxy[l][k] = make_short2(sx, sy);
    float z = QNAN;
    if (x < w - 5 && y < h - 1) {
        // rows y-2..y+1 and columns x-2..x+1 inside the image (the upper
        // bounds hold here), all 16 loads issued before any test
        bool valid = true;
#pragma unroll
        for (int du = -2; du < 2; ++du)
#pragma unroll
            for (int dv = -2; dv < 2; ++dv) {
                const uint8_t t = img[(size_t)max(y + du, 0) * w + max(x + dv, 0)];
                valid = valid & ((y + du < 0) | (x + dv < 0) | (t > 0));
            }
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

__global__ __launch_bounds__(256) void k_rgb_prep(PrepLevels a, Seg3 g)
{
    const int l = seg_level(g, blockIdx.x);
    rgb_prep_px(a, l, (blockIdx.x - g.b0[l]) * 256 + threadIdx.x);
}

void launch_rgb_prep(const PrepLevels& a, hipStream_t s)
{
    int n[kLevels];
    for (int l = 0; l < kLevels; ++l) n[l] = a.w[l] * a.h[l];
    const Seg3 g = make_seg(n, 256);
    hipLaunchKernelGGL(k_rgb_prep, dim3(g.b0[kLevels]), dim3(256), 0, s, a, g);
}

// ============================================================ reductions
// Wave64 sums by DPP: row_shr 1/2/4/8 (inclusive scans of the 16-lane rows)
// then row_bcast 15/31 carry the rows forward; lane 63 ends with the total.
// The order of additions is fixed, so the results are reproducible.
#define ODOM_DPP(x, ctrl, rmask) __builtin_amdgcn_update_dpp(0, (x), (ctrl), (rmask), 0xf, false)

__device__ __forceinline__ float wave_total_f(float v)
{
    v = v + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, v), 0x111, 0xf));
    v = v + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, v), 0x112, 0xf));
    v = v + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, v), 0x114, 0xf));
    v = v + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, v), 0x118, 0xf));
    v = v + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, v), 0x142, 0xa));
    v = v + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, v), 0x143, 0xc));
    return v;
}

__device__ __forceinline__ unsigned int wave_total_u(unsigned int v)
{
    v = v + (unsigned int)ODOM_DPP((int)v, 0x111, 0xf);
    v = v + (unsigned int)ODOM_DPP((int)v, 0x112, 0xf);
    v = v + (unsigned int)ODOM_DPP((int)v, 0x114, 0xf);
    v = v + (unsigned int)ODOM_DPP((int)v, 0x118, 0xf);
    v = v + (unsigned int)ODOM_DPP((int)v, 0x142, 0xa);
    v = v + (unsigned int)ODOM_DPP((int)v, 0x143, 0xc);
    return v;
}

// Wave totals of V <= 32 values per lane by reduce-scatter.  Each step pairs
// every lane with a lane in the other half of its group of 64, 32, 16, 8 or 4
// lanes; the lane keeps half of its values and adds the partner's copy of
// that half.  The groups of 64 and 32 exchange by v_permlane32_swap and
// v_permlane16_swap (one swap moves a value pair between the two halves), the
// others by DPP row_mirror, row_half_mirror and quad_perm.  After a last add
// across lanes 2c and 2c+1 both hold the total of value c.  About 70
// operations for 32 values, where a scan per value takes 6 dependent ones.
template <int V>
__device__ __forceinline__ float wave_scatter(const float (&a)[V])
{
    static_assert(V <= 32, "one value per lane pair");
    const int lane = threadIdx.x & 63;
    float s16[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        if (j >= V) {
            s16[j] = 0.0f;
            continue;
        }
        const float y = (j + 16 < V) ? a[j + 16] : 0.0f;
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(a[j]), __float_as_uint(y), false, false);
        s16[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    float s8[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(s16[j]), __float_as_uint(s16[j + 8]), false,
                                                        false);
        s8[j] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const bool h8 = (lane & 8) != 0, h4 = (lane & 4) != 0, h2 = (lane & 2) != 0;
    float s4[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float send = h8 ? s8[j] : s8[j + 4], keep = h8 ? s8[j + 4] : s8[j];
        s4[j] = keep + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, send), 0x140, 0xf));
    }
    float s2[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        const float send = h4 ? s4[j] : s4[j + 2], keep = h4 ? s4[j + 2] : s4[j];
        s2[j] = keep + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, send), 0x141, 0xf));
    }
    const float send = h2 ? s2[0] : s2[1], keep = h2 ? s2[1] : s2[0];
    const float s1 = keep + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, send), 0x4e, 0xf));
    return s1 + __builtin_bit_cast(float, ODOM_DPP(__builtin_bit_cast(int, s1), 0xb1, 0xf));
}

// Block partial of V floats: wave totals (value c in lanes 2c, 2c+1), then
// the waves in order.
template <int V, int NT>
__device__ __forceinline__ void block_partial(const float (&a)[V], float* __restrict__ out, float (*red)[V])
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const float t = wave_scatter<V>(a);
    if ((lane & 1) == 0 && (lane >> 1) < V) red[wid][lane >> 1] = t;
    __syncthreads();
    if (threadIdx.x < V) {
        float s = 0.0f;
        for (int wv = 0; wv < NT / 64; ++wv) s += red[wv][threadIdx.x];
        gp(out)[threadIdx.x] = s;
    }
}

// Integer block partial (count, wrapping int32 sum of diff^2).
template <int NT>
__device__ __forceinline__ void block_partial_int(unsigned int c0, unsigned int c1, unsigned int* __restrict__ out,
                                                  unsigned int (*red)[2])
{
    c0 = wave_total_u(c0);
    c1 = wave_total_u(c1);
    if ((threadIdx.x & 63) == 63) {
        red[threadIdx.x >> 6][0] = c0;
        red[threadIdx.x >> 6][1] = c1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned int s = 0;
        for (int wv = 0; wv < NT / 64; ++wv) s += red[wv][threadIdx.x];
        gp(out)[threadIdx.x] = s;
    }
}

// FP64 totals of the kSE3 columns of up to two partial buffers (pa over na
// blocks -> tot[0, kSE3), pb over nb blocks -> tot[kSE3, 2 kSE3)), in a fixed
// order; a null buffer gives zeros.  Every thread sees tot[] afterwards.  The
// loads are unconditional (clamped index) so each thread has all of them in
// flight at once.
template <int NT>
__device__ __forceinline__ void partials_total2(const float* __restrict__ pa, int na, const float* __restrict__ pb,
                                                int nb, double* tot, double (*red)[NT / 64])
{
    constexpr int G = NT / 64;
    constexpr int R = kMaxBlk / G;
    const int t = threadIdx.x;
    if (t < 2 * kSE3 * G) {
        const int col = t / G, gi = t - col * G;
        const bool second = col >= kSE3;
        const float* part = second ? pb : pa;
        const int nblk = second ? nb : na;
        const int k = second ? col - kSE3 : col;
        double s = 0.0;
        if (part) {
            float v[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int b = gi + r * G;
                v[r] = part[(size_t)(b < nblk ? b : 0) * kPartStride + k];
            }
#pragma unroll
            for (int r = 0; r < R; ++r) s += (gi + r * G < nblk) ? (double)v[r] : 0.0;
        }
        red[col][gi] = s;
    }
    __syncthreads();
    if (t < 2 * kSE3) {
        double s = 0.0;
        for (int gi = 0; gi < G; ++gi) s += red[t][gi];
        tot[t] = s;
    }
    __syncthreads();
}

// Wrapping uint32 totals of the integer partials; every thread sees tot[2].
template <int NT>
__device__ __forceinline__ void int_total(const unsigned int* __restrict__ part, int nblk, unsigned int* tot,
                                          unsigned int (*red)[2])
{
    unsigned int c0 = 0, c1 = 0;
    for (int b = threadIdx.x; b < nblk; b += NT) {
        c0 += part[2 * b];
        c1 += part[2 * b + 1];
    }
    c0 = wave_total_u(c0);
    c1 = wave_total_u(c1);
    if ((threadIdx.x & 63) == 63) {
        red[threadIdx.x >> 6][0] = c0;
        red[threadIdx.x >> 6][1] = c1;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        unsigned int s = 0;
        for (int wv = 0; wv < NT / 64; ++wv) s += red[wv][threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
}

// ======================================================= per-pixel steps
// a14 ICPReduction::search/getProducts, reduce.cu:282-387, split into a
// load phase, a projection phase, a gather phase and a test/product phase so
// one thread keeps several pixels' memory accesses in flight.
struct IcpC {
    float Rc[9], tc[3], Rpi[9], tp[3];
    float fx, fy, cx, cy, dist, angle;
};

struct Px6 {
    float v[3], n[3];
};

__device__ __forceinline__ void icp_load(const LevelDesc& L, int i, size_t P, Px6& s)
{
    s.v[0] = gp(L.vcur)[i];
    s.v[1] = gp(L.vcur)[i + P];
    s.v[2] = gp(L.vcur)[i + 2 * P];
    s.n[0] = gp(L.ncur)[i];
    s.n[1] = gp(L.ncur)[i + P];
    s.n[2] = gp(L.ncur)[i + 2 * P];
}

// Projective association (reduce.cu:290-298).  Returns the gather index or
// -1.  A pixel whose vertex x is NaN is rejected here: in the reference its
// NaN reaches `dist` and fails `dist <= distThres`, so the outcome is the same.
__device__ __forceinline__ int icp_project(const LevelDesc& L, const IcpC& c, const Px6& s, float* vg, float* vcp)
{
    if (isnan(s.v[0])) return -1;
    float d3[3];
    for (int r = 0; r < 3; ++r)
        vg[r] = c.Rc[3 * r] * s.v[0] + c.Rc[3 * r + 1] * s.v[1] + c.Rc[3 * r + 2] * s.v[2] + c.tc[r];
    for (int r = 0; r < 3; ++r) d3[r] = vg[r] - c.tp[r];
    for (int r = 0; r < 3; ++r) vcp[r] = c.Rpi[3 * r] * d3[0] + c.Rpi[3 * r + 1] * d3[1] + c.Rpi[3 * r + 2] * d3[2];
    const int ux = f2i_rn(vcp[0] * c.fx / vcp[2] + c.cx);
    const int uy = f2i_rn(vcp[1] * c.fy / vcp[2] + c.cy);
    if (ux < 0 || uy < 0 || ux >= L.w || uy >= L.h || vcp[2] < 0) return -1;
    return uy * L.w + ux;
}

__device__ __forceinline__ void icp_gather(const LevelDesc& L, int j, size_t P, Px6& d)
{
    d.v[0] = gp(L.vprev)[j];
    d.v[1] = gp(L.vprev)[j + P];
    d.v[2] = gp(L.vprev)[j + 2 * P];
    d.n[0] = gp(L.nprev)[j];
    d.n[1] = gp(L.nprev)[j + P];
    d.n[2] = gp(L.nprev)[j + 2 * P];
}

// reduce.cu:305-348: normal/distance tests and the point-to-plane row.
__device__ __forceinline__ bool icp_row(const IcpC& c, const Px6& s, const float* vg, const float* vcp,
                                       const Px6& d, float* row)
{
    const float* nc = s.n;
    const float* vp = d.v;
    const float* np = d.n;
    float ncg[3];
    for (int r = 0; r < 3; ++r) ncg[r] = c.Rc[3 * r] * nc[0] + c.Rc[3 * r + 1] * nc[1] + c.Rc[3 * r + 2] * nc[2];
    const float df[3] = {vp[0] - vg[0], vp[1] - vg[1], vp[2] - vg[2]};
    const float dist = sqrtf(df[0] * df[0] + df[1] * df[1] + df[2] * df[2]);
    const float cr[3] = {ncg[1] * np[2] - ncg[2] * np[1], ncg[2] * np[0] - ncg[0] * np[2],
                         ncg[0] * np[1] - ncg[1] * np[0]};
    const float sine = sqrtf(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    if (!(sine < c.angle && dist <= c.dist && !isnan(nc[0]) && !isnan(np[0]))) return false;
    float dd[3], dv[3], n[3];
    for (int r = 0; r < 3; ++r) dd[r] = vp[r] - c.tp[r];
    for (int r = 0; r < 3; ++r) dv[r] = c.Rpi[3 * r] * dd[0] + c.Rpi[3 * r + 1] * dd[1] + c.Rpi[3 * r + 2] * dd[2];
    for (int r = 0; r < 3; ++r) n[r] = c.Rpi[3 * r] * np[0] + c.Rpi[3 * r + 1] * np[1] + c.Rpi[3 * r + 2] * np[2];
    const float* sp = vcp;  // Rprev_inv * (vcurr_g - tprev), bit for bit
    row[0] = n[0]; row[1] = n[1]; row[2] = n[2];
    row[3] = sp[1] * n[2] - sp[2] * n[1];
    row[4] = sp[2] * n[0] - sp[0] * n[2];
    row[5] = sp[0] * n[1] - sp[1] * n[0];
    const float sd[3] = {sp[0] - dv[0], sp[1] - dv[1], sp[2] - dv[2]};
    row[6] = n[0] * sd[0] + n[1] * sd[1] + n[2] * sd[2];
    return true;
}

// a15 RGBResidual::getProducts, reduce.cu:808-830; the pose-independent
// tests are folded into zsrc by k_rgb_prep.
struct RgbC {
    float K[9], kt[3];
    float maxDepthDelta;
};

// A block-uniform value moved to a scalar register (the constants of an
// iteration are the same in every lane; kept in SGPRs they cost no VGPRs).
__device__ __forceinline__ float uni(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

__device__ __forceinline__ int rgb_warp(const LevelDesc& L, const RgbC& c, int x, int y, float d1, float& tz)
{
    if (isnan(d1)) return -1;
    const float xf = (float)x, yf = (float)y;
    tz = d1 * (c.K[6] * xf + c.K[7] * yf + c.K[8]) + c.kt[2];
    const int u0 = f2i_rn((d1 * (c.K[0] * xf + c.K[1] * yf + c.K[2]) + c.kt[0]) / tz);
    const int v0 = f2i_rn((d1 * (c.K[3] * xf + c.K[4] * yf + c.K[5]) + c.kt[1]) / tz);
    if (!(u0 >= 0 && v0 >= 0 && u0 < L.w && v0 < L.h)) return -1;
    return v0 * L.w + u0;
}

__device__ __forceinline__ bool rgb_test(const RgbC& c, float tz, float d0, uint8_t il)
{
    return d0 > 0 && fabsf(tz - d0) <= c.maxDepthDelta && il != 0;
}

// a16 RGBReduction::getProducts, reduce.cu:512-552; the cloud point is
// projectPointsKernel's (utils.cu:655-657) recomputed from lastDepth.
__device__ __forceinline__ void rgb_row(const LevelDesc& L, float sigma, float sobelScale, short2 gr, float diff,
                                        float Z, int m, float* row)
{
    const int v0 = m / L.w, u0 = m - v0 * L.w;
    float wt = sigma + fabsf(diff);
    wt = wt > FLT_EPSILON ? 1.0f / wt : 1.0f;
    if (sigma == -1) wt = 1;
    row[6] = -wt * diff;
    const float X = ((float)u0 - L.cx) * Z * L.invFx;
    const float Y = ((float)v0 - L.cy) * Z * L.invFy;
    const float invz = 1.0f / Z;
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

// =================================================== solve and constants
// The ICP statistics of an iteration (the residual[] of the ICP step and its
// errors, RGBDOdometryef.cpp:509-510), by one wave while another solves.
__device__ void store_icp_stats(IterState& S, const TrackParams& P, const double* icp_sums, int lane)
{
    if (P.icp && lane < kSE3) S.icp_sums[lane] = icp_sums[lane];
    if (lane == 0) {
        if (P.icp) {
            S.res0 = (float)icp_sums[27];
            S.res1 = (float)icp_sums[28];
        }
        S.lastICPError = sqrtf(S.res0) / S.res1;
        S.lastICPCount = S.res1;
    }
}

// Select of one of three uniform values by a lane-varying index (registers
// cannot be indexed per lane).
template <class T>
__device__ __forceinline__ T pick3(int i, T a, T b, T c)
{
    return i == 0 ? a : (i == 1 ? b : c);
}

// RGBDOdometryef.cpp:533-575 (combine, solve, computeUpdateSE3, pose update)
// by one wave: the 42 entries of the system one per lane (stored as lastA /
// lastb and read back), the factorisation and the increment on every lane
// (a chain of dependent FP64 operations, from registers), then the new
// resultRt entries, Rcurr and tcurr one per lane.  Each entry is formed by
// the same expression on whichever lane computes it, so the lanes only remove
// the serial chain of independent entries.  rgb_sums: FP64 totals of the
// photometric pass (unused without it).  level: the level of the iteration
// being finished; with the convergence policy (P.convT > 0) an update below
// both thresholds stops that level.
__device__ __forceinline__ void solve_update_wave(IterState& S, const TrackParams& P, const double* icp_sums,
                                                  const double* rgb_sums, const float* Rprev, const float* tprev,
                                                  int level)
{
    const int lane = threadIdx.x & 63;
    const bool lead = lane == 0;
    // the system (RGBDOdometryef.cpp:533-545), one entry per lane
    if (lane < 42) {
        const int i = lane < 36 ? lane / 6 : lane - 36, j = lane < 36 ? lane % 6 : 6;
        const int r = i < j ? i : j, c = i < j ? j : i;
        const int sidx = 7 * r - r * (r - 1) / 2 + (c - r);
        const double ai = P.icp ? icp_sums[sidx] : 0.0;
        const double ar = P.rgb ? rgb_sums[sidx] : 0.0;
        const double wgt = P.icpWeight;
        double v;
        if (P.icp && P.rgb) v = (c == 6) ? ar + wgt * ai : ar + wgt * wgt * ai;
        else v = P.icp ? ai : ar;
        if (lane < 36) S.lastA[lane] = v;
        else S.lastb[lane - 36] = v;
    }
    __builtin_amdgcn_wave_barrier();
    double A[36], b[6], x[6];
#pragma unroll
    for (int q = 0; q < 36; ++q) A[q] = S.lastA[q];
#pragma unroll
    for (int q = 0; q < 6; ++q) b[q] = S.lastb[q];
    int zeros = 0;
    if (!ldlt_spd<6>(A, b, x)) zeros = ldlt_solve<6>(A, b, x);
    double rot[9];
    rodrigues(x + 3, rot);
    // resultRt = [rot | x_t] * resultRt, entry (r, c) on lane 4r + c
    {
        const int r = (lane >> 2) < 3 ? (lane >> 2) : 2, c = lane & 3;
        const double o0 = S.resultRt[c], o1 = S.resultRt[4 + c], o2 = S.resultRt[8 + c], o3 = S.resultRt[12 + c];
        const double r0 = pick3(r, rot[0], rot[3], rot[6]);
        const double r1 = pick3(r, rot[1], rot[4], rot[7]);
        const double r2 = pick3(r, rot[2], rot[5], rot[8]);
        const double xr = pick3(r, x[0], x[1], x[2]);
        const double v = r0 * o0 + r1 * o1 + r2 * o2 + xr * o3;
        __builtin_amdgcn_wave_barrier();
        if (lane < 12) S.resultRt[lane] = v;
    }
    __builtin_amdgcn_wave_barrier();
    // Rcurr = Rprev * Rf^T (lanes 0-8), tcurr = Rprev * (-Rf^T tf) + tprev (lanes 9-11),
    // with Rf, tf the float rotation and translation of the new resultRt
    if (lane < 12) {
        float Rf[9], tf[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int c = 0; c < 3; ++c) Rf[a * 3 + c] = (float)S.resultRt[a * 4 + c];
            tf[a] = (float)S.resultRt[a * 4 + 3];
        }
        // Ri = Rf^T, ti = -Ri tf
        float Ri[9], ti[3];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int c = 0; c < 3; ++c) Ri[a * 3 + c] = Rf[c * 3 + a];
#pragma unroll
        for (int a = 0; a < 3; ++a) ti[a] = -(Ri[a * 3] * tf[0] + Ri[a * 3 + 1] * tf[1] + Ri[a * 3 + 2] * tf[2]);
        const int a = lane < 9 ? lane / 3 : lane - 9, c = lane < 9 ? lane % 3 : 0;
        const float p0 = pick3(a, Rprev[0], Rprev[3], Rprev[6]);
        const float p1 = pick3(a, Rprev[1], Rprev[4], Rprev[7]);
        const float p2 = pick3(a, Rprev[2], Rprev[5], Rprev[8]);
        if (lane < 9) {
            S.Rcurr[lane] = p0 * pick3(c, Ri[0], Ri[1], Ri[2]) + p1 * pick3(c, Ri[3], Ri[4], Ri[5]) +
                            p2 * pick3(c, Ri[6], Ri[7], Ri[8]);
        } else {
            S.tcurr[a] = p0 * ti[0] + p1 * ti[1] + p2 * ti[2] + pick3(a, tprev[0], tprev[1], tprev[2]);
        }
    }
    if (lead) {
        S.degenerate = (zeros == 6);
        S.iters_done++;
        if (P.convT > 0.0f) {
            const double nt = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
            const double nr = sqrt(x[3] * x[3] + x[4] * x[4] + x[5] * x[5]);
            if (nt < (double)P.convT && nr < (double)P.convR) S.break_level = level;
        }
    }
}

// Per-iteration warp constants from resultRt (RGBDOdometryef.cpp:422-432),
// Rt = resultRt^-1 taken as the rigid inverse [R^T | -R^T t], by one wave:
// krkinv entry e = (K Rinv Ki)[e] on lane e (lanes 0-8) and kt[a] = (K tinv)[a]
// on lane 9 + a; each lane forms the row of K Rinv it needs with the
// expressions of mm3d.  K, Ki: the level's intrinsics and their inverse.
__device__ __forceinline__ void iter_constants_wave(IterState& S, const double* K, const double* Ki)
{
    const int lane = threadIdx.x & 63;
    if (lane >= 12) return;
    const double* Rt = S.resultRt;
    double Rinv[9];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) Rinv[a * 3 + b] = Rt[b * 4 + a];
    const int a = lane < 9 ? lane / 3 : lane - 9, c = lane < 9 ? lane % 3 : 0;
    const double k0 = pick3(a, K[0], K[3], K[6]), k1 = pick3(a, K[1], K[4], K[7]), k2 = pick3(a, K[2], K[5], K[8]);
    if (lane < 9) {
        double KR[3];
#pragma unroll
        for (int j = 0; j < 3; ++j) KR[j] = k0 * Rinv[j] + k1 * Rinv[3 + j] + k2 * Rinv[6 + j];
        const double v = KR[0] * pick3(c, Ki[0], Ki[1], Ki[2]) + KR[1] * pick3(c, Ki[3], Ki[4], Ki[5]) +
                         KR[2] * pick3(c, Ki[6], Ki[7], Ki[8]);
        S.krkinv[lane] = (float)v;
    } else {
        double tinv[3];
#pragma unroll
        for (int m = 0; m < 3; ++m) tinv[m] = -(Rinv[m * 3] * Rt[3] + Rinv[m * 3 + 1] * Rt[7] + Rinv[m * 3 + 2] * Rt[11]);
        S.kt[a] = (float)(k0 * tinv[0] + k1 * tinv[1] + k2 * tinv[2]);
    }
}

// RGBDOdometryef.cpp:434-471: photometric bookkeeping after the association
// pass of an iteration.  cmp: the lastRGBError this iteration compares to.
// Sets S.broke (rgbOnly early exit) or S.pending (a solve is due).
// sigmaVal and the error of the photometric association from its count and
// sum of squared differences (with the precedence of RGBDOdometryef.cpp:457).
__device__ __forceinline__ void rgb_sigma_error(int count, int sigma, float& sigmaVal, float& rgbError)
{
    sigmaVal = (float)sqrt((double)((((float)sigma / (float)count) == 0.0f) ? 1 : count));
    rgbError = (float)(sqrt((double)sigma) / (double)(count == 0 ? 1 : count));
}

__device__ void assoc_rgb(IterState& S, const TrackParams& P, float cmp, int count, int sigma)
{
    float sigmaVal, rgbError;
    if (P.rgb) {
        S.rgb_count = count;
        S.rgb_sigma = sigma;
        rgb_sigma_error(count, sigma, sigmaVal, rgbError);
    } else {
        sigmaVal = 0.0f;
        rgbError = 0.0f;
        count = 0;
    }
    if (P.rgbOnly && rgbError > cmp) {
        S.broke = 1;
        S.pending = 0;
        return;
    }
    S.broke = 0;
    S.lastRGBError = rgbError;
    S.lastRGBCount = (float)count;
    if (P.rgbOnly) sigmaVal = -1;
    S.sigmaVal = sigmaVal;
    S.pending = 1;
}

// Load of a state record from global memory into LDS, word per thread.
template <int NT, class T>
__device__ __forceinline__ void copy_in(T* lds, const T* src)
{
    static_assert(sizeof(T) % 4 == 0, "state records are word aligned");
    const GLOBAL_AS unsigned int* s = gp((const unsigned int*)src);
    unsigned int* d = (unsigned int*)lds;
    for (int i = threadIdx.x; i < (int)(sizeof(T) / 4); i += NT) d[i] = s[i];
}

// Store of a record from LDS to global memory, word per thread.
template <int NT, class T>
__device__ __forceinline__ void copy_out(T* dst, const T* lds)
{
    static_assert(sizeof(T) % 4 == 0, "state records are word aligned");
    const unsigned int* s = (const unsigned int*)lds;
    GLOBAL_AS unsigned int* d = gp((unsigned int*)dst);
    for (int i = threadIdx.x; i < (int)(sizeof(T) / 4); i += NT) d[i] = s[i];
}

// Scratch shared by the prologues.
// The pose the call started from, held in registers from the first
// instructions of a step kernel so no later phase waits on its loads.
struct PrevPose {
    float R[9], t[3], Ri[9];
};

__device__ __forceinline__ void load_prev(const TrackState* st_, PrevPose& pp)
{
    const GLOBAL_AS TrackState* st = gp(st_);
#pragma unroll
    for (int q = 0; q < 9; ++q) {
        pp.R[q] = st->Rprev[q];
        pp.Ri[q] = st->Rprev_inv[q];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) pp.t[q] = st->tprev[q];
}

// The block partials of a finished pass, loaded as float4 chunks: thread t
// owns chunk t % 8 (floats 4c..4c+3 of a record) of the records t / 8,
// t / 8 + NT / 8, ...  The loads are issued together with the state copy so
// both complete in one memory round trip.
template <int NT>
struct PartLoad {
    static constexpr int RG = NT / 8;           // record groups
    static constexpr int IT = kMaxBlk / RG;     // records per thread and pass
    float4 v[2][IT];
    int n[2];                                   // records of each pass (0: not summed)
};

// Both passes' loads back to back and without branches, so that one wait
// covers them: records past a pass's count re-read record 0, and a pass that
// is off reads its (allocated) buffer; neither is summed.
template <int NT>
__device__ __forceinline__ void partials_issue(const float* __restrict__ pa, int na, const float* __restrict__ pb,
                                               int nb, PartLoad<NT>& L)
{
    constexpr int RG = PartLoad<NT>::RG, IT = PartLoad<NT>::IT;
    const int c = threadIdx.x & 7, rg = threadIdx.x >> 3;
    L.n[0] = na;
    L.n[1] = nb;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const float* part = a ? pb : pa;
        const int n = L.n[a];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const int r = rg + i * RG;
            L.v[a][i] = *gp((const float4*)(part + (size_t)(r < n ? r : 0) * kPartStride + 4 * c));
        }
    }
}

// Marks the point from which the partials may be waited for: the empty asm
// statements keep the compiler from consuming them (and so waiting for the
// whole load queue) before the loads issued after them.
template <int NT>
__device__ __forceinline__ void partials_fence(PartLoad<NT>& L)
{
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int i = 0; i < PartLoad<NT>::IT; ++i) {
            float x = L.v[a][i].x, y = L.v[a][i].y, z = L.v[a][i].z, w = L.v[a][i].w;
            asm volatile("" : "+v"(x), "+v"(y), "+v"(z), "+v"(w));
            L.v[a][i] = make_float4(x, y, z, w);
        }
}

// FP64 totals of the loaded partials in a fixed order: per thread over its
// records, then per column over the record groups in LDS (each wave a slice
// of the groups: the slice of wave w holds exactly the groups its own lanes
// stored, so this step needs no block barrier), then the waves' sums.  Those
// are formed by wave 0 into tot (for the solve) and by wave 1 into tot1 (for
// the system and statistics it stores), so one block barrier suffices.
// tot[0..28] pass A, tot[29..57] pass B.
template <int NT>
__device__ __forceinline__ void partials_sum(const PartLoad<NT>& L, double* tot, double* tot1, double (*red)[64],
                                             double (*fin)[64])
{
    constexpr int RG = PartLoad<NT>::RG, IT = PartLoad<NT>::IT, G = NT / 64, SL = RG / G;
    static_assert(SL * 8 == 64, "a wave's lanes store exactly its slice of record groups");
    const int c = threadIdx.x & 7, rg = threadIdx.x >> 3;
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int n = L.n[a];
        double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const bool ok = rg + i * RG < n;
            s0 += ok ? (double)L.v[a][i].x : 0.0;
            s1 += ok ? (double)L.v[a][i].y : 0.0;
            s2 += ok ? (double)L.v[a][i].z : 0.0;
            s3 += ok ? (double)L.v[a][i].w : 0.0;
        }
        red[rg][a * 32 + 4 * c] = s0;
        red[rg][a * 32 + 4 * c + 1] = s1;
        red[rg][a * 32 + 4 * c + 2] = s2;
        red[rg][a * 32 + 4 * c + 3] = s3;
    }
    __builtin_amdgcn_wave_barrier();
    const int col = threadIdx.x & 63, w = threadIdx.x >> 6;
    {
        double t = 0.0;
#pragma unroll
        for (int g = 0; g < SL; ++g) t += red[w * SL + g][col];
        fin[w][col] = t;
    }
    __syncthreads();
    if (w < 2) {
        const int a = col >> 5, k = col & 31;
        if (k < kSE3) {
            double t = 0.0;
#pragma unroll
            for (int q = 0; q < G; ++q) t += fin[q][col];
            (w == 0 ? tot : tot1)[a * kSE3 + k] = t;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

template <int NT>
struct Prolog {
    IterState S;
    double fin[2 * kSE3][NT / 64];
    double red[NT / 8][64];
    double fin2[NT / 64][64];
    double tot[2 * kSE3];
    double tot1[2 * kSE3];
    unsigned int redi[NT / 64][2];
    unsigned int toti[2];
};

// Shared memory of the unfused pass A: the state record and the count totals.
template <int NT>
struct StateScratch {
    IterState S;
    unsigned int redi[NT / 64][2];
};

// The prologue of an iteration: finish iteration k-1 and leave the state
// entering iteration k in W.S -- one fused FP64 reduction of its ICP partials
// (pass A) and photometric partials (pass B), the ICP statistics, the
// photometric bookkeeping when there is no pass B, the solve and, given the
// level of iteration k, its warp constants.  Every block computes the same
// state.  Split in two so that a kernel can issue its own loads in between:
// prologue_issue loads the state record (one word per thread) and the
// partials; prologue_finish waits for them only after those loads.
template <int NT>
struct PrologIn {
    unsigned int sw;                 // this thread's word of the state record
    PartLoad<NT> PL;
};

template <int NT>
__device__ __forceinline__ void prologue_issue(const PairDesc& D, const TrackParams& P, int k, int prev_level,
                                               PrologIn<NT>& I)
{
    constexpr int NW = (int)(sizeof(IterState) / 4);
    static_assert(NW <= NT, "one word of the state record per thread");
    // k == 0 starts from slot 1 (written by the call's set-up) and sums no
    // partials; its loads still go to allocated buffers, without branches
    const int slot = (k == 0) ? 1 : ((k - 1) & 1);
    const int plv = (k > 0) ? prev_level : 0;
    const GLOBAL_AS unsigned int* src = gp((const unsigned int*)&D.st->it[slot]);
    I.sw = src[(int)threadIdx.x < NW ? threadIdx.x : 0];
    partials_issue<NT>(D.part_a[slot], (k > 0 && P.icp) ? D.nblk_a[plv] : 0, D.part_b[slot],
                       (k > 0 && P.rgb) ? D.nblk_b[plv] : 0, I.PL);
}

// Orders one wave's LDS accesses across its lanes (a lane reads what another
// lane of the wave stored before this point).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// State entering iteration `iter` of `level` from the first wave: skipped
// when the rgbOnly test stopped this level, else the warp constants of the
// current estimate (RGBDOdometryef.cpp:418-432), and the lastRGBError the test
// compares to.  K, Ki: the level's intrinsics and their inverse.
__device__ __forceinline__ void enter_iteration_wave(IterState& S, const double* K, const double* Ki, int level,
                                                     int iter)
{
    if ((threadIdx.x & 63) == 0) {
        S.skip = (S.break_level == level);
        S.rgb_err_cmp = (iter == 0) ? FLT_MAX : S.lastRGBError;
    }
    wave_lds_sync();
    if (!S.skip) iter_constants_wave(S, K, Ki);
}

// K, Ki: intrinsics of the level of iteration k (its warp constants are set),
// or nullptr.
template <int NT>
__device__ __forceinline__ void prologue_finish(const PairDesc& D, const TrackParams& P, int k, int prev_level,
                                                const PrevPose& pp, Prolog<NT>& W, PrologIn<NT>& I, const double* K,
                                                const double* Ki, int level, int iter)
{
    constexpr int NW = (int)(sizeof(IterState) / 4);
    asm volatile("" : "+v"(I.sw));
    partials_fence<NT>(I.PL);
    if ((int)threadIdx.x < NW) ((unsigned int*)&W.S)[threadIdx.x] = I.sw;
    __syncthreads();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (k == 0) {
        if (K && wave == 0) enter_iteration_wave(W.S, K, Ki, level, iter);
        __syncthreads();
        return;
    }
    const bool run = P.rgb ? (W.S.pending != 0) : (W.S.skip == 0);
    if (run) partials_sum<NT>(I.PL, W.tot, W.tot1, W.red, W.fin2);
    // wave 0 solves, does the bookkeeping and sets the next iteration's
    // constants; wave 1 meanwhile stores the ICP statistics (disjoint fields)
    if (wave == 0) {
        if (run) solve_update_wave(W.S, P, W.tot, W.tot + kSE3, pp.R, pp.t, prev_level);
        wave_lds_sync();
        if (lane == 0) {
            if (run && !P.rgb) assoc_rgb(W.S, P, W.S.rgb_err_cmp, 0, 0);
            if (W.S.broke) W.S.break_level = prev_level;
            W.S.pending = 0;
            W.S.broke = 0;
        }
        wave_lds_sync();
        if (K) enter_iteration_wave(W.S, K, Ki, level, iter);
    } else if (wave == 1 && run) {
        store_icp_stats(W.S, P, W.tot1, lane);
    }
    __syncthreads();
}

template <int NT>
__device__ __forceinline__ void finish_previous(const PairDesc& D, const TrackParams& P, int k, int prev_level,
                                                const PrevPose& pp, Prolog<NT>& W, const double* K, const double* Ki,
                                                int level, int iter)
{
    PrologIn<NT> I;
    prologue_issue<NT>(D, P, k, prev_level, I);
    prologue_finish<NT>(D, P, k, prev_level, pp, W, I, K, Ki, level, iter);
}

// ======================================================== step kernel A
// ICP products (a14) and the photometric association count (a15) of one
// level of every pair, after finishing the previous iteration in the
// prologue.  Each pair's level is cut into tiles of NT*PPT pixels; its
// blocks (at most kMaxBlk) stride over the tiles.  mode 1: parity pass with
// the constants of state slot 0.
template <int NT, int PPT>
struct TileA {
    Px6 src[PPT];
    float zs[PPT];
    uint8_t in[PPT];
};

template <int NT, int PPT>
__device__ __forceinline__ void load_tile_a(const LevelDesc& L, const TrackParams& P, int base, int N, size_t P3,
                                            TileA<NT, PPT>& T)
{
    // Without branches, so that the prologue's waits (issued after these
    // loads) can count them exactly: a pixel past the end reads the last
    // pixel, a pass that is off reads valid memory (the vertex map, present
    // in every pair), and both are masked after the loads.
    const float* zsrc = P.rgb ? L.zsrc : L.vcur;
    const uint8_t* inext = P.rgb ? L.inext : (const uint8_t*)L.vcur;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = base + q * NT;
        const bool in_range = idx < N;
        const int i = in_range ? idx : N - 1;
        icp_load(L, i, P3, T.src[q]);
        const float z = gp(zsrc)[i];
        const uint8_t in = gp(inext)[i];
        if (!(in_range && P.icp)) T.src[q].v[0] = QNAN;
        T.zs[q] = (in_range && P.rgb) ? z : QNAN;
        T.in[q] = (in_range && P.rgb) ? in : (uint8_t)0;
    }
}

// The pose-dependent gathers of a tile: the ICP projections and their
// gathers, then the photometric warps (computed while those gathers fly) and
// theirs, all without branches (an index of -1 reads pixel 0, and a pass that
// is off reads the vertex map; both are masked in rows_tile_a).
template <int NT, int PPT>
struct GatherA {
    int j[PPT], jm[PPT];
    float vg[PPT][3], vcp[PPT][3], tz[PPT], d0[PPT];
    Px6 dst[PPT];
    uint8_t il[PPT];
};

template <int NT, int PPT>
__device__ __forceinline__ void gather_tile_a(const LevelDesc& L, const TrackParams& P, const IcpC& ic,
                                              const RgbC& rc, int base, int N, size_t P3, const TileA<NT, PPT>& T,
                                              GatherA<NT, PPT>& G)
{
    const float* vprev = P.icp ? L.vprev : L.vcur;
    const float* nprev = P.icp ? L.nprev : L.vcur;
    const float* zlast = P.rgb ? L.zlast : L.vcur;
    const uint8_t* ilast = P.rgb ? L.ilast : (const uint8_t*)L.vcur;
#pragma unroll
    for (int q = 0; q < PPT; ++q) G.j[q] = icp_project(L, ic, T.src[q], G.vg[q], G.vcp[q]);
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int j = G.j[q] >= 0 ? G.j[q] : 0;
        G.dst[q].v[0] = gp(vprev)[j];
        G.dst[q].v[1] = gp(vprev)[j + P3];
        G.dst[q].v[2] = gp(vprev)[j + 2 * P3];
        G.dst[q].n[0] = gp(nprev)[j];
        G.dst[q].n[1] = gp(nprev)[j + P3];
        G.dst[q].n[2] = gp(nprev)[j + 2 * P3];
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int idx = base + q * NT;
        const int y = idx / L.w, x = idx - y * L.w;
        G.jm[q] = (idx < N) ? rgb_warp(L, rc, x, y, T.zs[q], G.tz[q]) : -1;
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        const int j = G.jm[q] >= 0 ? G.jm[q] : 0;
        G.d0[q] = gp(zlast)[j];
        G.il[q] = gp(ilast)[j];
    }
}

template <int NT, int PPT>
__device__ __forceinline__ void rows_tile_a(const LevelDesc& L, const IcpC& ic, const RgbC& rc, int base, int N,
                                            const TileA<NT, PPT>& T, const GatherA<NT, PPT>& G,
                                            float (&acc)[kSE3], unsigned int& cnt, unsigned int& sig,
                                            uint8_t* __restrict__ dbg_mask)
{
    unsigned int m[PPT];
#pragma unroll
    for (int q = 0; q < PPT; ++q) m[q] = 0;
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        float row[7];
        if (G.j[q] >= 0 && icp_row(ic, T.src[q], G.vg[q], G.vcp[q], G.dst[q], row)) {
            acc_row7(acc, row);
            m[q] = 1u;
        }
    }
#pragma unroll
    for (int q = 0; q < PPT; ++q)
        if (G.jm[q] >= 0 && rgb_test(rc, G.tz[q], G.d0[q], G.il[q])) {
            const float diff = (float)T.in[q] - (float)G.il[q];
            cnt += 1u;
            sig += (unsigned int)(int)(diff * diff);
            m[q] |= 2u;
        }
    if (dbg_mask)
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int idx = base + q * NT;
            if (idx < N) dbg_mask[idx] = (uint8_t)m[q];
        }
}

template <int NT, int PPT, bool FUSED>
__global__ __launch_bounds__(NT) void k_step_a(const PairDesc* __restrict__ pairs, const int* __restrict__ blk_pair,
                                               int level, int iter, int k, int prev_level, TrackParams P,
                                               uint8_t* __restrict__ dbg_mask)
{
    // the fused prologue's scratch; the unfused pass reads the state only
    __shared__ std::conditional_t<FUSED, Prolog<NT>, StateScratch<NT>> W;
    __shared__ float s_red[NT / 64][kSE3];
    const int p = blk_pair ? blk_pair[blockIdx.x] : 0;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    const LevelDesc L = D.lv[level];
    const int lb = blockIdx.x - D.blk0_a[level];
    const int nblk = D.nblk_a[level];
    const int N = L.w * L.h;
    const size_t P3 = (size_t)N;
    const int ntiles = (N + NT * PPT - 1) / (NT * PPT);

    // the prologue's loads first, then the pose-independent loads of the
    // first tile, which stay in flight while the prologue sums and solves
    PrologIn<NT> PI;
    if constexpr (FUSED) prologue_issue<NT>(D, P, k, prev_level, PI);
    PrevPose pp;
    load_prev(st, pp);
    TileA<NT, PPT> T;
    load_tile_a<NT, PPT>(L, P, lb * NT * PPT + threadIdx.x, N, P3, T);

    if constexpr (FUSED) {
        prologue_finish<NT>(D, P, k, prev_level, pp, W, PI, L.K, L.Ki, level, iter);
        if (lb == 0) copy_out<NT>(&st->it[k & 1], &W.S);
    } else {
        copy_in<NT>(&W.S, &st->it[k & 1]);
        __syncthreads();
    }
    if (W.S.skip) return;

    IcpC ic;
    RgbC rc;
    for (int q = 0; q < 9; ++q) {
        ic.Rc[q] = uni(W.S.Rcurr[q]);
        ic.Rpi[q] = uni(pp.Ri[q]);
        rc.K[q] = uni(W.S.krkinv[q]);
    }
    for (int q = 0; q < 3; ++q) {
        ic.tc[q] = uni(W.S.tcurr[q]);
        ic.tp[q] = uni(pp.t[q]);
        rc.kt[q] = uni(W.S.kt[q]);
    }
    ic.fx = L.fx; ic.fy = L.fy; ic.cx = L.cx; ic.cy = L.cy;
    ic.dist = P.distThres; ic.angle = P.angleThres;
    rc.maxDepthDelta = P.maxDepthDelta;

    float acc[kSE3];
#pragma unroll
    for (int q = 0; q < kSE3; ++q) acc[q] = 0.0f;
    unsigned int cnt = 0, sig = 0;
    GatherA<NT, PPT> G;
    if (nblk >= ntiles) {
        // at most one tile per block (single pairs)
        if (lb < ntiles) {
            const int base = lb * NT * PPT + threadIdx.x;
            gather_tile_a<NT, PPT>(L, P, ic, rc, base, N, P3, T, G);
            rows_tile_a<NT, PPT>(L, ic, rc, base, N, T, G, acc, cnt, sig, dbg_mask);
        }
    } else {
        // several tiles per block (batches): the next tile's loads are
        // issued behind this tile's gathers and fly while its rows are formed
        for (int tile = lb; tile < ntiles; tile += nblk) {
            const int base = tile * NT * PPT + threadIdx.x;
            gather_tile_a<NT, PPT>(L, P, ic, rc, base, N, P3, T, G);
            asm volatile("" ::: "memory");   // the gathers stay ahead of the next tile's loads
            TileA<NT, PPT> Tn;
            load_tile_a<NT, PPT>(L, P, base + nblk * NT * PPT, N, P3, Tn);
            rows_tile_a<NT, PPT>(L, ic, rc, base, N, T, G, acc, cnt, sig, dbg_mask);
            T = Tn;
        }
    }
    const int pb = k & 1;
    block_partial<kSE3, NT>(acc, D.part_a[pb] + (size_t)lb * kPartStride, s_red);
    block_partial_int<NT>(cnt, sig, D.part_ai[pb] + (size_t)lb * 2, W.redi);
}

// The prologue of the fused pass A as a kernel of its own, one workgroup per
// pair: finish iteration k-1 and store the state entering iteration k, which
// the unfused pass A (FUSED = false) then reads.  Used by batches, where it
// keeps the solve's registers out of the pixel kernel.
// NT: the thread count of the pass A that follows, whose fused prologue
// would sum the partials in this same order (so the state is the same bits
// either way).
template <int NT>
__global__ __launch_bounds__(NT) void k_solve(const PairDesc* __restrict__ pairs, int level, int iter, int k,
                                              int prev_level, TrackParams P)
{
    __shared__ Prolog<NT> W;
    const PairDesc& D = pairs[blockIdx.x];
    TrackState* st = D.st;
    PrevPose pp;
    load_prev(st, pp);
    finish_previous<NT>(D, P, k, prev_level, pp, W, D.lv[level].K, D.lv[level].Ki, level, iter);
    copy_out<NT>(&st->it[k & 1], &W.S);
}

// ======================================================== step kernel B
// Photometric products (a16) with the robust weight of the global count.
// The prologue reduces pass A's partials of the same iteration: counts
// (every block: sigma) and, in the pair's block 0, the ICP sums it records.
template <int NT, int PPT>
struct TileB {
    float zs[PPT];
    uint8_t in[PPT];
    short2 gr[PPT];
};

template <int NT, int PPT>
__device__ __forceinline__ void load_tile_b(const LevelDesc& L, int base, int N, TileB<NT, PPT>& T)
{
#pragma unroll
    for (int q = 0; q < PPT; ++q) {
        // without branches (see load_tile_a): past the end, the last pixel, masked
        const int idx = base + q * NT;
        const bool in_range = idx < N;
        const int i = in_range ? idx : N - 1;
        const float z = gp(L.zsrc)[i];
        const uint8_t in = gp(L.inext)[i];
        const short2 g = gp(L.dxy)[i];
        T.zs[q] = in_range ? z : QNAN;
        T.in[q] = in_range ? in : (uint8_t)0;
        T.gr[q] = in_range ? g : make_short2(0, 0);
    }
}

template <int NT, int PPT>
__global__ __launch_bounds__(NT) void k_step_b(const PairDesc* __restrict__ pairs, const int* __restrict__ blk_pair,
                                               int level, int iter, int k, TrackParams P, int mode)
{
    __shared__ unsigned int s_redi[NT / 64][2];
    __shared__ float s_red[NT / 64][kSE3];
    const int p = blk_pair ? blk_pair[blockIdx.x] : 0;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    const LevelDesc L = D.lv[level];
    const int lb = blockIdx.x - D.blk0_b[level];
    const int nblk = D.nblk_b[level];
    const int N = L.w * L.h;
    const int ntiles = (N + NT * PPT - 1) / (NT * PPT);
    const int pb = k & 1;
    const GLOBAL_AS IterState& Sg = *gp(&st->it[mode == 0 ? pb : 0]);
    static_assert(NT >= kMaxBlk, "one count partial per thread");

    // Every load of the prologue is issued before its first wait, without
    // branches: pass A's count partials, the state fields of the iteration,
    // then the first tile (which stays in flight while the counts are summed).
    const int na = D.nblk_a[level];
    const uint2 cp = gp((const uint2*)D.part_ai[pb])[(int)threadIdx.x < na ? threadIdx.x : 0];
    const int skip = Sg.skip;
    const float err_cmp = Sg.rgb_err_cmp;
    const float sigma_g = Sg.sigmaVal;
    RgbC rc;
    for (int q = 0; q < 9; ++q) rc.K[q] = uni(Sg.krkinv[q]);
    for (int q = 0; q < 3; ++q) rc.kt[q] = uni(Sg.kt[q]);
    rc.maxDepthDelta = P.maxDepthDelta;
    TileB<NT, PPT> T;
    load_tile_b<NT, PPT>(L, lb * NT * PPT + threadIdx.x, N, T);
    unsigned int c0 = cp.x, c1 = cp.y;
    asm volatile("" : "+v"(c0), "+v"(c1));   // consumed only after the tile loads are issued
    if (!(mode == 0 && (int)threadIdx.x < na)) c0 = c1 = 0;
    if (mode == 0 && skip) return;

    float sigma;
    if (mode == 0) {
        // the waves' count totals, then every thread sums them and takes the
        // association decision of assoc_rgb itself (one block barrier)
        c0 = wave_total_u(c0);
        c1 = wave_total_u(c1);
        if ((threadIdx.x & 63) == 63) {
            s_redi[threadIdx.x >> 6][0] = c0;
            s_redi[threadIdx.x >> 6][1] = c1;
        }
        __syncthreads();
        unsigned int count = 0, sum2 = 0;
#pragma unroll
        for (int wv = 0; wv < NT / 64; ++wv) {
            count += s_redi[wv][0];
            sum2 += s_redi[wv][1];
        }
        float sv, err;
        rgb_sigma_error((int)count, (int)sum2, sv, err);
        const bool broke = P.rgbOnly && err > err_cmp;
        if (P.rgbOnly) sv = -1;
        if (lb == 0 && threadIdx.x == 0) {
            GLOBAL_AS IterState* G = gp(&st->it[pb]);
            G->rgb_count = (int)count;
            G->rgb_sigma = (int)sum2;
            if (!broke) {
                G->lastRGBError = err;
                G->lastRGBCount = (float)(int)count;
                G->sigmaVal = sv;
            }
            G->pending = broke ? 0 : 1;
            G->broke = broke ? 1 : 0;
        }
        if (broke) return;
        sigma = sv;
    } else {
        sigma = sigma_g;
    }

    float acc[kSE3];
#pragma unroll
    for (int q = 0; q < kSE3; ++q) acc[q] = 0.0f;
    int jm[PPT];
    float tz[PPT], d0[PPT];
    uint8_t il[PPT];
    auto warp_tile = [&](int base) {
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const int idx = base + q * NT;
            const int y = idx / L.w, x = idx - y * L.w;
            jm[q] = (idx < N) ? rgb_warp(L, rc, x, y, T.zs[q], tz[q]) : -1;
        }
    };
    auto rows_tile = [&]() {
#pragma unroll
        for (int q = 0; q < PPT; ++q)
            if (jm[q] >= 0 && rgb_test(rc, tz[q], d0[q], il[q])) {
                const float diff = (float)T.in[q] - (float)il[q];
                float row[7];
                rgb_row(L, sigma, P.sobelScale, T.gr[q], diff, d0[q], jm[q], row);
                acc_row7(acc, row);
            }
    };
    if (nblk >= ntiles) {
        // at most one tile per block (single pairs)
        if (lb < ntiles) {
            warp_tile(lb * NT * PPT + threadIdx.x);
#pragma unroll
            for (int q = 0; q < PPT; ++q)
                if (jm[q] >= 0) {
                    d0[q] = gp(L.zlast)[jm[q]];
                    il[q] = gp(L.ilast)[jm[q]];
                }
            rows_tile();
        }
    } else {
        // several tiles per block (batches): the next tile's loads are
        // issued behind this tile's gathers, without branches (past the end
        // they read the last pixel), and fly while this tile's rows are formed
        for (int tile = lb; tile < ntiles; tile += nblk) {
            const int base = tile * NT * PPT + threadIdx.x;
            warp_tile(base);
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int j = jm[q] >= 0 ? jm[q] : 0;
                d0[q] = gp(L.zlast)[j];
                il[q] = gp(L.ilast)[j];
            }
            asm volatile("" ::: "memory");   // the gathers stay ahead of the next tile's loads
            TileB<NT, PPT> Tn;
            load_tile_b<NT, PPT>(L, base + nblk * NT * PPT, N, Tn);
            rows_tile();
            T = Tn;
        }
    }
    block_partial<kSE3, NT>(acc, D.part_b[pb] + (size_t)lb * kPartStride, s_red);
}

// Totals of one parity pass (mode 1) into the state's dbg fields.
__global__ __launch_bounds__(256) void k_reduce_dbg(const PairDesc* __restrict__ pairs, int which, int level)
{
    __shared__ Prolog<256> W;
    const PairDesc& D = pairs[0];
    TrackState* st = D.st;
    const int nblk = which == 0 ? D.nblk_a[level] : D.nblk_b[level];
    partials_total2<256>(which == 0 ? D.part_a[0] : D.part_b[0], nblk, nullptr, 1, W.tot, W.fin);
    if (which == 0) int_total<256>(D.part_ai[0], nblk, W.toti, W.redi);
    if (threadIdx.x < kSE3) st->dbg_sums[threadIdx.x] = W.tot[threadIdx.x];
    if (threadIdx.x == 0 && which == 0) {
        st->dbg_count = (int)W.toti[0];
        st->dbg_sigma = (int)W.toti[1];
    }
}

void launch_reduce_dbg(const PairDesc* pairs, int which, int level, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_dbg, dim3(1), dim3(256), 0, s, pairs, which, level);
}

int step_shape_supported(StepShape sh)
{
    const bool t = sh.threads == 256 || sh.threads == 512;
    const bool p = sh.ppt >= 1 && sh.ppt <= 4;
    return t && p;
}

#define STEP_DISPATCH(KER, ...)                                                                           \
    do {                                                                                                 \
        const dim3 grid(L.total_blocks);                                                                 \
        switch (L.shape.threads * 100 + L.shape.ppt) {                                                   \
            case 25601: hipLaunchKernelGGL((KER<256, 1>), grid, dim3(256), 0, s, __VA_ARGS__); break;    \
            case 25602: hipLaunchKernelGGL((KER<256, 2>), grid, dim3(256), 0, s, __VA_ARGS__); break;    \
            case 25603: hipLaunchKernelGGL((KER<256, 3>), grid, dim3(256), 0, s, __VA_ARGS__); break;    \
            case 25604: hipLaunchKernelGGL((KER<256, 4>), grid, dim3(256), 0, s, __VA_ARGS__); break;    \
            case 51201: hipLaunchKernelGGL((KER<512, 1>), grid, dim3(512), 0, s, __VA_ARGS__); break;    \
            case 51202: hipLaunchKernelGGL((KER<512, 2>), grid, dim3(512), 0, s, __VA_ARGS__); break;    \
            case 51203: hipLaunchKernelGGL((KER<512, 3>), grid, dim3(512), 0, s, __VA_ARGS__); break;    \
            case 51204: hipLaunchKernelGGL((KER<512, 4>), grid, dim3(512), 0, s, __VA_ARGS__); break;    \
            default: break;                                                                              \
        }                                                                                                \
    } while (0)

#define STEP_DISPATCH_A(FUS, ...)                                                                        \
    do {                                                                                                 \
        const dim3 grid(L.total_blocks);                                                                 \
        switch (L.shape.threads * 100 + L.shape.ppt) {                                                   \
            case 25601: hipLaunchKernelGGL((k_step_a<256, 1, FUS>), grid, dim3(256), 0, s, __VA_ARGS__); break; \
            case 25602: hipLaunchKernelGGL((k_step_a<256, 2, FUS>), grid, dim3(256), 0, s, __VA_ARGS__); break; \
            case 25603: hipLaunchKernelGGL((k_step_a<256, 3, FUS>), grid, dim3(256), 0, s, __VA_ARGS__); break; \
            case 25604: hipLaunchKernelGGL((k_step_a<256, 4, FUS>), grid, dim3(256), 0, s, __VA_ARGS__); break; \
            case 51201: hipLaunchKernelGGL((k_step_a<512, 1, FUS>), grid, dim3(512), 0, s, __VA_ARGS__); break; \
            case 51202: hipLaunchKernelGGL((k_step_a<512, 2, FUS>), grid, dim3(512), 0, s, __VA_ARGS__); break; \
            case 51203: hipLaunchKernelGGL((k_step_a<512, 3, FUS>), grid, dim3(512), 0, s, __VA_ARGS__); break; \
            case 51204: hipLaunchKernelGGL((k_step_a<512, 4, FUS>), grid, dim3(512), 0, s, __VA_ARGS__); break; \
            default: break;                                                                              \
        }                                                                                                \
    } while (0)

// mode 0: fused prologue (finish iteration k-1 in every block); mode 1: the
// state entering iteration k is already in slot k & 1 (k_solve, or the
// parity-check entry points with k = 0).
void launch_step_a(const StepLaunch& L, int level, int iter, int k, int prev_level, const TrackParams& P, int mode,
                   uint8_t* dbg_mask, hipStream_t s)
{
    if (mode == 0) STEP_DISPATCH_A(true, L.pairs, L.blk_pair, level, iter, k, prev_level, P, dbg_mask);
    else STEP_DISPATCH_A(false, L.pairs, L.blk_pair, level, iter, k, prev_level, P, dbg_mask);
}

void launch_solve(const PairDesc* pairs, int npairs, int level, int iter, int k, int prev_level, const TrackParams& P,
                  int threads, hipStream_t s)
{
    if (threads == 512)
        hipLaunchKernelGGL((k_solve<512>), dim3(npairs), dim3(512), 0, s, pairs, level, iter, k, prev_level, P);
    else
        hipLaunchKernelGGL((k_solve<256>), dim3(npairs), dim3(256), 0, s, pairs, level, iter, k, prev_level, P);
}

void launch_step_b(const StepLaunch& L, int level, int iter, int k, const TrackParams& P, int mode, hipStream_t s)
{
    STEP_DISPATCH(k_step_b, L.pairs, L.blk_pair, level, iter, k, P, mode);
}

// ======================================================= begin / end
// RGBDOdometryef.cpp:278-282, 386-401: pose prior and its inverse, and the
// state before the first iteration (the increment seeded with the SO3
// rotation) in slot 1.
// current = [Rprev | tprev] * resultRt^-1 with resultRt taken to float
// (RGBDOdometryef.cpp:563-575), as the solve's pose update.
__device__ void pose_from_result(const double* Rt, const float* Rprev, const float* tprev, float* Rcurr,
                                 float* tcurr)
{
    float Rf[9], tf[3], Ri[9], ti[3];
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) Rf[r * 3 + c] = (float)Rt[r * 4 + c];
        tf[r] = (float)Rt[r * 4 + 3];
    }
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) Ri[r * 3 + c] = Rf[c * 3 + r];
    for (int r = 0; r < 3; ++r) ti[r] = -(Ri[r * 3] * tf[0] + Ri[r * 3 + 1] * tf[1] + Ri[r * 3 + 2] * tf[2]);
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c)
            Rcurr[r * 3 + c] = Rprev[r * 3] * Ri[c] + Rprev[r * 3 + 1] * Ri[3 + c] + Rprev[r * 3 + 2] * Ri[6 + c];
        tcurr[r] = Rprev[r * 3] * ti[0] + Rprev[r * 3 + 1] * ti[1] + Rprev[r * 3 + 2] * ti[2] + tprev[r];
    }
}

// Warm start.  The predicted camera motion (pred: scaled translation and
// rotation vector, 6 floats per pair) becomes the candidate seed
// resultRt = M^-1, stored by block 0 of the pair.  With `select`, both starts
// -- the prediction and the plain one (identity increment) -- are scored by
// the ICP association at level 2 (inlier count and sum of squared
// point-to-plane residuals), each of kSeedBlocks workgroups per pair storing
// its 4 sums as a block partial; k_track_begin decides.
constexpr int kSeedBlocks = 32;

__global__ __launch_bounds__(256) void k_seed_eval(const PairDesc* __restrict__ pairs, const float* __restrict__ pred,
                                                   float t_scale, float r_scale, int select, TrackParams P)
{
    __shared__ double sSeed[16];
    __shared__ float sRc[2][9], sTc[2][3], sRpi[9], sTp[3];
    __shared__ float sRed[4][4];
    const int p = blockIdx.x / kSeedBlocks, lb = blockIdx.x - p * kSeedBlocks;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    if (threadIdx.x == 0) {
        float Rprev[9], tprev[3];
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) Rprev[r * 3 + c] = st->pose_out[r * 4 + c];
            tprev[r] = st->pose_out[r * 4 + 3];
        }
        inv3f(Rprev, sRpi);
        for (int r = 0; r < 3; ++r) sTp[r] = tprev[r];
        const float* q = pred + 6 * p;
        double w[3], R[9], t[3];
        for (int r = 0; r < 3; ++r) {
            t[r] = (double)q[r] * (double)t_scale;
            w[r] = (double)q[3 + r] * (double)r_scale;
        }
        rodrigues(w, R);
        // M = [R | t] (prev -> current camera); seed = M^-1 = [R^T | -R^T t]
        for (int r = 0; r < 3; ++r) {
            for (int c = 0; c < 3; ++c) sSeed[r * 4 + c] = R[c * 3 + r];
            sSeed[r * 4 + 3] = -(R[0 * 3 + r] * t[0] + R[1 * 3 + r] * t[1] + R[2 * 3 + r] * t[2]);
        }
        sSeed[12] = sSeed[13] = sSeed[14] = 0.0;
        sSeed[15] = 1.0;
        for (int a = 0; a < 9; ++a) sRc[0][a] = Rprev[a];
        for (int a = 0; a < 3; ++a) sTc[0][a] = tprev[a];
        pose_from_result(sSeed, Rprev, tprev, sRc[1], sTc[1]);
        if (lb == 0)
            for (int a = 0; a < 16; ++a) st->seedRt[a] = sSeed[a];
    }
    __syncthreads();
    if (!select) return;
    const LevelDesc& L = D.lv[2];
    const int N = L.w * L.h;
    const size_t P3 = (size_t)N;
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    for (int cand = 0; cand < 2; ++cand) {
        IcpC ic;
        for (int a = 0; a < 9; ++a) { ic.Rc[a] = sRc[cand][a]; ic.Rpi[a] = sRpi[a]; }
        for (int a = 0; a < 3; ++a) { ic.tc[a] = sTc[cand][a]; ic.tp[a] = sTp[a]; }
        ic.fx = L.fx; ic.fy = L.fy; ic.cx = L.cx; ic.cy = L.cy;
        ic.dist = P.distThres; ic.angle = P.angleThres;
        for (int i = lb * 256 + threadIdx.x; i < N; i += kSeedBlocks * 256) {
            Px6 src, dst;
            float vg[3], vcp[3], row[7];
            icp_load(L, i, P3, src);
            const int j = icp_project(L, ic, src, vg, vcp);
            if (j < 0) continue;
            icp_gather(L, j, P3, dst);
            if (icp_row(ic, src, vg, vcp, dst, row)) {
                v[2 * cand] += 1.0f;
                v[2 * cand + 1] += row[6] * row[6];
            }
        }
    }
    for (int q = 0; q < 4; ++q) v[q] = wave_total_f(v[q]);
    if ((threadIdx.x & 63) == 63)
        for (int q = 0; q < 4; ++q) sRed[threadIdx.x >> 6][q] = v[q];
    __syncthreads();
    if (threadIdx.x < 4) {
        float t = 0.0f;
        for (int wv = 0; wv < 4; ++wv) t += sRed[wv][threadIdx.x];
        D.part_b[0][(size_t)lb * kPartStride + threadIdx.x] = t;
    }
}

void launch_seed_select(const PairDesc* pairs, int npairs, const float* pred, float t_scale, float r_scale,
                        int select, const TrackParams& P, hipStream_t s)
{
    hipLaunchKernelGGL(k_seed_eval, dim3(npairs * kSeedBlocks), dim3(256), 0, s, pairs, pred, t_scale, r_scale,
                       select, P);
}

// The keep-or-fall-back rule: the prediction is kept when its level-2
// association has at least 99 % of the plain start's inliers and a mean
// squared residual at most 1 % above it.
__device__ int seed_keep(const PairDesc& D)
{
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = 0; b < kSeedBlocks; ++b)
        for (int q = 0; q < 4; ++q) t[q] += (double)gp(D.part_b[0])[(size_t)b * kPartStride + q];
    const double m0 = t[1] / fmax(t[0], 1.0), m1 = t[3] / fmax(t[2], 1.0);
    return (t[2] >= 0.99 * t[0]) && (m1 <= 1.01 * m0);
}

// State entering the first iteration of a tracking call of pair D.  so3R:
// the SO3 result when P.so3 (else unused).
__device__ void track_begin_pair(const PairDesc& D, const TrackParams& P, const double* so3R)
{
    GLOBAL_AS TrackState* st = gp(D.st);
    // every load first, then the stores
    float pose[12], errs[4];
    double lastA[36], lastb[6], seed[16];
    for (int a = 0; a < 12; ++a) pose[a] = st->pose_out[a];
    for (int a = 0; a < 36; ++a) lastA[a] = st->lastA[a];
    for (int a = 0; a < 6; ++a) lastb[a] = st->lastb[a];
    for (int a = 0; a < 16; ++a) seed[a] = P.seeded ? st->seedRt[a] : 0.0;
    errs[0] = st->lastICPError;
    errs[1] = st->lastICPCount;
    errs[2] = st->lastRGBError;
    errs[3] = st->lastRGBCount;
    const int keep = (P.seeded == 2) ? seed_keep(D) : 1;

    float Rprev[9], tprev[3], Rprev_inv[9];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) Rprev[a * 3 + b] = pose[a * 4 + b];
        tprev[a] = pose[a * 4 + 3];
    }
    inv3f(Rprev, Rprev_inv);
    double resultRt[16];
    float Rcurr[9], tcurr[3];
    for (int a = 0; a < 9; ++a) Rcurr[a] = Rprev[a];
    for (int a = 0; a < 3; ++a) tcurr[a] = tprev[a];
    for (int a = 0; a < 16; ++a) resultRt[a] = (a % 5 == 0) ? 1.0 : 0.0;
    if (P.so3)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) resultRt[a * 4 + b] = so3R[a * 3 + b];
    if (P.seeded) {
        // the warm start (seeded 2: only if the level-2 scores keep it):
        // resultRt = seed, and (unlike the SO3 seed of the reference) the
        // current estimate prev * seed^-1 from the first pass
        for (int a = 0; a < 16; ++a) resultRt[a] = keep ? seed[a] : ((a % 5 == 0) ? 1.0 : 0.0);
        pose_from_result(resultRt, Rprev, tprev, Rcurr, tcurr);
    }

    for (int a = 0; a < 9; ++a) {
        st->Rprev[a] = Rprev[a];
        st->Rprev_inv[a] = Rprev_inv[a];
    }
    for (int a = 0; a < 3; ++a) st->tprev[a] = tprev[a];
    st->alt_iters = 0;
    st->judge_level = -1;
    if (P.seeded == 1 || P.seeded == 2) {
        st->seed_used = keep;
        st->seed_calls = st->seed_calls + 1;
        st->seed_kept = st->seed_kept + keep;
    }
    GLOBAL_AS IterState& G = st->it[1];
    for (int a = 0; a < 16; ++a) G.resultRt[a] = resultRt[a];
    for (int a = 0; a < 9; ++a) G.Rcurr[a] = Rcurr[a];
    for (int a = 0; a < 3; ++a) G.tcurr[a] = tcurr[a];
    for (int a = 0; a < 36; ++a) G.lastA[a] = lastA[a];
    for (int a = 0; a < 6; ++a) G.lastb[a] = lastb[a];
    G.lastICPError = errs[0];
    G.lastICPCount = errs[1];
    G.lastRGBError = errs[2];
    G.lastRGBCount = errs[3];
    G.res0 = 0.0f;
    G.res1 = 0.0f;
    G.skip = 0;
    G.break_level = -1;
    G.pending = 0;
    G.broke = 0;
    G.degenerate = 0;
    G.iters_done = 0;
    G.rgb_err_cmp = FLT_MAX;
}

__global__ __launch_bounds__(64) void k_track_begin(const PairDesc* __restrict__ pairs, int npairs, TrackParams P)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npairs) return;
    double so3R[9];
    for (int a = 0; a < 9; ++a) so3R[a] = P.so3 ? gp(pairs[p].st)->so3R[a] : 0.0;
    track_begin_pair(pairs[p], P, so3R);
}

void launch_track_begin(const PairDesc* pairs, int npairs, const TrackParams& P, hipStream_t s)
{
    hipLaunchKernelGGL(k_track_begin, dim3(cdiv(npairs, 64)), dim3(64), 0, s, pairs, npairs, P);
}

// Finish the last iteration, then RGBDOdometryef.cpp:579-583, 593-594: the
// 0.3 m revert and the result.  One workgroup per pair.
__global__ __launch_bounds__(256) void k_track_end(const PairDesc* __restrict__ pairs, TrackParams P, int K,
                                                   int last_level)
{
    __shared__ Prolog<256> W;
    const PairDesc& D = pairs[blockIdx.x];
    TrackState* st = D.st;
    PrevPose pp;
    load_prev(st, pp);
    // read before the prologue, so the stores below wait on no load
    float* const plog = gp(st)->pose_log;
    const int plog_cap = gp(st)->pose_log_cap;
    const int frame = gp(st)->frame_count;
    finish_previous<256>(D, P, K, last_level, pp, W, nullptr, nullptr, 0, 0);
    if (threadIdx.x != 0) return;
    GLOBAL_AS TrackState* g = gp(st);
    const IterState& S = W.S;
    float R[9], t[3];
    for (int a = 0; a < 9; ++a) R[a] = S.Rcurr[a];
    for (int a = 0; a < 3; ++a) t[a] = S.tcurr[a];
    int reverted = 0;
    if (P.rgb) {
        const float d0 = t[0] - pp.t[0], d1 = t[1] - pp.t[1], d2 = t[2] - pp.t[2];
        if (sqrtf(d0 * d0 + d1 * d1 + d2 * d2) > 0.3f) {
            for (int a = 0; a < 9; ++a) R[a] = pp.R[a];
            for (int a = 0; a < 3; ++a) t[a] = pp.t[a];
            reverted = 1;
        }
    }
    g->reverted = reverted;
    float pose[12];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) pose[a * 4 + b] = R[a * 3 + b];
        pose[a * 4 + 3] = t[a];
    }
    for (int a = 0; a < 12; ++a) g->pose_out[a] = pose[a];
    for (int a = 0; a < 36; ++a) g->lastA[a] = S.lastA[a];
    for (int a = 0; a < 6; ++a) g->lastb[a] = S.lastb[a];
    g->lastICPError = S.lastICPError;
    g->lastICPCount = S.lastICPCount;
    g->lastRGBError = S.lastRGBError;
    g->lastRGBCount = S.lastRGBCount;
    g->iters_done = S.iters_done;
    g->degenerate = S.degenerate;
    for (int a = 0; a < 16; ++a) g->lastRt[a] = reverted ? ((a % 5 == 0) ? 1.0 : 0.0) : S.resultRt[a];
    if (plog && plog_cap > 0) {
        GLOBAL_AS float* dst = gp(plog) + (size_t)(frame % plog_cap) * 12;
        for (int a = 0; a < 12; ++a) dst[a] = pose[a];
    }
    g->frame_count = frame + 1;
}

void launch_track_end(const PairDesc* pairs, int npairs, const TrackParams& P, int K, int last_level, hipStream_t s)
{
    hipLaunchKernelGGL(k_track_end, dim3(npairs), dim3(256), 0, s, pairs, P, K, last_level);
}

// ============================================== starts and the judge
// The start of a policy call, one workgroup per entry (hyp entries per
// handle, entry hyp*i + h).  Entry h = 0 is the handle's own state: it starts
// from the prediction (pred, 6 floats per handle: t / t_scale, rotation
// vector / r_scale) when there is one, else from the constant-velocity
// extrapolation.  Entry h = 1 is the second refinement of a judged call: it
// takes the handle's prior (pose and the statistics a call carries) and
// starts from the constant-velocity extrapolation.  The extrapolation repeats
// the last call's increment (TrackState::lastRt).  The seed is stored in the
// entry's seedRt for k_track_begin (seeded 3).
__device__ void seed_from_motion(const float* q, float t_scale, float r_scale, double* seed)
{
    double w[3], R[9], t[3];
    for (int r = 0; r < 3; ++r) {
        t[r] = (double)q[r] * (double)t_scale;
        w[r] = (double)q[3 + r] * (double)r_scale;
    }
    rodrigues(w, R);
    // M = [R | t] (previous -> current camera); seed = M^-1 = [R^T | -R^T t]
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) seed[r * 4 + c] = R[c * 3 + r];
        seed[r * 4 + 3] = -(R[0 * 3 + r] * t[0] + R[1 * 3 + r] * t[1] + R[2 * 3 + r] * t[2]);
    }
    seed[12] = seed[13] = seed[14] = 0.0;
    seed[15] = 1.0;
}

__global__ __launch_bounds__(64) void k_seed_policy(const PairDesc* __restrict__ pairs, int hyp,
                                                    const float* __restrict__ pred, float t_scale, float r_scale)
{
    if (threadIdx.x != 0) return;
    const int e = blockIdx.x, i = e / hyp, h = e - i * hyp;
    GLOBAL_AS TrackState* st = gp(pairs[e].st);
    const GLOBAL_AS TrackState* own = gp(pairs[e - h].st);   // the handle's state
    double seed[16];
    if (h == 0 && pred) {
        seed_from_motion(pred + 6 * i, t_scale, r_scale, seed);
    } else {
        for (int a = 0; a < 16; ++a) seed[a] = own->lastRt[a];
    }
    if (h > 0) {
        // the handle's prior and carried statistics (RGBDOdometryef.cpp:
        // 278-282: the call starts from the pose and keeps lastA/lastb and
        // the errors of the previous call until it overwrites them)
        float pose[12], errs[4];
        double A[36], b[6], lr[16];
        for (int a = 0; a < 12; ++a) pose[a] = own->pose_out[a];
        for (int a = 0; a < 36; ++a) A[a] = own->lastA[a];
        for (int a = 0; a < 6; ++a) b[a] = own->lastb[a];
        for (int a = 0; a < 16; ++a) lr[a] = own->lastRt[a];
        errs[0] = own->lastICPError;
        errs[1] = own->lastICPCount;
        errs[2] = own->lastRGBError;
        errs[3] = own->lastRGBCount;
        for (int a = 0; a < 12; ++a) st->pose_out[a] = pose[a];
        for (int a = 0; a < 36; ++a) st->lastA[a] = A[a];
        for (int a = 0; a < 6; ++a) st->lastb[a] = b[a];
        for (int a = 0; a < 16; ++a) st->lastRt[a] = lr[a];
        st->lastICPError = errs[0];
        st->lastICPCount = errs[1];
        st->lastRGBError = errs[2];
        st->lastRGBCount = errs[3];
        st->pose_log = nullptr;
        st->pose_log_cap = 0;
    }
    for (int a = 0; a < 16; ++a) st->seedRt[a] = seed[a];
    if (h == 0 && pred && hyp == 1) {
        // an unjudged prediction is always kept
        st->seed_used = 1;
        st->seed_calls = st->seed_calls + 1;
        st->seed_kept = st->seed_kept + 1;
    }
}

void launch_seed_policy(const PairDesc* pairs, int nentries, int hyp, const float* pred, float t_scale, float r_scale,
                        hipStream_t s)
{
    hipLaunchKernelGGL(k_seed_policy, dim3(nentries), dim3(64), 0, s, pairs, hyp, pred, t_scale, r_scale);
}

// True when refinement a (error, count) is worse than refinement b by more
// than the relative margin tol.
__device__ __forceinline__ bool refinement_worse(float err_a, float cnt_a, float err_b, float cnt_b, float tol)
{
    return (err_a > err_b * (1.0f + tol)) || (cnt_a < cnt_b * (1.0f - tol));
}

// The judge of a call with two refinements, one workgroup per handle
// (entries 2i: the prediction's, 2i+1: the constant-velocity start's).  The
// prediction's refinement is kept unless its ICP residual or inlier count
// (the photometric ones for RGB-only calls) is worse than the other's.
// mid (slot >= 0): after the coarse levels, from the statistics of the last
// finished iteration in state slot `slot`; on a fall-back the other state of
// that slot replaces the handle's, and the finer levels continue from it.
// final (slot < 0): after the call; on a fall-back the other call's results
// replace the handle's (pose, system, statistics, the logged pose).
__global__ __launch_bounds__(256) void k_judge(const PairDesc* __restrict__ pairs, TrackParams P, int slot,
                                               int level)
{
    __shared__ int s_fall;
    GLOBAL_AS TrackState* own = gp(pairs[2 * blockIdx.x].st);
    GLOBAL_AS TrackState* alt = gp(pairs[2 * blockIdx.x + 1].st);
    if (threadIdx.x == 0) {
        float ea, ca, eb, cb;
        if (slot >= 0) {
            const GLOBAL_AS IterState& A = own->it[slot];
            const GLOBAL_AS IterState& B = alt->it[slot];
            ea = P.icp ? A.lastICPError : A.lastRGBError;
            ca = P.icp ? A.lastICPCount : A.lastRGBCount;
            eb = P.icp ? B.lastICPError : B.lastRGBError;
            cb = P.icp ? B.lastICPCount : B.lastRGBCount;
        } else {
            ea = P.icp ? own->lastICPError : own->lastRGBError;
            ca = P.icp ? own->lastICPCount : own->lastRGBCount;
            eb = P.icp ? alt->lastICPError : alt->lastRGBError;
            cb = P.icp ? alt->lastICPCount : alt->lastRGBCount;
        }
        const int fall = refinement_worse(ea, ca, eb, cb, P.judgeTol) ? 1 : 0;
        s_fall = fall;
        const int iters_own = (slot >= 0) ? own->it[slot].iters_done : own->iters_done;
        const int iters_alt = (slot >= 0) ? alt->it[slot].iters_done : alt->iters_done;
        own->alt_iters = fall ? iters_own : iters_alt;
        own->seed_used = fall ? 0 : 1;
        own->seed_calls = own->seed_calls + 1;
        own->seed_kept = own->seed_kept + (fall ? 0 : 1);
        own->judge_level = level;
    }
    __syncthreads();
    if (!s_fall) return;
    if (slot >= 0) {
        const GLOBAL_AS unsigned int* src = (const GLOBAL_AS unsigned int*)&alt->it[slot];
        GLOBAL_AS unsigned int* dst = (GLOBAL_AS unsigned int*)&own->it[slot];
        for (int i = threadIdx.x; i < (int)(sizeof(IterState) / 4); i += 256) dst[i] = src[i];
        return;
    }
    if (threadIdx.x != 0) return;
    float pose[12];
    for (int a = 0; a < 12; ++a) pose[a] = alt->pose_out[a];
    for (int a = 0; a < 12; ++a) own->pose_out[a] = pose[a];
    for (int a = 0; a < 36; ++a) own->lastA[a] = alt->lastA[a];
    for (int a = 0; a < 6; ++a) own->lastb[a] = alt->lastb[a];
    for (int a = 0; a < 16; ++a) own->lastRt[a] = alt->lastRt[a];
    own->lastICPError = alt->lastICPError;
    own->lastICPCount = alt->lastICPCount;
    own->lastRGBError = alt->lastRGBError;
    own->lastRGBCount = alt->lastRGBCount;
    own->iters_done = alt->iters_done;
    own->degenerate = alt->degenerate;
    own->reverted = alt->reverted;
    float* const plog = own->pose_log;
    const int cap = own->pose_log_cap;
    if (plog && cap > 0) {
        GLOBAL_AS float* d = gp(plog) + (size_t)((own->frame_count - 1) % cap) * 12;
        for (int a = 0; a < 12; ++a) d[a] = pose[a];
    }
}

void launch_judge(const PairDesc* pairs, int nhandles, const TrackParams& P, int slot, int level, hipStream_t s)
{
    hipLaunchKernelGGL(k_judge, dim3(nhandles), dim3(256), 0, s, pairs, P, slot, level);
}

// ================================================================= SO3
// a17 SO3Reduction (reduce.cu:937-1075) per level-2 pixel, split into the
// part that does not depend on the rotation (the last image's gradient and
// value at the pixel, loaded before a launch's prologue), the warp, and the
// part at the warped pixel of the next image.
struct So3Src {
    float g[2], v;
};

template <class Ptr>
__device__ __forceinline__ void so3_grad(Ptr img, int w, int x, int y, float* g)
{
    const float actu = (float)img[(size_t)y * w + x];
    float back = (float)img[(size_t)y * w + x - 1];
    float fore = (float)img[(size_t)y * w + x + 1];
    g[0] = ((back + actu) / 2.0f) - ((fore + actu) / 2.0f);
    back = (float)img[(size_t)(y - 1) * w + x];
    fore = (float)img[(size_t)(y + 1) * w + x];
    g[1] = ((back + actu) / 2.0f) - ((fore + actu) / 2.0f);
}

__device__ __forceinline__ bool so3_interior(int w, int h, int x, int y)
{
    return x >= 1 && x < w - 1 && y >= 1 && y < h - 1;
}

template <class Ptr>
__device__ __forceinline__ void so3_src(Ptr last, int w, int h, int x, int y, So3Src& s)
{
    s.g[0] = s.g[1] = s.v = 0.0f;
    if (so3_interior(w, h, x, y)) {
        so3_grad(last, w, x, y, s.g);
        s.v = (float)last[(size_t)y * w + x];
    }
}

// The warped pixel (wx, wy) of (x, y); false where the reference skips it.
__device__ __forceinline__ bool so3_warp(const float* H, int w, int h, int x, int y, int& wx, int& wy)
{
    const float xf = (float)x, yf = (float)y;
    const float wp0 = H[0] * xf + H[1] * yf + H[2] * 1.0f;
    const float wp1 = H[3] * xf + H[4] * yf + H[5] * 1.0f;
    const float wp2 = H[6] * xf + H[7] * yf + H[8] * 1.0f;
    wx = f2i_rn(wp0 / wp2);
    wy = f2i_rn(wp1 / wp2);
    return wx >= 1 && wx < w - 1 && wy >= 1 && wy < h - 1 && so3_interior(w, h, x, y);
}

template <class Ptr>
__device__ __forceinline__ void so3_row(Ptr next, int w, int wx, int wy, const So3Src& src, int x, int y,
                                        const float* Ki, const float* KR, float* row)
{
    const float xf = (float)x, yf = (float)y;
    float gn[2];
    so3_grad(next, w, wx, wy, gn);
    const float gx = (gn[0] + src.g[0]) / 2.0f;
    const float gy = (gn[1] + src.g[1]) / 2.0f;
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
    row[3] = -((float)next[(size_t)wy * w + wx] - src.v);
}

// One pixel in full (the parity-check kernel).
template <class Ptr>
__device__ __forceinline__ bool so3_pixel(Ptr last, Ptr next, int w, int h, int x, int y, const float* H,
                                          const float* Ki, const float* KR, float* row)
{
    int wx, wy;
    if (!so3_warp(H, w, h, x, y, wx, wy)) return false;
    So3Src src;
    so3_src(last, w, h, x, y, src);
    so3_row(next, w, wx, wy, src, x, y, Ki, KR, row);
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
    const int lane = threadIdx.x & 63;
    const float t = wave_scatter<kSO3>(acc);
    if ((lane & 1) == 0 && (lane >> 1) < kSO3) red[threadIdx.x >> 6][lane >> 1] = t;
    __syncthreads();
    if (threadIdx.x < kSO3) {
        double s = 0.0;
        for (int wv = 0; wv < NT / 64; ++wv) s += (double)red[wv][threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
}

constexpr int kSO3Threads = 512;

// The SO3 loop as up to 11 launches of kSO3Blocks workgroups per pair.
// Launch `it` (< 10) first evaluates iteration it-1 from its block partials
// -- FP64 totals in a fixed order, the convergence / divergence tests and the
// 3x3 solve of the loop body, identically in every block -- then, unless the
// loop has ended, sums iteration it over its share of the level-2 pixels.
// Launch 10 only evaluates iteration 9 and stores the result.  Block 0 of a
// pair stores the loop state of each launch (slot it & 1).
constexpr int kSO3Blocks = 16;
// pixels per thread whose rotation-independent part is loaded before the
// prologue (all of a level 2 of up to 3 * 16 * 512 pixels)
constexpr int kSO3Pre = 3;

__device__ void so3_evaluate(So3Slot& S, const double* tot)
{
    ++S.iters;
    const float res0 = (float)tot[9], res1 = (float)tot[10];
    const float err = sqrtf(res0) / res1;
    S.err = err;
    S.cnt = res1;
    if (err < S.lastError && S.lastCount == res1) {
        S.done = 1;
    } else if (err > S.lastError + 0.001f) {
        S.err = S.lastError;
        S.cnt = S.lastCount;
        for (int k = 0; k < 9; ++k) S.resultR[k] = S.lastResultR[k];
        S.done = 1;
    }
    if (S.done) return;
    S.lastError = err;
    S.lastCount = res1;
    for (int k = 0; k < 9; ++k) S.lastResultR[k] = S.resultR[k];
    double A3[9], b3[3], delta[3], rot[9];
    unpack_so3(tot, A3, b3);
    ldlt_solve<3>(A3, b3, delta);
    rodrigues(delta, rot);
    float rotf[9], nR[9];
    for (int k = 0; k < 9; ++k) rotf[k] = (float)rot[k];
    mm3f(rotf, S.R_lr, nR);
    for (int k = 0; k < 9; ++k) {
        S.R_lr[k] = nR[k];
        S.resultR[k] = (double)nR[k];
    }
}

// With `begin`, block 0 of each pair of launch 10 also sets up the
// Gauss-Newton loop from the SO3 result (the work of k_track_begin).
__global__ __launch_bounds__(kSO3Threads) void k_so3_iter(const PairDesc* __restrict__ pairs, int it, int begin,
                                                          TrackParams P)
{
    __shared__ So3Slot S;
    __shared__ double s_tot[kSO3];
    __shared__ float s_red[kSO3Threads / 64][kSO3];
    __shared__ float sH[9], sKi[9], sKR[9];
    const int p = blockIdx.x / kSO3Blocks, lb = blockIdx.x - p * kSO3Blocks;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    // once an earlier launch has ended the loop, the launches before the last
    // have nothing to do (the state is in so3_final)
    const int ended = (it > 0) ? gp(st)->so3_done : 0;
    if (ended && it < 10) return;
    const LevelDesc& L = D.lv[2];
    const int w = L.w, h = L.h, n = w * h;
    constexpr int stride = kSO3Blocks * kSO3Threads;
    const int first = lb * kSO3Threads + threadIdx.x;
    const GLOBAL_AS uint8_t* last = gp(D.so3_last);
    const GLOBAL_AS uint8_t* next = gp(D.so3_next);

    // the rotation-independent loads of this thread's pixels overlap the prologue
    So3Src src[kSO3Pre];
    if (it < 10) {
#pragma unroll
        for (int j = 0; j < kSO3Pre; ++j) {
            const int idx = first + j * stride;
            if (idx < n) {
                const int y = idx / w, x = idx - y * w;
                so3_src(last, w, h, x, y, src[j]);
            }
        }
    }
    if (it == 0) {
        if (threadIdx.x == 0) {
            for (int k = 0; k < 9; ++k) {
                S.resultR[k] = S.lastResultR[k] = (k % 4 == 0) ? 1.0 : 0.0;
                S.R_lr[k] = (k % 4 == 0) ? 1.0f : 0.0f;
            }
            S.lastError = S.lastCount = FLT_MAX / 2;
            S.err = S.cnt = 0.0f;
            S.iters = 0;
            S.done = 0;
        }
        __syncthreads();
    } else if (ended) {
        copy_in<kSO3Threads>(&S, &st->so3_final);
        __syncthreads();
    } else {
        // issue the partial loads with the state copy (one memory round trip)
        const GLOBAL_AS float* part = gp(D.part_a[(it - 1) & 1]);
        float pv[kSO3Blocks] = {};
        const int col = threadIdx.x;
        if (col < kSO3) {
#pragma unroll
            for (int r = 0; r < kSO3Blocks; ++r) pv[r] = part[r * kPartStride + col];
        }
        copy_in<kSO3Threads>(&S, &st->so3s[(it - 1) & 1]);
        __syncthreads();
        if (!S.done) {
            if (col < kSO3) {
                double t = 0.0;
#pragma unroll
                for (int r = 0; r < kSO3Blocks; ++r) t += (double)pv[r];
                s_tot[col] = t;
            }
            __syncthreads();
            if (threadIdx.x == 0) so3_evaluate(S, s_tot);
            __syncthreads();
        }
    }
    if (lb == 0) {
        GLOBAL_AS TrackState* g = gp(st);
        if (it < 10) copy_out<kSO3Threads>(&st->so3s[it & 1], &S);
        if (it == 0 && threadIdx.x == 0) g->so3_done = 0;
        if (it > 0 && it < 10 && S.done) {
            copy_out<kSO3Threads>(&st->so3_final, &S);
            if (threadIdx.x == 0) g->so3_done = 1;
        }
        if (it == 10 && threadIdx.x == 0) {
            for (int k = 0; k < 9; ++k) g->so3R[k] = S.resultR[k];
            g->so3_iters = S.iters;
            g->lastSO3Error = S.err;
            g->lastSO3Count = S.cnt;
            if (begin) track_begin_pair(D, P, S.resultR);
        }
    }
    if (S.done || it == 10) return;
    if (threadIdx.x == 0) {
        double KR[9], Hd[9];
        mm3d(L.K, S.resultR, KR);
        mm3d(KR, L.Ki, Hd);
        for (int k = 0; k < 9; ++k) {
            sH[k] = (float)Hd[k];
            sKi[k] = (float)L.Ki[k];
            sKR[k] = (float)KR[k];
        }
    }
    __syncthreads();
    float Hr[9], Kir[9], KRr[9];
    for (int k = 0; k < 9; ++k) {
        Hr[k] = uni(sH[k]);
        Kir[k] = uni(sKi[k]);
        KRr[k] = uni(sKR[k]);
    }
    float acc[kSO3];
#pragma unroll
    for (int k = 0; k < kSO3; ++k) acc[k] = 0.0f;
    // the preloaded pixels: all warps, then all next-image loads, then the
    // rows in pixel order
    int wx[kSO3Pre], wy[kSO3Pre];
    bool ok[kSO3Pre];
#pragma unroll
    for (int j = 0; j < kSO3Pre; ++j) {
        const int idx = first + j * stride;
        const int y = idx / w, x = idx - y * w;
        ok[j] = idx < n && so3_warp(Hr, w, h, x, y, wx[j], wy[j]);
    }
#pragma unroll
    for (int j = 0; j < kSO3Pre; ++j)
        if (ok[j]) {
            const int idx = first + j * stride;
            const int y = idx / w, x = idx - y * w;
            float row[4];
            so3_row(next, w, wx[j], wy[j], src[j], x, y, Kir, KRr, row);
            acc_row4(acc, row);
        }
    for (int idx = first + kSO3Pre * stride; idx < n; idx += stride) {
        const int y = idx / w, x = idx - y * w;
        float row[4];
        if (so3_pixel(last, next, w, h, x, y, Hr, Kir, KRr, row)) acc_row4(acc, row);
    }
    block_partial<kSO3, kSO3Threads>(acc, D.part_a[it & 1] + (size_t)lb * kPartStride, s_red);
}

// The SO3 loop in one launch: the kSO3Blocks workgroups of a pair run the
// iterations of k_so3_iter back to back and meet at a grid barrier after each
// iteration's block partials (an agent-scope release add on the pair's arrival
// counter, then acquire loads until every workgroup has arrived).  Each
// workgroup then sums the partials and evaluates the iteration itself, as
// k_so3_iter's prologue does, so the state, the sums and their order are
// those of the launch-per-iteration form.  Every wait is bounded by
// spin_limit polls; a wait that gives up sets the sticky so3_fail flag and
// its workgroup leaves the loop, so every workgroup exits.  The last
// workgroup to exit sets the counters back to 0 for the next call.  Only
// launched when the pairs' workgroups fit on the device at once.  Blocks from
// nloop on compute the derivative images and photometric masks of one pair
// (k_rgb_prep's work, `prep`, segments `pseg` of kSO3Threads pixels), which
// the loop does not read; they come after the loop's blocks in dispatch order
// and take no part in its barrier.
__global__ __launch_bounds__(kSO3Threads) void k_so3_loop(const PairDesc* __restrict__ pairs, int begin,
                                                          TrackParams P, int spin_limit, int nloop,
                                                          PrepLevels prep, Seg3 pseg)
{
    if ((int)blockIdx.x >= nloop) {
        const int b = (int)blockIdx.x - nloop, l = seg_level(pseg, b);
        rgb_prep_px(prep, l, (b - pseg.b0[l]) * kSO3Threads + threadIdx.x);
        return;
    }
    __shared__ So3Slot S;
    __shared__ double s_tot[kSO3];
    __shared__ float s_red[kSO3Threads / 64][kSO3];
    __shared__ float sH[9], sKi[9], sKR[9];
    __shared__ int s_fail;
    const int p = blockIdx.x / kSO3Blocks, lb = blockIdx.x - p * kSO3Blocks;
    const PairDesc& D = pairs[p];
    TrackState* st = D.st;
    const LevelDesc& L = D.lv[2];
    const int w = L.w, h = L.h, n = w * h;
    constexpr int stride = kSO3Blocks * kSO3Threads;
    const int first = lb * kSO3Threads + threadIdx.x;
    const GLOBAL_AS uint8_t* last = gp(D.so3_last);
    const GLOBAL_AS uint8_t* next = gp(D.so3_next);
    unsigned int* arrive = &st->so3_arrive;

    // the rotation-independent part of this thread's first pixels, once
    So3Src src[kSO3Pre];
#pragma unroll
    for (int j = 0; j < kSO3Pre; ++j) {
        const int idx = first + j * stride;
        if (idx < n) {
            const int y = idx / w, x = idx - y * w;
            so3_src(last, w, h, x, y, src[j]);
        }
    }
    if (threadIdx.x == 0) {
        for (int k = 0; k < 9; ++k) {
            S.resultR[k] = S.lastResultR[k] = (k % 4 == 0) ? 1.0 : 0.0;
            S.R_lr[k] = (k % 4 == 0) ? 1.0f : 0.0f;
        }
        S.lastError = S.lastCount = FLT_MAX / 2;
        S.err = S.cnt = 0.0f;
        S.iters = 0;
        S.done = 0;
        s_fail = 0;
    }
    __syncthreads();
    for (int it = 0; it < 10; ++it) {
        if (threadIdx.x == 0) {
            double KR[9], Hd[9];
            mm3d(L.K, S.resultR, KR);
            mm3d(KR, L.Ki, Hd);
            for (int k = 0; k < 9; ++k) {
                sH[k] = (float)Hd[k];
                sKi[k] = (float)L.Ki[k];
                sKR[k] = (float)KR[k];
            }
        }
        __syncthreads();
        float Hr[9], Kir[9], KRr[9];
        for (int k = 0; k < 9; ++k) {
            Hr[k] = uni(sH[k]);
            Kir[k] = uni(sKi[k]);
            KRr[k] = uni(sKR[k]);
        }
        float acc[kSO3];
#pragma unroll
        for (int k = 0; k < kSO3; ++k) acc[k] = 0.0f;
        int wx[kSO3Pre], wy[kSO3Pre];
        bool ok[kSO3Pre];
#pragma unroll
        for (int j = 0; j < kSO3Pre; ++j) {
            const int idx = first + j * stride;
            const int y = idx / w, x = idx - y * w;
            ok[j] = idx < n && so3_warp(Hr, w, h, x, y, wx[j], wy[j]);
        }
#pragma unroll
        for (int j = 0; j < kSO3Pre; ++j)
            if (ok[j]) {
                const int idx = first + j * stride;
                const int y = idx / w, x = idx - y * w;
                float row[4];
                so3_row(next, w, wx[j], wy[j], src[j], x, y, Kir, KRr, row);
                acc_row4(acc, row);
            }
        for (int idx = first + kSO3Pre * stride; idx < n; idx += stride) {
            const int y = idx / w, x = idx - y * w;
            float row[4];
            if (so3_pixel(last, next, w, h, x, y, Hr, Kir, KRr, row)) acc_row4(acc, row);
        }
        const float* part = D.part_a[it & 1];
        block_partial<kSO3, kSO3Threads>(acc, (float*)part + (size_t)lb * kPartStride, s_red);
        // grid barrier of the pair's workgroups
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned int target = (unsigned int)(it + 1) * kSO3Blocks;
            __hip_atomic_fetch_add(arrive, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int polls = 0;
            // spin_limit 0 gives up at once (the failure path, forced from the host)
            while (spin_limit == 0 ||
                   __hip_atomic_load(arrive, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                if (++polls > spin_limit) {
                    s_fail = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (s_fail) break;
        // the iteration's totals (FP64, rows in order) and its evaluation
        if (threadIdx.x < kSO3) {
            double t = 0.0;
#pragma unroll
            for (int r = 0; r < kSO3Blocks; ++r) t += (double)gp(part)[r * kPartStride + threadIdx.x];
            s_tot[threadIdx.x] = t;
        }
        __syncthreads();
        if (threadIdx.x == 0) so3_evaluate(S, s_tot);
        __syncthreads();
        if (S.done) break;
    }
    if (threadIdx.x == 0) {
        GLOBAL_AS TrackState* g = gp(st);
        if (s_fail) {
            g->so3_fail = 1;
        } else if (lb == 0) {
            for (int k = 0; k < 9; ++k) g->so3R[k] = S.resultR[k];
            g->so3_iters = S.iters;
            g->lastSO3Error = S.err;
            g->lastSO3Count = S.cnt;
            if (begin) track_begin_pair(D, P, S.resultR);
        }
        // the last workgroup out resets the counters
        if (__hip_atomic_fetch_add(&st->so3_exit, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
            kSO3Blocks - 1) {
            __hip_atomic_store(arrive, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&st->so3_exit, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// Workgroups of the one-launch SO3 loop that are sure to be resident at once
// on an otherwise idle device (one per CU of a 256-CU part, each 512 threads
// and ~2 KB of LDS); larger batches take the launch-per-iteration form.
constexpr int kSO3LoopMaxBlocks = 256;

void launch_so3(const PairDesc* pairs, int npairs, const TrackParams* begin, int spin_limit, hipStream_t s,
                const PrepLevels* prep)
{
    const TrackParams P = begin ? *begin : TrackParams{};
    if (spin_limit >= 0 && npairs * kSO3Blocks <= kSO3LoopMaxBlocks) {
        const int nloop = npairs * kSO3Blocks;
        PrepLevels a{};
        Seg3 g{};
        if (prep) {
            a = *prep;
            int n[kLevels];
            for (int l = 0; l < kLevels; ++l) n[l] = a.w[l] * a.h[l];
            g = make_seg(n, kSO3Threads);
        }
        hipLaunchKernelGGL(k_so3_loop, dim3(nloop + (prep ? g.b0[kLevels] : 0)), dim3(kSO3Threads), 0, s, pairs,
                           begin ? 1 : 0, P, spin_limit, nloop, a, g);
        return;
    }
    if (prep) launch_rgb_prep(*prep, s);
    for (int it = 0; it <= 10; ++it)
        hipLaunchKernelGGL(k_so3_iter, dim3(npairs * kSO3Blocks), dim3(kSO3Threads), 0, s, pairs, it,
                           (it == 10 && begin) ? 1 : 0, P);
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
        const bool f = so3_pixel(gp(D.so3_last), gp(D.so3_next), w, h, x, y, H, Ki, KR, row);
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
