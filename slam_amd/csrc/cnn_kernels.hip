// cnn_kernels.hip -- the learned warm start: a small BF16 convnet on MFMA.
//
// Input: 2 x (80x60) depth + intensity, NHWC [N][60][80][4] bf16 in [0, 1]
// (model depth, model intensity, current depth, current intensity).
// Six 3x3 stride-2 convolutions (pad 1) with ReLU, 4 -> 32 -> 64 -> 96 ->
// 128 -> 128 -> 128 channels (60x80 -> 30x40 -> 15x20 -> 8x10 -> 4x5 -> 2x3
// -> 1x2), global average pooling, FC 128 -> 3840 (ReLU) -> 6: a scaled se(3)
// increment (1.0 M parameters).  No counterpart in the reference (its only network is the
// unrelated YOLACT segmenter); the tracker uses the prediction as the initial
// increment of a Gauss-Newton solve.
//
// Every convolution and linear layer runs as one GEMM on MFMA
// (v_mfma_f32_16x16x16_bf16: bf16 operands, fp32 accumulation): forward
// Y = im2col(X) W, weight gradient dW = im2col(X)^T dY (split over the
// reduction and summed in a fixed order), input gradient dX = col2im(dY W^T).
// Master weights and Adam moments are fp32; the GEMMs read bf16 copies.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "slam_cnn.h"

namespace cnn {

typedef uint16_t bf16;
typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__host__ __device__ inline bf16 f2bf(float x)
{
    uint32_t u = __builtin_bit_cast(uint32_t, x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (bf16)(u >> 16);  // inf / nan
    u += 0x7fffu + ((u >> 16) & 1u);                              // round to nearest even
    return (bf16)(u >> 16);
}

__host__ __device__ inline float bf2f(bf16 h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// ============================================================= GEMM
// Eight consecutive bf16 from p[i..i+7] when `row_ok`, of which the first
// `avail` exist (zero-filled beyond); one 16-byte load when all eight do.
// Every operand's leading dimension and base offset are multiples of 8
// elements, so full segments are 16-byte aligned.
__device__ __forceinline__ void load8(const bf16* __restrict__ p, size_t i, bool row_ok, int avail, bf16 (&v)[8])
{
    if (row_ok && avail >= 8) {
        *(uint4*)v = *(const uint4*)(p + i);
        return;
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = (row_ok && q < avail) ? p[i + q] : (bf16)0;
}

// C[m][n] = sum_k A(m, k) B(k, n); A(m, k) = TA ? A[k][m] : A[m][k],
// B(k, n) = TB ? B[n][k] : B[k][n].  64x64 block tile, 32-deep k tiles, four
// waves each owning a 32x32 quarter as 2x2 MFMA 16x16 tiles.  Both operands
// are staged in LDS k-contiguous ([m][k] and [n][k]) so each lane's MFMA
// fragment (4 consecutive k) is one 8-byte read.
//   splits == 1: C = acc (+ bias[n]) (ReLU) stored as fp32 (Cf) or bf16 (Cb);
//   splits  > 1: the k range is cut into gridDim.z chunks of kchunk and chunk z
//                stores its partial sum to Cf + z * M * ldc (no bias).
constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void k_gemm(const bf16* __restrict__ A, int lda, const bf16* __restrict__ B, int ldb,
                                              float* __restrict__ Cf, bf16* __restrict__ Cb, int ldc, int M, int N,
                                              int K, int kchunk, const float* __restrict__ bias, int relu)
{
    __shared__ __align__(16) bf16 As[BM][LDK];
    __shared__ __align__(16) bf16 Bs[BN][LDK];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int wm = w >> 1, wn = w & 1;
    const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
    const int kbeg = blockIdx.z * kchunk;
    const int kend = min(K, kbeg + kchunk);
    v4f acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4f{0.f, 0.f, 0.f, 0.f};

    for (int k0 = kbeg; k0 < kend; k0 += BK) {
        // ---- A tile: rows m0..m0+63, k0..k0+31
        if (!TA) {
            const int r = t >> 2, ks = (t & 3) * 8;
            const int m = m0 + r;
            bf16 v[8];
            load8(A, (size_t)m * lda + k0 + ks, m < M, kend - (k0 + ks), v);
            *(uint4*)&As[r][ks] = *(const uint4*)v;
        } else {
            const int kr = t >> 3, ms = (t & 7) * 8;
            const int k = k0 + kr;
            bf16 v[8];
            load8(A, (size_t)k * lda + m0 + ms, k < kend, M - (m0 + ms), v);
#pragma unroll
            for (int q = 0; q < 8; ++q) As[ms + q][kr] = v[q];
        }
        // ---- B tile: k0..k0+31, cols n0..n0+63
        if (!TB) {
            const int kr = t >> 3, ns = (t & 7) * 8;
            const int k = k0 + kr;
            bf16 v[8];
            load8(B, (size_t)k * ldb + n0 + ns, k < kend, N - (n0 + ns), v);
#pragma unroll
            for (int q = 0; q < 8; ++q) Bs[ns + q][kr] = v[q];
        } else {
            const int r = t >> 2, ks = (t & 3) * 8;
            const int n = n0 + r;
            bf16 v[8];
            load8(B, (size_t)n * ldb + k0 + ks, n < N, kend - (k0 + ks), v);
            *(uint4*)&Bs[r][ks] = *(const uint4*)v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 16) {
            v4s a[2], b[2];
            const int kq = kk + 4 * (lane >> 4);
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = *(const v4s*)&As[wm * 32 + i * 16 + (lane & 15)][kq];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = *(const v4s*)&Bs[wn * 32 + j * 16 + (lane & 15)][kq];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        __syncthreads();
    }
    // ---- epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each 16x16 tile
    const bool split = gridDim.z > 1;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int n = n0 + wn * 32 + j * 16 + (lane & 15);
                if (m >= M || n >= N) continue;
                float v = acc[i][j][r];
                if (split) {
                    Cf[((size_t)blockIdx.z * M + m) * ldc + n] = v;
                    continue;
                }
                if (bias) v += bias[n];
                if (relu) v = fmaxf(v, 0.0f);
                if (Cb) Cb[(size_t)m * ldc + n] = f2bf(v);
                else Cf[(size_t)m * ldc + n] = v;
            }
}

// Sum of split partials in split order: out[i] = sum_z part[z * n + i].
__global__ __launch_bounds__(256) void k_split_sum(const float* __restrict__ part, int splits, size_t n,
                                                   float* __restrict__ out)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(size_t)z * n + i];
    out[i] = s;
}

// The same with the GEMM epilogue: + bias[col], ReLU, stored as fp32 or bf16.
__global__ __launch_bounds__(256) void k_split_sum_epi(const float* __restrict__ part, int splits, int M, int N,
                                                       int ldc, const float* __restrict__ bias, int relu,
                                                       float* __restrict__ Cf, bf16* __restrict__ Cb)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= M * N) return;
    const int m = i / N, n = i - m * N;
    const size_t o = (size_t)m * ldc + n;
    float s = 0.0f;
    for (int z = 0; z < splits; ++z) s += part[(size_t)z * M * ldc + o];
    if (bias) s += bias[n];
    if (relu) s = fmaxf(s, 0.0f);
    if (Cb) Cb[o] = f2bf(s);
    else Cf[o] = s;
}

// ============================================================= conv helpers
// im2col of a 3x3 / stride 2 / pad 1 convolution on NHWC bf16:
// Xcol[m][k], m = (n, ho, wo), k = (kh * 3 + kw) * C + c, zero outside the
// image and for k >= 9 C (padding of the reduction to Kp).
__global__ __launch_bounds__(256) void k_im2col(const bf16* __restrict__ X, int Nb, int H, int W, int C, int Ho,
                                                int Wo, int Kp, bf16* __restrict__ Xcol)
{
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)Nb * Ho * Wo * Kp;
    if (idx >= total) return;
    const int k = (int)(idx % Kp);
    const size_t m = idx / Kp;
    const int wo = (int)(m % Wo);
    const int ho = (int)((m / Wo) % Ho);
    const int n = (int)(m / ((size_t)Wo * Ho));
    bf16 v = 0;
    if (k < 9 * C) {
        const int c = k % C, kw = (k / C) % 3, kh = k / (3 * C);
        const int hi = 2 * ho - 1 + kh, wi = 2 * wo - 1 + kw;
        if (hi >= 0 && hi < H && wi >= 0 && wi < W) v = X[(((size_t)n * H + hi) * W + wi) * C + c];
    }
    Xcol[idx] = v;
}

// col2im (the input gradient), as a gather: dX[n][h][w][c] = sum over the
// taps (kh, kw) with h = 2 ho - 1 + kh, w = 2 wo - 1 + kw in range of
// dXcol[(n, ho, wo)][(kh * 3 + kw) * C + c], in tap order.
__global__ __launch_bounds__(256) void k_col2im(const float* __restrict__ dXcol, int Nb, int H, int W, int C, int Ho,
                                                int Wo, int Kp, float* __restrict__ dX)
{
    const size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)Nb * H * W * C;
    if (idx >= total) return;
    const int c = (int)(idx % C);
    const int w = (int)((idx / C) % W);
    const int h = (int)((idx / ((size_t)C * W)) % H);
    const int n = (int)(idx / ((size_t)C * W * H));
    float s = 0.0f;
    for (int kh = 0; kh < 3; ++kh) {
        const int hh = h + 1 - kh;
        if (hh < 0 || (hh & 1) || (hh >> 1) >= Ho) continue;
        for (int kw = 0; kw < 3; ++kw) {
            const int ww = w + 1 - kw;
            if (ww < 0 || (ww & 1) || (ww >> 1) >= Wo) continue;
            const size_t m = ((size_t)n * Ho + (hh >> 1)) * Wo + (ww >> 1);
            s += dXcol[m * Kp + (kh * 3 + kw) * C + c];
        }
    }
    dX[idx] = s;
}

// dZ = dY where the layer's output was positive (ReLU), else 0; bf16.
__global__ __launch_bounds__(256) void k_relu_grad(const float* __restrict__ dY, const bf16* __restrict__ Y, size_t n,
                                                   bf16* __restrict__ dZ, float* __restrict__ dZf)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float g = (bf2f(Y[i]) > 0.0f) ? dY[i] : 0.0f;
    dZ[i] = f2bf(g);
    if (dZf) dZf[i] = g;
}

// Column sums of a [M][N] fp32 matrix (bias gradients) in two fixed-order
// stages: chunk partials, then their sum.
__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ G, int M, int N, int rows_per,
                                                     float* __restrict__ part)
{
    const int n = blockIdx.y * 256 + threadIdx.x;
    if (n >= N) return;
    const int r0 = blockIdx.x * rows_per, r1 = min(M, r0 + rows_per);
    float s = 0.0f;
    for (int r = r0; r < r1; ++r) s += G[(size_t)r * N + n];
    part[(size_t)blockIdx.x * N + n] = s;
}

// Global average pooling over P positions: G[n][c] = mean_p Y[n][p][c].
__global__ __launch_bounds__(256) void k_gap_fwd(const bf16* __restrict__ Y, int Nb, int P, int C, bf16* __restrict__ G)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Nb * C) return;
    const int n = i / C, c = i - n * C;
    float s = 0.0f;
    for (int p = 0; p < P; ++p) s += bf2f(Y[((size_t)n * P + p) * C + c]);
    G[i] = f2bf(s / (float)P);
}

__global__ __launch_bounds__(256) void k_gap_bwd(const float* __restrict__ dG, int Nb, int P, int C,
                                                 float* __restrict__ dY)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Nb * P * C) return;
    const int c = i % C, n = i / (P * C);
    dY[i] = dG[n * C + c] / (float)P;
}

// Weighted squared error over the first D of each row of O (stride ldo):
// loss_n = sum_j w_j (O - T)^2; dO = 2 w_j (O - T) / N as bf16 (and fp32).
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ O, int ldo, const float* __restrict__ T,
                                              const float* __restrict__ wts, int Nb, int D, float* __restrict__ loss_n,
                                              bf16* __restrict__ dO, float* __restrict__ dOf)
{
    const int n = blockIdx.x * 256 + threadIdx.x;
    if (n >= Nb) return;
    float l = 0.0f;
    for (int j = 0; j < ldo; ++j) {
        float g = 0.0f;
        if (j < D) {
            const float e = O[n * ldo + j] - T[n * D + j];
            l += wts[j] * e * e;
            g = 2.0f * wts[j] * e / (float)Nb;
        }
        dO[n * ldo + j] = f2bf(g);
        dOf[n * ldo + j] = g;
    }
    loss_n[n] = l;
}

// Adam with decoupled weight decay on fp32 master weights; writes the bf16
// copy the GEMMs read.  Entries with mask 0 (reduction padding) stay 0.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v, bf16* __restrict__ pb,
                                              const uint8_t* __restrict__ mask, size_t n, float lr, float b1, float b2,
                                              float eps, float bc1, float bc2, float wd)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[i]) {
        p[i] = 0.0f;
        pb[i] = 0;
        return;
    }
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float upd = (mi / bc1) / (sqrtf(vi / bc2) + eps);
    const float pi = p[i] * (1.0f - lr * wd) - lr * upd;
    p[i] = pi;
    pb[i] = f2bf(pi);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ a, size_t n, bf16* __restrict__ b)
{
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = f2bf(a[i]);
}

// Network input from frame data at twice the input resolution (2W x 2H):
// each channel is the mean of the valid values of a 2x2 block (depth valid
// when positive and finite), depth in metres / 5 clamped to [0, 1], intensity
// / 255.  Channels: model depth (float metres), model intensity, current
// depth (uint16 millimetres), current intensity.
__global__ __launch_bounds__(256) void k_prepare(const float* __restrict__ dA, const uint8_t* __restrict__ iA,
                                                 const uint16_t* __restrict__ dB, const uint8_t* __restrict__ iB,
                                                 int Nb, int W, int H, size_t frame_stride, bf16* __restrict__ X)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= Nb * W * H) return;
    const int n = i / (W * H), p = i - n * (W * H);
    const int y = p / W, x = p - y * W;
    const int W2 = 2 * W;
    const size_t base = (size_t)n * frame_stride;
    float sa = 0.f, sb = 0.f, ia = 0.f, ib = 0.f;
    int ca = 0, cb = 0;
    for (int dy = 0; dy < 2; ++dy)
        for (int dx = 0; dx < 2; ++dx) {
            const size_t q = base + (size_t)(2 * y + dy) * W2 + (2 * x + dx);
            const float za = dA[q];
            if (za > 0.0f && za < 1e30f) { sa += za; ++ca; }
            const float zb = (float)dB[q] * 0.001f;
            if (zb > 0.0f) { sb += zb; ++cb; }
            ia += (float)iA[q];
            ib += (float)iB[q];
        }
    float v[4];
    v[0] = ca ? fminf(sa / (float)ca / 5.0f, 1.0f) : 0.0f;
    v[1] = ia / (4.0f * 255.0f);
    v[2] = cb ? fminf(sb / (float)cb / 5.0f, 1.0f) : 0.0f;
    v[3] = ib / (4.0f * 255.0f);
    for (int c = 0; c < 4; ++c) X[(size_t)i * 4 + c] = f2bf(v[c]);
}

}  // namespace cnn

namespace slam {
void cnn_launch_prepare(const float* depth_a, const uint8_t* image_a, const uint16_t* depth_b, const uint8_t* image_b,
                        int n, size_t frame_stride, void* x, hipStream_t s)
{
    hipLaunchKernelGGL(cnn::k_prepare, dim3((unsigned)((n * 80 * 60 + 255) / 256)), dim3(256), 0, s, depth_a, image_a,
                       depth_b, image_b, n, 80, 60, frame_stride, (cnn::bf16*)x);
}
}  // namespace slam

namespace cnn {

// ============================================================= network
struct ConvSpec {
    int cin, cout, hi, wi, ho, wo;
};
constexpr int kConvs = 6;
const ConvSpec kSpec[kConvs] = {{4, 32, 60, 80, 30, 40},  {32, 64, 30, 40, 15, 20},  {64, 96, 15, 20, 8, 10},
                                {96, 128, 8, 10, 4, 5},   {128, 128, 4, 5, 2, 3},  {128, 128, 2, 3, 1, 2}};
constexpr int kFc1In = 128, kFc1Out = 3840, kOut = 6, kOutPad = 8;

struct Param {
    size_t off, n;
};

}  // namespace cnn

using namespace cnn;

struct cnn_net_s {
    int max_batch = 0;
    hipStream_t stream = nullptr;    // null: the legacy default stream
    // parameters (flat): per conv W [Kp][Cout] + b [Cout]; fc1 W [128][3840] + b; fc2 W [3840][8] + b [8]
    Param convW[kConvs], convB[kConvs], fc1W, fc1B, fc2W, fc2B;
    int Kp[kConvs];
    size_t nparam = 0;
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16* pb = nullptr;
    // the conv weights in the inference kernel's fragment order (k_pack_infer),
    // repacked when the weights have changed since the last pack
    bf16* pw = nullptr;
    size_t pw_off[kConvs] = {};
    unsigned long long wver = 1, pver = 0;
    // the split inference head: pooled features, per-workgroup FC2 partials
    // and the per-sample arrival tickets (back to 0 after every launch)
    int infer_split = 1;
    float *infer_gap = nullptr, *infer_part = nullptr;
    unsigned int* infer_ticket = nullptr;
    uint8_t* mask = nullptr;
    int step = 0;
    // activations of the last forward (for backward), sized for max_batch
    bf16* xcol[kConvs] = {};
    bf16* y[kConvs] = {};
    bf16 *gap = nullptr, *h1 = nullptr;
    float* out = nullptr;            // [N][8]
    // backward scratch
    float *dy = nullptr, *dx = nullptr, *dcol = nullptr, *split = nullptr, *colpart = nullptr;
    bf16 *dz = nullptr, *dob = nullptr;
    float *dof = nullptr, *lossn = nullptr, *wts = nullptr;
    size_t split_elems = 0, colpart_elems = 0;
    std::vector<void*> allocs;
};

namespace {

int fail(hipError_t e, const char* what)
{
    fprintf(stderr, "slam_cnn: %s failed: %s\n", what, hipGetErrorString(e));
    return CNN_ERR_HIP;
}

#define CK(x)                                \
    do {                                     \
        hipError_t e_ = (x);                 \
        if (e_ != hipSuccess) return fail(e_, #x); \
    } while (0)

template <class T>
int dalloc(cnn_net_s* c, T** p, size_t n)
{
    CK(hipMalloc((void**)p, n * sizeof(T) + 16));
    c->allocs.push_back(*p);
    return CNN_OK;
}

int launch_check(const char* what)
{
    hipError_t e = hipPeekAtLastError();
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(e, what);
    }
    return CNN_OK;
}

inline dim3 g1(size_t n) { return dim3((unsigned)((n + 255) / 256)); }

// C = op(A) op(B) with the conventions of k_gemm; splits > 1 reduces through
// c->split (fp32 [splits][M][ldc]) into Cf.
int gemm(cnn_net_s* c, bool ta, const bf16* A, int lda, bool tb, const bf16* B, int ldb, float* Cf, bf16* Cb,
         int ldc, int M, int N, int K, const float* bias, bool relu, int splits)
{
    if (splits < 1) splits = 1;
    const int kchunk = ((cdiv(K, splits) + BK - 1) / BK) * BK;
    splits = cdiv(K, kchunk);
    if (splits > 1 && (size_t)splits * M * ldc > c->split_elems) return CNN_ERR_ARG;
    dim3 grid(cdiv(M, BM), cdiv(N, BN), splits);
    float* out = splits > 1 ? c->split : Cf;
    bf16* outb = splits > 1 ? nullptr : Cb;
    const float* bb = splits > 1 ? nullptr : bias;
    const int rl = (splits > 1) ? 0 : (relu ? 1 : 0);
    if (!ta && !tb)
        hipLaunchKernelGGL((k_gemm<false, false>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, out, outb, ldc, M, N, K,
                           kchunk, bb, rl);
    else if (!ta && tb)
        hipLaunchKernelGGL((k_gemm<false, true>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, out, outb, ldc, M, N, K,
                           kchunk, bb, rl);
    else if (ta && !tb)
        hipLaunchKernelGGL((k_gemm<true, false>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, out, outb, ldc, M, N, K,
                           kchunk, bb, rl);
    else
        hipLaunchKernelGGL((k_gemm<true, true>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, out, outb, ldc, M, N, K,
                           kchunk, bb, rl);
    if (splits > 1) {
        if (!bias && !relu && Cf) {
            const size_t n = (size_t)M * ldc;
            hipLaunchKernelGGL(k_split_sum, g1(n), dim3(256), 0, c->stream, c->split, splits, n, Cf);
        } else {
            hipLaunchKernelGGL(k_split_sum_epi, g1((size_t)M * N), dim3(256), 0, c->stream, c->split, splits, M, N,
                               ldc, bias, relu ? 1 : 0, Cf, Cb);
        }
    }
    return launch_check("gemm");
}

int colsum(cnn_net_s* c, const float* G, int M, int N, float* out)
{
    const int chunks = std::min(256, std::max(1, M / 64));
    const int rows_per = cdiv(M, chunks);
    if ((size_t)chunks * N > c->colpart_elems) return CNN_ERR_ARG;
    hipLaunchKernelGGL(k_colsum_part, dim3(chunks, cdiv(N, 256)), dim3(256), 0, c->stream, G, M, N, rows_per,
                       c->colpart);
    hipLaunchKernelGGL(k_split_sum, g1(N), dim3(256), 0, c->stream, c->colpart, chunks, (size_t)N, out);
    return launch_check("colsum");
}

// Forward GEMMs with few output tiles (small batches) split the reduction so
// more than a handful of workgroups share it: at most K / 128 chunks, aiming
// at 64 workgroups.
int fwd_splits(int M, int N, int K)
{
    const int tiles = cdiv(M, BM) * cdiv(N, BN);
    if (tiles >= 64) return 1;
    return std::max(1, std::min(64 / tiles, K / 128));
}

// The split count of a forward GEMM of N samples with `rows` output rows
// each.  For inference-sized batches (N <= 64) it is chosen from one
// sample's rows, so the k-chunks and their summation order -- and with them
// a sample's output, bit for bit -- do not depend on the batch the sample
// runs in; training batches choose from all rows.
int fwd_splits_n(int N, int rows, int cols, int K)
{
    return fwd_splits(N <= 64 ? rows : N * rows, cols, K);
}

int splits_for(int K, int MN)
{
    // enough blocks along the reduction when the output tile count is small
    const int tiles = std::max(1, MN / (BM * BN));
    int s = std::max(1, 512 / tiles);
    s = std::min(s, std::max(1, K / 512));
    return s;
}

}  // namespace

// ============================================================= fused inference
// The whole forward pass of one sample in one workgroup: the activations stay
// in LDS (at most 76.8 KB + 38.4 KB, the first layer's output and input), every
// convolution is an implicit-GEMM on MFMA (the im2col operand is gathered
// from the LDS activations as it is consumed, 4 consecutive channels of one
// input pixel per lane), the weights are read from the bf16 copy, and the
// epilogue (bias, ReLU, bf16) writes the next layer's input.  The rounding
// points are those of cnn_forward (bf16 activations, fp32 accumulation), and a
// sample's result does not depend on the other samples of the launch.  One
// launch for the whole network; no activations are kept.
constexpr int kInferThreads = 1024;
constexpr int kHeadGroups = kInferThreads / 64;   // k_cnn_head's groups per sample: one per head wave
// Activations in LDS are [H][W][C + kActPad] (the input: [H][W][4]): a pixel
// stride of C + 4 elements puts the 16 pixels of a stride-2 fragment load
// 16 bytes apart modulo the 128-byte bank period (2-way), where C alone
// (a multiple of 32 from layer 2 on) would put all of them on one bank.
constexpr int kActPad = 4;
constexpr int kInferBufA = 30 * 40 * (32 + kActPad);   // elements: the largest output (layer 1)
constexpr int kInferBufB = 15 * 20 * (64 + kActPad);   // layer 2's output; also holds the input
constexpr int kInferIn = 60 * 80 * 4;                  // the input sample
static_assert(kInferIn <= kInferBufB, "the input fits its buffer");

struct InferParams {
    const bf16* W[kConvs];                   // fragment order (k_pack_infer) bf16
    const float* b[kConvs];
    int Kp[kConvs];
    const bf16* fc1W;                        // [128][3840]
    const float* fc1b;
    const bf16* fc2W;                        // [256][8]
    const float* fc2b;
};

// The conv weights [Kp][Cout] of every layer in the inference kernel's
// fragment order (grid.y = layer).
struct PackArgs {
    const bf16* src[kConvs];
    bf16* dst[kConvs];
    size_t n[kConvs];
    int nt[kConvs], cout[kConvs];
};

__global__ __launch_bounds__(256) void k_pack_infer(PackArgs A)
{
    const int l = blockIdx.y;
    const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= A.n[l]) return;
    const int q = (int)(e & 3), lane = (int)((e >> 2) & 63);
    const size_t r = e >> 8;
    const int nt = (int)(r % A.nt[l]), ks = (int)(r / A.nt[l]);
    const int k = ks * 16 + 4 * (lane >> 4) + q, n = nt * 16 + (lane & 15);
    A.dst[l][e] = A.src[l][(size_t)k * A.cout[l] + n];
}

// One convolution layer (3x3, stride 2, pad 1) from X (LDS, [HI][WI][CIN])
// into Y (LDS, [HO][WO][COUT]).  Each wave takes (column tile, up to four row
// tiles) work items.  The weight fragments of up to kGroup consecutive
// 16-deep k steps are loaded together before their MFMAs, so an item waits
// for one memory latency per group rather than per step; each fragment serves
// all the item's row tiles.  W is in fragment order: step ks, column tile nt,
// lane L hold W[16 ks + 4 (L >> 4) + q][16 nt + (L & 15)], q = 0..3.
constexpr int kGroup = 24;

// The k steps [k0, k1) of one work item (column tile nt, row tiles from
// ch * CH) into acc.
template <int CIN, int COUT, int HI, int WI, int HO, int WO, int XS, int CH, int GR, int k0, int k1>
__device__ __forceinline__ void infer_conv_item(const bf16* __restrict__ X, const bf16* __restrict__ W, int nt,
                                                int ch, v4f (&acc)[CH])
{
    constexpr int M = HO * WO, MT = (M + 15) / 16, NT = COUT / 16;
    constexpr int KREAL = 9 * CIN;
    const int lane = threadIdx.x & 63;
    const int rq = 4 * (lane >> 4);           // this lane's k offset within a 16-deep step
    // the rows of this lane's A fragments (one per row tile of the chunk)
    int hob[CH], wob[CH];
    bool mok[CH];
#pragma unroll
    for (int i = 0; i < CH; ++i) {
        const int m = (ch * CH + i) * 16 + (lane & 15);
        mok[i] = (ch * CH + i) < MT && m < M;
        hob[i] = 2 * (m / WO) - 1;
        wob[i] = 2 * (m % WO) - 1;
    }
#pragma unroll
    for (int i = 0; i < CH; ++i) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
    for (int g0 = k0; g0 < k1; g0 += GR) {
        // one 8-byte fragment per lane and step (the wave reads 512
        // contiguous bytes), all of the group's steps in flight together
        v4s b[GR];
#pragma unroll
        for (int g = 0; g < GR; ++g)
            if (g0 + g < k1) b[g] = *(const v4s*)(W + ((size_t)((g0 + g) * NT + nt) * 64 + lane) * 4);
#pragma unroll
        for (int g = 0; g < GR; ++g) {
            if (g0 + g >= k1) break;
            const int kq = (g0 + g) * 16 + rq;
            const int c = kq % CIN, tap = kq / CIN;
            const int kh = tap / 3, kw = tap - 3 * kh;
#pragma unroll
            for (int i = 0; i < CH; ++i) {
                v4s a = v4s{0, 0, 0, 0};
                const int hi = hob[i] + kh, wi = wob[i] + kw;
                if (mok[i] && kq < KREAL && hi >= 0 && hi < HI && wi >= 0 && wi < WI)
                    a = *(const v4s*)&X[(hi * WI + wi) * XS + c];
                acc[i] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[g], acc[i], 0, 0, 0);
            }
        }
    }
}

// Bias, ReLU and the bf16 store of one work item.
template <int COUT, int HO, int WO, int YS, int CH>
__device__ __forceinline__ void infer_conv_store(bf16* __restrict__ Y, const float* __restrict__ bias, int nt, int ch,
                                                 const v4f (&acc)[CH])
{
    constexpr int M = HO * WO, MT = (M + 15) / 16;
    const int lane = threadIdx.x & 63;
    const int rq = 4 * (lane >> 4);
    const int n = nt * 16 + (lane & 15);
    const float bn = bias[n];
#pragma unroll
    for (int i = 0; i < CH; ++i)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int m = (ch * CH + i) * 16 + rq + r;
            if ((ch * CH + i) >= MT || m >= M) continue;
            float v = acc[i][r] + bn;
            v = fmaxf(v, 0.0f);
            Y[m * YS + n] = f2bf(v);
        }
}

// XS, YS: the pixel strides (elements) of X and Y.  CH: row tiles per work
// item (fewer when the layer has fewer, so no MFMA runs on an all-padding
// tile).  Layers with at most half as many work items as waves (the deep,
// small ones) split each item's k steps in two halves on two waves; the
// second half's sums go through LDS past the layer's output in Y (at
// kKPartOff elements; Y is larger than the small outputs that split) and are
// added to the first half's, in that order.
constexpr int kKPartOff = 4096;

template <int CIN, int COUT, int HI, int WI, int HO, int WO, int XS, int YS>
__device__ __forceinline__ void infer_conv(const bf16* __restrict__ X, bf16* __restrict__ Y, const bf16* __restrict__ W,
                                           const float* __restrict__ bias)
{
    constexpr int M = HO * WO, MT = (M + 15) / 16, NT = COUT / 16, CH = MT < 4 ? MT : 4;
    constexpr int KREAL = 9 * CIN, KP = ((KREAL + BK - 1) / BK) * BK, KS = KP / 16;
    static_assert(COUT % 16 == 0 && CIN % 4 == 0 && KP % 16 == 0, "tile shapes");
    const int wave = threadIdx.x >> 6;
    constexpr int WAVES = kInferThreads / 64;
    constexpr int CHUNKS = (MT + CH - 1) / CH, ITEMS = NT * CHUNKS;
    constexpr bool SPLIT = 2 * ITEMS <= WAVES && KS >= 2 * kGroup;
    if constexpr (!SPLIT) {
        constexpr int GR = KS < kGroup ? KS : kGroup;
        for (int item = wave; item < ITEMS; item += WAVES) {
            const int nt = item % NT, ch = item / NT;
            v4f acc[CH];
            infer_conv_item<CIN, COUT, HI, WI, HO, WO, XS, CH, GR, 0, KS>(X, W, nt, ch, acc);
            infer_conv_store<COUT, HO, WO, YS, CH>(Y, bias, nt, ch, acc);
        }
    } else {
        constexpr int KH = (KS + 1) / 2;
        constexpr int GR = KH < kGroup ? KH : kGroup;
        static_assert(M * YS <= kKPartOff && kKPartOff * 2 + ITEMS * CH * 256 * 4 <= kInferBufB * 2 &&
                          kKPartOff * 2 + ITEMS * CH * 256 * 4 <= kInferBufA * 2,
                      "the second-half sums fit past the output");
        const int lane = threadIdx.x & 63;
        const bool live = wave < 2 * ITEMS;
        const int item = wave % ITEMS, half = wave / ITEMS;
        const int nt = item % NT, ch = item / NT;
        v4f acc[CH];
        if (!live) {
#pragma unroll
            for (int i = 0; i < CH; ++i) acc[i] = v4f{0.f, 0.f, 0.f, 0.f};
        } else if (half == 0) {
            infer_conv_item<CIN, COUT, HI, WI, HO, WO, XS, CH, GR, 0, KH>(X, W, nt, ch, acc);
        } else {
            infer_conv_item<CIN, COUT, HI, WI, HO, WO, XS, CH, GR, KH, KS>(X, W, nt, ch, acc);
        }
        float* kp = (float*)(Y + kKPartOff) + ((size_t)item * CH * 64 + lane) * 4;
        if (live && half == 1)
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) kp[i * 256 + r] = acc[i][r];
        __syncthreads();
        if (live && half == 0) {
#pragma unroll
            for (int i = 0; i < CH; ++i)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[i][r] = acc[i][r] + kp[i * 256 + r];
            infer_conv_store<COUT, HO, WO, YS, CH>(Y, bias, nt, ch, acc);
        }
    }
}

__global__ __launch_bounds__(kInferThreads) void k_cnn_infer(const bf16* __restrict__ x, InferParams P,
                                                             float* __restrict__ out, float* __restrict__ gap_out)
{
    __shared__ __align__(16) bf16 bufA[kInferBufA];
    __shared__ __align__(16) bf16 bufB[kInferBufB];
    __shared__ float s_h1[kFc1Out];
    __shared__ float s_gap[kFc1In];
    __shared__ float s_fc2[kInferThreads / 64][kOut];
    const int t = threadIdx.x;
    // the input, 16 bytes per thread and step
    const uint4* src = (const uint4*)(x + (size_t)blockIdx.x * kInferIn);
    for (int i = t; i < kInferIn / 8; i += kInferThreads) ((uint4*)bufB)[i] = src[i];
    __syncthreads();
    constexpr int P4 = 4 + 0, P32 = 32 + kActPad, P64 = 64 + kActPad, P96 = 96 + kActPad, P128 = 128 + kActPad;
    infer_conv<4, 32, 60, 80, 30, 40, P4, P32>(bufB, bufA, P.W[0], P.b[0]);
    __syncthreads();
    infer_conv<32, 64, 30, 40, 15, 20, P32, P64>(bufA, bufB, P.W[1], P.b[1]);
    __syncthreads();
    infer_conv<64, 96, 15, 20, 8, 10, P64, P96>(bufB, bufA, P.W[2], P.b[2]);
    __syncthreads();
    infer_conv<96, 128, 8, 10, 4, 5, P96, P128>(bufA, bufB, P.W[3], P.b[3]);
    __syncthreads();
    infer_conv<128, 128, 4, 5, 2, 3, P128, P128>(bufB, bufA, P.W[4], P.b[4]);
    __syncthreads();
    infer_conv<128, 128, 2, 3, 1, 2, P128, P128>(bufA, bufB, P.W[5], P.b[5]);
    __syncthreads();
    // global average pooling over the 1 x 2 output (k_gap_fwd's arithmetic)
    if (t < kFc1In) {
        const float sgap = bf2f(bufB[t]) + bf2f(bufB[P128 + t]);
        s_gap[t] = bf2f(f2bf(sgap / 2.0f));
        if (gap_out) gap_out[(size_t)blockIdx.x * kFc1In + t] = s_gap[t];
    }
    if (gap_out) return;   // the FC layers run in k_cnn_head
    __syncthreads();
    // fc1 (ReLU, bf16 output) and fc2, fp32 accumulation of bf16 operands.
    // fc1 is a matrix-vector product over 983 KB of weights: each thread
    // forms 4 consecutive hidden units from 8-byte weight loads, 32 rows of
    // them in flight at a time; every unit sums its 128 terms in row order.
    static_assert(kFc1Out % 4 == 0 && kFc1Out / 4 <= kInferThreads && kFc1In % 32 == 0, "fc1 blocking");
    if (t < kFc1Out / 4) {
        constexpr int G = kFc1Out / 4;
        const uint2* Wv = (const uint2*)P.fc1W;
        float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int k0 = 0; k0 < kFc1In; k0 += 32) {
            uint2 wv[32];
#pragma unroll
            for (int kk = 0; kk < 32; ++kk) wv[kk] = Wv[(size_t)(k0 + kk) * G + t];
#pragma unroll
            for (int kk = 0; kk < 32; ++kk) {
                const float g = s_gap[k0 + kk];
                acc[0] += g * bf2f((bf16)(wv[kk].x & 0xffffu));
                acc[1] += g * bf2f((bf16)(wv[kk].x >> 16));
                acc[2] += g * bf2f((bf16)(wv[kk].y & 0xffffu));
                acc[3] += g * bf2f((bf16)(wv[kk].y >> 16));
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) s_h1[4 * t + q] = bf2f(f2bf(fmaxf(acc[q] + P.fc1b[4 * t + q], 0.0f)));
    }
    __syncthreads();
    // fc2: every wave sums its slice of the hidden units for all six outputs
    // (one 16-byte weight row per lane and step, all of a lane's rows in
    // flight together), then the waves' sums in wave order
    {
        constexpr int NW = kInferThreads / 64, RPW = kFc1Out / NW, STEPS = (RPW + 63) / 64;
        static_assert(kFc1Out % NW == 0 && kOutPad == 8 && kOut <= 8, "fc2 blocking");
        const int wv = t >> 6, lane = t & 63;
        const uint4* W8 = (const uint4*)P.fc2W;   // row k: the 8 (padded) outputs' bf16 weights
        uint4 w[STEPS];
#pragma unroll
        for (int q = 0; q < STEPS; ++q) {
            const int r = lane + 64 * q;
            w[q] = W8[wv * RPW + (r < RPW ? r : 0)];
        }
        float acc[kOut];
#pragma unroll
        for (int j = 0; j < kOut; ++j) acc[j] = 0.0f;
#pragma unroll
        for (int q = 0; q < STEPS; ++q) {
            const int r = lane + 64 * q;
            if (r < RPW) {
                const float h = s_h1[wv * RPW + r];
                const uint32_t wq[4] = {w[q].x, w[q].y, w[q].z, w[q].w};
#pragma unroll
                for (int j = 0; j < kOut; ++j)
                    acc[j] += h * bf2f((bf16)((j & 1) ? (wq[j >> 1] >> 16) : (wq[j >> 1] & 0xffffu)));
            }
        }
#pragma unroll
        for (int j = 0; j < kOut; ++j) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
        }
        if (lane == 0)
#pragma unroll
            for (int j = 0; j < kOut; ++j) s_fc2[wv][j] = acc[j];
        __syncthreads();
        if (t < kOut) {
            float o = 0.0f;
            for (int q = 0; q < NW; ++q) o += s_fc2[q][t];
            out[(size_t)blockIdx.x * kOut + t] = o + P.fc2b[t];
        }
    }
}

extern "C" {

int cnn_create(int max_batch, cnn_net* out)
{
    if (!out || max_batch < 1 || max_batch > 4096) return CNN_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return CNN_ERR_NODEV;
    cnn_net_s* c = new cnn_net_s();
    c->max_batch = max_batch;
    size_t off = 0;
    auto take = [&](Param& p, size_t n) {
        p.off = off;
        p.n = n;
        off += (n + 63) & ~(size_t)63;
    };
    for (int l = 0; l < kConvs; ++l) {
        c->Kp[l] = ((9 * kSpec[l].cin + BK - 1) / BK) * BK;
        take(c->convW[l], (size_t)c->Kp[l] * kSpec[l].cout);
        take(c->convB[l], kSpec[l].cout);
    }
    take(c->fc1W, (size_t)kFc1In * kFc1Out);
    take(c->fc1B, kFc1Out);
    take(c->fc2W, (size_t)kFc1Out * kOutPad);
    take(c->fc2B, kOutPad);
    c->nparam = off;
    const int N = max_batch;
    size_t big = 0, split = 0;
    int rc = CNN_OK;
    rc = rc ? rc : dalloc(c, &c->p, off);
    rc = rc ? rc : dalloc(c, &c->g, off);
    rc = rc ? rc : dalloc(c, &c->m, off);
    rc = rc ? rc : dalloc(c, &c->v, off);
    rc = rc ? rc : dalloc(c, &c->pb, off);
    rc = rc ? rc : dalloc(c, &c->mask, off);
    size_t pw_n = 0;
    for (int l = 0; l < kConvs; ++l) {
        c->pw_off[l] = pw_n;
        pw_n += (size_t)c->Kp[l] * kSpec[l].cout;
    }
    rc = rc ? rc : dalloc(c, &c->pw, pw_n);
    rc = rc ? rc : dalloc(c, &c->infer_gap, (size_t)N * kFc1In);
    rc = rc ? rc : dalloc(c, &c->infer_part, (size_t)N * kHeadGroups * kOutPad);
    rc = rc ? rc : dalloc(c, &c->infer_ticket, (size_t)N);
    for (int l = 0; l < kConvs && !rc; ++l) {
        const size_t M = (size_t)N * kSpec[l].ho * kSpec[l].wo;
        rc = dalloc(c, &c->xcol[l], M * c->Kp[l]);
        rc = rc ? rc : dalloc(c, &c->y[l], M * kSpec[l].cout);
        big = std::max(big, M * (size_t)c->Kp[l]);
        big = std::max(big, (size_t)N * kSpec[l].hi * kSpec[l].wi * kSpec[l].cin);
        split = std::max(split, (size_t)c->Kp[l] * kSpec[l].cout);
    }
    split = std::max(split, (size_t)kFc1In * kFc1Out);
    c->split_elems = split * 64;
    c->colpart_elems = (size_t)256 * 256;
    rc = rc ? rc : dalloc(c, &c->gap, (size_t)N * kFc1In);
    rc = rc ? rc : dalloc(c, &c->h1, (size_t)N * kFc1Out);
    rc = rc ? rc : dalloc(c, &c->out, (size_t)N * kOutPad);
    rc = rc ? rc : dalloc(c, &c->dy, big);
    rc = rc ? rc : dalloc(c, &c->dx, big);
    rc = rc ? rc : dalloc(c, &c->dcol, big);
    rc = rc ? rc : dalloc(c, &c->dz, big);
    rc = rc ? rc : dalloc(c, &c->split, c->split_elems);
    rc = rc ? rc : dalloc(c, &c->colpart, c->colpart_elems);
    rc = rc ? rc : dalloc(c, &c->dob, (size_t)N * kOutPad);
    rc = rc ? rc : dalloc(c, &c->dof, (size_t)N * kOutPad);
    rc = rc ? rc : dalloc(c, &c->lossn, (size_t)N);
    rc = rc ? rc : dalloc(c, &c->wts, (size_t)kOutPad);
    if (rc) {
        cnn_destroy(c);
        return rc;
    }
    // parameter mask: 0 on the reduction padding rows of each conv weight and
    // on the padding columns of fc2
    std::vector<uint8_t> mk(off, 1);
    for (int l = 0; l < kConvs; ++l)
        for (int k = 9 * kSpec[l].cin; k < c->Kp[l]; ++k)
            for (int n = 0; n < kSpec[l].cout; ++n) mk[c->convW[l].off + (size_t)k * kSpec[l].cout + n] = 0;
    for (int k = 0; k < kFc1Out; ++k)
        for (int n = kOut; n < kOutPad; ++n) mk[c->fc2W.off + (size_t)k * kOutPad + n] = 0;
    for (int n = kOut; n < kOutPad; ++n) mk[c->fc2B.off + n] = 0;
    if (hipMemcpy(c->mask, mk.data(), off, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(c->p, 0, off * 4) != hipSuccess || hipMemset(c->g, 0, off * 4) != hipSuccess ||
        hipMemset(c->m, 0, off * 4) != hipSuccess || hipMemset(c->v, 0, off * 4) != hipSuccess ||
        hipMemset(c->pb, 0, off * 2) != hipSuccess ||
        hipMemset(c->infer_ticket, 0, (size_t)N * sizeof(unsigned int)) != hipSuccess) {
        cnn_destroy(c);
        return CNN_ERR_HIP;
    }
    *out = c;
    return CNN_OK;
}

int cnn_destroy(cnn_net c)
{
    if (!c) return CNN_ERR_ARG;
    (void)hipDeviceSynchronize();
    for (void* p : c->allocs) (void)hipFree(p);
    delete c;
    return CNN_OK;
}

int cnn_set_stream(cnn_net c, void* stream)
{
    if (!c) return CNN_ERR_ARG;
    c->stream = (hipStream_t)stream;
    return CNN_OK;
}

int cnn_num_params(cnn_net c, size_t* n)
{
    if (!c || !n) return CNN_ERR_ARG;
    *n = c->nparam;
    return CNN_OK;
}

int cnn_param_layout(cnn_net c, size_t* offsets, size_t* sizes, int cap)
{
    if (!c || !offsets || !sizes || cap < 2 * kConvs + 4) return CNN_ERR_ARG;
    int q = 0;
    for (int l = 0; l < kConvs; ++l) {
        offsets[q] = c->convW[l].off; sizes[q++] = c->convW[l].n;
        offsets[q] = c->convB[l].off; sizes[q++] = c->convB[l].n;
    }
    offsets[q] = c->fc1W.off; sizes[q++] = c->fc1W.n;
    offsets[q] = c->fc1B.off; sizes[q++] = c->fc1B.n;
    offsets[q] = c->fc2W.off; sizes[q++] = c->fc2W.n;
    offsets[q] = c->fc2B.off; sizes[q++] = c->fc2B.n;
    return q;
}

int cnn_set_params(cnn_net c, const float* host, size_t n)
{
    if (!c || !host || n != c->nparam) return CNN_ERR_ARG;
    CK(hipMemcpy(c->p, host, n * 4, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_to_bf16, g1(n), dim3(256), 0, c->stream, c->p, n, c->pb);
    c->wver++;
    CK(hipMemset(c->m, 0, n * 4));
    CK(hipMemset(c->v, 0, n * 4));
    c->step = 0;
    CK(hipDeviceSynchronize());
    return launch_check("set_params");
}

int cnn_get_params(cnn_net c, float* host, size_t n)
{
    if (!c || !host || n != c->nparam) return CNN_ERR_ARG;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(host, c->p, n * 4, hipMemcpyDeviceToHost));
    return CNN_OK;
}

int cnn_get_grads(cnn_net c, float* host, size_t n)
{
    if (!c || !host || n != c->nparam) return CNN_ERR_ARG;
    CK(hipDeviceSynchronize());
    CK(hipMemcpy(host, c->g, n * 4, hipMemcpyDeviceToHost));
    return CNN_OK;
}

int cnn_prepare_input(cnn_net c, const float* depth_a, const uint8_t* image_a, const uint16_t* depth_b,
                      const uint8_t* image_b, int n, size_t frame_stride, void* x)
{
    if (!c || !depth_a || !image_a || !depth_b || !image_b || !x || n < 1) return CNN_ERR_ARG;
    hipLaunchKernelGGL(k_prepare, g1((size_t)n * 80 * 60), dim3(256), 0, c->stream, depth_a, image_a, depth_b,
                       image_b, n, 80, 60, frame_stride, (bf16*)x);
    return launch_check("prepare_input");
}

// Forward of N samples (x: NHWC bf16 [N][60][80][4], device); out (device,
// fp32 [N][6]) may be null.  Keeps the activations for cnn_backward.
int cnn_forward(cnn_net c, const void* x, int N, float* out)
{
    if (!c || !x || N < 1 || N > c->max_batch) return CNN_ERR_ARG;
    const bf16* in = (const bf16*)x;
    int rc;
    for (int l = 0; l < kConvs; ++l) {
        const ConvSpec& s = kSpec[l];
        const int M = N * s.ho * s.wo;
        const size_t tot = (size_t)M * c->Kp[l];
        hipLaunchKernelGGL(k_im2col, g1(tot), dim3(256), 0, c->stream, in, N, s.hi, s.wi, s.cin, s.ho, s.wo, c->Kp[l],
                           c->xcol[l]);
        if ((rc = gemm(c, false, c->xcol[l], c->Kp[l], false, c->pb + c->convW[l].off, s.cout, nullptr, c->y[l], s.cout,
                       M, s.cout, c->Kp[l], c->p + c->convB[l].off, true,
                       fwd_splits_n(N, s.ho * s.wo, s.cout, c->Kp[l]))))
            return rc;
        in = c->y[l];
    }
    hipLaunchKernelGGL(k_gap_fwd, g1((size_t)N * kFc1In), dim3(256), 0, c->stream, c->y[kConvs - 1], N, 2, kFc1In,
                       c->gap);
    if ((rc = gemm(c, false, c->gap, kFc1In, false, c->pb + c->fc1W.off, kFc1Out, nullptr, c->h1, kFc1Out, N, kFc1Out,
                   kFc1In, c->p + c->fc1B.off, true, fwd_splits_n(N, 1, kFc1Out, kFc1In))))
        return rc;
    if ((rc = gemm(c, false, c->h1, kFc1Out, false, c->pb + c->fc2W.off, kOutPad, c->out, nullptr, kOutPad, N, kOutPad,
                   kFc1Out, c->p + c->fc2B.off, false, fwd_splits_n(N, 1, kOutPad, kFc1Out))))
        return rc;
    if (out) CK(hipMemcpy2DAsync(out, kOut * 4, c->out, kOutPad * 4, kOut * 4, N, hipMemcpyDeviceToDevice, c->stream));
    return launch_check("forward");
}

// The FC layers of the split inference: kHeadGroups workgroups per sample
// (blockIdx.x = sample * kHeadGroups + group).  Group g forms the 240 hidden
// units [240 g, 240 g + 240) with all 128 of a unit's weights in flight at
// once, each unit summing its terms in row order, and their FC2 sums exactly
// as wave g of k_cnn_infer's head does (so the same bits); the group that
// arrives last at the sample's ticket adds the groups' sums in group order and
// the bias, and sets the ticket back to 0.
__global__ __launch_bounds__(256) void k_cnn_head(const float* __restrict__ gap, InferParams P,
                                                  float* __restrict__ part, unsigned int* __restrict__ ticket,
                                                  float* __restrict__ out)
{
    constexpr int RPW = kFc1Out / kHeadGroups, STEPS = (RPW + 63) / 64;
    static_assert(kFc1Out % kHeadGroups == 0 && RPW <= 256 && kOutPad == 8 && kOut <= 8, "head blocking");
    __shared__ float s_gap[kFc1In];
    __shared__ float s_h1[RPW];
    __shared__ unsigned int s_last;
    const int t = threadIdx.x;
    const int smp = blockIdx.x / kHeadGroups, grp = blockIdx.x - smp * kHeadGroups;
    const int unit = grp * RPW + t;
    // the unit's 128 fc1 weights (coalesced over the units) and the group's
    // fc2 rows, all loads issued before the first wait
    bf16 w[kFc1In];
    if (t < RPW) {
#pragma unroll
        for (int k = 0; k < kFc1In; ++k) w[k] = P.fc1W[(size_t)k * kFc1Out + unit];
    }
    const int lane = t & 63;
    uint4 w2[STEPS];
    const uint4* W8 = (const uint4*)P.fc2W;
    if (t < 64) {
#pragma unroll
        for (int q = 0; q < STEPS; ++q) {
            const int r = lane + 64 * q;
            w2[q] = W8[grp * RPW + (r < RPW ? r : 0)];
        }
    }
    if (t < kFc1In) s_gap[t] = gap[(size_t)smp * kFc1In + t];
    __syncthreads();
    if (t < RPW) {
        float acc = 0.0f;
#pragma unroll
        for (int k = 0; k < kFc1In; ++k) acc += s_gap[k] * bf2f(w[k]);
        s_h1[t] = bf2f(f2bf(fmaxf(acc + P.fc1b[unit], 0.0f)));
    }
    __syncthreads();
    if (t < 64) {
        float acc[kOut];
#pragma unroll
        for (int j = 0; j < kOut; ++j) acc[j] = 0.0f;
#pragma unroll
        for (int q = 0; q < STEPS; ++q) {
            const int r = lane + 64 * q;
            if (r < RPW) {
                const float h = s_h1[r];
                const uint32_t wq[4] = {w2[q].x, w2[q].y, w2[q].z, w2[q].w};
#pragma unroll
                for (int j = 0; j < kOut; ++j)
                    acc[j] += h * bf2f((bf16)((j & 1) ? (wq[j >> 1] >> 16) : (wq[j >> 1] & 0xffffu)));
            }
        }
#pragma unroll
        for (int j = 0; j < kOut; ++j) {
#pragma unroll
            for (int off = 32; off >= 1; off >>= 1) acc[j] += __shfl_xor(acc[j], off, 64);
        }
        if (lane == 0) {
#pragma unroll
            for (int j = 0; j < kOut; ++j) part[((size_t)smp * kHeadGroups + grp) * kOutPad + j] = acc[j];
            // the partial sums first, then the arrival (release); the last
            // arrival reads every group's sums (acquire)
            s_last = __hip_atomic_fetch_add(ticket + smp, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT) ==
                     (unsigned)(kHeadGroups - 1);
        }
    }
    __syncthreads();
    if (s_last && t < kOut) {
        const float* ps = part + (size_t)smp * kHeadGroups * kOutPad;
        float o = 0.0f;
        for (int q = 0; q < kHeadGroups; ++q)
            o += __hip_atomic_load(ps + (size_t)q * kOutPad + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        out[(size_t)smp * kOut + t] = o + P.fc2b[t];
        if (t == 0) __hip_atomic_store(ticket + smp, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Inference of N samples in one launch (k_cnn_infer): out (device, fp32
// [N][6]).  Keeps no activations.
int cnn_infer(cnn_net c, const void* x, int N, float* out)
{
    if (!c || !x || !out || N < 1 || N > c->max_batch) return CNN_ERR_ARG;
    if ((uintptr_t)x % 16 != 0) return CNN_ERR_ARG;
    if (c->pver != c->wver) {
        PackArgs A;
        size_t n = 0;
        for (int l = 0; l < kConvs; ++l) {
            A.src[l] = c->pb + c->convW[l].off;
            A.dst[l] = c->pw + c->pw_off[l];
            A.n[l] = (size_t)c->Kp[l] * kSpec[l].cout;
            A.nt[l] = kSpec[l].cout / 16;
            A.cout[l] = kSpec[l].cout;
            n = std::max(n, A.n[l]);
        }
        hipLaunchKernelGGL(k_pack_infer, dim3((unsigned)((n + 255) / 256), kConvs), dim3(256), 0, c->stream, A);
        c->pver = c->wver;
    }
    InferParams P;
    for (int l = 0; l < kConvs; ++l) {
        P.W[l] = c->pw + c->pw_off[l];
        P.b[l] = c->p + c->convB[l].off;
        P.Kp[l] = c->Kp[l];
    }
    P.fc1W = c->pb + c->fc1W.off;
    P.fc1b = c->p + c->fc1B.off;
    P.fc2W = c->pb + c->fc2W.off;
    P.fc2b = c->p + c->fc2B.off;
    if (c->infer_split) {
        hipLaunchKernelGGL(k_cnn_infer, dim3(N), dim3(kInferThreads), 0, c->stream, (const bf16*)x, P, out,
                           c->infer_gap);
        hipLaunchKernelGGL(k_cnn_head, dim3(N * kHeadGroups), dim3(256), 0, c->stream, c->infer_gap, P,
                           c->infer_part, c->infer_ticket, out);
    } else {
        hipLaunchKernelGGL(k_cnn_infer, dim3(N), dim3(kInferThreads), 0, c->stream, (const bf16*)x, P, out,
                           (float*)nullptr);
    }
    return launch_check("infer");
}

// The inference head's layout: 1 (default) the FC layers in k_cnn_head's
// workgroups, 0 in the convolution workgroup; the same bits either way.
int cnn_set_infer_split(cnn_net c, int split)
{
    if (!c || (split != 0 && split != 1)) return CNN_ERR_ARG;
    c->infer_split = split;
    return CNN_OK;
}

// Gradients of the weighted squared error against target (device fp32
// [N][6]) with weights w (host, 6) for the last forward of N samples: loss
// per sample into loss_n (device, optional), parameter gradients into the
// net's gradient buffer.
int cnn_backward(cnn_net c, int N, const float* target, const float* w, float* loss_n)
{
    if (!c || !target || !w || N < 1 || N > c->max_batch) return CNN_ERR_ARG;
    int rc;
    float wp[kOutPad] = {0};
    for (int j = 0; j < kOut; ++j) wp[j] = w[j];
    CK(hipMemcpyAsync(c->wts, wp, sizeof wp, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_loss, g1(N), dim3(256), 0, c->stream, c->out, kOutPad, target, c->wts, N, kOut, c->lossn,
                       c->dob, c->dof);
    if (loss_n) CK(hipMemcpyAsync(loss_n, c->lossn, N * 4, hipMemcpyDeviceToDevice, c->stream));
    // fc2: dW2 = H1^T dO, db2 = colsum dO, dH1 = dO W2^T
    if ((rc = gemm(c, true, c->h1, kFc1Out, false, c->dob, kOutPad, c->g + c->fc2W.off, nullptr, kOutPad, kFc1Out,
                   kOutPad, N, nullptr, false, splits_for(N, kFc1Out * kOutPad))))
        return rc;
    if ((rc = colsum(c, c->dof, N, kOutPad, c->g + c->fc2B.off))) return rc;
    if ((rc = gemm(c, false, c->dob, kOutPad, true, c->pb + c->fc2W.off, kOutPad, c->dy, nullptr, kFc1Out, N, kFc1Out,
                   kOutPad, nullptr, false, 1)))
        return rc;
    // fc1 (ReLU): dZ1 = dH1 * (H1 > 0); dW1 = G^T dZ1; db1; dG = dZ1 W1^T
    hipLaunchKernelGGL(k_relu_grad, g1((size_t)N * kFc1Out), dim3(256), 0, c->stream, c->dy, c->h1,
                       (size_t)N * kFc1Out, c->dz, c->dx);
    if ((rc = gemm(c, true, c->gap, kFc1In, false, c->dz, kFc1Out, c->g + c->fc1W.off, nullptr, kFc1Out, kFc1In,
                   kFc1Out, N, nullptr, false, splits_for(N, kFc1In * kFc1Out))))
        return rc;
    if ((rc = colsum(c, c->dx, N, kFc1Out, c->g + c->fc1B.off))) return rc;
    if ((rc = gemm(c, false, c->dz, kFc1Out, true, c->pb + c->fc1W.off, kFc1Out, c->dcol, nullptr, kFc1In, N, kFc1In,
                   kFc1Out, nullptr, false, 1)))
        return rc;
    // GAP backward into dy of the last conv
    hipLaunchKernelGGL(k_gap_bwd, g1((size_t)N * 2 * kFc1In), dim3(256), 0, c->stream, c->dcol, N, 2, kFc1In, c->dy);
    for (int l = kConvs - 1; l >= 0; --l) {
        const ConvSpec& s = kSpec[l];
        const int M = N * s.ho * s.wo;
        const size_t ny = (size_t)M * s.cout;
        // dZ = dY * (Y > 0); fp32 copy in dx for the bias sums
        hipLaunchKernelGGL(k_relu_grad, g1(ny), dim3(256), 0, c->stream, c->dy, c->y[l], ny, c->dz, c->dx);
        if ((rc = gemm(c, true, c->xcol[l], c->Kp[l], false, c->dz, s.cout, c->g + c->convW[l].off, nullptr, s.cout,
                       c->Kp[l], s.cout, M, nullptr, false, splits_for(M, c->Kp[l] * s.cout))))
            return rc;
        if ((rc = colsum(c, c->dx, M, s.cout, c->g + c->convB[l].off))) return rc;
        if (l == 0) break;
        // dXcol = dZ W^T, then col2im into dy of the layer below
        if ((rc = gemm(c, false, c->dz, s.cout, true, c->pb + c->convW[l].off, s.cout, c->dcol, nullptr, c->Kp[l], M,
                       c->Kp[l], s.cout, nullptr, false, 1)))
            return rc;
        const size_t nx = (size_t)N * s.hi * s.wi * s.cin;
        hipLaunchKernelGGL(k_col2im, g1(nx), dim3(256), 0, c->stream, c->dcol, N, s.hi, s.wi, s.cin, s.ho, s.wo,
                           c->Kp[l], c->dy);
    }
    return launch_check("backward");
}

int cnn_adam_step(cnn_net c, float lr, float wd)
{
    if (!c || !(lr >= 0.0f)) return CNN_ERR_ARG;
    c->step++;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - powf(b1, (float)c->step), bc2 = 1.0f - powf(b2, (float)c->step);
    hipLaunchKernelGGL(k_adam, g1(c->nparam), dim3(256), 0, c->stream, c->p, c->g, c->m, c->v, c->pb, c->mask,
                       c->nparam, lr, b1, b2, eps, bc1, bc2, wd);
    c->wver++;
    return launch_check("adam");
}

}  // extern "C"
