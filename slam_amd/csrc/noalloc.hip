// noalloc.hip -- odom_icp_step_noalloc (include/slam_odom.h): the ICP step of
// the odometry on caller-provided device maps, with every buffer its kernels
// need beyond the inputs and the output in one file-scope device array.
//
// The array holds the step's descriptor (one pair, one level), the state
// record the pass reads (the two poses), the block partials of the pass and
// the FP64 totals.  Its device address is taken once per device with
// hipGetSymbolAddress and used wherever host code needs the array's memory;
// no call allocates or frees anything.  All calls on one device share the
// array: two calls that overlap in time on the same device race.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstddef>
#include <mutex>

#include "odom_internal.h"
#include "odom_math.h"
#include "slam_odom.h"

namespace {

using namespace slam;

constexpr int kNaShape = 256;            // threads of the pass
constexpr int kNaPpt = 2;                // pixels per thread per tile
constexpr size_t kOffPair = 0;
constexpr size_t kOffState = 4096;
constexpr size_t kOffPartA = kOffState + 16384;
constexpr size_t kOffPartAi = kOffPartA + (size_t)kMaxBlk * kPartStride * 4;
constexpr size_t kScratchBytes = kOffPartAi + (size_t)kMaxBlk * 2 * 4;
static_assert(sizeof(PairDesc) <= kOffState - kOffPair, "descriptor slot");
static_assert(sizeof(TrackState) <= kOffPartA - kOffState, "state slot");
static_assert(kScratchBytes == ODOM_NOALLOC_SCRATCH_BYTES, "size stated in slam_odom.h");

constexpr int kMaxDevices = 64;

}  // namespace

// The scratch array.  In host code this name is only a handle for
// hipGetSymbolAddress: device addresses come from scratch_ptr().
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

// The array's address on the current device, looked up on the first call
// made with that device current.
int scratch_ptr(unsigned char** out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return ODOM_ERR_NODEV;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            fprintf(stderr, "slam_odom: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return ODOM_ERR_HIP;
        }
        g_ptr[dev] = p;
    }
    *out = (unsigned char*)g_ptr[dev];
    return ODOM_OK;
}

#define NACHK(x)                                                                                       \
    do {                                                                                               \
        const hipError_t e_ = (x);                                                                     \
        if (e_ != hipSuccess) {                                                                        \
            fprintf(stderr, "slam_odom: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__,  \
                    __LINE__);                                                                         \
            return ODOM_ERR_HIP;                                                                       \
        }                                                                                              \
    } while (0)

}  // namespace

extern "C" int odom_icp_step_noalloc(const float* vcur, const float* ncur, const float* vprev, const float* nprev,
                                     int width, int height, float fx, float fy, float cx, float cy,
                                     const float Rcurr[9], const float tcurr[3], const float Rprev_inv[9],
                                     const float tprev[3], float dist_thresh, float angle_thresh, double out[29],
                                     void* stream)
{
    if (!vcur || !ncur || !vprev || !nprev || !Rcurr || !tcurr || !Rprev_inv || !tprev || !out) return ODOM_ERR_ARG;
    if (width < 1 || height < 1 || width > ODOM_NOALLOC_MAX_DIM || height > ODOM_NOALLOC_MAX_DIM) {
        fprintf(stderr, "slam_odom: odom_icp_step_noalloc: %dx%d maps exceed the %dx%d maximum\n", width, height,
                ODOM_NOALLOC_MAX_DIM, ODOM_NOALLOC_MAX_DIM);
        return ODOM_ERR_ARG;
    }
    if (!(fx != 0.0f) || !(fy != 0.0f)) return ODOM_ERR_ARG;
    unsigned char* base = nullptr;
    if (int rc = scratch_ptr(&base)) return rc;
    hipStream_t s = (hipStream_t)stream;

    PairDesc D;
    memset(&D, 0, sizeof D);
    LevelDesc& L = D.lv[0];
    L.w = width;
    L.h = height;
    L.fx = fx;
    L.fy = fy;
    L.cx = cx;
    L.cy = cy;
    L.invFx = 1.0f / fx;
    L.invFy = 1.0f / fy;
    L.vcur = vcur;
    L.ncur = ncur;
    L.vprev = vprev;
    L.nprev = nprev;
    kmat(fx, fy, cx, cy, L.K, L.Ki);
    const int n = width * height;
    const int nblk = std::min(kMaxBlk, (n + kNaShape * kNaPpt - 1) / (kNaShape * kNaPpt));
    D.nblk_a[0] = nblk;
    D.nblk_b[0] = nblk;
    D.st = (TrackState*)(base + kOffState);
    for (int i = 0; i < 2; ++i) {
        D.part_a[i] = (float*)(base + kOffPartA);
        D.part_b[i] = (float*)(base + kOffPartA);
        D.part_ai[i] = (unsigned int*)(base + kOffPartAi);
    }

    // the state the pass reads: the current estimate in slot 0 and the
    // previous pose's inverse rotation and translation
    TrackState S;
    memset(&S, 0, sizeof S);
    memcpy(S.it[0].Rcurr, Rcurr, sizeof(float) * 9);
    memcpy(S.it[0].tcurr, tcurr, sizeof(float) * 3);
    memcpy(S.Rprev_inv, Rprev_inv, sizeof(float) * 9);
    memcpy(S.tprev, tprev, sizeof(float) * 3);
    S.it[0].skip = 0;
    S.it[0].break_level = -1;

    NACHK(hipMemcpyAsync(base + kOffPair, &D, sizeof D, hipMemcpyHostToDevice, s));
    NACHK(hipMemcpyAsync(base + kOffState, &S, sizeof S, hipMemcpyHostToDevice, s));

    TrackParams P{};
    P.icp = 1;
    P.rgb = 0;
    P.icpWeight = 100.0f;
    P.distThres = dist_thresh;
    P.angleThres = angle_thresh;
    P.sobelScale = 0.125f;
    P.maxDepthDelta = 0.07f;
    StepLaunch La;
    La.pairs = (const PairDesc*)(base + kOffPair);
    La.npairs = 1;
    La.blk_pair = nullptr;
    La.total_blocks = nblk;
    La.shape = StepShape{kNaShape, kNaPpt};
    launch_step_a(La, 0, 0, 0, -1, P, 1, nullptr, s);
    launch_reduce_dbg(La.pairs, 0, 0, s);
    NACHK(hipGetLastError());
    NACHK(hipMemcpyAsync(out, base + kOffState + offsetof(TrackState, dbg_sums), sizeof(double) * kSE3,
                         hipMemcpyDeviceToHost, s));
    NACHK(hipStreamSynchronize(s));
    return ODOM_OK;
}
