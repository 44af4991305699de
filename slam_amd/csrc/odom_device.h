// odom_device.h -- data layout shared by the host class and the HIP kernels.
//
// One "pair" is one tracking problem (a current frame against a model
// prediction).  The single-frame API is a batch of one pair; the batched API
// runs many pairs of uneven size through the same launches.  Every step
// kernel covers one pyramid level of every pair in one launch: a block-to-pair
// table maps each workgroup to (pair, tile).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace slam {

constexpr int kLevels = 3;
constexpr int kSE3 = 29;   // 27 upper-triangular products of [J|r], r^2, inliers
constexpr int kSO3 = 11;   // 9 products of [J|r] (3x3+3), r^2, count
constexpr int kPartStride = 32;  // floats per block partial record

// Maps of one pyramid level of one pair.  Vertex/normal maps are planar
// (x plane, y plane, z plane of w*h floats each); a pixel is invalid iff its x
// value is NaN.
struct LevelDesc {
    int w, h;
    float fx, fy, cx, cy;
    float invFx, invFy;              // 1.0f / f, as projectToPointCloud
    const float* vcur;               // current frame, camera frame
    const float* ncur;
    const float* vprev;              // model prediction, global frame
    const float* nprev;
    const float* zsrc;               // next depth, NaN where the photometric
                                     // source pixel is statically invalid
    const short2* dxy;               // next image derivatives (dx, dy)
    const uint8_t* inext;            // next intensity
    const float* zlast;              // model (last) depth
    const uint8_t* ilast;            // model (last) intensity
    float minScale;                  // (minGrad / sobelScale)^2
    int pad_;
};

// Device-resident tracking state of one pair (one getIncrementalTransformation
// call).  Written by single-thread begin/finalize code, read uniformly by the
// step kernels.
struct TrackState {
    // pose prior and running estimate (float, as the reference's Eigen state)
    float Rprev[9], tprev[3], Rprev_inv[9];
    float Rcurr[9], tcurr[3];
    float pose_out[12];              // [R | t] row-major result of the call
    // accumulated increment (double 4x4 row-major) and SO3 pre-alignment
    double resultRt[16];
    double so3R[9];
    // per-iteration constants of the photometric warp
    float krkinv[9], kt[3];
    float sigmaVal;
    int rgb_count, rgb_sigma;
    // FP64 totals of the current iteration
    double icp_sums[kSE3];
    double rgb_sums[kSE3];
    double lastA[36], lastb[6];
    double lastx[6];
    // reported errors (RGBDOdometryef.h:62-67)
    float lastICPError, lastICPCount, lastRGBError, lastRGBCount, lastSO3Error, lastSO3Count;
    float res0, res1;                // residual[2] of the ICP step
    int so3_iters;
    int break_level;                 // rgbOnly early exit: level that stopped, else -1
    int iters_done;
    int degenerate;                  // last solve had zero information
    int reverted;                    // 0.3 m revert taken
    unsigned int ticket_a, ticket_b; // last-block-done counters (reset by the last block)
    int pad_[2];
};

// Uniform options of one track call.
struct TrackParams {
    int icp, rgb, rgbOnly, so3;
    float icpWeight;
    float distThres, angleThres;
    float sobelScale;
    float maxDepthDelta;
    int iters[kLevels];
};

// Per pair: level maps, state, scratch.
struct PairDesc {
    LevelDesc lv[kLevels];
    TrackState* st;
    float* part_a;                   // block partials of the ICP/association pass
    float* part_b;                   // block partials of the photometric pass
    unsigned int* part_ai;           // integer partials (count, sum diff^2)
    int nblk_a[kLevels];             // blocks of this pair per level, ICP/association pass
    int blk0_a[kLevels];             // first global block of this pair per level
    int nblk_b[kLevels];             // same for the photometric pass
    int blk0_b[kLevels];
    const uint8_t* so3_last;         // lastNextImage level 2
    const uint8_t* so3_next;         // nextImage level 2
    int pad_[2];
};

}  // namespace slam
