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
constexpr int kMaxBlk = 256;     // block partials per pair, level and pass: the step
                                 // grids stride over their tiles with at most this many blocks
// Polls of the one-launch SO3 loop's barrier waits before one gives up (about
// a second); the default of a handle (odom_set_so3_loop).
constexpr int kSO3SpinDefault = 1 << 20;

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
    double K[9], Ki[9];              // level intrinsics and closed-form inverse, in double
};

// State of the Gauss-Newton loop between two iterations.  Two slots
// alternate (iteration k lives in slot k & 1): the ICP/association pass A_k
// reads slot k-1 plus the photometric partials of iteration k-1, solves, and
// its block 0 writes slot k; the photometric pass B_k reads the pose fields of
// slot k and its block 0 writes the association fields of slot k.  No launch
// reads a field that a block of the same launch writes.
struct IterState {
    // written by A_k
    double resultRt[16];             // accumulated increment (row-major 4x4)
    double lastA[36], lastb[6];      // system of the last solve
    float Rcurr[9], tcurr[3];        // current estimate (float, as the reference)
    float krkinv[9], kt[3];          // photometric warp constants of iteration k
    float rgb_err_cmp;               // lastRGBError the rgbOnly test of iteration k compares to
    int skip;                        // iteration k skipped (its level already stopped)
    int break_level;                 // level stopped by the rgbOnly test, else -1
    int degenerate;                  // last solve had zero information
    int iters_done;
    // ICP statistics of iteration k-1, written by A_k from the fused
    // reduction of its prologue (carried otherwise)
    double icp_sums[kSE3];
    float res0, res1;                // residual[2] of the ICP step
    float lastICPError, lastICPCount;
    // written by B_k (by A_{k+1} when there is no photometric pass)
    float lastRGBError, lastRGBCount;
    float sigmaVal;
    int rgb_count, rgb_sigma;
    int pending;                     // a solve of iteration k is due
    int broke;                       // the rgbOnly test fired at iteration k
    int pad_;
};

// Hypotheses of one tracking call: a pair's main state (the call's result)
// and, when a start is judged against the constant-velocity start, the state
// of that second refinement (same inputs, same prior).
constexpr int kMaxHyp = 2;

// State of the SO3 pre-alignment loop (RGBDOdometryef.cpp:294-380) between
// its launches; iteration k lives in slot k & 1.
struct So3Slot {
    double resultR[9], lastResultR[9];
    float R_lr[9];
    float lastError, lastCount;      // the loop's comparison values
    float err, cnt;                  // lastSO3Error / lastSO3Count so far
    int iters, done;
};

// Device-resident tracking state of one pair (one getIncrementalTransformation
// call).
struct TrackState {
    float Rprev[9], tprev[3], Rprev_inv[9];
    float pose_out[12];              // [R | t] row-major result of the call
    double so3R[9];                  // SO3 pre-alignment
    So3Slot so3s[2];
    So3Slot so3_final;               // the SO3 loop state once it has ended
    int so3_done;                    // set by the launch that ended the loop
    IterState it[2];
    // results of the call, copied by the end kernel (RGBDOdometryef.h:62-70)
    double lastA[36], lastb[6];
    float lastICPError, lastICPCount, lastRGBError, lastRGBCount, lastSO3Error, lastSO3Count;
    int so3_iters;
    int iters_done;
    int degenerate;
    int reverted;                    // 0.3 m revert taken
    // kernel-level entry points (parity checks)
    double dbg_sums[kSE3];
    int dbg_count, dbg_sigma;
    float* pose_log;                 // optional device ring of [R|t] results
    int pose_log_cap;
    int frame_count;                 // track calls completed
    // warm start: the increment the seeded call starts from (resultRt
    // convention), whether the prediction was kept over the plain start, and
    // counters of seeded calls and kept predictions
    double seedRt[16];
    int seed_used, seed_calls, seed_kept;
    // the increment of the last completed call (resultRt convention: the
    // pose moved from prev to prev * lastRt^-1; identity after a revert or
    // a first frame), from which the constant-velocity start extrapolates
    double lastRt[16];
    int alt_iters;                   // iterations of the judged-away hypothesis
    int judge_level;                 // level after which the last judge ran (-1 none)
    // the SO3 loop in one launch: arrivals at its grid barrier and exits of
    // its workgroups (both back to 0 when the launch ends), and a sticky flag
    // set when a barrier wait gave up
    unsigned int so3_arrive, so3_exit;
    int so3_fail;
};

// Uniform options of one track call.
struct TrackParams {
    int icp, rgb, rgbOnly, so3;
    float icpWeight;
    float distThres, angleThres;
    float sobelScale;
    float maxDepthDelta;
    int iters[kLevels];
    int seeded;                      // start from TrackState::seedRt: 1 the warm start,
                                     // 2 the warm start if its level-2 association keeps it,
                                     // 3 a policy call's start (k_seed_policy; counted there
                                     // or by the judge)
    // convergence policy: a level stops after a solve whose update has
    // |dt| < convT (m) and |dw| < convR (rad); 0 = the fixed schedule
    float convT, convR;
    float judgeTol;                  // a judged start falls back when its ICP residual is
                                     // more than this fraction above, or its inlier count
                                     // this fraction below, the constant-velocity start's
};

// Per pair: level maps, state, scratch.
struct PairDesc {
    LevelDesc lv[kLevels];
    TrackState* st;
    float* part_a[2];                // block partials of the ICP/association pass, by iteration parity
    float* part_b[2];                // block partials of the photometric pass
    unsigned int* part_ai[2];        // integer partials (count, sum diff^2)
    int nblk_a[kLevels];             // blocks of this pair per level, ICP/association pass
    int blk0_a[kLevels];             // first global block of this pair per level
    int nblk_b[kLevels];             // same for the photometric pass
    int blk0_b[kLevels];
    const uint8_t* so3_last;         // lastNextImage level 2
    const uint8_t* so3_next;         // nextImage level 2
    int pad_[2];
};

}  // namespace slam
