// odom_internal.h -- launchers of the HIP kernels (odom_kernels.hip), called by
// the host class (odom_host.cpp).  All launchers are asynchronous on `s`.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "odom_device.h"
#include "slam_odom.h"

namespace slam {

// The statistics of the state last read back to the host (by odom_track,
// odom_get_pose or odom_get_stats), without a device round trip; for a caller
// that has just made one of those calls.  ODOM_ERR_ARG before any readback.
int odom_stats_last_read(odom_handle c, odom_stats* s);

// Launch shape of a step kernel at one level (replaces GPUConfig.h's
// threads/blocks table; chosen by the autotuner).
struct StepShape {
    int threads;    // 64, 128, 256 or 512
    int ppt;        // pixels per thread: 1, 2, 4, 8 or 16
};

// Kernel names and the shapes they support, for the autotuner.
int step_shape_supported(StepShape sh);

// ---- preprocessing (a1-a13) ----
void launch_bilateral(const uint16_t* src, int src_pitch, uint16_t* dst, int w, int h, float maxD,
                      hipStream_t s);
// depth pyramid -> vertex+normal maps of all levels in one launch
struct VNLevels {
    uint16_t* depth[kLevels];       // level 0 is read; launch_depth_pyramid_maps writes levels 1, 2
    float* vmap[kLevels];
    float* nmap[kLevels];
    int w[kLevels], h[kLevels];
    float fx_inv[kLevels], fy_inv[kLevels], cx[kLevels], cy[kLevels];
    float cutoff;
};
// initICP's depth levels 1 and 2 from a.depth[0] and the three levels'
// vertex and normal maps, in one launch.
void launch_depth_pyramid_maps(const VNLevels& a, hipStream_t s);
// RGBA32F vertex/normal maps -> planar pyramids (copyMaps + resize, optional
// rigid transform), plus the vertex z plane kept for verticesToDepth.
struct ModelLevels {
    const float4* v4;
    const float4* n4;
    int v_pitch, n_pitch;           // in float4 elements
    float* vmap[kLevels];
    float* nmap[kLevels];
    int w[kLevels], h[kLevels];
    float* vz;                      // w0*h0 z values (may be null)
    int transform;
    float R[9], t[3];
    const float* pose_dev;          // if set: [R|t] (3x4 row-major) read on the device
};
void launch_model_maps(const ModelLevels& a, hipStream_t s);
// Levels 0 (when generated), 1 and 2 of up to two pyramids in one launch.
// kind: 0 = float Gaussian (a9), 1 = uint8 Gaussian (a11), 2 = uint16
// bilateral-gated (a2).  gen: 0 = level 0 given (src0); 1 = level-0 depth
// from the vertex z plane vz (verticesToDepth, cutoff); 2 = level-0
// intensity from colour (rgb, pitch, ch).  dst[0] is written when gen != 0.
struct PyrSrc {
    int kind, gen;
    const void* src0;
    const float* vz;
    float cutoff;
    const uint8_t* rgb;
    int pitch, ch;
    void* dst[3];
};
struct PyrJob {
    PyrSrc p[2];
    int count;
    int w0, h0;
};
void launch_pyramid(const PyrJob& j, hipStream_t s);
// derivatives (a12) + static photometric source mask, all levels of all pairs
struct PrepLevels {
    const uint8_t* img[kLevels];
    const float* depth[kLevels];
    short2* dxy[kLevels];
    float* zsrc[kLevels];
    int w[kLevels], h[kLevels];
    float minScale[kLevels];
};
void launch_rgb_prep(const PrepLevels& a, hipStream_t s);

// ---- tracking (a14-a21) ----
struct StepLaunch {
    const PairDesc* pairs;          // device array
    int npairs;
    const int* blk_pair;            // device table, per level (null when npairs == 1)
    int total_blocks;               // blocks of this level over all pairs
    StepShape shape;
};
// SO3 pre-alignment; max_px_l2: largest level-2 pixel count over the pairs
// The SO3 loop; with begin, its last launch also does launch_track_begin's work.
// spin_limit < 0: one launch per iteration; else the one-launch loop (when
// the pairs fit), whose barrier waits give up after spin_limit polls.  With
// prep (one pair), launch_rgb_prep's work is done too: by extra blocks of the
// one-launch loop, or by a launch of its own before the per-iteration form.
void launch_so3(const PairDesc* pairs, int npairs, const TrackParams* begin, int spin_limit, hipStream_t s,
                const PrepLevels* prep = nullptr);
void launch_track_begin(const PairDesc* pairs, int npairs, const TrackParams& P, hipStream_t s);
// warm start: seed from the network prediction (6 floats per pair, device),
// optionally kept only when the level-2 ICP association is not worse than plain
void launch_seed_select(const PairDesc* pairs, int npairs, const float* pred, float t_scale, float r_scale,
                        int select, const TrackParams& P, hipStream_t s);
// network input (cnn_kernels.hip) from 160x120 level-2 buffers
void cnn_launch_prepare(const float* depth_a, const uint8_t* image_a, const uint16_t* depth_b, const uint8_t* image_b,
                        int n, size_t frame_stride, void* x, hipStream_t s);
// Gauss-Newton iteration k (global index over the call) at (level, iter).
// mode 0: tracking (the prologue finishes iteration k-1, which ran at
// prev_level); mode 1: parity pass with the constants of state slot 0,
// partials into parity 0 (totals by launch_reduce_dbg).
void launch_step_a(const StepLaunch& L, int level, int iter, int k, int prev_level, const TrackParams& P, int mode,
                   uint8_t* dbg_mask, hipStream_t s);
void launch_step_b(const StepLaunch& L, int level, int iter, int k, const TrackParams& P, int mode, hipStream_t s);
// Finish iteration k-1 and store the state of iteration k, one workgroup of
// `threads` (256 or 512: those of the pass A that follows) per pair
// (followed by launch_step_a with mode 1).
void launch_solve(const PairDesc* pairs, int npairs, int level, int iter, int k, int prev_level, const TrackParams& P,
                  int threads, hipStream_t s);
// Policy calls: the start of every entry (hyp entries per handle; the
// prediction or the constant-velocity extrapolation), and the judge of two
// refinements per handle (slot >= 0: after the coarse levels, from state
// slot `slot`; slot < 0: after the call).
void launch_seed_policy(const PairDesc* pairs, int nentries, int hyp, const float* pred, float t_scale, float r_scale,
                        hipStream_t s);
void launch_judge(const PairDesc* pairs, int nhandles, const TrackParams& P, int slot, int level, hipStream_t s);
// finishes the last iteration (K-1, at last_level), revert, result
void launch_track_end(const PairDesc* pairs, int npairs, const TrackParams& P, int K, int last_level, hipStream_t s);
// FP64 totals of a parity pass: which 0 = pass A (ICP sums, count, sigma), 1 = pass B
void launch_reduce_dbg(const PairDesc* pairs, int which, int level, hipStream_t s);
void launch_so3_step_dbg(const PairDesc* pairs, const float* consts27, double* out11, uint8_t* mask,
                         hipStream_t s);

// Constant tables of the kernels (the bilateral's spatial factors), per device.
void init_device_tables();

// Host copy of the device exp (same operation sequence): tables and tests.
float bexpf_host(float x);

}  // namespace slam
