// frame_loop.cpp -- the reference's frame loop without loop closure, behind
// include/slam_loop.h.  Host code only: every stage is a call into the
// odometry, splat and surfel entry points of this library, all on the
// odometry handle's stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>

#include "odom_internal.h"
#include "slam_loop.h"

struct efloop_ctx {
    efloop_params p{};
    int w = 0, h = 0;
    odom_handle odo = nullptr;
    splat_handle pred = nullptr;
    surfel_handle model = nullptr;
    uint16_t* filtered = nullptr;    // device, the bilateral output
    float pose[16];
    int tick = 0;
    float last_icp_error = 0.0f;
    float weighting = 1.0f;
};

namespace {

void free_loop(efloop_ctx* c)
{
    if (c->model) surfel_destroy(c->model);
    if (c->pred) splat_destroy(c->pred);
    if (c->odo) odom_destroy(c->odo);
    if (c->filtered) (void)hipFree(c->filtered);
    delete c;
}

odom_frame_view device_view(const void* p, int w, int h, int bpp, int fmt)
{
    return odom_frame_view{p, w * bpp, w, h, fmt, ODOM_MEM_DEVICE};
}

// max(|t|, angle(R)) of the motion between two rigid poses, in double
// (elastic_fusion_file.cpp:377-398: diff = currPose^-1 * lastPose, the norm
// of its translation and of its rotation's axis-angle vector).
double motion(const float cur[16], const float last[16])
{
    double Ri[9], ti[3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) Ri[3 * r + k] = cur[4 * k + r];
    for (int r = 0; r < 3; ++r)
        ti[r] = -(Ri[3 * r] * cur[3] + Ri[3 * r + 1] * cur[7] + Ri[3 * r + 2] * cur[11]);
    double D[9], t[3];
    for (int r = 0; r < 3; ++r) {
        for (int k = 0; k < 3; ++k)
            D[3 * r + k] = Ri[3 * r] * last[k] + Ri[3 * r + 1] * last[4 + k] + Ri[3 * r + 2] * last[8 + k];
        t[r] = Ri[3 * r] * last[3] + Ri[3 * r + 1] * last[7] + Ri[3 * r + 2] * last[11] + ti[r];
    }
    const double tn = sqrt(t[0] * t[0] + t[1] * t[1] + t[2] * t[2]);
    const double c = std::min(1.0, std::max(-1.0, (D[0] + D[4] + D[8] - 1.0) / 2.0));
    return std::max(tn, acos(c));
}

}  // namespace

extern "C" {

int efloop_default_params(efloop_params* p)
{
    if (!p) return ODOM_ERR_ARG;
    p->depth_cutoff = 3.0f;
    p->max_depth = 20.0f;
    p->confidence = 10.0f;
    p->time_delta = 200;
    p->icp_weight = 10.0f;
    p->so3 = 1;
    p->max_surfels = 4000000;
    p->use_graph = 1;
    p->autotune = 1;
    return ODOM_OK;
}

int efloop_create(int width, int height, float cx, float cy, float fx, float fy, const efloop_params* p,
                  efloop_handle* out)
{
    if (!out || !p) return ODOM_ERR_ARG;
    *out = nullptr;
    efloop_ctx* c = new efloop_ctx();
    c->p = *p;
    c->w = width;
    c->h = height;
    odom_params op;
    int rc = odom_default_params(&op, width, height, cx, cy, fx, fy);
    if (!rc) rc = odom_create(&op, &c->odo);
    if (!rc) rc = splat_create(width, height, cx, cy, fx, fy, 0, &c->pred);
    if (!rc) rc = surfel_create(width, height, cx, cy, fx, fy, p->max_surfels, &c->model);
    if (rc) {
        free_loop(c);
        return rc;
    }
    void* stream = nullptr;
    rc = odom_get_stream(c->odo, &stream);
    if (!rc) rc = splat_set_stream(c->pred, stream);
    if (!rc) rc = surfel_set_stream(c->model, stream);
    if (!rc && p->use_graph) rc = odom_set_graph(c->odo, 1);
    if (!rc && hipMalloc((void**)&c->filtered, (size_t)width * height * 2) != hipSuccess) rc = ODOM_ERR_HIP;
    for (int k = 0; k < 16; ++k) c->pose[k] = (k % 5 == 0) ? 1.0f : 0.0f;
    if (!rc) {
        const float t[3] = {0.0f, 0.0f, 0.0f};
        const float R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
        rc = odom_set_pose(c->odo, t, R);
    }
    if (rc) {
        free_loop(c);
        return rc;
    }
    *out = c;
    return ODOM_OK;
}

int efloop_destroy(efloop_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    odom_synchronize(c->odo);
    free_loop(c);
    return ODOM_OK;
}

int efloop_process(efloop_handle c, const odom_frame_view* rgb, const odom_frame_view* depth, float pose_out[16])
{
    if (!c || !rgb || !depth || !pose_out) return ODOM_ERR_ARG;
    const efloop_params& p = c->p;
    int rc = odom_bilateral_depth(c->odo, depth, p.depth_cutoff, c->filtered, ODOM_MEM_DEVICE);
    if (rc) return rc;
    const odom_frame_view filt = device_view(c->filtered, c->w, c->h, 2, ODOM_FMT_DEPTH_U16);
    c->tick += 1;
    const int tick = c->tick;
    if (tick == 1) {
        rc = surfel_initialise(c->model, rgb, depth, &filt, p.depth_cutoff, tick, p.max_depth);
        if (!rc) rc = odom_init_first_rgb(c->odo, rgb);
        if (rc) return rc;
    } else {
        float last[16];
        memcpy(last, c->pose, sizeof last);
        const bool tracking_ok = c->last_icp_error < 1e-4f;  // the previous call's, as the loop reads it
        void *fv = nullptr, *fn = nullptr, *fi = nullptr;
        rc = splat_get_buffer(c->pred, SPLAT_FILL_VERTEX, &fv);
        if (!rc) rc = splat_get_buffer(c->pred, SPLAT_FILL_NORMAL, &fn);
        if (!rc) rc = splat_get_buffer(c->pred, SPLAT_FILL_IMAGE, &fi);
        if (rc) return rc;
        const odom_frame_view vv = device_view(fv, c->w, c->h, 16, ODOM_FMT_RGBA32F);
        const odom_frame_view nv = device_view(fn, c->w, c->h, 16, ODOM_FMT_RGBA32F);
        const odom_frame_view iv = device_view(fi, c->w, c->h, 4, ODOM_FMT_RGBA8);
        rc = odom_init_icp_model(c->odo, &vv, &nv, p.max_depth, c->pose);
        if (!rc) rc = odom_init_rgb_model(c->odo, &iv);
        if (!rc) rc = odom_init_icp_depth(c->odo, &filt, p.max_depth);
        if (!rc) rc = odom_init_rgb(c->odo, rgb);
        float t[3], R[9];
        for (int r = 0; r < 3; ++r) {
            t[r] = c->pose[4 * r + 3];
            for (int k = 0; k < 3; ++k) R[3 * r + k] = c->pose[4 * r + k];
        }
        if (!rc) rc = odom_track(c->odo, t, R, 0, p.icp_weight, 1, 0, p.so3);
        odom_stats st;  // read back with the pose by odom_track
        if (!rc) rc = slam::odom_stats_last_read(c->odo, &st);
        // launch shapes from the first tracking call's inputs; the timing
        // passes write block partials only, which the next call rewrites
        if (!rc && tick == 2 && p.autotune) rc = odom_autotune(c->odo, 0, p.icp_weight);
        if (rc) return rc;
        c->last_icp_error = st.last_icp_error;
        for (int r = 0; r < 3; ++r) {
            c->pose[4 * r + 3] = t[r];
            for (int k = 0; k < 3; ++k) c->pose[4 * r + k] = R[3 * r + k];
        }
        // weight by velocity (elastic_fusion_file.cpp:390-398)
        const float largest = 0.01f, min_weight = 0.5f;
        float wgt = (float)motion(c->pose, last);
        if (wgt > largest) wgt = largest;
        c->weighting = std::max(1.0f - (wgt / largest), min_weight) * 1.0f;
        if (tracking_ok) {
            rc = surfel_predict_indices(c->model, c->pose, tick, p.max_depth, p.time_delta);
            if (!rc)
                rc = surfel_fuse(c->model, c->pose, tick, rgb, depth, &filt, p.depth_cutoff, p.max_depth,
                                 c->weighting);
            if (!rc) rc = surfel_predict_indices(c->model, c->pose, tick, p.max_depth, p.time_delta);
            if (!rc) rc = surfel_clean(c->model, c->pose, tick, p.confidence, p.time_delta);
            if (rc) return rc;
        }
    }
    // predict() (elastic_fusion_file.cpp:17-44): the active model at the
    // current pose, then the fill-in from the filtered depth and the colour
    const float* mp;
    int n = 0;
    rc = surfel_get_model(c->model, &mp);
    if (!rc) rc = surfel_count(c->model, &n);
    if (!rc)
        rc = splat_combined_predict(c->pred, c->pose, mp, n, ODOM_MEM_DEVICE, p.max_depth, p.confidence, tick, tick,
                                    p.time_delta);
    if (!rc) rc = splat_fill_in(c->pred, &filt, rgb, 0, 0);
    if (rc) return rc;
    memcpy(pose_out, c->pose, sizeof c->pose);
    return ODOM_OK;
}

int efloop_get_handles(efloop_handle c, odom_handle* odom, splat_handle* prediction, surfel_handle* model)
{
    if (!c) return ODOM_ERR_ARG;
    if (odom) *odom = c->odo;
    if (prediction) *prediction = c->pred;
    if (model) *model = c->model;
    return ODOM_OK;
}

int efloop_get_state(efloop_handle c, int* tick, float* last_weighting, float* last_icp_error)
{
    if (!c) return ODOM_ERR_ARG;
    if (tick) *tick = c->tick;
    if (last_weighting) *last_weighting = c->weighting;
    if (last_icp_error) *last_icp_error = c->last_icp_error;
    return ODOM_OK;
}

}  // extern "C"
