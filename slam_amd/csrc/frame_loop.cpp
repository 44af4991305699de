// frame_loop.cpp -- the reference's frame loop without loop closure, behind
// include/slam_loop.h.  Host code only: every stage is a call into the
// odometry, splat and surfel entry points of this library, all on the
// odometry handle's stream.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "odom_internal.h"
#include "slam_cnn.h"
#include "slam_loop.h"

struct efloop_ctx {
    efloop_params p{};
    int w = 0, h = 0;
    odom_handle odo = nullptr;
    splat_handle pred = nullptr;
    surfel_handle model = nullptr;
    uint16_t* filtered = nullptr;    // device, the bilateral output
    float pose[16];
    float last_pose[16];
    int tick = 0;
    float last_icp_error = 0.0f;
    float weighting = 1.0f;
    bool tracking_ok = false;
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

// The frame's work up to tracking: the bilateral filter, then on the first
// frame the model's initialisation and initFirstRGB, on later frames the
// init calls of frame-to-model tracking against the filled-in prediction
// (elastic_fusion_file.cpp:342-374).  *first: this was the first frame.
int loop_begin_frame(efloop_ctx* c, const odom_frame_view* rgb, const odom_frame_view* depth, bool* first)
{
    const efloop_params& p = c->p;
    int rc = odom_bilateral_depth(c->odo, depth, p.depth_cutoff, c->filtered, ODOM_MEM_DEVICE);
    if (rc) return rc;
    const odom_frame_view filt = device_view(c->filtered, c->w, c->h, 2, ODOM_FMT_DEPTH_U16);
    c->tick += 1;
    *first = c->tick == 1;
    if (*first) {
        rc = surfel_initialise(c->model, rgb, depth, &filt, p.depth_cutoff, c->tick, p.max_depth);
        if (!rc) rc = odom_init_first_rgb(c->odo, rgb);
        return rc;
    }
    c->tracking_ok = c->last_icp_error < 1e-4f;  // the previous call's, as the loop reads it
    memcpy(c->last_pose, c->pose, sizeof c->pose);
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
    return rc;
}

// After tracking: the pose, the velocity weighting and, when the previous
// call's ICP error was low enough, the model update
// (elastic_fusion_file.cpp:377-420).
int loop_after_tracking(efloop_ctx* c, const odom_frame_view* rgb, const odom_frame_view* depth, const float t[3],
                        const float R[9], const odom_stats& st)
{
    const efloop_params& p = c->p;
    c->last_icp_error = st.last_icp_error;
    for (int r = 0; r < 3; ++r) {
        c->pose[4 * r + 3] = t[r];
        for (int k = 0; k < 3; ++k) c->pose[4 * r + k] = R[3 * r + k];
    }
    // weight by velocity (elastic_fusion_file.cpp:390-398)
    const float largest = 0.01f, min_weight = 0.5f;
    float wgt = (float)motion(c->pose, c->last_pose);
    if (wgt > largest) wgt = largest;
    c->weighting = std::max(1.0f - (wgt / largest), min_weight) * 1.0f;
    if (!c->tracking_ok) return ODOM_OK;
    const odom_frame_view filt = device_view(c->filtered, c->w, c->h, 2, ODOM_FMT_DEPTH_U16);
    const int tick = c->tick;
    int rc = surfel_predict_indices(c->model, c->pose, tick, p.max_depth, p.time_delta);
    if (!rc)
        rc = surfel_fuse(c->model, c->pose, tick, rgb, depth, &filt, p.depth_cutoff, p.max_depth, c->weighting);
    if (!rc) rc = surfel_predict_indices(c->model, c->pose, tick, p.max_depth, p.time_delta);
    if (!rc) rc = surfel_clean(c->model, c->pose, tick, p.confidence, p.time_delta);
    return rc;
}

// predict() (elastic_fusion_file.cpp:17-44): the active model at the current
// pose, then the fill-in from the filtered depth and the colour.
int loop_predict(efloop_ctx* c, const odom_frame_view* rgb)
{
    const efloop_params& p = c->p;
    const odom_frame_view filt = device_view(c->filtered, c->w, c->h, 2, ODOM_FMT_DEPTH_U16);
    const float* mp;
    int n = 0;
    int rc = surfel_get_model(c->model, &mp);
    if (!rc) rc = surfel_count(c->model, &n);
    if (!rc)
        rc = splat_combined_predict(c->pred, c->pose, mp, n, ODOM_MEM_DEVICE, p.max_depth, p.confidence, c->tick,
                                    c->tick, p.time_delta);
    if (!rc) rc = splat_fill_in(c->pred, &filt, rgb, 0, 0);
    return rc;
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
    p->autotune = 0;
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
    bool first = false;
    int rc = loop_begin_frame(c, rgb, depth, &first);
    if (rc) return rc;
    if (!first) {
        float t[3], R[9];
        for (int r = 0; r < 3; ++r) {
            t[r] = c->pose[4 * r + 3];
            for (int k = 0; k < 3; ++k) R[3 * r + k] = c->pose[4 * r + k];
        }
        rc = odom_track(c->odo, t, R, 0, c->p.icp_weight, 1, 0, c->p.so3);
        odom_stats st;  // read back with the pose by odom_track
        if (!rc) rc = slam::odom_stats_last_read(c->odo, &st);
        // launch shapes from the first tracking call's inputs; the timing
        // passes write block partials only, which the next call rewrites
        if (!rc && c->tick == 2 && c->p.autotune) rc = odom_autotune(c->odo, 0, c->p.icp_weight);
        if (!rc) rc = loop_after_tracking(c, rgb, depth, t, R, st);
        if (rc) return rc;
    }
    rc = loop_predict(c, rgb);
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

// ================================================================ batches of loops
// n sequences of one image size in lockstep: each keeps its own loop (model,
// prediction, odometry handle and stream); their tracking calls run as one
// batched call per frame (odom_track_batch_policy), with the learned start's
// network input and forward for all n sequences in one launch each on the
// batch's stream.  Every sequence's launches are those of the same sequence
// tracked alone (a batch of one): the batch gives each pair the block count
// of a single-pair call and the handles' launch shapes, and the network's
// inference (one workgroup per sample) is the same for a sample in any batch.
struct efloop_batch_s {
    std::vector<efloop_ctx*> loops;
    odom_batch batch = nullptr;
    odom_policy pol{};
    cnn_net net = nullptr;
    float t_scale = 0.0f, r_scale = 0.0f;
    void* x = nullptr;                 // network input, bf16 [n][60][80][4]
    float* pred = nullptr;             // network output, fp32 [n][6]
    std::vector<efloop_counters> counters;
};

namespace {

void free_batch_loop(efloop_batch_s* b)
{
    if (b->batch) odom_batch_destroy(b->batch);
    for (auto* c : b->loops)
        if (c) efloop_destroy(c);
    if (b->net) cnn_destroy(b->net);
    if (b->x) (void)hipFree(b->x);
    if (b->pred) (void)hipFree(b->pred);
    delete b;
}

}  // namespace

extern "C" {

int efloop_batch_create(int n, int width, int height, float cx, float cy, float fx, float fy, const efloop_params* p,
                        const odom_policy* pol, efloop_batch* out)
{
    if (!out || !p || !pol || n < 1 || n > 256) return ODOM_ERR_ARG;
    *out = nullptr;
    efloop_batch_s* b = new efloop_batch_s();
    b->pol = *pol;
    b->counters.assign(n, efloop_counters{});
    b->loops.assign(n, nullptr);
    int rc = ODOM_OK;
    for (int i = 0; i < n && !rc; ++i) rc = efloop_create(width, height, cx, cy, fx, fy, p, &b->loops[i]);
    std::vector<odom_handle> hs(n);
    for (int i = 0; i < n && !rc; ++i) hs[i] = b->loops[i]->odo;
    if (!rc) rc = odom_batch_create(hs.data(), n, &b->batch);
    // the single-pair call's block counts, shapes and fused prologue
    if (!rc) rc = odom_batch_set_blocks(b->batch, 0);
    for (int k = 0; k < 2 && !rc; ++k)
        for (int l = 0; l < 3 && !rc; ++l) {
            int t = 0, q = 0;
            rc = odom_get_step_shape(hs[0], k, l, &t, &q);
            if (!rc) rc = odom_batch_set_step_shape(b->batch, k, l, t, q);
        }
    if (!rc) rc = odom_batch_set_split_solve(b->batch, 0);
    if (!rc) rc = odom_batch_set_graph(b->batch, p->use_graph);
    if (rc) {
        free_batch_loop(b);
        return rc;
    }
    *out = b;
    return ODOM_OK;
}

int efloop_batch_destroy(efloop_batch b)
{
    if (!b) return ODOM_ERR_ARG;
    if (b->batch) odom_batch_synchronize(b->batch);
    free_batch_loop(b);
    return ODOM_OK;
}

int efloop_batch_set_net(efloop_batch b, const float* params, size_t nparams, float t_scale, float r_scale)
{
    if (!b || !params || !(t_scale > 0.0f) || !(r_scale > 0.0f)) return ODOM_ERR_ARG;
    const int n = (int)b->loops.size();
    if (b->loops[0]->w != 640 || b->loops[0]->h != 480) return ODOM_ERR_ARG;  // the network's input size
    int rc = ODOM_OK;
    if (!b->net) {
        rc = cnn_create(n, &b->net);
        if (rc) return rc == CNN_ERR_NODEV ? ODOM_ERR_NODEV : ODOM_ERR_HIP;
        if (hipMalloc(&b->x, (size_t)n * 80 * 60 * 4 * 2) != hipSuccess ||
            hipMalloc((void**)&b->pred, (size_t)n * 6 * sizeof(float)) != hipSuccess)
            return ODOM_ERR_HIP;
        void* stream = nullptr;
        if ((rc = odom_batch_get_stream(b->batch, &stream))) return rc;
        if (cnn_set_stream(b->net, stream) != CNN_OK) return ODOM_ERR_HIP;
    }
    if (cnn_set_params(b->net, params, nparams) != CNN_OK) return ODOM_ERR_ARG;
    b->t_scale = t_scale;
    b->r_scale = r_scale;
    return ODOM_OK;
}

int efloop_batch_process(efloop_batch b, const odom_frame_view* rgb, const odom_frame_view* depth, float* poses_out)
{
    if (!b || !rgb || !depth || !poses_out) return ODOM_ERR_ARG;
    const int n = (int)b->loops.size();
    bool first = false;
    int rc = ODOM_OK;
    for (int i = 0; i < n && !rc; ++i) {
        efloop_ctx* c = b->loops[i];
        rc = loop_begin_frame(c, &rgb[i], &depth[i], &first);
        if (!rc && !first) {
            float t[3], R[9];
            for (int r = 0; r < 3; ++r) {
                t[r] = c->pose[4 * r + 3];
                for (int k = 0; k < 3; ++k) R[3 * r + k] = c->pose[4 * r + k];
            }
            rc = odom_set_pose(c->odo, t, R);
        }
    }
    if (rc) return rc;
    if (!first) {
        // one batched tracking call (the learned start: the network's input
        // and forward for every sequence first, on the batch's stream)
        const float* pred = nullptr;
        if (b->net) {
            rc = odom_batch_warmstart_input(b->batch, b->x);
            if (!rc && cnn_infer(b->net, b->x, n, b->pred) != CNN_OK) rc = ODOM_ERR_HIP;
            pred = b->pred;
        }
        if (!rc) rc = odom_track_batch_policy(b->batch, pred, b->t_scale, b->r_scale, &b->pol, 0,
                                              b->loops[0]->p.icp_weight);
        for (int i = 0; i < n && !rc; ++i) {
            efloop_ctx* c = b->loops[i];
            float t[3], R[9];
            odom_stats st;
            rc = odom_get_pose(c->odo, t, R);
            if (!rc) rc = slam::odom_stats_last_read(c->odo, &st);
            if (rc) break;
            efloop_counters& k = b->counters[i];
            k.frames_tracked += 1;
            k.gn_iterations += st.gn_iterations;
            k.alt_iterations += st.alt_iterations;
            k.predictions = st.seed_calls;
            k.predictions_kept = st.seed_kept;
            rc = loop_after_tracking(c, &rgb[i], &depth[i], t, R, st);
        }
    }
    for (int i = 0; i < n && !rc; ++i) rc = loop_predict(b->loops[i], &rgb[i]);
    if (rc) return rc;
    for (int i = 0; i < n; ++i) memcpy(poses_out + 16 * i, b->loops[i]->pose, sizeof(float) * 16);
    return ODOM_OK;
}

int efloop_batch_get_loop(efloop_batch b, int i, efloop_handle* out)
{
    if (!b || !out || i < 0 || i >= (int)b->loops.size()) return ODOM_ERR_ARG;
    *out = b->loops[i];
    return ODOM_OK;
}

int efloop_batch_get_counters(efloop_batch b, int i, efloop_counters* out)
{
    if (!b || !out || i < 0 || i >= (int)b->loops.size()) return ODOM_ERR_ARG;
    *out = b->counters[i];
    return ODOM_OK;
}

}  // extern "C"
