// odom_host.cpp -- the RGBDOdometry host object behind the C ABI of
// include/slam_odom.h.
//
// It keeps the reference's object model (RGBDOdometryef.{h,cpp}): one object
// owns every pyramid buffer, allocated once in the constructor
// (RGBDOdometryef.cpp:21-111), and keeps the cross-call state (vmaps_tmp,
// lastNextImage, lastA/lastb).  Unlike the reference it never returns to the
// host inside a frame: each init* call and the whole Gauss-Newton loop of
// getIncrementalTransformation are enqueued on the object's stream, the
// per-iteration solve runs on the device, and only the synchronous entry
// points wait.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "host_wait.h"
#include "timing_events.h"
#include "odom_internal.h"
#include "odom_math.h"
#include "slam_odom.h"

using namespace slam;

namespace {


int cdiv(int a, int b) { return (a + b - 1) / b; }

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

int fmt_bpp(int fmt)
{
    switch (fmt) {
        case ODOM_FMT_DEPTH_U16: return 2;
        case ODOM_FMT_RGB8: return 3;
        case ODOM_FMT_RGBA8: return 4;
        case ODOM_FMT_RGBA32F: return 16;
        default: return 0;
    }
}

// Launch shapes of the two Gauss-Newton step kernels per level (threads,
// pixels per thread): the role of the reference's GPUConfig table
// (GPUConfig.h:30-98), for gfx950, the only target of this library.  The
// rows were measured with odom_autotune / odom_batch_autotune on MI355X at
// the benchmark image sizes (the modal choice over repeated runs) and are
// fixed here, so a handle's shapes -- which set the order of its FP32 block
// sums -- depend on its image size only, never on run-time timing.  A size
// without a row takes the row of the nearest pixel count.
struct ShapeRow {
    int w, h;
    StepShape a[kLevels], b[kLevels];   // ICP/association pass, photometric pass; levels 0, 1, 2
};
const ShapeRow kShapeTable[] = {
    {320, 240, {{512, 1}, {256, 1}, {256, 1}}, {{512, 1}, {256, 1}, {256, 1}}},
    {640, 480, {{512, 3}, {512, 1}, {256, 1}}, {{512, 3}, {512, 1}, {256, 1}}},
    {848, 480, {{512, 4}, {512, 1}, {256, 1}}, {{512, 4}, {512, 1}, {256, 1}}},
};
// The batches' row: the 64-pair workload of uneven sizes (16 at 320x240, 32
// at 640x480, 16 at 848x480) with 4 blocks per compute unit.
const ShapeRow kBatchShapes = {0, 0, {{512, 2}, {512, 2}, {256, 1}}, {{512, 3}, {256, 3}, {512, 3}}};

StepShape table_shape(int w, int h, int kernel, int level)
{
    const ShapeRow* best = &kShapeTable[0];
    long long bd = -1;
    for (const ShapeRow& r : kShapeTable) {
        const long long d = std::llabs((long long)r.w * r.h - (long long)w * h);
        if (bd < 0 || d < bd) {
            bd = d;
            best = &r;
        }
    }
    return kernel == 0 ? best->a[level] : best->b[level];
}

StepShape batch_shape(int kernel, int level)
{
    return kernel == 0 ? kBatchShapes.a[level] : kBatchShapes.b[level];
}

}  // namespace

struct odom_ctx {
    odom_params prm{};
    int w[kLevels], h[kLevels], n[kLevels];
    float fx[kLevels], fy[kLevels], cx[kLevels], cy[kLevels];
    hipStream_t stream = nullptr;
    char* arena = nullptr;
    size_t arena_bytes = 0;

    uint16_t* depth[kLevels];
    float *vcur[kLevels], *ncur[kLevels], *vprev[kLevels], *nprev[kLevels];
    float *lastDepth[kLevels], *nextDepth[kLevels];
    uint8_t* lastImage[kLevels];
    uint8_t* img[2][kLevels];         // nextImage = img[parity], lastNextImage = img[parity ^ 1]
    int parity = 0;                   // flipped by the swap of RGBDOdometryef.cpp:585-591
    short2* dxy[kLevels];
    float* zsrc[kLevels];
    float* vz = nullptr;              // z plane of the last vertex input (vmaps_tmp)
    uint16_t* stage_u16 = nullptr;
    uint8_t* stage_rgb = nullptr;
    float4* stage_v4 = nullptr;
    float4* stage_n4 = nullptr;
    uint8_t* dbg_mask = nullptr;
    TrackState* st = nullptr;
    PairDesc* d_pair[2] = {nullptr, nullptr};  // one descriptor per image parity
    float* part_a[2] = {nullptr, nullptr};
    float* part_b[2] = {nullptr, nullptr};
    unsigned int* part_ai[2] = {nullptr, nullptr};
    // the second refinement of a judged call (odom_track_policy): its own
    // state and block partials over the same pyramids
    TrackState* st_alt = nullptr;
    float* alt_a[2] = {nullptr, nullptr};
    float* alt_b[2] = {nullptr, nullptr};
    unsigned int* alt_ai[2] = {nullptr, nullptr};
    PairDesc h_pair_alt[2]{};
    odom_batch_s* self_batch = nullptr;   // this handle's policy calls, as a batch of one
    double* d_out11 = nullptr;
    float* d_c27 = nullptr;
    int maxblk = 0;

    PairDesc h_pair[2]{};
    StepShape shape_a[kLevels], shape_b[kLevels];
    int so3_spin = kSO3SpinDefault;  // polls before an SO3-loop barrier wait gives up; < 0: one launch per iteration
    TrackState h_st{};
    TrackState* h_pin = nullptr;     // pinned landing buffer of the state readback
    bool have_state = false;
    // call-order bookkeeping: which inputs have been provided at least once.
    // vtmp: the model vertex buffer that initRGB/initRGBModel read
    // (vmaps_tmp, RGBDOdometryef.cpp:208-235) has been filled.
    bool icp_cur = false, icp_model = false, rgb_next = false, rgb_last = false, vtmp = false;
    // captured tracking calls (odom_set_graph)
    bool use_graph = false;
    struct Graph { std::vector<int> key; hipGraphExec_t exec; };
    std::vector<Graph> graphs;
    hipError_t last_hip = hipSuccess;
};

namespace {

#define HIPCHK(ctx, x)                                    \
    do {                                                  \
        hipError_t e_ = (x);                              \
        if (e_ != hipSuccess) {                           \
            (ctx)->last_hip = e_;                         \
            fprintf(stderr, "slam_odom: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return ODOM_ERR_HIP;                          \
        }                                                 \
    } while (0)

int check_launch(odom_ctx* c)
{
    hipError_t e = hipPeekAtLastError();
    if (e != hipSuccess) {
        (void)hipGetLastError();
        c->last_hip = e;
        fprintf(stderr, "slam_odom: kernel launch failed: %s\n", hipGetErrorString(e));
        return ODOM_ERR_HIP;
    }
    return ODOM_OK;
}

// Rebuild the host descriptor from the buffers and shapes, and upload it.
uint8_t* next_img(odom_ctx* c, int l) { return c->img[c->parity][l]; }
uint8_t* last_next_img(odom_ctx* c, int l) { return c->img[c->parity ^ 1][l]; }

void drop_graphs(odom_ctx* c);
void drop_self_batch(odom_ctx* c);

int upload_pair(odom_ctx* c)
{
    drop_graphs(c);
    for (int q = 0; q < 2; ++q) {
        PairDesc& D = c->h_pair[q];
        for (int l = 0; l < kLevels; ++l) {
            LevelDesc& L = D.lv[l];
            L.w = c->w[l]; L.h = c->h[l];
            L.fx = c->fx[l]; L.fy = c->fy[l]; L.cx = c->cx[l]; L.cy = c->cy[l];
            L.invFx = 1.0f / c->fx[l];
            L.invFy = 1.0f / c->fy[l];
            L.vcur = c->vcur[l]; L.ncur = c->ncur[l]; L.vprev = c->vprev[l]; L.nprev = c->nprev[l];
            L.zsrc = c->zsrc[l]; L.dxy = c->dxy[l]; L.inext = c->img[q][l];
            L.zlast = c->lastDepth[l]; L.ilast = c->lastImage[l];
            const float sobelScale = (float)(1.0 / pow(2.0, 3));
            const double minGrad[kLevels] = {5, 3, 1};
            L.minScale = (float)(pow(minGrad[l], 2.0) / pow((double)sobelScale, 2.0));
            kmat(L.fx, L.fy, L.cx, L.cy, L.K, L.Ki);
            const StepShape& sa = c->shape_a[l];
            const StepShape& sb = c->shape_b[l];
            D.nblk_a[l] = std::min(kMaxBlk, cdiv(c->n[l], sa.threads * sa.ppt));
            D.nblk_b[l] = std::min(kMaxBlk, cdiv(c->n[l], sb.threads * sb.ppt));
            D.blk0_a[l] = 0;
            D.blk0_b[l] = 0;
        }
        D.st = c->st;
        for (int i = 0; i < 2; ++i) {
            D.part_a[i] = c->part_a[i];
            D.part_b[i] = c->part_b[i];
            D.part_ai[i] = c->part_ai[i];
        }
        D.so3_last = c->img[q ^ 1][2];
        D.so3_next = c->img[q][2];
        HIPCHK(c, hipMemcpyAsync(c->d_pair[q], &D, sizeof(PairDesc), hipMemcpyHostToDevice, c->stream));
        PairDesc& E = c->h_pair_alt[q];
        E = D;
        E.st = c->st_alt;
        for (int i = 0; i < 2; ++i) {
            E.part_a[i] = c->alt_a[i];
            E.part_b[i] = c->alt_b[i];
            E.part_ai[i] = c->alt_ai[i];
        }
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ODOM_OK;
}

// Resolve a frame view to a device pointer and row pitch (bytes); host data
// is copied into `staging` first.
int resolve(odom_ctx* c, const odom_frame_view* v, int fmt_a, int fmt_b, void* staging, const void** dptr,
            int* pitch, int* fmt_out)
{
    if (!v || !v->ptr) return ODOM_ERR_ARG;
    if (v->format != fmt_a && v->format != fmt_b) return ODOM_ERR_ARG;
    if (v->width != c->w[0] || v->height != c->h[0]) return ODOM_ERR_ARG;
    const int bpp = fmt_bpp(v->format);
    const int row = v->width * bpp;
    const int p = v->pitch_bytes ? v->pitch_bytes : row;
    if (p < row) return ODOM_ERR_ARG;
    if (v->format == ODOM_FMT_RGBA32F && (p % 16) != 0) return ODOM_ERR_ARG;
    if (v->format == ODOM_FMT_DEPTH_U16 && (p % 2) != 0) return ODOM_ERR_ARG;
    *fmt_out = v->format;
    if (v->mem == ODOM_MEM_HOST) {
        HIPCHK(c, hipMemcpy2DAsync(staging, row, v->ptr, p, row, v->height, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *dptr = staging;
        *pitch = row;
    } else if (v->mem == ODOM_MEM_DEVICE) {
        *dptr = v->ptr;
        *pitch = p;
    } else {
        return ODOM_ERR_ARG;
    }
    return ODOM_OK;
}

int enqueue_depth_pyramid_maps(odom_ctx* c, float cutoff)
{
    VNLevels a{};
    for (int l = 0; l < kLevels; ++l) {
        a.depth[l] = c->depth[l];
        a.vmap[l] = c->vcur[l];
        a.nmap[l] = c->ncur[l];
        a.w[l] = c->w[l];
        a.h[l] = c->h[l];
        a.fx_inv[l] = 1.f / c->fx[l];
        a.fy_inv[l] = 1.f / c->fy[l];
        a.cx[l] = c->cx[l];
        a.cy[l] = c->cy[l];
    }
    a.cutoff = cutoff;
    launch_depth_pyramid_maps(a, c->stream);
    return check_launch(c);
}

int enqueue_model(odom_ctx* c, const odom_frame_view* v, const odom_frame_view* n, const float* pose, bool model,
                  bool pose_on_device = false)
{
    const void *pv, *pn;
    int pitch_v, pitch_n, f;
    int rc = resolve(c, v, ODOM_FMT_RGBA32F, ODOM_FMT_RGBA32F, c->stage_v4, &pv, &pitch_v, &f);
    if (rc) return rc;
    rc = resolve(c, n, ODOM_FMT_RGBA32F, ODOM_FMT_RGBA32F, c->stage_n4, &pn, &pitch_n, &f);
    if (rc) return rc;
    ModelLevels a{};
    a.v4 = (const float4*)pv;
    a.n4 = (const float4*)pn;
    a.v_pitch = pitch_v / 16;
    a.n_pitch = pitch_n / 16;
    for (int l = 0; l < kLevels; ++l) {
        a.vmap[l] = model ? c->vprev[l] : c->vcur[l];
        a.nmap[l] = model ? c->nprev[l] : c->ncur[l];
        a.w[l] = c->w[l];
        a.h[l] = c->h[l];
    }
    a.vz = c->vz;
    a.transform = model ? 1 : 0;
    a.pose_dev = pose_on_device ? c->st->pose_out : nullptr;
    if (model && !pose_on_device) {
        // Rcam = modelPose.topLeftCorner(3,3), tcam = topRightCorner (RGBDOdometryef.cpp:194-198)
        const float R[9] = {pose[0], pose[1], pose[2], pose[4], pose[5], pose[6], pose[8], pose[9], pose[10]};
        const float t[3] = {pose[3], pose[7], pose[11]};
        memcpy(a.R, R, sizeof R);
        memcpy(a.t, t, sizeof t);
    }
    launch_model_maps(a, c->stream);
    return check_launch(c);
}

// populateRGBDData, RGBDOdometryef.cpp:208-235 (maxDepthRGB 6.0, :37)
int enqueue_populate(odom_ctx* c, const odom_frame_view* rgb, float** dstDepth, uint8_t** dstImage)
{
    const void* p;
    int pitch, fmt;
    int rc = resolve(c, rgb, ODOM_FMT_RGB8, ODOM_FMT_RGBA8, c->stage_rgb, &p, &pitch, &fmt);
    if (rc) return rc;
    PyrJob py{};
    py.w0 = c->w[0];
    py.h0 = c->h[0];
    if (dstDepth) {
        PyrSrc& q = py.p[py.count++];
        q.kind = 0;
        q.gen = 1;
        q.vz = c->vz;
        q.cutoff = 6.0f;
        for (int l = 0; l < kLevels; ++l) q.dst[l] = dstDepth[l];
    }
    PyrSrc& q = py.p[py.count++];
    q.kind = 1;
    q.gen = 2;
    q.rgb = (const uint8_t*)p;
    q.pitch = pitch;
    q.ch = fmt == ODOM_FMT_RGB8 ? 3 : 4;
    for (int l = 0; l < kLevels; ++l) q.dst[l] = dstImage[l];
    launch_pyramid(py, c->stream);
    return check_launch(c);
}

TrackParams make_params(odom_ctx* c, int rgbOnly, float icpWeight, int pyramid, int fastOdom, int so3)
{
    TrackParams P{};
    P.icp = (!rgbOnly && icpWeight > 0) ? 1 : 0;
    P.rgb = (rgbOnly || icpWeight < 100) ? 1 : 0;
    P.rgbOnly = rgbOnly ? 1 : 0;
    P.so3 = so3 ? 1 : 0;
    P.icpWeight = icpWeight;
    P.distThres = c->prm.dist_thresh;
    P.angleThres = c->prm.angle_thresh;
    P.sobelScale = (float)(1.0 / pow(2.0, 3));
    P.maxDepthDelta = 0.07f;
    P.iters[0] = fastOdom ? 3 : 10;
    P.iters[1] = pyramid ? 5 : 0;
    P.iters[2] = pyramid ? 4 : 0;
    return P;
}

// The inputs a track call with these options reads have all been provided.
bool inputs_ready(const odom_ctx* c, const TrackParams& P)
{
    return (!P.icp || (c->icp_cur && c->icp_model)) && (!P.rgb || (c->rgb_next && c->rgb_last));
}

StepLaunch step_launch(odom_ctx* c, int kernel, int level)
{
    StepLaunch L;
    L.pairs = c->d_pair[c->parity];
    L.npairs = 1;
    L.blk_pair = nullptr;
    L.shape = kernel == 0 ? c->shape_a[level] : c->shape_b[level];
    L.total_blocks = kernel == 0 ? c->h_pair[0].nblk_a[level] : c->h_pair[0].nblk_b[level];
    return L;
}

void prep_levels(odom_ctx* c, PrepLevels& a)
{
    for (int l = 0; l < kLevels; ++l) {
        a.img[l] = next_img(c, l);
        a.depth[l] = c->nextDepth[l];
        a.dxy[l] = c->dxy[l];
        a.zsrc[l] = c->zsrc[l];
        a.w[l] = c->w[l];
        a.h[l] = c->h[l];
        a.minScale[l] = c->h_pair[0].lv[l].minScale;
    }
}

// getIncrementalTransformation (RGBDOdometryef.cpp:267-595) without the
// host round trips: derivatives, SO3 loop, begin, 2 passes per Gauss-Newton
// iteration (the solve of iteration k runs in the prologue of pass A of
// iteration k+1), end, image swap.  With `ev`, an event pair brackets every
// step launch (ev[2*i], ev[2*i+1] for launch i, A and B alternating).
void track_launches(odom_ctx* c, const TrackParams& P, hipEvent_t* ev)
{
    PrepLevels a{};
    if (P.rgb) prep_levels(c, a);
    PairDesc* dp = c->d_pair[c->parity];
    if (P.so3) {
        // begin rides on the SO3 loop's end, the derivatives on its spare blocks
        launch_so3(dp, 1, &P, c->so3_spin, c->stream, P.rgb ? &a : nullptr);
    } else {
        if (P.rgb) launch_rgb_prep(a, c->stream);
        launch_track_begin(dp, 1, P, c->stream);
    }
    std::vector<int> sched;
    for (int l = kLevels - 1; l >= 0; --l)
        for (int j = 0; j < P.iters[l]; ++j) sched.push_back(l * 64 + j);
    int e = 0;
    for (size_t k = 0; k < sched.size(); ++k) {
        const int level = sched[k] / 64, iter = sched[k] % 64;
        const int prev = k > 0 ? sched[k - 1] / 64 : -1;
        if (ev) (void)hipEventRecord(ev[e++], c->stream);
        launch_step_a(step_launch(c, 0, level), level, iter, (int)k, prev, P, 0, nullptr, c->stream);
        if (ev) (void)hipEventRecord(ev[e++], c->stream);
        if (P.rgb) {
            if (ev) (void)hipEventRecord(ev[e++], c->stream);
            launch_step_b(step_launch(c, 1, level), level, iter, (int)k, P, 0, c->stream);
            if (ev) (void)hipEventRecord(ev[e++], c->stream);
        }
    }
    const int K = (int)sched.size();
    launch_track_end(dp, 1, P, K, K > 0 ? sched[K - 1] / 64 : 0, c->stream);
}

std::vector<int> track_key(const odom_ctx* c, const TrackParams& P)
{
    int w;
    memcpy(&w, &P.icpWeight, 4);
    return {P.icp, P.rgb, P.rgbOnly, P.so3, P.seeded, P.iters[0], P.iters[1], P.iters[2], w, c->parity};
}

void drop_graphs(odom_ctx* c)
{
    if (!c->graphs.empty()) (void)hipStreamSynchronize(c->stream);
    for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
    c->graphs.clear();
}

// One getIncrementalTransformation's launches, replayed from a hipGraph when
// the handle has graphs on (one graph per option set and image parity,
// captured on first use), else launched one by one.
int enqueue_track(odom_ctx* c, const TrackParams& P, hipEvent_t* ev = nullptr)
{
    if (c->use_graph && !ev) {
        const std::vector<int> key = track_key(c, P);
        hipGraphExec_t exec = nullptr;
        for (auto& g : c->graphs)
            if (g.key == key) exec = g.exec;
        if (!exec) {
            hipGraph_t graph = nullptr;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
            track_launches(c, P, nullptr);
            HIPCHK(c, hipStreamEndCapture(c->stream, &graph));
            const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPCHK(c, e);
            c->graphs.push_back({key, exec});
        }
        HIPCHK(c, hipGraphLaunch(exec, c->stream));
    } else {
        track_launches(c, P, ev);
    }
    // std::swap(lastNextImage[i], nextImage[i]) (RGBDOdometryef.cpp:585-591):
    // later launches take the other descriptor
    if (P.so3) c->parity ^= 1;
    return check_launch(c);
}

int read_state(odom_ctx* c)
{
    HIPCHK(c, hipMemcpyAsync(c->h_pin, c->st, sizeof(TrackState), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, poll_stream(c->stream));
    memcpy(&c->h_st, c->h_pin, sizeof(TrackState));
    c->have_state = true;
    return ODOM_OK;
}

template <class T>
int write_field(odom_ctx* c, T* dev_field, const T* host, size_t count)
{
    HIPCHK(c, hipMemcpyAsync(dev_field, host, sizeof(T) * count, hipMemcpyHostToDevice, c->stream));
    return ODOM_OK;
}

}  // namespace

extern "C" {

int odom_default_params(odom_params* p, int width, int height, float cx, float cy, float fx, float fy)
{
    if (!p) return ODOM_ERR_ARG;
    p->width = width;
    p->height = height;
    p->cx = cx;
    p->cy = cy;
    p->fx = fx;
    p->fy = fy;
    p->dist_thresh = 0.10f;
    p->angle_thresh = sinf(20.f * 3.14159254f / 180.f);
    return ODOM_OK;
}

int odom_device_info(char* name, int name_len, char* arch, int arch_len)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return ODOM_ERR_NODEV;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return ODOM_ERR_NODEV;
    if (name && name_len > 0) snprintf(name, name_len, "%s", prop.name);
    if (arch && arch_len > 0) snprintf(arch, arch_len, "%s", prop.gcnArchName);
    return ODOM_OK;
}

int odom_create(const odom_params* p, odom_handle* out)
{
    if (!p || !out) return ODOM_ERR_ARG;
    if (p->width < 16 || p->height < 16 || p->width > 8192 || p->height > 8192) return ODOM_ERR_ARG;
    char arch[128] = {0};
    if (odom_device_info(nullptr, 0, arch, sizeof arch) != ODOM_OK) return ODOM_ERR_NODEV;
    if (strncmp(arch, "gfx950", 6) != 0) {
        fprintf(stderr, "slam_odom: device is %s; this library is built for gfx950 only\n", arch);
        return ODOM_ERR_NODEV;
    }
    init_device_tables();
    odom_ctx* c = new odom_ctx();
    c->prm = *p;
    for (int l = 0; l < kLevels; ++l) {
        c->w[l] = p->width >> l;
        c->h[l] = p->height >> l;
        c->n[l] = c->w[l] * c->h[l];
        // CameraModel::operator()(level), Camera.h:14-18
        const int div = 1 << l;
        c->fx[l] = p->fx / (float)div;
        c->fy[l] = p->fy / (float)div;
        c->cx[l] = p->cx / (float)div;
        c->cy[l] = p->cy / (float)div;
        c->shape_a[l] = table_shape(p->width, p->height, 0, l);
        c->shape_b[l] = table_shape(p->width, p->height, 1, l);
    }
    c->maxblk = kMaxBlk;
    // one arena for every device buffer
    const size_t N0 = (size_t)c->n[0];
    struct Req { void** dst; size_t bytes; };
    std::vector<Req> reqs;
    for (int l = 0; l < kLevels; ++l) {
        const size_t n = (size_t)c->n[l];
        reqs.push_back({(void**)&c->depth[l], n * 2});
        reqs.push_back({(void**)&c->vcur[l], n * 12});
        reqs.push_back({(void**)&c->ncur[l], n * 12});
        reqs.push_back({(void**)&c->vprev[l], n * 12});
        reqs.push_back({(void**)&c->nprev[l], n * 12});
        reqs.push_back({(void**)&c->lastDepth[l], n * 4});
        reqs.push_back({(void**)&c->nextDepth[l], n * 4});
        reqs.push_back({(void**)&c->lastImage[l], n});
        reqs.push_back({(void**)&c->img[0][l], n});
        reqs.push_back({(void**)&c->img[1][l], n});
        reqs.push_back({(void**)&c->dxy[l], n * 4});
        reqs.push_back({(void**)&c->zsrc[l], n * 4});
    }
    reqs.push_back({(void**)&c->vz, N0 * 4});
    reqs.push_back({(void**)&c->stage_u16, N0 * 2});
    reqs.push_back({(void**)&c->stage_rgb, N0 * 4});
    reqs.push_back({(void**)&c->stage_v4, N0 * 16});
    reqs.push_back({(void**)&c->stage_n4, N0 * 16});
    reqs.push_back({(void**)&c->dbg_mask, N0});
    reqs.push_back({(void**)&c->st, sizeof(TrackState)});
    reqs.push_back({(void**)&c->d_pair[0], sizeof(PairDesc)});
    reqs.push_back({(void**)&c->d_pair[1], sizeof(PairDesc)});
    for (int i = 0; i < 2; ++i) {
        reqs.push_back({(void**)&c->part_a[i], (size_t)c->maxblk * kPartStride * 4});
        reqs.push_back({(void**)&c->part_b[i], (size_t)c->maxblk * kPartStride * 4});
        reqs.push_back({(void**)&c->part_ai[i], (size_t)c->maxblk * 2 * 4});
        reqs.push_back({(void**)&c->alt_a[i], (size_t)c->maxblk * kPartStride * 4});
        reqs.push_back({(void**)&c->alt_b[i], (size_t)c->maxblk * kPartStride * 4});
        reqs.push_back({(void**)&c->alt_ai[i], (size_t)c->maxblk * 2 * 4});
    }
    reqs.push_back({(void**)&c->st_alt, sizeof(TrackState)});
    reqs.push_back({(void**)&c->d_out11, 16 * 8});
    reqs.push_back({(void**)&c->d_c27, 32 * 4});
    size_t total = 0;
    for (auto& r : reqs) total += align_up(r.bytes, 256);
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc((void**)&c->arena, total) != hipSuccess ||
        hipHostMalloc((void**)&c->h_pin, sizeof(TrackState), 0) != hipSuccess) {
        odom_destroy(c);
        return ODOM_ERR_HIP;
    }
    c->arena_bytes = total;
    size_t off = 0;
    for (auto& r : reqs) {
        *r.dst = c->arena + off;
        off += align_up(r.bytes, 256);
    }
    // zero everything once (the reference leaves buffers uninitialised; zero
    // makes a call-order mistake deterministic)
    if (hipMemsetAsync(c->arena, 0, total, c->stream) != hipSuccess) {
        odom_destroy(c);
        return ODOM_ERR_HIP;
    }
    // initial stats (RGBDOdometryef.cpp:26-33) and identity pose
    TrackState s{};
    s.lastICPCount = (float)(p->width * p->height);
    s.lastRGBCount = (float)(p->width * p->height);
    s.lastSO3Count = (float)(p->width * p->height);
    s.pose_out[0] = s.pose_out[5] = s.pose_out[10] = 1.0f;
    for (int k = 0; k < 9; ++k) s.so3R[k] = (k % 4 == 0) ? 1.0 : 0.0;
    for (int k = 0; k < 16; ++k) s.lastRt[k] = (k % 5 == 0) ? 1.0 : 0.0;
    s.judge_level = -1;
    if (hipMemcpyAsync(c->st, &s, sizeof s, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipMemcpyAsync(c->st_alt, &s, sizeof s, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        upload_pair(c) != ODOM_OK) {
        odom_destroy(c);
        return ODOM_ERR_HIP;
    }
    c->h_st = s;
    *out = c;
    return ODOM_OK;
}

int odom_destroy(odom_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_self_batch(c);
    drop_graphs(c);
    if (c->arena) (void)hipFree(c->arena);
    if (c->h_pin) (void)hipHostFree(c->h_pin);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return ODOM_OK;
}

int odom_bilateral_depth(odom_handle c, const odom_frame_view* raw, float maxD, void* dst, int dst_mem)
{
    if (!c || !dst) return ODOM_ERR_ARG;
    const void* p;
    int pitch, fmt;
    int rc = resolve(c, raw, ODOM_FMT_DEPTH_U16, ODOM_FMT_DEPTH_U16, c->stage_u16, &p, &pitch, &fmt);
    if (rc) return rc;
    if (dst_mem == ODOM_MEM_DEVICE) {
        launch_bilateral((const uint16_t*)p, pitch / 2, (uint16_t*)dst, c->w[0], c->h[0], maxD, c->stream);
        return check_launch(c);
    }
    // host destination: filter into scratch (the RGB staging area), then copy out
    uint16_t* tmp = (uint16_t*)c->stage_rgb;
    launch_bilateral((const uint16_t*)p, pitch / 2, tmp, c->w[0], c->h[0], maxD, c->stream);
    rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(dst, tmp, (size_t)c->n[0] * 2, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ODOM_OK;
}

int odom_init_icp_depth(odom_handle c, const odom_frame_view* depth, float cutoff)
{
    if (!c) return ODOM_ERR_ARG;
    const void* p;
    int pitch, fmt;
    int rc = resolve(c, depth, ODOM_FMT_DEPTH_U16, ODOM_FMT_DEPTH_U16, c->stage_u16, &p, &pitch, &fmt);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy2DAsync(c->depth[0], (size_t)c->w[0] * 2, p, pitch, (size_t)c->w[0] * 2, c->h[0],
                               hipMemcpyDeviceToDevice, c->stream));
    rc = enqueue_depth_pyramid_maps(c, cutoff);
    if (rc == ODOM_OK) c->icp_cur = true;
    return rc;
}

int odom_init_icp_raw_depth(odom_handle c, const odom_frame_view* raw, float maxD, float cutoff)
{
    if (!c) return ODOM_ERR_ARG;
    const void* p;
    int pitch, fmt;
    int rc = resolve(c, raw, ODOM_FMT_DEPTH_U16, ODOM_FMT_DEPTH_U16, c->stage_u16, &p, &pitch, &fmt);
    if (rc) return rc;
    launch_bilateral((const uint16_t*)p, pitch / 2, c->depth[0], c->w[0], c->h[0], maxD, c->stream);
    rc = enqueue_depth_pyramid_maps(c, cutoff);
    if (rc == ODOM_OK) c->icp_cur = true;
    return rc;
}

int odom_init_icp_maps(odom_handle c, const odom_frame_view* v, const odom_frame_view* n, float cutoff)
{
    (void)cutoff;
    if (!c) return ODOM_ERR_ARG;
    const int rc = enqueue_model(c, v, n, nullptr, false);
    if (rc == ODOM_OK) c->icp_cur = c->vtmp = true;
    return rc;
}

int odom_init_icp_model(odom_handle c, const odom_frame_view* v, const odom_frame_view* n, float cutoff,
                        const float pose[16])
{
    (void)cutoff;
    if (!c || !pose) return ODOM_ERR_ARG;
    const int rc = enqueue_model(c, v, n, pose, true);
    if (rc == ODOM_OK) c->icp_model = c->vtmp = true;
    return rc;
}

int odom_init_icp_model_current(odom_handle c, const odom_frame_view* v, const odom_frame_view* n, float cutoff)
{
    (void)cutoff;
    if (!c) return ODOM_ERR_ARG;
    const int rc = enqueue_model(c, v, n, nullptr, true, true);
    if (rc == ODOM_OK) c->icp_model = c->vtmp = true;
    return rc;
}

int odom_init_rgb(odom_handle c, const odom_frame_view* rgb)
{
    if (!c) return ODOM_ERR_ARG;
    if (!c->vtmp) return ODOM_ERR_STATE;  // "initICP* must be called before initRGB*"
    uint8_t* dst[kLevels] = {next_img(c, 0), next_img(c, 1), next_img(c, 2)};
    const int rc = enqueue_populate(c, rgb, c->nextDepth, dst);
    if (rc == ODOM_OK) c->rgb_next = true;
    return rc;
}

int odom_init_rgb_model(odom_handle c, const odom_frame_view* rgb)
{
    if (!c) return ODOM_ERR_ARG;
    if (!c->vtmp) return ODOM_ERR_STATE;
    const int rc = enqueue_populate(c, rgb, c->lastDepth, c->lastImage);
    if (rc == ODOM_OK) c->rgb_last = true;
    return rc;
}

int odom_init_first_rgb(odom_handle c, const odom_frame_view* rgb)
{
    if (!c) return ODOM_ERR_ARG;
    uint8_t* dst[kLevels] = {last_next_img(c, 0), last_next_img(c, 1), last_next_img(c, 2)};
    // a sequence starts: no motion yet to extrapolate
    double eye[16];
    for (int k = 0; k < 16; ++k) eye[k] = (k % 5 == 0) ? 1.0 : 0.0;
    if (int rc = write_field(c, c->st->lastRt, eye, 16)) return rc;
    const int zero = 0;
    if (int rc = write_field(c, &c->st->so3_fail, &zero, 1)) return rc;
    return enqueue_populate(c, rgb, nullptr, dst);
}

int odom_set_pose(odom_handle c, const float trans[3], const float rot[9])
{
    if (!c || !trans || !rot) return ODOM_ERR_ARG;
    float pose[12];
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) pose[a * 4 + b] = rot[a * 3 + b];
        pose[a * 4 + 3] = trans[a];
    }
    return write_field(c, c->st->pose_out, pose, 12);
}

int odom_get_pose(odom_handle c, float trans[3], float rot[9])
{
    if (!c || !trans || !rot) return ODOM_ERR_ARG;
    int rc = read_state(c);
    if (rc) return rc;
    if (c->h_st.so3_fail) return ODOM_ERR_HIP;
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) rot[a * 3 + b] = c->h_st.pose_out[a * 4 + b];
        trans[a] = c->h_st.pose_out[a * 4 + 3];
    }
    return ODOM_OK;
}

int odom_set_pose_log(odom_handle c, float* dev_ptr, int capacity)
{
    if (!c || capacity < 0 || (capacity > 0 && !dev_ptr)) return ODOM_ERR_ARG;
    int rc;
    const int zero = 0;
    if ((rc = write_field(c, &c->st->pose_log, &dev_ptr, 1)) || (rc = write_field(c, &c->st->pose_log_cap, &capacity, 1)) ||
        (rc = write_field(c, &c->st->frame_count, &zero, 1)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ODOM_OK;
}

int odom_synchronize(odom_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return ODOM_OK;
}

int odom_loop_check(odom_handle c, double cov_thresh, float icp_count_thresh, float icp_err_thresh, int* ok)
{
    if (!c || !ok) return ODOM_ERR_ARG;
    *ok = 0;
    double cov[36];
    const int rc = odom_get_covariance(c, cov);
    if (rc == ODOM_ERR_STATE) return ODOM_OK;  // singular last system: not accepted
    if (rc) return rc;
    bool cov_ok = true;
    for (int i = 0; i < 6; ++i)
        if (cov[i * 7] > cov_thresh) {
            cov_ok = false;
            break;
        }
    const TrackState& s = c->h_st;             // read by odom_get_covariance
    *ok = (cov_ok && s.lastICPCount > icp_count_thresh && s.lastICPError < icp_err_thresh) ? 1 : 0;
    return ODOM_OK;
}

int odom_track_async(odom_handle c, int rgbOnly, float icpWeight, int pyramid, int fastOdom, int so3)
{
    if (!c) return ODOM_ERR_ARG;
    const TrackParams P = make_params(c, rgbOnly, icpWeight, pyramid, fastOdom, so3);
    if (!P.icp && !P.rgb) return ODOM_ERR_ARG;  // RGBDOdometryef.cpp:558-561
    if (!inputs_ready(c, P)) return ODOM_ERR_STATE;
    return enqueue_track(c, P);
}

int odom_track(odom_handle c, float trans[3], float rot[9], int rgbOnly, float icpWeight, int pyramid,
               int fastOdom, int so3)
{
    if (!c || !trans || !rot) return ODOM_ERR_ARG;
    int rc = odom_set_pose(c, trans, rot);
    if (rc) return rc;
    rc = odom_track_async(c, rgbOnly, icpWeight, pyramid, fastOdom, so3);
    if (rc) return rc;
    return odom_get_pose(c, trans, rot);
}

int odom_set_so3_loop(odom_handle c, int spin_limit)
{
    if (!c) return ODOM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->so3_spin = spin_limit;
    drop_graphs(c);
    drop_self_batch(c);
    return ODOM_OK;
}

int odom_set_graph(odom_handle c, int enable)
{
    if (!c) return ODOM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->use_graph = enable != 0;
    if (!c->use_graph) drop_graphs(c);
    drop_self_batch(c);
    return ODOM_OK;
}

int odom_get_stream(odom_handle c, void** stream)
{
    if (!c || !stream) return ODOM_ERR_ARG;
    *stream = (void*)c->stream;
    return ODOM_OK;
}

int odom_init_warmstart_input(odom_handle c, void* x)
{
    if (!c || !x) return ODOM_ERR_ARG;
    if (c->w[2] != 2 * 80 || c->h[2] != 2 * 60) return ODOM_ERR_ARG;  // 640x480 frames
    if (!(c->icp_cur && c->rgb_last && c->rgb_next)) return ODOM_ERR_STATE;
    cnn_launch_prepare(c->lastDepth[2], c->lastImage[2], c->depth[2], next_img(c, 2), 1, (size_t)c->n[2], x,
                       c->stream);
    return check_launch(c);
}

int odom_track_seeded(odom_handle c, const float* pred, float t_scale, float r_scale, const int iters[3], int rgbOnly,
                      float icpWeight, int select)
{
    if (!c || !pred || !iters || !(t_scale > 0.0f) || !(r_scale > 0.0f)) return ODOM_ERR_ARG;
    TrackParams P = make_params(c, rgbOnly, icpWeight, 1, 0, 0);
    if (!P.icp && !P.rgb) return ODOM_ERR_ARG;
    for (int l = 0; l < kLevels; ++l) {
        if (iters[l] < 0 || iters[l] > 63) return ODOM_ERR_ARG;
        P.iters[l] = iters[l];
    }
    P.seeded = select ? 2 : 1;
    if (!inputs_ready(c, P)) return ODOM_ERR_STATE;
    if (select && !(c->icp_cur && c->icp_model)) return ODOM_ERR_STATE;
    launch_seed_select(c->d_pair[c->parity], 1, pred, t_scale, r_scale, select, P, c->stream);
    return enqueue_track(c, P);
}

int odom_get_stats(odom_handle c, odom_stats* s)
{
    if (!c || !s) return ODOM_ERR_ARG;
    const int rc = read_state(c);
    if (rc) return rc;
    if (c->h_st.so3_fail) return ODOM_ERR_HIP;
    return slam::odom_stats_last_read(c, s);
}

}  // extern "C"

int slam::odom_stats_last_read(odom_handle c, odom_stats* s)
{
    if (!c || !s || !c->have_state) return ODOM_ERR_ARG;
    const TrackState& t = c->h_st;
    s->last_icp_error = t.lastICPError;
    s->last_icp_count = t.lastICPCount;
    s->last_rgb_error = t.lastRGBError;
    s->last_rgb_count = t.lastRGBCount;
    s->last_so3_error = t.lastSO3Error;
    s->last_so3_count = t.lastSO3Count;
    s->so3_iterations = t.so3_iters;
    s->gn_iterations = t.iters_done;
    s->degenerate = t.degenerate;
    s->reverted = t.reverted;
    s->seed_used = t.seed_used;
    s->seed_calls = t.seed_calls;
    s->seed_kept = t.seed_kept;
    s->alt_iterations = t.alt_iters;
    return ODOM_OK;
}

extern "C" {

int odom_get_last_system(odom_handle c, double A[36], double b[6])
{
    if (!c || !A || !b) return ODOM_ERR_ARG;
    int rc = read_state(c);
    if (rc) return rc;
    memcpy(A, c->h_st.lastA, sizeof(double) * 36);
    memcpy(b, c->h_st.lastb, sizeof(double) * 6);
    return ODOM_OK;
}

// lastA^-1 by Gauss-Jordan elimination with partial pivoting (the LU inverse
// of RGBDOdometryef.cpp:599).
int odom_get_covariance(odom_handle c, double cov[36])
{
    if (!c || !cov) return ODOM_ERR_ARG;
    int rc = read_state(c);
    if (rc) return rc;
    double a[36], inv[36];
    memcpy(a, c->h_st.lastA, sizeof a);
    for (int i = 0; i < 36; ++i) inv[i] = (i % 7 == 0) ? 1.0 : 0.0;
    for (int col = 0; col < 6; ++col) {
        int p = col;
        for (int r = col + 1; r < 6; ++r)
            if (fabs(a[r * 6 + col]) > fabs(a[p * 6 + col])) p = r;
        if (a[p * 6 + col] == 0.0) return ODOM_ERR_STATE;
        if (p != col)
            for (int k = 0; k < 6; ++k) {
                std::swap(a[col * 6 + k], a[p * 6 + k]);
                std::swap(inv[col * 6 + k], inv[p * 6 + k]);
            }
        const double d = a[col * 6 + col];
        for (int k = 0; k < 6; ++k) { a[col * 6 + k] /= d; inv[col * 6 + k] /= d; }
        for (int r = 0; r < 6; ++r) {
            if (r == col) continue;
            const double f = a[r * 6 + col];
            for (int k = 0; k < 6; ++k) { a[r * 6 + k] -= f * a[col * 6 + k]; inv[r * 6 + k] -= f * inv[col * 6 + k]; }
        }
    }
    memcpy(cov, inv, sizeof inv);
    return ODOM_OK;
}

// ---------------------------------------------------------------- kernels

int odom_icp_step(odom_handle c, int level, const float Rc[9], const float tc[3], const float Rpi[9],
                  const float tp[3], double out[29], uint8_t* mask)
{
    if (!c || level < 0 || level >= kLevels || !Rc || !tc || !Rpi || !tp || !out) return ODOM_ERR_ARG;
    int rc;
    IterState* S0 = &c->st->it[0];
    const int zero = 0;
    if ((rc = write_field(c, S0->Rcurr, Rc, 9)) || (rc = write_field(c, S0->tcurr, tc, 3)) ||
        (rc = write_field(c, c->st->Rprev_inv, Rpi, 9)) || (rc = write_field(c, c->st->tprev, tp, 3)) ||
        (rc = write_field(c, &S0->skip, &zero, 1)))
        return rc;
    TrackParams P = make_params(c, 0, 100.0f, 1, 0, 0);
    P.icp = 1;
    P.rgb = 0;
    launch_step_a(step_launch(c, 0, level), level, 0, 0, -1, P, 1, mask ? c->dbg_mask : nullptr, c->stream);
    launch_reduce_dbg(c->d_pair[c->parity], 0, level, c->stream);
    if ((rc = check_launch(c))) return rc;
    if ((rc = read_state(c))) return rc;
    memcpy(out, c->h_st.dbg_sums, sizeof(double) * kSE3);
    if (mask) {
        HIPCHK(c, hipMemcpy(mask, c->dbg_mask, (size_t)c->n[level], hipMemcpyDeviceToHost));
        for (int i = 0; i < c->n[level]; ++i) mask[i] = mask[i] & 1u;
    }
    return ODOM_OK;
}

int odom_compute_derivatives(odom_handle c)
{
    if (!c) return ODOM_ERR_ARG;
    PrepLevels a{};
    prep_levels(c, a);
    launch_rgb_prep(a, c->stream);
    return check_launch(c);
}

int odom_rgb_residual(odom_handle c, int level, const float K[9], const float kt[3], int32_t* count,
                      int32_t* sigma, uint8_t* mask)
{
    if (!c || level < 0 || level >= kLevels || !K || !kt || !count || !sigma) return ODOM_ERR_ARG;
    int rc;
    IterState* S0 = &c->st->it[0];
    const int zero = 0;
    if ((rc = write_field(c, S0->krkinv, K, 9)) || (rc = write_field(c, S0->kt, kt, 3)) ||
        (rc = write_field(c, &S0->skip, &zero, 1)))
        return rc;
    TrackParams P = make_params(c, 0, 10.0f, 1, 0, 0);
    P.icp = 0;
    P.rgb = 1;
    launch_step_a(step_launch(c, 0, level), level, 0, 0, -1, P, 1, mask ? c->dbg_mask : nullptr, c->stream);
    launch_reduce_dbg(c->d_pair[c->parity], 0, level, c->stream);
    if ((rc = check_launch(c))) return rc;
    if ((rc = read_state(c))) return rc;
    *count = c->h_st.dbg_count;
    *sigma = c->h_st.dbg_sigma;
    if (mask) {
        HIPCHK(c, hipMemcpy(mask, c->dbg_mask, (size_t)c->n[level], hipMemcpyDeviceToHost));
        for (int i = 0; i < c->n[level]; ++i) mask[i] = (mask[i] >> 1) & 1u;
    }
    return ODOM_OK;
}

int odom_rgb_step(odom_handle c, int level, float sigma, const float K[9], const float kt[3], double out[29])
{
    if (!c || level < 0 || level >= kLevels || !K || !kt || !out) return ODOM_ERR_ARG;
    int rc;
    IterState* S0 = &c->st->it[0];
    if ((rc = write_field(c, S0->krkinv, K, 9)) || (rc = write_field(c, S0->kt, kt, 3)) ||
        (rc = write_field(c, &S0->sigmaVal, &sigma, 1)))
        return rc;
    TrackParams P = make_params(c, 0, 10.0f, 1, 0, 0);
    P.icp = 0;
    P.rgb = 1;
    launch_step_b(step_launch(c, 1, level), level, 0, 0, P, 1, c->stream);
    launch_reduce_dbg(c->d_pair[c->parity], 1, level, c->stream);
    if ((rc = check_launch(c))) return rc;
    if ((rc = read_state(c))) return rc;
    memcpy(out, c->h_st.dbg_sums, sizeof(double) * kSE3);
    return ODOM_OK;
}

int odom_so3_step(odom_handle c, const float H[9], const float Ki[9], const float KR[9], double out[11],
                  uint8_t* mask)
{
    if (!c || !H || !Ki || !KR || !out) return ODOM_ERR_ARG;
    float c27[27];
    memcpy(c27, H, 36);
    memcpy(c27 + 9, Ki, 36);
    memcpy(c27 + 18, KR, 36);
    HIPCHK(c, hipMemcpyAsync(c->d_c27, c27, sizeof c27, hipMemcpyHostToDevice, c->stream));
    launch_so3_step_dbg(c->d_pair[c->parity], c->d_c27, c->d_out11, mask ? c->dbg_mask : nullptr, c->stream);
    int rc = check_launch(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->d_out11, sizeof(double) * kSO3, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (mask) HIPCHK(c, hipMemcpy(mask, c->dbg_mask, (size_t)c->n[2], hipMemcpyDeviceToHost));
    return ODOM_OK;
}

int odom_read_buffer(odom_handle c, const char* name, int level, void* dst, size_t bytes)
{
    if (!c || !name || !dst || level < 0 || level >= kLevels) return ODOM_ERR_ARG;
    const std::string s(name);
    const size_t n = (size_t)c->n[level];
    const void* src = nullptr;
    size_t need = 0;
    if (s == "depth") { src = c->depth[level]; need = n * 2; }
    else if (s == "vmap_curr") { src = c->vcur[level]; need = n * 12; }
    else if (s == "nmap_curr") { src = c->ncur[level]; need = n * 12; }
    else if (s == "vmap_g_prev") { src = c->vprev[level]; need = n * 12; }
    else if (s == "nmap_g_prev") { src = c->nprev[level]; need = n * 12; }
    else if (s == "last_depth") { src = c->lastDepth[level]; need = n * 4; }
    else if (s == "next_depth") { src = c->nextDepth[level]; need = n * 4; }
    else if (s == "last_image") { src = c->lastImage[level]; need = n; }
    else if (s == "next_image") { src = next_img(c, level); need = n; }
    else if (s == "last_next_image") { src = last_next_img(c, level); need = n; }
    else if (s == "dxy") { src = c->dxy[level]; need = n * 4; }
    else if (s == "zsrc") { src = c->zsrc[level]; need = n * 4; }
    else return ODOM_ERR_ARG;
    if (bytes < need) return ODOM_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(dst, src, need, hipMemcpyDeviceToHost));
    return ODOM_OK;
}

int odom_set_step_shape(odom_handle c, int kernel, int level, int threads, int ppt)
{
    if (!c || kernel < 0 || kernel > 1 || level < 0 || level >= kLevels) return ODOM_ERR_ARG;
    const StepShape sh{threads, ppt};
    if (!step_shape_supported(sh)) return ODOM_ERR_ARG;
    if (kernel == 0) c->shape_a[level] = sh;
    else c->shape_b[level] = sh;
    drop_self_batch(c);
    return upload_pair(c);
}

int odom_get_step_shape(odom_handle c, int kernel, int level, int* threads, int* ppt)
{
    if (!c || kernel < 0 || kernel > 1 || level < 0 || level >= kLevels || !threads || !ppt) return ODOM_ERR_ARG;
    const StepShape sh = kernel == 0 ? c->shape_a[level] : c->shape_b[level];
    *threads = sh.threads;
    *ppt = sh.ppt;
    return ODOM_OK;
}

int odom_profile_track(odom_handle c, int rgbOnly, float icpWeight, int pyramid, int fastOdom, int so3,
                       float us_a[3], float us_b[3], int n_a[3], int n_b[3], float* us_total)
{
    if (!c || !us_a || !us_b || !n_a || !n_b || !us_total) return ODOM_ERR_ARG;
    const TrackParams P = make_params(c, rgbOnly, icpWeight, pyramid, fastOdom, so3);
    if (!P.icp && !P.rgb) return ODOM_ERR_ARG;
    if (!inputs_ready(c, P)) return ODOM_ERR_STATE;
    std::vector<int> sched;
    for (int l = kLevels - 1; l >= 0; --l)
        for (int j = 0; j < P.iters[l]; ++j) sched.push_back(l);
    const int nev = (int)sched.size() * (P.rgb ? 4 : 2);
    TimingEvents ev(nev + 2);
    HIPCHK(c, ev.create());
    hipEvent_t t0 = ev[nev], t1 = ev[nev + 1];
    HIPCHK(c, hipEventRecord(t0, c->stream));
    int rc = enqueue_track(c, P, ev.data());
    HIPCHK(c, hipEventRecord(t1, c->stream));
    HIPCHK(c, hipEventSynchronize(t1));
    for (int l = 0; l < kLevels; ++l) { us_a[l] = 0; us_b[l] = 0; n_a[l] = 0; n_b[l] = 0; }
    int e = 0;
    for (size_t k = 0; k < sched.size() && rc == ODOM_OK; ++k) {
        const int l = sched[k];
        float ms = 0;
        HIPCHK(c, hipEventElapsedTime(&ms, ev[e], ev[e + 1]));
        us_a[l] += ms * 1000.0f;
        n_a[l]++;
        e += 2;
        if (P.rgb) {
            HIPCHK(c, hipEventElapsedTime(&ms, ev[e], ev[e + 1]));
            us_b[l] += ms * 1000.0f;
            n_b[l]++;
            e += 2;
        }
    }
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, t0, t1));
    *us_total = ms * 1000.0f;
    for (int l = 0; l < kLevels; ++l) {
        if (n_a[l]) us_a[l] /= (float)n_a[l];
        if (n_b[l]) us_b[l] /= (float)n_b[l];
    }
    return rc;
}

// Mean duration of the step kernels as they run inside tracking calls, per
// level: after a completed call with the same options, the call's last
// iteration at each level is launched `reps` times back to back between one
// event pair.  A relaunch of iteration k reads state slot (k-1) & 1 and the
// partials of parity (k-1) & 1 and writes the other parity, so every relaunch
// repeats the same work.
int odom_profile_steps(odom_handle c, int rgbOnly, float icpWeight, int reps, float us_a[3], float us_b[3])
{
    if (!c || !us_a || !us_b || reps < 1 || reps > 10000) return ODOM_ERR_ARG;
    const TrackParams P = make_params(c, rgbOnly, icpWeight, 1, 0, 1);
    if (!P.icp && !P.rgb) return ODOM_ERR_ARG;
    if (!inputs_ready(c, P)) return ODOM_ERR_STATE;
    std::vector<int> sched;
    for (int l = kLevels - 1; l >= 0; --l)
        for (int j = 0; j < P.iters[l]; ++j) sched.push_back(l * 64 + j);
    TimingEvents ev(2);
    HIPCHK(c, ev.create());
    hipEvent_t e0 = ev[0], e1 = ev[1];
    int rc = ODOM_OK;
    for (int level = 0; level < kLevels && rc == ODOM_OK; ++level) {
        int k = -1;
        for (int q = 0; q < (int)sched.size(); ++q)
            if (sched[q] / 64 == level) k = q;
        us_a[level] = us_b[level] = 0.0f;
        if (k <= 0) continue;
        const int iter = sched[k] % 64, prev = sched[k - 1] / 64;
        for (int pass = 0; pass < (P.rgb ? 2 : 1) && rc == ODOM_OK; ++pass) {
            // the relaunches are captured as one graph and replayed twice (the
            // first replay warms up), so that the host's per-launch cost, a
            // tracing profiler's included, leaves no idle time between them
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
            if (e == hipSuccess) {
                for (int r = 0; r < reps; ++r) {
                    if (pass == 0) launch_step_a(step_launch(c, 0, level), level, iter, k, prev, P, 0, nullptr, c->stream);
                    else launch_step_b(step_launch(c, 1, level), level, iter, k, P, 0, c->stream);
                }
                e = hipStreamEndCapture(c->stream, &graph);
            }
            if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (graph) (void)hipGraphDestroy(graph);
            if (e == hipSuccess) e = hipGraphLaunch(exec, c->stream);
            if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
            if (e == hipSuccess) e = hipGraphLaunch(exec, c->stream);
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            float ms = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (exec) (void)hipGraphExecDestroy(exec);
            if (e != hipSuccess) {
                c->last_hip = e;
                fprintf(stderr, "slam_odom: odom_profile_steps: %s\n", hipGetErrorString(e));
                rc = ODOM_ERR_HIP;
                break;
            }
            (pass == 0 ? us_a : us_b)[level] = ms * 1000.0f / reps;
        }
        if (rc == ODOM_OK) rc = check_launch(c);
    }
    return rc;
}

int odom_autotune(odom_handle c, int rgbOnly, float icpWeight)
{
    if (!c) return ODOM_ERR_ARG;
    TrackParams P = make_params(c, rgbOnly, icpWeight, 1, 0, 0);
    if (!P.icp && !P.rgb) return ODOM_ERR_ARG;
    if (!inputs_ready(c, P)) return ODOM_ERR_STATE;
    const int threads[] = {256, 512};
    const int ppts[] = {1, 2, 3, 4};
    // the events live until the function returns, and on any failure the
    // shapes in place when the call started are restored
    struct Guard {
        odom_ctx* c;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        StepShape a[kLevels], b[kLevels];
        bool done = false;
        ~Guard()
        {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            if (!done) {
                for (int l = 0; l < kLevels; ++l) {
                    c->shape_a[l] = a[l];
                    c->shape_b[l] = b[l];
                }
                drop_self_batch(c);
                (void)upload_pair(c);
            }
        }
    } g{c};
    for (int l = 0; l < kLevels; ++l) {
        g.a[l] = c->shape_a[l];
        g.b[l] = c->shape_b[l];
    }
    HIPCHK(c, hipEventCreate(&g.e0));
    HIPCHK(c, hipEventCreate(&g.e1));
    int rc = odom_compute_derivatives(c);
    if (rc) return rc;
    const int zero = 0;
    if ((rc = write_field(c, &c->st->it[0].skip, &zero, 1))) return rc;
    for (int kernel = 0; kernel < 2; ++kernel) {
        if (kernel == 1 && !P.rgb) break;
        for (int level = 0; level < kLevels; ++level) {
            float best = 1e30f;
            StepShape best_sh = kernel == 0 ? c->shape_a[level] : c->shape_b[level];
            for (int t : threads)
                for (int q : ppts) {
                    if ((rc = odom_set_step_shape(c, kernel, level, t, q))) return rc;
                    const int reps = 20;
                    for (int r = 0; r < reps + 3; ++r) {
                        if (r == 3) HIPCHK(c, hipEventRecord(g.e0, c->stream));
                        if (kernel == 0) launch_step_a(step_launch(c, 0, level), level, 0, 0, -1, P, 1, nullptr, c->stream);
                        else launch_step_b(step_launch(c, 1, level), level, 0, 0, P, 1, c->stream);
                    }
                    HIPCHK(c, hipEventRecord(g.e1, c->stream));
                    HIPCHK(c, hipEventSynchronize(g.e1));
                    float ms = 0.0f;
                    HIPCHK(c, hipEventElapsedTime(&ms, g.e0, g.e1));
                    if ((rc = check_launch(c))) return rc;
                    if (ms < best) { best = ms; best_sh = StepShape{t, q}; }
                }
            if ((rc = odom_set_step_shape(c, kernel, level, best_sh.threads, best_sh.ppt))) return rc;
        }
    }
    g.done = true;
    return ODOM_OK;
}

}  // extern "C"

// ================================================================ batches
// Many handles (frame pairs of any size) tracked by the same launches: one
// launch per pass and Gauss-Newton iteration covers every pair.  A call runs
// one refinement per handle, or two (a judged policy call: the prediction's
// and the constant-velocity start's) through the levels before its judge;
// each refinement is an entry with its own state and block partials.  The
// entries share a launch's blocks in proportion to their pixel counts (at
// most kMaxBlk each), or, with target 0, each gets the block count of a
// single-pair call, which makes a batched entry's sums bit-identical to the
// same pair tracked alone.  The batch holds the descriptor arrays with its
// block split, the block-to-entry tables and a stream (a handle's own policy
// calls run as a batch of one on the handle's stream).
struct BatchPlan {
    int* d_blk[2][kLevels] = {};
    int total[2][kLevels] = {};
    std::vector<int> nblk[2][kLevels], blk0[2][kLevels];   // per entry
    std::map<std::vector<int>, PairDesc*> slots;           // descriptor arrays by the pairs' image parities
    bool built = false;
};

struct odom_batch_s {
    std::vector<odom_ctx*> h;
    hipStream_t stream = nullptr;
    bool own_stream = true;          // false: the single handle's stream, no cross-stream events
    hipEvent_t done = nullptr;
    std::vector<hipEvent_t> ready;
    int target_blocks = 0;           // 0: each entry gets its single-call block count
    StepShape shape[2][kLevels];
    BatchPlan plan[kMaxHyp];         // plan[k]: k + 1 entries per handle
    PairDesc* h_stage = nullptr;     // pinned upload buffer (kMaxHyp entries per handle)
    bool use_graph = false;
    bool split_solve = true;         // k_solve + unfused pass A (else the fused prologue in every block)
    int so3_spin = kSO3SpinDefault;  // the SO3 loop setting of the batch's first handle
    struct G { std::vector<int> key; hipGraphExec_t exec; };
    std::vector<G> graphs;
    hipError_t last_hip = hipSuccess;
};

namespace {

// One batched tracking call.
struct CallSpec {
    TrackParams P{};
    int hyp = 1;                     // refinements per handle before the judge
    int judge = ODOM_JUDGE_NONE;
    bool policy_start = false;       // k_seed_policy sets each entry's start
    const float* pred = nullptr;     // device, 6 floats per handle
    float t_scale = 1.0f, r_scale = 1.0f;
};

void batch_drop_graphs(odom_batch_s* b)
{
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    b->graphs.clear();
}

void plan_free(BatchPlan& pl)
{
    for (auto& kv : pl.slots) (void)hipFree(kv.second);
    pl.slots.clear();
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < kLevels; ++l) {
            if (pl.d_blk[k][l]) (void)hipFree(pl.d_blk[k][l]);
            pl.d_blk[k][l] = nullptr;
            pl.total[k][l] = 0;
            pl.nblk[k][l].clear();
            pl.blk0[k][l].clear();
        }
    pl.built = false;
}

// Split each pass and level's tiles over the entries of plan `hyp` and
// upload the tables.
int plan_build(odom_batch_s* b, int hyp)
{
    BatchPlan& pl = b->plan[hyp - 1];
    plan_free(pl);
    const int n = (int)b->h.size(), ne = n * hyp;
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < kLevels; ++l) {
            const int tile = b->shape[k][l].threads * b->shape[k][l].ppt;
            std::vector<int> tiles(ne);
            long long T = 0;
            for (int e = 0; e < ne; ++e) {
                tiles[e] = cdiv(b->h[e / hyp]->n[l], tile);
                T += tiles[e];
            }
            auto& nb = pl.nblk[k][l];
            auto& b0 = pl.blk0[k][l];
            nb.assign(ne, 1);
            b0.assign(ne, 0);
            int tot = 0;
            for (int e = 0; e < ne; ++e) {
                if (b->target_blocks > 0) {
                    const double share = (double)tiles[e] * b->target_blocks / (double)std::max(T, 1LL);
                    nb[e] = std::max(1, std::min({(int)(share + 0.5), kMaxBlk, tiles[e]}));
                } else {
                    nb[e] = std::max(1, std::min(kMaxBlk, tiles[e]));
                }
                b0[e] = tot;
                tot += nb[e];
            }
            pl.total[k][l] = tot;
            std::vector<int> table(tot);
            for (int e = 0; e < ne; ++e)
                for (int j = 0; j < nb[e]; ++j) table[b0[e] + j] = e;
            hipError_t err = hipMalloc((void**)&pl.d_blk[k][l], sizeof(int) * tot);
            if (err == hipSuccess)
                err = hipMemcpyAsync(pl.d_blk[k][l], table.data(), sizeof(int) * tot, hipMemcpyHostToDevice,
                                     b->stream);
            if (err == hipSuccess) err = hipStreamSynchronize(b->stream);
            if (err != hipSuccess) {
                b->last_hip = err;
                return ODOM_ERR_HIP;
            }
        }
    pl.built = true;
    return ODOM_OK;
}

// After a change of shapes or block target: graphs and plans are rebuilt
// (the one-refinement plan now, the two-refinement plan on first use).
int batch_plan(odom_batch_s* b)
{
    batch_drop_graphs(b);
    for (auto& pl : b->plan) plan_free(pl);
    return plan_build(b, 1);
}

// The descriptor array of plan `hyp` for the pairs' current image parities
// (built once per parity pattern: a batch tracked with so3 alternates
// between two).  Entry hyp * i + j is handle i's state j (0: its own, 1: the
// second refinement's).
int batch_slot(odom_batch_s* b, int hyp, PairDesc** out)
{
    BatchPlan& pl = b->plan[hyp - 1];
    if (!pl.built)
        if (int rc = plan_build(b, hyp)) return rc;
    const int n = (int)b->h.size();
    std::vector<int> key(n);
    for (int i = 0; i < n; ++i) key[i] = b->h[i]->parity;
    auto it = pl.slots.find(key);
    if (it != pl.slots.end()) {
        *out = it->second;
        return ODOM_OK;
    }
    for (int e = 0; e < n * hyp; ++e) {
        const odom_ctx* c = b->h[e / hyp];
        PairDesc D = (e % hyp == 0) ? c->h_pair[key[e / hyp]] : c->h_pair_alt[key[e / hyp]];
        for (int l = 0; l < kLevels; ++l) {
            D.nblk_a[l] = pl.nblk[0][l][e];
            D.blk0_a[l] = pl.blk0[0][l][e];
            D.nblk_b[l] = pl.nblk[1][l][e];
            D.blk0_b[l] = pl.blk0[1][l][e];
        }
        b->h_stage[e] = D;
    }
    PairDesc* d = nullptr;
    hipError_t e = hipMalloc((void**)&d, sizeof(PairDesc) * n * hyp);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d, b->h_stage, sizeof(PairDesc) * n * hyp, hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        if (d) (void)hipFree(d);
        b->last_hip = e;
        return ODOM_ERR_HIP;
    }
    pl.slots[key] = d;
    *out = d;
    return ODOM_OK;
}

StepLaunch plan_launch(odom_batch_s* b, int hyp, PairDesc* dp, int kernel, int level)
{
    const BatchPlan& pl = b->plan[hyp - 1];
    StepLaunch L;
    L.pairs = dp;
    L.npairs = (int)b->h.size() * hyp;
    L.blk_pair = pl.d_blk[kernel][level];
    L.total_blocks = pl.total[kernel][level];
    L.shape = b->shape[kernel][level];
    return L;
}

// The launches of one batched getIncrementalTransformation (the sequence of
// enqueue_track over every pair at once), with the start and judge of a
// policy call.  dp[k]: the descriptor array of plan k + 1 (dp[1] only for
// two refinements).  A coarse judge runs when the level-2 iterations are
// done: the last of them is finished for both refinements (k_solve), the
// judge keeps one state per handle, and the finer levels continue with one
// entry per handle (their first pass A reads the state the solve stored).
void batch_launches(odom_batch_s* b, const CallSpec& cs, PairDesc* const* dp, hipEvent_t* ev)
{
    const TrackParams& P = cs.P;
    const int n = (int)b->h.size();
    int hyp = cs.hyp;
    PairDesc* cur = dp[hyp - 1];
    if (cs.policy_start) launch_seed_policy(cur, n * hyp, hyp, cs.pred, cs.t_scale, cs.r_scale, b->stream);
    if (P.so3) launch_so3(cur, n * hyp, &P, b->so3_spin, b->stream);  // begin rides on the SO3 loop's end
    else launch_track_begin(cur, n * hyp, P, b->stream);
    std::vector<int> sched;
    for (int l = kLevels - 1; l >= 0; --l)
        for (int j = 0; j < P.iters[l]; ++j) sched.push_back(l * 64 + j);
    int e = 0;
    for (size_t k = 0; k < sched.size(); ++k) {
        const int level = sched[k] / 64, iter = sched[k] % 64;
        const int prev = k > 0 ? sched[k - 1] / 64 : -1;
        bool state_ready = false;   // the state entering iteration k is in its slot already
        if (hyp == 2 && cs.judge == ODOM_JUDGE_COARSE && prev == 2 && level < 2) {
            launch_solve(cur, n * 2, level, iter, (int)k, prev, P, b->shape[0][level].threads, b->stream);
            launch_judge(cur, n, P, (int)(k & 1), 2, b->stream);
            hyp = 1;
            cur = dp[0];
            state_ready = true;
        }
        const StepLaunch La = plan_launch(b, hyp, cur, 0, level);
        if (ev) (void)hipEventRecord(ev[e++], b->stream);
        if (state_ready) {
            launch_step_a(La, level, iter, (int)k, prev, P, 1, nullptr, b->stream);
        } else if (b->split_solve) {
            launch_solve(cur, n * hyp, level, iter, (int)k, prev, P, La.shape.threads, b->stream);
            launch_step_a(La, level, iter, (int)k, prev, P, 1, nullptr, b->stream);
        } else {
            launch_step_a(La, level, iter, (int)k, prev, P, 0, nullptr, b->stream);
        }
        if (ev) (void)hipEventRecord(ev[e++], b->stream);
        if (P.rgb) {
            const StepLaunch Lb = plan_launch(b, hyp, cur, 1, level);
            if (ev) (void)hipEventRecord(ev[e++], b->stream);
            launch_step_b(Lb, level, iter, (int)k, P, 0, b->stream);
            if (ev) (void)hipEventRecord(ev[e++], b->stream);
        }
    }
    const int K = (int)sched.size();
    launch_track_end(cur, n * hyp, P, K, K > 0 ? sched[K - 1] / 64 : 0, b->stream);
    // a full judge, or a coarse one whose schedule has no finer level
    if (hyp == 2) launch_judge(cur, n, P, -1, 0, b->stream);
}

#define BCHK(b, x)                                                                                       \
    do {                                                                                                 \
        hipError_t e_ = (x);                                                                             \
        if (e_ != hipSuccess) {                                                                          \
            (b)->last_hip = e_;                                                                          \
            fprintf(stderr, "slam_odom: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), __FILE__, __LINE__); \
            return ODOM_ERR_HIP;                                                                         \
        }                                                                                                \
    } while (0)

// Everything before the batch's own launches: checks, the per-pair
// derivative pass on each handle's stream, the cross-stream waits, and the
// descriptor arrays the call uses.
int batch_prepare(odom_batch_s* b, const CallSpec& cs, PairDesc** dp)
{
    for (auto* c : b->h)
        if (!inputs_ready(c, cs.P)) return ODOM_ERR_STATE;
    for (size_t i = 0; i < b->h.size(); ++i) {
        odom_ctx* c = b->h[i];
        if (cs.P.rgb) {
            PrepLevels a{};
            prep_levels(c, a);
            launch_rgb_prep(a, c->stream);
            if (int rc = check_launch(c)) return rc;
        }
        if (b->own_stream) {
            BCHK(b, hipEventRecord(b->ready[i], c->stream));
            BCHK(b, hipStreamWaitEvent(b->stream, b->ready[i], 0));
        }
    }
    dp[0] = dp[1] = nullptr;
    if (int rc = batch_slot(b, 1, &dp[0])) return rc;
    if (cs.hyp == 2)
        if (int rc = batch_slot(b, 2, &dp[1])) return rc;
    return ODOM_OK;
}

// Handles continue after the batch: their streams wait for it, and the image
// swap of RGBDOdometryef.cpp:585-591 flips each pair's parity.
int batch_finish(odom_batch_s* b, const TrackParams& P)
{
    if (b->own_stream) BCHK(b, hipEventRecord(b->done, b->stream));
    for (auto* c : b->h) {
        if (b->own_stream) BCHK(b, hipStreamWaitEvent(c->stream, b->done, 0));
        if (P.so3) c->parity ^= 1;
        c->have_state = false;
    }
    return ODOM_OK;
}

template <class T>
int bits(T v)
{
    static_assert(sizeof(T) == 4, "32-bit field");
    int w;
    memcpy(&w, &v, 4);
    return w;
}

std::vector<int> graph_key(const CallSpec& cs, const std::vector<int>& parities)
{
    const TrackParams& P = cs.P;
    std::vector<int> k = {P.icp, P.rgb, P.rgbOnly, P.so3, P.iters[0], P.iters[1], P.iters[2], bits(P.icpWeight),
                          P.seeded, bits(P.convT), bits(P.convR), bits(P.judgeTol), cs.hyp, cs.judge,
                          cs.policy_start ? 1 : 0, bits(cs.t_scale), bits(cs.r_scale)};
    uint64_t pp = (uint64_t)(uintptr_t)cs.pred;
    k.push_back((int)(pp & 0xffffffffu));
    k.push_back((int)(pp >> 32));
    k.insert(k.end(), parities.begin(), parities.end());
    return k;
}

// One batched call: prepared, then launched directly or replayed from the
// hipGraph captured on its first use.
int batch_call(odom_batch_s* b, const CallSpec& cs)
{
    PairDesc* dp[2];
    int rc = batch_prepare(b, cs, dp);
    if (rc) return rc;
    if (!b->use_graph) {
        batch_launches(b, cs, dp, nullptr);
        hipError_t e = hipPeekAtLastError();
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b->last_hip = e;
            return ODOM_ERR_HIP;
        }
    } else {
        std::vector<int> par;
        for (auto* c : b->h) par.push_back(c->parity);
        const std::vector<int> key = graph_key(cs, par);
        hipGraphExec_t exec = nullptr;
        for (auto& g : b->graphs)
            if (g.key == key) exec = g.exec;
        if (!exec) {
            hipGraph_t graph = nullptr;
            BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
            batch_launches(b, cs, dp, nullptr);
            BCHK(b, hipStreamEndCapture(b->stream, &graph));
            hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            BCHK(b, e);
            b->graphs.push_back({key, exec});
        }
        BCHK(b, hipGraphLaunch(exec, b->stream));
    }
    return batch_finish(b, cs.P);
}

// The TrackParams and call shape of a policy call.
int policy_spec(odom_ctx* c0, const float* pred, float t_scale, float r_scale, const odom_policy* pol, int rgbOnly,
                float icpWeight, CallSpec& cs)
{
    if (!pol) return ODOM_ERR_ARG;
    if (pol->start != ODOM_START_PREVIOUS && pol->start != ODOM_START_VELOCITY) return ODOM_ERR_ARG;
    if (pol->judge < ODOM_JUDGE_NONE || pol->judge > ODOM_JUDGE_FULL) return ODOM_ERR_ARG;
    if (pred && !(t_scale > 0.0f && r_scale > 0.0f)) return ODOM_ERR_ARG;
    if (!(pol->conv_t >= 0.0f) || !(pol->conv_r >= 0.0f) || !(pol->judge_tol >= 0.0f) || pol->judge_tol >= 1.0f)
        return ODOM_ERR_ARG;
    // the SO3 pre-alignment seeds the reference's start only
    if (pol->so3 && (pred || pol->start != ODOM_START_PREVIOUS)) return ODOM_ERR_ARG;
    cs.P = make_params(c0, rgbOnly, icpWeight, 1, 0, pol->so3);
    if (!cs.P.icp && !cs.P.rgb) return ODOM_ERR_ARG;
    for (int l = 0; l < kLevels; ++l) {
        if (pol->iters[l] < 0 || pol->iters[l] > 63) return ODOM_ERR_ARG;
        cs.P.iters[l] = pol->iters[l];
    }
    cs.P.convT = pol->conv_t;
    cs.P.convR = pol->conv_t > 0.0f ? pol->conv_r : 0.0f;
    cs.P.judgeTol = pol->judge_tol;
    cs.policy_start = pred != nullptr || pol->start == ODOM_START_VELOCITY;
    cs.P.seeded = cs.policy_start ? 3 : 0;
    cs.pred = pred;
    cs.t_scale = pred ? t_scale : 1.0f;
    cs.r_scale = pred ? r_scale : 1.0f;
    cs.judge = pred ? pol->judge : ODOM_JUDGE_NONE;
    cs.hyp = (pred && pol->judge != ODOM_JUDGE_NONE) ? 2 : 1;
    return ODOM_OK;
}

int batch_new(const odom_handle* handles, int n, bool own_stream, odom_batch_s** out);

// The handle's own batch of one for its policy calls: the handle's stream,
// shapes and graph setting, the single-call block counts and the fused
// prologue, so its launches are those of odom_track_async.
int self_batch(odom_ctx* c, odom_batch_s** out)
{
    if (!c->self_batch) {
        odom_batch_s* b = nullptr;
        if (int rc = batch_new(&c, 1, false, &b)) return rc;
        c->self_batch = b;
    }
    *out = c->self_batch;
    return ODOM_OK;
}

}  // namespace

namespace {

void drop_self_batch(odom_ctx* c)
{
    if (c->self_batch) {
        odom_batch_destroy(c->self_batch);
        c->self_batch = nullptr;
    }
}

int batch_new(const odom_handle* handles, int n, bool own_stream, odom_batch_s** out)
{
    odom_batch_s* b = new odom_batch_s();
    b->h.assign(handles, handles + n);
    b->own_stream = own_stream;
    b->so3_spin = handles[0]->so3_spin;
    if (own_stream) {
        int dev = 0, cus = 256;
        if (hipGetDevice(&dev) == hipSuccess) {
            hipDeviceProp_t prop;
            if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
                cus = prop.multiProcessorCount;
        }
        b->target_blocks = 4 * cus;
        for (int k = 0; k < 2; ++k)
            for (int l = 0; l < kLevels; ++l) b->shape[k][l] = batch_shape(k, l);
    } else {
        odom_ctx* c = handles[0];
        b->target_blocks = 0;
        b->split_solve = false;
        b->use_graph = c->use_graph;
        b->stream = c->stream;
        for (int l = 0; l < kLevels; ++l) {
            b->shape[0][l] = c->shape_a[l];
            b->shape[1][l] = c->shape_b[l];
        }
    }
    b->ready.resize(n, nullptr);
    bool ok = hipHostMalloc((void**)&b->h_stage, sizeof(PairDesc) * n * kMaxHyp, 0) == hipSuccess;
    if (ok && own_stream)
        ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->done, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && own_stream && i < n; ++i)
        ok = hipEventCreateWithFlags(&b->ready[i], hipEventDisableTiming) == hipSuccess;
    if (!ok || batch_plan(b) != ODOM_OK) {
        odom_batch_destroy(b);
        return ODOM_ERR_HIP;
    }
    *out = b;
    return ODOM_OK;
}

}  // namespace

extern "C" {

int odom_default_policy(odom_policy* p)
{
    if (!p) return ODOM_ERR_ARG;
    p->iters[0] = 10;
    p->iters[1] = 5;
    p->iters[2] = 4;
    p->conv_t = 0.0f;
    p->conv_r = 0.0f;
    p->start = ODOM_START_PREVIOUS;
    p->so3 = 1;
    p->judge = ODOM_JUDGE_FULL;
    p->judge_tol = 0.01f;
    return ODOM_OK;
}

int odom_track_policy(odom_handle c, const float* pred, float t_scale, float r_scale, const odom_policy* pol,
                      int rgbOnly, float icpWeight)
{
    if (!c) return ODOM_ERR_ARG;
    CallSpec cs;
    int rc = policy_spec(c, pred, t_scale, r_scale, pol, rgbOnly, icpWeight, cs);
    if (rc) return rc;
    if (!inputs_ready(c, cs.P)) return ODOM_ERR_STATE;
    odom_batch_s* b = nullptr;
    if ((rc = self_batch(c, &b))) return rc;
    if ((rc = batch_call(b, cs))) return rc;
    return check_launch(c);
}

int odom_batch_create(const odom_handle* handles, int n, odom_batch* out)
{
    if (!handles || !out || n <= 0 || n > 4096) return ODOM_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < n; ++i) {
        if (!handles[i]) return ODOM_ERR_ARG;
        if (handles[i]->prm.dist_thresh != handles[0]->prm.dist_thresh ||
            handles[i]->prm.angle_thresh != handles[0]->prm.angle_thresh)
            return ODOM_ERR_ARG;  // one TrackParams serves every pair of a launch
        for (int j = 0; j < i; ++j)
            if (handles[j] == handles[i]) return ODOM_ERR_ARG;
    }
    return batch_new(handles, n, true, out);
}

int odom_batch_destroy(odom_batch b)
{
    if (!b) return ODOM_ERR_ARG;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graphs(b);
    for (auto& pl : b->plan) plan_free(pl);
    for (auto e : b->ready)
        if (e) (void)hipEventDestroy(e);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->h_stage) (void)hipHostFree(b->h_stage);
    if (b->stream && b->own_stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return ODOM_OK;
}

int odom_batch_set_blocks(odom_batch b, int target_blocks)
{
    if (!b || target_blocks < 0 || target_blocks > 1 << 20) return ODOM_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->stream));
    b->target_blocks = target_blocks;
    return batch_plan(b);
}

int odom_batch_set_step_shape(odom_batch b, int kernel, int level, int threads, int ppt)
{
    if (!b || kernel < 0 || kernel > 1 || level < 0 || level >= kLevels) return ODOM_ERR_ARG;
    if (!step_shape_supported(StepShape{threads, ppt})) return ODOM_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->stream));
    b->shape[kernel][level] = StepShape{threads, ppt};
    return batch_plan(b);
}

int odom_batch_get_blocks(odom_batch b, int kernel, int level, int* total_blocks)
{
    if (!b || !total_blocks || kernel < 0 || kernel > 1 || level < 0 || level >= kLevels) return ODOM_ERR_ARG;
    *total_blocks = b->plan[0].total[kernel][level];
    return ODOM_OK;
}

int odom_batch_set_split_solve(odom_batch b, int enable)
{
    if (!b) return ODOM_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->stream));
    b->split_solve = enable != 0;
    batch_drop_graphs(b);
    return ODOM_OK;
}

int odom_batch_set_graph(odom_batch b, int enable)
{
    if (!b) return ODOM_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->stream));
    b->use_graph = enable != 0;
    if (!b->use_graph) batch_drop_graphs(b);
    return ODOM_OK;
}

int odom_batch_get_stream(odom_batch b, void** stream)
{
    if (!b || !stream) return ODOM_ERR_ARG;
    *stream = (void*)b->stream;
    return ODOM_OK;
}

int odom_track_batch(odom_batch b, int rgbOnly, float icpWeight, int pyramid, int fastOdom, int so3)
{
    if (!b) return ODOM_ERR_ARG;
    CallSpec cs;
    cs.P = make_params(b->h[0], rgbOnly, icpWeight, pyramid, fastOdom, so3);
    if (!cs.P.icp && !cs.P.rgb) return ODOM_ERR_ARG;  // RGBDOdometryef.cpp:558-561
    return batch_call(b, cs);
}

int odom_track_batch_policy(odom_batch b, const float* pred, float t_scale, float r_scale, const odom_policy* pol,
                            int rgbOnly, float icpWeight)
{
    if (!b) return ODOM_ERR_ARG;
    CallSpec cs;
    const int rc = policy_spec(b->h[0], pred, t_scale, r_scale, pol, rgbOnly, icpWeight, cs);
    if (rc) return rc;
    return batch_call(b, cs);
}

int odom_batch_warmstart_input(odom_batch b, void* x)
{
    if (!b || !x) return ODOM_ERR_ARG;
    for (size_t i = 0; i < b->h.size(); ++i) {
        odom_ctx* c = b->h[i];
        if (c->w[2] != 2 * 80 || c->h[2] != 2 * 60) return ODOM_ERR_ARG;  // 640x480 frames
        if (!(c->icp_cur && c->rgb_last && c->rgb_next)) return ODOM_ERR_STATE;
    }
    for (size_t i = 0; i < b->h.size(); ++i) {
        odom_ctx* c = b->h[i];
        if (b->own_stream) {
            BCHK(b, hipEventRecord(b->ready[i], c->stream));
            BCHK(b, hipStreamWaitEvent(b->stream, b->ready[i], 0));
        }
        cnn_launch_prepare(c->lastDepth[2], c->lastImage[2], c->depth[2], next_img(c, 2), 1, (size_t)c->n[2],
                           (char*)x + i * (size_t)(80 * 60 * 4 * 2), b->stream);
    }
    hipError_t e = hipPeekAtLastError();
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b->last_hip = e;
        return ODOM_ERR_HIP;
    }
    return ODOM_OK;
}

// Times every supported shape of the two step kernels per level on the
// batch's current inputs (pixel passes only, with the state of the last
// batched call) and keeps the fastest.
int odom_batch_autotune(odom_batch b, int rgbOnly, float icpWeight)
{
    if (!b) return ODOM_ERR_ARG;
    TrackParams P = make_params(b->h[0], rgbOnly, icpWeight, 1, 0, 0);
    if (!P.icp && !P.rgb) return ODOM_ERR_ARG;
    for (auto* c : b->h)
        if (!inputs_ready(c, P)) return ODOM_ERR_STATE;
    BCHK(b, hipStreamSynchronize(b->stream));
    for (auto* c : b->h) {
        BCHK(b, hipStreamSynchronize(c->stream));
        const int zero = 0;
        if (int rc = write_field(c, &c->st->it[0].skip, &zero, 1)) return rc;
        BCHK(b, hipStreamSynchronize(c->stream));
    }
    TimingEvents ev(2);
    BCHK(b, ev.create());
    hipEvent_t e0 = ev[0], e1 = ev[1];
    const int threads[] = {256, 512};
    const int ppts[] = {1, 2, 3, 4};
    int rc = ODOM_OK;
    for (int kernel = 0; kernel < 2 && rc == ODOM_OK; ++kernel) {
        if (kernel == 1 && !P.rgb) break;
        for (int level = 0; level < kLevels && rc == ODOM_OK; ++level) {
            float best = 1e30f;
            StepShape best_sh = b->shape[kernel][level];
            for (int t : threads)
                for (int q : ppts) {
                    b->shape[kernel][level] = StepShape{t, q};
                    if ((rc = batch_plan(b))) break;
                    PairDesc* dp = nullptr;
                    if ((rc = batch_slot(b, 1, &dp))) break;
                    const StepLaunch Ls = plan_launch(b, 1, dp, kernel, level);
                    const int reps = 10;
                    for (int r = 0; r < reps + 2; ++r) {
                        if (r == 2) (void)hipEventRecord(e0, b->stream);
                        if (kernel == 0) launch_step_a(Ls, level, 0, 0, -1, P, 1, nullptr, b->stream);
                        else launch_step_b(Ls, level, 0, 0, P, 1, b->stream);
                    }
                    (void)hipEventRecord(e1, b->stream);
                    hipError_t e = hipEventSynchronize(e1);
                    if (e == hipSuccess) e = hipPeekAtLastError();
                    if (e != hipSuccess) {
                        (void)hipGetLastError();
                        b->last_hip = e;
                        rc = ODOM_ERR_HIP;
                        break;
                    }
                    float ms = 0.0f;
                    (void)hipEventElapsedTime(&ms, e0, e1);
                    if (ms < best) {
                        best = ms;
                        best_sh = StepShape{t, q};
                    }
                }
            b->shape[kernel][level] = best_sh;
        }
    }
    if (rc) return rc;
    return batch_plan(b);
}

int odom_batch_get_step_shape(odom_batch b, int kernel, int level, int* threads, int* ppt)
{
    if (!b || !threads || !ppt || kernel < 0 || kernel > 1 || level < 0 || level >= kLevels) return ODOM_ERR_ARG;
    *threads = b->shape[kernel][level].threads;
    *ppt = b->shape[kernel][level].ppt;
    return ODOM_OK;
}

int odom_batch_synchronize(odom_batch b)
{
    if (!b) return ODOM_ERR_ARG;
    BCHK(b, hipStreamSynchronize(b->stream));
    return ODOM_OK;
}

int odom_batch_profile_track(odom_batch b, int rgbOnly, float icpWeight, int pyramid, int fastOdom, int so3,
                             float us_a[3], float us_b[3], int n_a[3], int n_b[3], float* us_total)
{
    if (!b || !us_a || !us_b || !n_a || !n_b || !us_total) return ODOM_ERR_ARG;
    CallSpec cs;
    cs.P = make_params(b->h[0], rgbOnly, icpWeight, pyramid, fastOdom, so3);
    if (!cs.P.icp && !cs.P.rgb) return ODOM_ERR_ARG;
    PairDesc* dp[2];
    int rc = batch_prepare(b, cs, dp);
    if (rc) return rc;
    const TrackParams& P = cs.P;
    std::vector<int> sched;
    for (int l = kLevels - 1; l >= 0; --l)
        for (int j = 0; j < P.iters[l]; ++j) sched.push_back(l);
    TimingEvents ev(sched.size() * (P.rgb ? 4 : 2) + 2);
    BCHK(b, ev.create());
    BCHK(b, hipEventRecord(ev[ev.size() - 2], b->stream));
    batch_launches(b, cs, dp, ev.data());
    BCHK(b, hipEventRecord(ev[ev.size() - 1], b->stream));
    BCHK(b, hipEventSynchronize(ev[ev.size() - 1]));
    for (int l = 0; l < kLevels; ++l) { us_a[l] = 0; us_b[l] = 0; n_a[l] = 0; n_b[l] = 0; }
    int e = 0;
    for (size_t k = 0; k < sched.size(); ++k) {
        const int l = sched[k];
        float ms = 0;
        BCHK(b, hipEventElapsedTime(&ms, ev[e], ev[e + 1]));
        us_a[l] += ms * 1000.0f;
        n_a[l]++;
        e += 2;
        if (P.rgb) {
            BCHK(b, hipEventElapsedTime(&ms, ev[e], ev[e + 1]));
            us_b[l] += ms * 1000.0f;
            n_b[l]++;
            e += 2;
        }
    }
    for (int l = 0; l < kLevels; ++l) {
        if (n_a[l]) us_a[l] /= n_a[l];
        if (n_b[l]) us_b[l] /= n_b[l];
    }
    float ms = 0;
    BCHK(b, hipEventElapsedTime(&ms, ev[ev.size() - 2], ev[ev.size() - 1]));
    *us_total = ms * 1000.0f;
    return batch_finish(b, P);
}

}  // extern "C"
