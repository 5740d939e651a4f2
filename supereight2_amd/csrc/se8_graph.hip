// Frame graphs: one frame of the tracking loop of the reference app (main.cpp:98-162) —
// Tracker::track from the previous pose, MapIntegrator::integrateDepth and raycast_volume at the
// tracked pose — captured once as a HIP graph of kernel nodes and replayed per frame, for callers
// that capture or replay work (a frame is 25 launches on maps with the dense tables).  Measured on
// the 640x480 tracking workload, a replayed frame takes as long on the device as the same frame's
// stream launches: empty kernels replay about 1 us faster per dependent launch, but a launch of this
// frame's kernels costs the same either way.
//
// The capture runs the same entry points as the eager calls (se8_track_dev, se8_integrate_depth_dev,
// se8_raycast_dev) with placeholder inputs.  The six launches whose arguments change per frame report
// their node (graph_note_launch); a launch then sets those arguments on the instantiated graph and
// replays it.  Two instantiations alternate by frame parity: instantiation p raycasts into vertex /
// normal pair p and tracks against pair 1 - p, and before its arguments are set the previous replay
// of the same instantiation has finished (the host runs at most two frames ahead).
#include <vector>

#include "se8_internal.hpp"

namespace se8 {

// arguments of the per-frame fields inside the pyramid's and the final ICP launch's argument
// structs (se8_track.hip)
size_t graph_track_args_bytes(int kind);
void graph_track_args_fields(int kind, void* args, void** f0, void** f1);

namespace {

struct GraphRecorder {
    hipGraphNode_t node[kNumGraphNodes]{};
    int count[kNumGraphNodes]{};
    bool failed = false;
};

thread_local GraphRecorder* t_recorder = nullptr;

// kernel arguments of the multi-argument nodes and the position of their per-frame arguments
constexpr int kArgsFrameBegin = 4, kArgsFrustumCarve = 9, kArgsUpdate = 7, kArgsRaycast = 7;
constexpr int kPosFrameBeginPose = 2, kPosCarveDepth = 3, kPosUpdateDepth = 4, kPosUpdateTs = 5, kPosRaycastPose = 3;
constexpr int kMaxArgs = 9;
constexpr int kTsPlaceholder = 0x5e8c0de;

int arg_count(int kind)
{
    switch (kind) {
    case kNodeFrameBegin:
        return kArgsFrameBegin;
    case kNodeFrustumCarve:
        return kArgsFrustumCarve;
    case kNodeUpdate:
        return kArgsUpdate;
    case kNodeRaycast:
        return kArgsRaycast;
    default:
        return 1;
    }
}

} // namespace

void graph_note_launch(int kind, hipStream_t s)
{
    GraphRecorder* r = t_recorder;
    if (!r)
        return;
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    const hipGraphNode_t* deps = nullptr;
    size_t n = 0;
    if (hipStreamGetCaptureInfo_v2(s, &st, nullptr, nullptr, &deps, &n) != hipSuccess
        || st != hipStreamCaptureStatusActive || n != 1 || !deps) {
        r->failed = true;
        return;
    }
    r->node[kind] = deps[0];
    r->count[kind]++;
}

} // namespace se8

using namespace se8;

struct se8_frame_graph {
    se8_tracker* tr = nullptr;
    se8_map* map = nullptr;
    se8_pinhole sensor{};
    float* points[2]{};
    float* normals[2]{};
    DeviceBuffer ph_depth, ph_pose, ph_ok; // the inputs the capture was taken with
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph[2]{};
    hipGraphExec_t exec[2]{};
    hipGraphNode_t node[2][kNumGraphNodes]{};
    hipKernelNodeParams tmpl[2][kNumGraphNodes]{};
    std::vector<uint8_t> struct_args[2][kNumGraphNodes]; // pyramid / final ICP argument structs
    hipEvent_t done[2]{};
    bool pending[2]{};
    int64_t launches = 0;
};

namespace {

void destroy_graph(se8_frame_graph* g)
{
    for (int p = 0; p < 2; p++) {
        if (g->pending[p])
            (void) hipEventSynchronize(g->done[p]);
        if (g->exec[p])
            (void) hipGraphExecDestroy(g->exec[p]);
        if (g->graph[p])
            (void) hipGraphDestroy(g->graph[p]);
        if (g->done[p])
            (void) hipEventDestroy(g->done[p]);
    }
    if (g->cap_stream)
        (void) hipStreamDestroy(g->cap_stream);
    dev_free(g->ph_depth);
    dev_free(g->ph_pose);
    dev_free(g->ph_ok);
    delete g;
}

const float* pose_in_ph(const se8_frame_graph* g)
{
    return (const float*) g->ph_pose.ptr;
}

float* pose_out_ph(const se8_frame_graph* g)
{
    return (float*) g->ph_pose.ptr + 16;
}

// Checks that each per-frame argument of the captured nodes holds the placeholder it was captured
// with, so that a launch replaces exactly those arguments.
bool check_capture(se8_frame_graph* g, int p)
{
    const float* depth = (const float*) g->ph_depth.ptr;
    for (int k = 0; k < kNumGraphNodes; k++) {
        const hipKernelNodeParams& t = g->tmpl[p][k];
        if (!t.kernelParams || t.extra)
            return false;
        void** a = t.kernelParams;
        switch (k) {
        case kNodePyramid:
        case kNodeIcpFinish: {
            const size_t bytes = graph_track_args_bytes(k);
            g->struct_args[p][k].assign((const uint8_t*) a[0], (const uint8_t*) a[0] + bytes);
            void *f0, *f1;
            graph_track_args_fields(k, g->struct_args[p][k].data(), &f0, &f1);
            if (k == kNodePyramid && (*(const float**) f0 != depth || *(const float**) f1 != pose_in_ph(g)))
                return false;
            if (k == kNodeIcpFinish && (*(float**) f0 != pose_out_ph(g) || *(int**) f1 != (int*) g->ph_ok.ptr))
                return false;
            break;
        }
        case kNodeFrameBegin:
            if (*(const float**) a[kPosFrameBeginPose] != pose_out_ph(g))
                return false;
            break;
        case kNodeFrustumCarve:
            if (*(const float**) a[kPosCarveDepth] != depth)
                return false;
            break;
        case kNodeUpdate:
            if (*(const float**) a[kPosUpdateDepth] != depth || *(const int*) a[kPosUpdateTs] != kTsPlaceholder)
                return false;
            break;
        case kNodeRaycast:
            if (*(const float**) a[kPosRaycastPose] != pose_out_ph(g))
                return false;
            break;
        }
    }
    return true;
}

// Captures instantiation p: track against pair 1 - p, integrate, raycast into pair p.
int capture(se8_frame_graph* g, int p)
{
    hipStream_t s = g->cap_stream;
    const float* depth = (const float*) g->ph_depth.ptr;
    SE8_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    GraphRecorder rec;
    t_recorder = &rec;
    int rc = se8_track_dev(g->tr, depth, pose_in_ph(g), pose_out_ph(g), g->points[1 - p], g->normals[1 - p],
                           (int*) g->ph_ok.ptr, s);
    if (rc == SE8_OK)
        rc = se8_integrate_depth_dev(g->map, kTsPlaceholder, depth, &g->sensor, pose_out_ph(g), s);
    if (rc == SE8_OK)
        rc = se8_raycast_dev(g->map, &g->sensor, pose_out_ph(g), g->points[p], g->normals[p], nullptr, s);
    t_recorder = nullptr;
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    if (rc != SE8_OK) {
        if (graph)
            (void) hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess || !graph) {
        set_error(std::string("se8_frame_graph_create: capture failed: ") + hipGetErrorString(e));
        return SE8_ERR_HIP;
    }
    g->graph[p] = graph;
    for (int k = 0; k < kNumGraphNodes; k++) {
        if (rec.failed || rec.count[k] != 1) {
            set_error("se8_frame_graph_create: the frame's launches were not captured as expected");
            return SE8_ERR_INVALID;
        }
        g->node[p][k] = rec.node[k];
        SE8_HIP_CHECK(hipGraphKernelNodeGetParams(rec.node[k], &g->tmpl[p][k]));
    }
    if (!check_capture(g, p)) {
        set_error("se8_frame_graph_create: captured kernel arguments are not in the expected places");
        return SE8_ERR_INVALID;
    }
    SE8_HIP_CHECK(hipGraphInstantiate(&g->exec[p], graph, nullptr, nullptr, 0));
    return SE8_OK;
}

// Sets the per-frame arguments of instantiation p.
int set_frame_args(se8_frame_graph* g, int p, int ts, const float* depth, const float* T_in, float* T_out, int* ok)
{
    for (int k = 0; k < kNumGraphNodes; k++) {
        hipKernelNodeParams np = g->tmpl[p][k];
        void* args[kMaxArgs];
        std::vector<uint8_t> sa;
        const float* pose_out_c = T_out;
        if (k == kNodePyramid || k == kNodeIcpFinish) {
            sa = g->struct_args[p][k];
            void *f0, *f1;
            graph_track_args_fields(k, sa.data(), &f0, &f1);
            if (k == kNodePyramid) {
                *(const float**) f0 = depth;
                *(const float**) f1 = T_in;
            }
            else {
                *(float**) f0 = T_out;
                *(int**) f1 = ok;
            }
            args[0] = sa.data();
        }
        else {
            const int n = arg_count(k);
            for (int i = 0; i < n; i++)
                args[i] = g->tmpl[p][k].kernelParams[i];
            if (k == kNodeFrameBegin)
                args[kPosFrameBeginPose] = &pose_out_c;
            else if (k == kNodeFrustumCarve)
                args[kPosCarveDepth] = &depth;
            else if (k == kNodeUpdate) {
                args[kPosUpdateDepth] = &depth;
                args[kPosUpdateTs] = &ts;
            }
            else
                args[kPosRaycastPose] = &pose_out_c;
        }
        np.kernelParams = args;
        SE8_HIP_CHECK(hipGraphExecKernelNodeSetParams(g->exec[p], g->node[p][k], &np));
    }
    return SE8_OK;
}

} // namespace

extern "C" int se8_frame_graph_create(se8_tracker* tracker, se8_map* map, const se8_pinhole* sensor,
                                      float* points0_dev, float* normals0_dev, float* points1_dev,
                                      float* normals1_dev, se8_frame_graph** out)
{
    SE8_REQUIRE(tracker && map && sensor && out && points0_dev && normals0_dev && points1_dev && normals1_dev,
                "se8_frame_graph_create: null argument");
    *out = nullptr;
    SE8_REQUIRE(!is_multires(map), "se8_frame_graph_create: not available for a multi-resolution map");
    SE8_REQUIRE(map->view.block_dense && map->view.node_dense,
                "se8_frame_graph_create: the map resolves blocks through the brick index (more than 4096 voxels a "
                "side); use the per-call entry points");
    SE8_REQUIRE(map->timing == SE8_TIMING_OFF && tracker->timing == SE8_TIMING_OFF,
                "se8_frame_graph_create: stage timing must be off (se8_set_stage_timing)");
    SE8_REQUIRE(sensor->width == tracker->sensor.width && sensor->height == tracker->sensor.height,
                "se8_frame_graph_create: sensor size differs from the tracker's");
    se8_frame_graph* g = new se8_frame_graph();
    g->tr = tracker;
    g->map = map;
    g->sensor = *sensor;
    g->points[0] = points0_dev;
    g->normals[0] = normals0_dev;
    g->points[1] = points1_dev;
    g->normals[1] = normals1_dev;
    int rc = dev_alloc(g->ph_depth, sizeof(float) * (size_t) sensor->width * sensor->height);
    rc = rc ? rc : dev_alloc(g->ph_pose, sizeof(float) * 32);
    rc = rc ? rc : dev_alloc(g->ph_ok, sizeof(int));
    if (rc == SE8_OK && (hipStreamCreateWithFlags(&g->cap_stream, hipStreamNonBlocking) != hipSuccess
                         || hipEventCreateWithFlags(&g->done[0], hipEventDisableTiming) != hipSuccess
                         || hipEventCreateWithFlags(&g->done[1], hipEventDisableTiming) != hipSuccess)) {
        set_error("se8_frame_graph_create: stream or event creation failed");
        rc = SE8_ERR_HIP;
    }
    for (int p = 0; p < 2 && rc == SE8_OK; p++)
        rc = capture(g, p);
    if (rc != SE8_OK) {
        destroy_graph(g);
        return rc;
    }
    *out = g;
    return SE8_OK;
}

extern "C" int se8_frame_graph_launch(se8_frame_graph* g, int timestamp, const float* depth_dev,
                                      const float* T_in_dev, float* T_out_dev, int* ok_dev, se8_stream stream)
{
    SE8_REQUIRE(g && depth_dev && T_in_dev && T_out_dev, "se8_frame_graph_launch: null argument");
    SE8_REQUIRE(g->map->timing == SE8_TIMING_OFF && g->tr->timing == SE8_TIMING_OFF,
                "se8_frame_graph_launch: stage timing must be off (se8_set_stage_timing)");
    if (g->map->overflowed) {
        set_error("se8_frame_graph_launch: map capacity exhausted (raise se8_tsdf_config.max_blocks)");
        return SE8_ERR_CAPACITY;
    }
    hipStream_t s = (hipStream_t) stream;
    g->tr->prepared_depth = g->tr->prepared_T = nullptr; // the replay's pyramid overwrites a prepared one
    g->map->prepared_T = nullptr;                         // and its frame start the prepared integration
    const int p = (int) (g->launches & 1);
    // the previous replay of this instantiation (two frames back) must have finished before its
    // arguments change
    if (g->pending[p]) {
        SE8_HIP_CHECK(hipEventSynchronize(g->done[p]));
        g->pending[p] = false;
    }
    int rc = set_frame_args(g, p, timestamp, depth_dev, T_in_dev, T_out_dev, ok_dev ? ok_dev : (int*) g->ph_ok.ptr);
    if (rc != SE8_OK)
        return rc;
    SE8_HIP_CHECK(hipGraphLaunch(g->exec[p], s));
    SE8_HIP_CHECK(hipEventRecord(g->done[p], s));
    g->pending[p] = true;
    g->launches++;
    return SE8_OK;
}

extern "C" int se8_frame_graph_destroy(se8_frame_graph* g)
{
    if (g)
        destroy_graph(g);
    return SE8_OK;
}
