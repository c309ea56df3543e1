// rc_internal.h — context object and kernel launch interfaces shared by the
// translation units of libretrocam_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "rc_device.h"
#include "retrocam_hip.h"

namespace rc {

// Camera + lens arguments of a path-trace launch (rc_camera in device form).
struct CamArgs {
    float pos[3], right[3], up[3], fwd[3];
    float focalDist, lensRadius, tanHalf, saStrength;
    float lca[3];
    int blades, width, height;
};
CamArgs cam_args(const rc_camera &cam);  // rc_pathtrace.hip

// Device scene (see rc_device.h for the layout).
struct DeviceScene {
    float *inner = nullptr;  // 14 floats per inner node (NodeRec)
    float *triPos = nullptr;  // 9 floats per triangle: v0, e1, e2
    float4 *triShade = nullptr;
    float4 *mats = nullptr;
    float4 *lights = nullptr;
    float *cdf = nullptr;
    int *leafBigCount = nullptr;
    int nTris = 0, nInner = 0, nLeaves = 0, nMats = 0, nLights = 0;
    float totalLightPower = 0.0f;
    float rootMin[3] = {0, 0, 0}, rootMax[3] = {0, 0, 0};
    int rootRef = 0;
    int maxDepth = 0;
    bool valid = false;
    int cdfSorted = 0;  // the light CDF is non-decreasing and NaN-free
};

struct Images {
    int width = 0, height = 0;
    float4 *accum = nullptr;       // RGBA32F
    double4 *accum64 = nullptr;    // FP64 mode only
    ushort4 *gbuf = nullptr;       // RGBA16F bits
    ushort4 *galbedo = nullptr;    // RGBA16F bits
    float *variance = nullptr;     // R32F
    int *primId = nullptr;
    int *gbufId = nullptr;
    // post chain (allocated on first rc_post / rc_post_image)
    int postW = 0, postH = 0;
    float4 *postA = nullptr, *postB = nullptr;  // RGBA32F ping-pong
    float4 *postIn = nullptr;                   // rc_post_image input
    uchar4 *display = nullptr;                  // RGBA8 display output
    const float4 *postOut = nullptr;            // which of postA/postB holds the last chain output
};

inline SceneView scene_view(const DeviceScene &D) {
    SceneView S;
    S.inner = D.inner;
    S.triPos = D.triPos;
    S.triShade = D.triShade;
    S.mats = D.mats;
    S.lights = D.lights;
    S.cdf = D.cdf;
    S.leafBigCount = D.leafBigCount;
    for (int k = 0; k < 3; ++k) { S.rootMin[k] = D.rootMin[k]; S.rootMax[k] = D.rootMax[k]; }
    S.rootRef = D.rootRef;
    S.nLights = D.nLights;
    S.cdfSorted = D.cdfSorted;
    S.totalLightPower = D.totalLightPower;
    return S;
}

// SPPM state (SPPMManager + PhotonMap, src/main/java/com/retrocam/scene/).
struct Sppm {
    float radius = 0.05f;      // searchRadius (RenderSettings.sppmInitRadius default)
    double radius64 = 0.05;    // FP64 radius schedule (rc_sppm_set_precision)
    int fp64Radius = 0;
    int iteration = 0;
    int photonCount = 0;       // photons of the last trace
    int hashSize = 0;          // hash cells of the last trace
    int capPhotons = 0, capHash = 0;
    float4 *photons = nullptr; // 3 float4 per photon: position, power, direction
    int *heads = nullptr;
    int *next = nullptr;
    unsigned int *counter = nullptr;
    unsigned int *occupied = nullptr;  // hashSize bits: cells holding a photon (gather skips the rest)
};

// Wavefront path-tracing buffers (rc_wavefront.hip), sized in 256-slot tiles.
struct Wavefront {
    int slots = 0;
    float4 *stThr = nullptr, *stRad = nullptr, *stNee = nullptr, *stSmp = nullptr;
    int *vis = nullptr;
    int *list[2] = {nullptr, nullptr};
    float4 *rayO[2] = {nullptr, nullptr}, *rayD[2] = {nullptr, nullptr};
    float4 *hits = nullptr;
    int *shSlot = nullptr;
    float4 *shO = nullptr, *shD = nullptr;
    int *cnt = nullptr;
    int *hostCnt = nullptr;  // pinned
};

// Sample buffer and unit counters of the lane-scheduler path tracer
// (rc_pathtrace_lanes.hip); grown on demand, kept for the context's lifetime.
struct LaneBuffers {
    float4 *samples = nullptr;
    size_t sampleBytes = 0;
    int *counters = nullptr;
    int nCounters = 0;
    float4 *cont = nullptr;  // 3 float4 per resident lane: shadow-ray continuation
    size_t contLanes = 0;
};

struct Net;    // learned denoiser (rc_net.hip)
struct Views;  // batched views (rc_batch.hip)

// Inputs and buffers of one post-chain run.
struct ChainIO {
    const float4 *src = nullptr;      // p00 input (accumulated sums)
    const ushort4 *gbuf = nullptr;    // RGBA16F normal + depth (null: no G-buffer)
    const ushort4 *galbedo = nullptr; // RGBA16F first-hit albedo
    const float *variance = nullptr;  // R32F sum of luma^2
    int w = 0, h = 0, total = 0;
    float4 *a = nullptr, *b = nullptr;  // RGBA32F ping-pong
    uchar4 *display = nullptr;          // RGBA8 output
    const float4 *out = nullptr;        // set to whichever of a/b holds the chain output
};

}  // namespace rc

struct rc_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    rc::DeviceScene scene;                                // scene slot 0
    rc::DeviceScene extraScenes[RC_MAX_SCENE_SLOTS - 1];  // slots 1..
    rc::Views *views = nullptr;
    rc::Images img;
    int32_t totalSamples = 0;
    int32_t accumMode = RC_ACCUM_FP32;
    rc::Sppm sppm;
    rc::Wavefront wf;
    int32_t ptMode = RC_PT_LANES;
    rc::LaneBuffers lanes;
    int32_t laneThresh = 32;  // lanes waiting before a wave shades (rc_pathtrace_lanes.hip)
    int32_t laneVariant = 9;  // lane-kernel variant (rc_pathtrace_lanes.hip)
    // Upper bound of the lane scheduler's sample buffer (16 B per pixel and
    // sample): a launch with more samples x pixels runs in chunks. Every chunk
    // ends with each wave draining its lanes once per view, so a batch of views
    // runs best as one chunk: 24 GiB holds 1.6 G pixel-samples (the buffer
    // grows only to what a launch needs).
    int32_t laneBufferMiB = 24576;
    int32_t laneTileShift = 3;  // unit tiles of 8 x 8 pixels
    rc::Net *net = nullptr;
    int32_t denoiser = RC_DENOISE_ATROUS;
    int32_t netFlips = 1;  // network passes per denoise (rc_set_net_flips)
    int32_t netKernels = RC_NET_KERNELS_TILED;  // forward convolution kernels (rc_set_net_kernels)
    int32_t postMode = RC_POST_FUSED;  // post chain as fused sweeps or one launch per pass
    int32_t postLaunches = 0;          // launches of the last post chain (excluding the denoiser network)
    unsigned long long *counters = nullptr;  // device work counters (bench instrumentation)
    int countersEnabled = 0;
    // captured frame (rc_frame_*): the graph, its instance, and the temporal
    // inputs of its post chains, one device slot per recorded chain
    int capturing = 0;
    int captureResetSeen = 0;
    hipGraph_t frameGraph = nullptr;
    hipGraphExec_t frameExec = nullptr;
    rc_temporal *dTemporal = nullptr;  // kMaxTemporalSlots device slots
    rc_temporal *hTemporal = nullptr;  // pinned: the values recorded at capture
    int nTemporal = 0;                 // slots used by the recorded chains
    uint64_t viewTablesSig = 0;        // content signature of the view tables on the device
    uint64_t graphViewTablesSig = 0;   // signature the recorded frame reads
    int graphUsesViews = 0;
    // bumped before any device buffer of the context is freed (reallocation): a
    // recorded frame whose generation differs may reference freed memory
    uint64_t freeGen = 0;
    uint64_t graphFreeGen = 0;
    int captureViewsResetSeen = 0;
    rc::CamArgs *dCams = nullptr;  // kMaxCameraSlots capture slots + 1 eager slot
    rc::CamArgs *hCams = nullptr;  // pinned: the cameras recorded at capture
    int nCams = 0;
    int replayCamSet = 0;          // rc_frame_set_camera: replays use replayCam
    rc::CamArgs replayCam{};
    int slotsExhausted = 0;        // capture ran out of temporal slots
    // host-side state at the end of the recorded calls, restored by each replay
    struct {
        int32_t totalSamples = 0;
        float radius = 0.0f;
        double radius64 = 0.0;
        int iteration = 0, photonCount = 0, hashSize = 0;
        std::vector<int> viewTotals;
    } captured;
};

namespace rc {

// Post chains one captured frame can record (one temporal slot each), and
// camera-using launches (path trace, SPPM gather; one camera slot each).
constexpr int kMaxTemporalSlots = 4096;
constexpr int kMaxCameraSlots = 4096;

int set_error(rc_ctx *ctx, int code, const std::string &msg);
// RC_ERR_STATE if a frame is being captured: `what` would allocate, wait for
// the device or copy to the host, none of which a recorded graph can hold.
int capture_block(rc_ctx *ctx, const char *what);
// Temporal input of a post chain: by value when run eagerly (null slot); while
// capturing, the next device slot, which rc_frame_replay refills before each
// launch.
const rc_temporal *temporal_slot(rc_ctx *ctx, const rc_temporal &t);
// The device copy of a launch's camera, in *out. Eagerly, the context's eager
// slot, written by a fill kernel on the stream before the launch; while
// capturing, the next camera slot, which rc_frame_replay refills (with the
// recorded camera, or rc_frame_set_camera's). Returns an RC_* code.
int camera_slot(rc_ctx *ctx, const CamArgs &c, const CamArgs **out);
int hip_fail(rc_ctx *ctx, hipError_t e, const char *what);

// Launches nSamples progressive passes starting at frame ctx->totalSamples.
hipError_t launch_pathtrace(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples,
                            const CamArgs *dcam);

// Releases the SPPM buffers (rc_sppm.hip).
void free_sppm(Sppm &s);

// Wavefront form of launch_pathtrace (rc_wavefront.hip); same results.
hipError_t launch_wavefront(rc_ctx *ctx, const rc_camera &cam, const rc_pt_params &p, int nSamples);
void free_wavefront(Wavefront &w);

// Learned denoiser (rc_net.hip).
void free_net(Net *n);
bool net_ready(const rc_ctx *ctx);
// Denoises the w x h RGBA32F post image src into dst with the frame's G-buffers.
int run_net_denoise(rc_ctx *ctx, const float4 *src, const ushort4 *gbuf, const ushort4 *galbedo,
                    const float *variance, int total, float4 *dst, int w, int h);

// Looks up rc_post_image_noalloc's scratch address on `device` (current) once
// per device (rc_noalloc.hip). Returns an RC_* code.
int noalloc_setup(rc_ctx *ctx, int device);

// Post chain on explicit buffers (rc_capi.hip).
int run_chain(rc_ctx *ctx, ChainIO &io, const rc_post_params &p, const rc_temporal &t, bool withGbuf);
// Scene of a slot (0 = rc_upload_scene's), or null when out of range.
DeviceScene *scene_slot(rc_ctx *ctx, int slot);
// Batched views (rc_batch.hip).
void free_views(rc_ctx *ctx);
// Sample counts of the current views (captured frames keep and restore them).
std::vector<int> views_totals(const rc_ctx *ctx);
void views_set_totals(rc_ctx *ctx, const std::vector<int> &totals);

// The first input of a fused post chain: a float4 image, read either as it is
// or as the p00 output of an accumulation (sum / max(total, 1) * exposure).
struct SweepSource {
    const float4 *p = nullptr;
    bool p00 = false;
};
// Runs the chain ops[0..nOps) (pass ids: 100 = a-trous with steps[i], 1..16 =
// VHS pN; nOps = 0: the source copied) over a w x h image as fused sweeps,
// ping-ponging between bufA and bufB (neither may be src.p unless src is the
// previous output), with the display encode of the result into rgba8 (may be
// null). *outFinal receives the buffer holding the chain output, *launches the
// number of sweeps (rc_post.hip).
// Workgroups of kBlock threads of `kernel` resident at once on the current
// device, and a persistent launch's grid: that many, but no more than `units`
// work items need (rc_pathtrace_lanes.hip).
int resident_blocks(const void *kernel);
int persistent_blocks(const void *kernel, int units);

hipError_t launch_post_sweeps(hipStream_t stream, const SweepSource &src, const int *ops, const int *steps, int nOps,
                              int w, int h, float4 *bufA, float4 *bufB, uchar4 *rgba8, const rc_post_params &p,
                              const rc_temporal &t, const rc_temporal *dt, const ushort4 *gbuf,
                              const ushort4 *galbedo, const float *variance, int total, const float4 **outFinal,
                              int *launches);

// One post-chain pass (rc_post.hip): pass 0 = p00, 1..16 = VHS pN, 100 = a-trous
// with the given step, 200 = display encode into rgba8.
// dt: when non-null, the kernel reads the temporal inputs from this device slot
// instead of t (captured frames).
hipError_t launch_post_pass(hipStream_t stream, int pass, const float4 *src, float4 *dst, int w, int h,
                            const rc_post_params &p, const rc_temporal &t, const ushort4 *gbuf,
                            const ushort4 *galbedo, const float *variance, int total, int step, uchar4 *rgba8,
                            const rc_temporal *dt = nullptr);

}  // namespace rc
