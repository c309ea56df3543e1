/*
 * retrocam_hip.h — C ABI of libretrocam_hip.so, the MI355X (gfx950) render hot path
 * of RetroCam: progressive path tracing, SPPM photon trace/insert/gather, the
 * p00 normalize + a-trous denoise + VHS post chain and the display encode.
 *
 * The library replaces the OpenGL work behind the reference's Java classes; the
 * Java host keeps the scene, the SAH BVH builder, the struct writers and the
 * settings, and passes the exact float streams those writers produce:
 *
 *   GpuTriangle.write  (src/main/java/com/retrocam/scene/GpuTriangle.java:68-79)  32 floats
 *   BVHNode.write      (src/main/java/com/retrocam/scene/BVHNode.java:32-40)      12 floats
 *   Material.write     (src/main/java/com/retrocam/scene/Material.java:61-66)     16 floats
 *   light entries      (src/main/java/com/retrocam/scene/SceneUploader.java:173-205) 20 floats
 *
 * Integers inside those streams are stored as float bit patterns
 * (Float.intBitsToFloat), exactly as the Java writers emit them.
 *
 * Conventions
 *   - Every entry point returns an int status (RC_OK = 0). On failure,
 *     rc_last_error(ctx) returns a message; nothing aborts the host process.
 *   - One context owns one HIP stream and all device buffers; calls on one
 *     context must be serialised by the caller (the reference runs on a single
 *     GL thread). Separate contexts are independent.
 *   - Images are row-major with row 0 = the BOTTOM row (GL texture order,
 *     pathtrace.comp:406-413), pixel (x, y) at index y*width + x.
 *   - Host arrays passed in are copied before the call returns; the caller may
 *     free them at once (SceneUploader.java:65-71 frees its buffers after upload).
 */
#ifndef RETROCAM_HIP_H
#define RETROCAM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
enum {
    RC_OK = 0,
    RC_ERR_INVALID_ARG = 1,   /* bad pointer, size, index or enum value            */
    RC_ERR_NO_DEVICE = 2,     /* HIP device missing or not gfx950-capable          */
    RC_ERR_HIP = 3,           /* a HIP runtime call failed (message has the code)  */
    RC_ERR_OOM = 4,           /* device allocation failed                          */
    RC_ERR_STATE = 5,         /* call out of order (e.g. render before upload)     */
    RC_ERR_LIMIT = 6          /* size beyond a documented library limit            */
};

/* ---- wire layouts (std430 records as written by the Java host) ---------- */
typedef struct rc_gpu_triangle {     /* 128 B, GpuTriangle.java:68-79          */
    float v0[3]; float pad0;
    float v1[3]; float pad1;
    float v2[3]; float pad2;
    float n0[3]; float pad3;
    float n1[3]; float pad4;
    float n2[3]; float pad5;
    float uv0[2]; float uv1[2];
    float uv2[2]; int32_t matIndex; float pad6;
} rc_gpu_triangle;

typedef struct rc_bvh_node {         /* 48 B, BVHNode.java:32-40; leaf iff primCount > 0 */
    float bmin[3]; float pad0;
    float bmax[3]; float pad1;
    int32_t leftChild, rightChild, primStart, primCount;
} rc_bvh_node;

typedef struct rc_material {         /* 64 B, Material.java:61-66              */
    float albedo[3]; float metallic;
    float emission[3]; float emissionStrength;
    float roughness, ior, transmission; int32_t albedoTex;
    float pad[4];
} rc_material;

typedef struct rc_light {            /* 80 B, SceneUploader.java:173-205       */
    float v0[3]; float area;
    float v1[3]; float pad1;
    float v2[3]; float pad2;
    float emission[3]; float power;  /* emission already scaled by strength    */
    float normal[3]; float cdf;      /* cdf: cumulative power fraction, last = 1 */
} rc_light;

/* The record sizes the Java host writes (std430 strides). */
#ifdef __cplusplus
#define RC_SIZE_CHECK(t, n) static_assert(sizeof(t) == (n), #t " must be " #n " bytes")
#else
#define RC_SIZE_CHECK(t, n) _Static_assert(sizeof(t) == (n), #t " must be " #n " bytes")
#endif
RC_SIZE_CHECK(rc_gpu_triangle, 128);
RC_SIZE_CHECK(rc_bvh_node, 48);
RC_SIZE_CHECK(rc_material, 64);
RC_SIZE_CHECK(rc_light, 80);
#undef RC_SIZE_CHECK

/* Camera + lens uniforms of one pass (ThinLensCamera.uploadTo,
 * src/main/java/com/retrocam/camera/ThinLensCamera.java:27-58, plus the render
 * resolution that Renderer.java:100-101 sets). */
typedef struct rc_camera {
    float camPos[3];
    float camRight[3];
    float camUp[3];
    float camForward[3];
    float focalDist;        /* TemporalState.focalDist (metres)                   */
    float lensRadius;       /* focalLength / (2 N), 0 = pinhole                   */
    float fovTanHalf;       /* ThinLensCamera.tanHalfFov                          */
    int32_t apertureBlades; /* N-gon bokeh blades (>= 3 when lensRadius > 0)      */
    float saStrength;       /* spherical aberration                               */
    float lcaDelta[3];      /* per-channel focal offsets (R, G, B)                */
    int32_t width, height;  /* render resolution in pixels                        */
} rc_camera;

/* Path-trace switches (Renderer.java:93-96). */
typedef struct rc_pt_params {
    int32_t neeEnabled;     /* 1 = next-event estimation at diffuse hits          */
    float neeFireflyClamp;  /* per-channel NEE clamp, multiplier of light luma    */
    int32_t maxBounces;     /* 6 in the reference (pathtrace.comp:465); 1..16     */
} rc_pt_params;

/* Post-chain settings: every RenderSettings field read by
 * PostProcessStack.runChain (src/main/java/com/retrocam/post/PostProcessStack.java:140-280)
 * and AtrousDenoiser.denoise (src/main/java/com/retrocam/post/AtrousDenoiser.java:60-125).
 * Defaults are those of RenderSettings.java:26-191 (rc_post_defaults). */
typedef struct rc_post_params {
    float exposure;                   /* p00 EV multiplier                      */
    int32_t atrousEnabled;            /* a-trous denoiser on/off                */
    int32_t atrousIterations;         /* passes, clamped to 1..5                */
    int32_t atrousMaxSpp;             /* bypass above this spp (0 = never)      */
    float atrousSigmaColor, atrousSigmaNormal, atrousSigmaDepth;
    int32_t enable[17];               /* enable[i] switches pass p<i>, i = 1..16;
                                         p05/p06 are pass-through in the reference */
    float lumaBlurSigma;                                                   /* p01 */
    float chromaResSigma;                                                  /* p02 */
    float colorTempBias, colorHueRot, colorSaturation, colorDriftSpeed;   /* p03 */
    float drBlackLift, drShadowCrush, drHighlightKnee, drHighlightClip;   /* p04 */
    float ccdNoiseLumaBase, ccdNoiseChromaScale;                          /* p07 */
    float ccdSmearThreshold, ccdSmearIntensity; int32_t ccdSmearLength;   /* p08 */
    float chromaBleedFactor;                                               /* p09 */
    float timebaseAmplitudePx, timebaseFreq, timebaseSpeed;               /* p10 */
    float dotCrawlIntensity, dotCrawlSubcarrier, dotCrawlEdgeThresh;      /* p11 */
    float tapeAge, dropoutBright, dropoutBurstRate;                       /* p12 */
    int32_t headSwitchLines;
    float headSwitchJitter, headSwitchRipple, headSwitchChroma;           /* p13 */
    float trackingSeverity, trackingMaxDisplacePx, trackingFringeWidthPx; /* p14 */
    float edgeEnhanceAmount, edgeEnhanceCoreThreshold;                    /* p15 */
    float interlaceFieldOffsetPx, interlaceCombStrength,
          interlaceCombEdgeThresh, interlaceLineWeighting;                /* p16 */
} rc_post_params;

/* Time inputs of the post chain: TemporalState.time / agcGain
 * (src/main/java/com/retrocam/core/TemporalState.java:96-111) and
 * RenderSettings.frameIndex (incremented per displayed frame, Main.java:313). */
typedef struct rc_temporal {
    float time;
    float agcGain;
    int32_t frameIndex;
} rc_temporal;

/* Readable buffers for rc_read. */
enum {
    RC_BUF_ACCUM = 0,       /* float4 per px: sum of radiance, sample count (RGBA32F) */
    RC_BUF_GBUFFER = 1,     /* half4 per px: shading normal, hit t (RGBA16F)       */
    RC_BUF_GALBEDO = 2,     /* half4 per px: first-hit albedo, 1 (RGBA16F)         */
    RC_BUF_VARIANCE = 3,    /* float per px: sum of luma^2 (R32F)                  */
    RC_BUF_PRIMARY_ID = 4,  /* int32 per px: triangle index of the camera ray's first
                               hit in the most recent pass (-1 = miss); triangles are
                               numbered in upload (BVH leaf) order                 */
    RC_BUF_GBUFFER_ID = 5,  /* int32 per px: triangle index hit by the G-buffer ray */
    RC_BUF_ACCUM64 = 6,     /* double4 per px: FP64 accumulation (RC_ACCUM_FP64)   */
    RC_BUF_POST_HDR = 7,    /* float4 per px: post-chain output before display     */
    RC_BUF_DISPLAY = 8      /* uint8x4 per px: display-encoded RGBA8 (ACES + sRGB)  */
};

/* Accumulation precision (rc_set_accum_mode). */
enum {
    RC_ACCUM_FP32 = 0,      /* reference behaviour: RGBA32F running sum            */
    RC_ACCUM_FP64 = 1       /* FP64 running sum; RC_BUF_ACCUM returns it rounded   */
};

/* Path-tracer kernel organisation (rc_set_pt_mode); all give identical results. */
enum {
    RC_PT_MEGAKERNEL = 0,   /* one kernel, one lane per pixel for all samples     */
    RC_PT_WAVEFRONT = 1,    /* extend / shadow / shade kernels over ray queues    */
    RC_PT_LANES = 2         /* default: per-lane state machines pulling single
                               samples from a work counter; used for scenes whose
                               BVH is shallower than 30 levels, the megakernel
                               otherwise                                         */
};

/* Denoiser of the post chain (rc_set_denoiser). */
enum {
    RC_DENOISE_ATROUS = 0,  /* the reference's a-trous filter (default)          */
    RC_DENOISE_NET = 1      /* the learned U-Net (rc_net_*), a-trous kept as the
                               fallback when no network is loaded                */
};

/* Organisation of the post chain (rc_set_post_mode); both give identical images. */
enum {
    RC_POST_FUSED = 0,      /* default: the chain in a few full-frame sweeps, each
                               running consecutive passes over a tile in LDS   */
    RC_POST_PER_PASS = 1    /* one full-frame launch per pass, as the reference's
                               FBO chain draws them                             */
};

typedef struct rc_ctx rc_ctx;

/* Scene slots for batched views (slot 0 is the scene of rc_upload_scene). */
#define RC_MAX_SCENE_SLOTS 64

/* ---- context ------------------------------------------------------------ */
/* Creates a context on HIP device `device`. Replaces the GL objects the
 * Renderer constructor allocates (Renderer.java:55-62). */
int rc_create(int device, rc_ctx **out);
int rc_destroy(rc_ctx *ctx);
/* Message of the last failure on ctx (or of the last rc_create when ctx is NULL). */
const char *rc_last_error(const rc_ctx *ctx);
/* The HIP stream (hipStream_t) all work of ctx is issued on. */
void *rc_stream(rc_ctx *ctx);
/* Waits for all work issued on ctx. */
int rc_sync(rc_ctx *ctx);

/* ---- scene -------------------------------------------------------------- */
/* SceneUploader.upload/bind (SceneUploader.java:54-115). Streams in the wire
 * layouts above; nLights may be 0 (the reference uploads a zero stub,
 * SceneUploader.java:188-196, and every light path early-outs). Indices are
 * validated: child/prim/material indices out of range return RC_ERR_INVALID_ARG. */
int rc_upload_scene(rc_ctx *ctx,
                    const float *tris, int32_t nTris,
                    const float *nodes, int32_t nNodes,
                    const float *mats, int32_t nMats,
                    const float *lights, int32_t nLights,
                    float totalLightPower);

/* ---- progressive path tracing ------------------------------------------- */
/* Renderer.render (Renderer.java:72-116): one sample per pixel with
 * frameIndex = total samples so far, then total samples += 1. A change of
 * camera->width/height reallocates the images and resets the count. */
int rc_render_pass(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *params);
/* nSamples consecutive rc_render_pass calls fused into one launch; results are
 * identical to the loop. */
int rc_render_passes(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *params,
                     int32_t nSamples);
/* Renderer.reset (Renderer.java:120-130): drop the accumulation, count = 0. */
int rc_reset(rc_ctx *ctx);
/* Renderer.getTotalSamples (Renderer.java:138). */
int32_t rc_total_samples(const rc_ctx *ctx);
/* RC_PT_LANES (default), RC_PT_MEGAKERNEL or RC_PT_WAVEFRONT for later passes;
   all give bit-identical images. */
int rc_set_pt_mode(rc_ctx *ctx, int32_t mode);
/* RC_ACCUM_FP32 (default) or RC_ACCUM_FP64; resets the accumulation. */
int rc_set_accum_mode(rc_ctx *ctx, int32_t mode);
/* Copies a device buffer (RC_BUF_*) of the current resolution to host memory;
 * `bytes` must equal the buffer size. Replaces glGetTexImage. */
int rc_read(rc_ctx *ctx, int32_t which, void *dst, size_t bytes);
/* Size in bytes of a readable buffer at the current resolution. */
size_t rc_buffer_bytes(const rc_ctx *ctx, int32_t which);

/* ---- SPPM caustics ------------------------------------------------------- */
/* SPPMManager.reset (src/main/java/com/retrocam/scene/SPPMManager.java:180-183):
 * radius := initRadius, iteration := 0. */
int rc_sppm_reset(rc_ctx *ctx, float initRadius);
/* 0 (default): the radius is a float updated as the reference does; 1: the
 * schedule runs in FP64 and the kernels use its rounding to float. */
int rc_sppm_set_precision(rc_ctx *ctx, int32_t fp64Radius);
/* SPPMManager.tracePhotons (SPPMManager.java:76-101): clears the hash heads to -1
 * and the counter to 0, then emits photonCount photons with seed
 * iteration*7919 + 6271 at the current radius and inserts caustic photons into
 * the hash (hashSize cells, a power of two). photonCount <= 32,000,000,
 * hashSize <= 2^26. A scene without lights stores nothing. */
int rc_sppm_trace(rc_ctx *ctx, int32_t photonCount, int32_t hashSize);
/* SPPMManager.gatherRadiance (SPPMManager.java:118-160): pinhole ray per pixel
 * from cam's position/basis/fovTanHalf (lens ignored, aspect = width/height),
 * through specular surfaces to the first diffuse hit, 3x3x3-cell density
 * estimate, accumulation += (L, 0). cam's resolution must equal the
 * accumulation's. */
int rc_sppm_gather(rc_ctx *ctx, const rc_camera *cam);
/* SPPMManager.updateRadius (SPPMManager.java:170-173), then iteration += 1. */
int rc_sppm_update_radius(rc_ctx *ctx, float alpha);
/* Current radius (float the kernels use), iteration and the FP64 schedule value. */
int rc_sppm_get_state(const rc_ctx *ctx, float *radius, int32_t *iteration, double *radius64);
/* Copies the photon map of the last trace: photons (photonCount x 12 floats:
 * position, pad, power, pad, direction, pad), heads (hashSize), next
 * (photonCount) and the raw counter; any pointer may be NULL. Slots at or above
 * the counter are not written by the trace. */
int rc_sppm_read(rc_ctx *ctx, float *photons, int32_t *heads, int32_t *next, uint32_t *counter);

/* ---- post chain ---------------------------------------------------------- */
/* Fills p with the RenderSettings defaults (RenderSettings.java:26-191):
 * a-trous off (4 iterations), every VHS pass off. */
void rc_post_defaults(rc_post_params *p);
/* PostProcessStack.runOnAccum + the display pass (PostProcessStack.java:109-118,
 * RenderPipeline.java:228-261) on the context's accumulation: p00 divides by the
 * total sample count and applies p->exposure, then the a-trous passes (if
 * enabled and the count is within atrousMaxSpp), the enabled VHS passes p01..p16
 * in order, and the display encode. Read RC_BUF_POST_HDR / RC_BUF_DISPLAY. */
int rc_post(rc_ctx *ctx, const rc_post_params *p, const rc_temporal *t);
/* PostProcessStack.runOnImage (PostProcessStack.java:130-136): the chain on a
 * host RGBA32F image (rows bottom-up) with N = 1, exposure 1 and no a-trous. */
int rc_post_image(rc_ctx *ctx, const float *rgba, int32_t width, int32_t height, const rc_post_params *p,
                  const rc_temporal *t);

/* No counterpart in the reference: rc_post_image for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Device pointers in and out: rgba is a w x h
 * RGBA32F image (rows bottom-up), hdrOut (may be NULL) receives the chain output
 * (RGBA32F) and rgba8Out the display encode (RGBA8); the chain is
 * rc_post_image's (N = 1, exposure 1, no a-trous), enqueued on the context's
 * stream (rc_sync waits for it). It allocates nothing: its two intermediate
 * images live in a library-wide device array of 1280 x 720 pixels each per GPU
 * (29.5 MB, held until the process ends); larger images return RC_ERR_LIMIT.
 * Every call on a GPU uses that GPU's one array: calls that overlap on one GPU
 * (other contexts' streams, other threads) race. Not allowed while capturing. */
int rc_post_image_noalloc(rc_ctx *ctx, const float *rgba, int32_t width, int32_t height, const rc_post_params *p,
                          const rc_temporal *t, float *hdrOut, uint8_t *rgba8Out);

/* No counterpart in the reference: RC_POST_FUSED (default) or RC_POST_PER_PASS
 * for later post chains (rc_post, rc_post_image, rc_views_post, ...); the images
 * are bit-identical. *launches (may be NULL) receives the kernel launches of the
 * last post chain, the denoiser network's own excluded. */
int rc_set_post_mode(rc_ctx *ctx, int32_t mode);
int rc_post_launches(const rc_ctx *ctx, int32_t *launches);

/* ---- frame export ------------------------------------------------------- */
/* FrameExporter.exportFrame (src/main/java/com/retrocam/export/FrameExporter.java:37-65):
 * the display image of the last rc_post (RGBA8, ACES + sRGB) with its rows flipped to
 * top-down order on the device, read back once, and written to `path` as a PNG (8-bit
 * RGBA) when path is not NULL. rgbaTopDown (may be NULL) receives the flipped pixels
 * (width x height x 4 bytes). The reference's JPEG option is not provided. */
int rc_export_frame(rc_ctx *ctx, const char *path, uint8_t *rgbaTopDown);
/* Size of the display image of the last rc_post (RC_ERR_STATE before one). */
int rc_display_size(const rc_ctx *ctx, int32_t *width, int32_t *height);
/* The PNG writer alone (host only, no device): width x height RGBA8, rows bottom-up
 * (GL order) when flipRows = 1, top-down otherwise. */
int rc_encode_png(const uint8_t *rgba, int32_t width, int32_t height, int32_t flipRows, const char *path);

/* ---- batched views (no counterpart in the reference) --------------------- */
/* Serving many render requests at once: each view names a scene slot, a
 * camera (with its own resolution), path-trace switches and a sample count.
 * rc_views_render traces all views in ONE persistent launch; each view keeps
 * its own accumulation and G-buffers, and its pixels are bit-identical to a
 * single-view rc_render_passes of the same camera and samples. */
typedef struct rc_view {
    int32_t scene;          /* scene slot, 0..RC_MAX_SCENE_SLOTS-1            */
    int32_t nSamples;       /* samples this call adds                         */
    rc_camera cam;
    rc_pt_params params;
} rc_view;
enum {
    RC_VIEWS_RESET = 1,     /* restart every view's accumulation at frame 0   */
    RC_VIEWS_FP64 = 2       /* FP64 accumulation (also kept in RGBA32F)       */
};
/* SceneUploader.upload into a scene slot (slot 0 = rc_upload_scene). */
int rc_upload_scene_slot(rc_ctx *ctx, int32_t slot, const float *tris, int32_t nTris, const float *nodes,
                         int32_t nNodes, const float *mats, int32_t nMats, const float *lights, int32_t nLights,
                         float totalLightPower);
/* A new set of views (different count or sizes) restarts accumulation. */
int rc_views_render(rc_ctx *ctx, const rc_view *views, int32_t n, int32_t flags);
/* The post chain (rc_post) on one view's accumulation and G-buffers. */
int rc_views_post(rc_ctx *ctx, int32_t view, const rc_post_params *p, const rc_temporal *t);
size_t rc_views_buffer_bytes(const rc_ctx *ctx, int32_t view, int32_t which);
int rc_views_read(rc_ctx *ctx, int32_t view, int32_t which, void *dst, size_t bytes);
int32_t rc_views_total_samples(const rc_ctx *ctx, int32_t view);

/* ---- learned denoiser (no counterpart in the reference) ------------------ */
/* A 5-level U-Net (3x3 convolutions, 32..128 channels) on bfloat16 MFMA with
 * FP32 accumulation, trained on the GPU by its own forward/backward kernels.
 * Inputs per pixel ("raw", 12 floats): radiance rgb, albedo rgb, normal xyz,
 * depth, mean of luma^2 over the samples, sample count — the render outputs of
 * pathtrace.comp:6-7, 644-654 (accumulation, G-buffers, variance). Targets: 3 floats
 * (high-spp radiance). Images are n x h x w, h and w multiples of 16. The flat
 * parameter vector holds, per layer, W[cout][ky][kx][cin] then b[cout]. */
int32_t rc_net_param_count(void);
/* He-normal initialisation from a seed; resets the optimizer state. */
int rc_net_init(rc_ctx *ctx, uint32_t seed);
int rc_net_set_params(rc_ctx *ctx, const float *params, int32_t count);
int rc_net_get_params(rc_ctx *ctx, float *params, int32_t count);
/* Copies a training set to the device: raw n x h x w x 12, target n x h x w x 3;
 * n x h x w <= 2^29 pixels (60 bytes each on the device). */
int rc_net_set_dataset(rc_ctx *ctx, const float *raw, const float *target, int32_t n, int32_t h, int32_t w);
/* `steps` Adam steps (b1 0.9, b2 0.999) on batches of `batch` crops drawn
 * from the dataset with random flips; loss = mean (x-r)^2/(r^2+0.01). Writes
 * the loss of every step to losses[steps] (may be NULL). */
int rc_net_train(rc_ctx *ctx, int32_t steps, int32_t batch, float lr, uint32_t seed, float *losses);
/* One forward + backward pass without an update: loss and the gradient
 * (rc_net_param_count floats). */
int rc_net_grad(rc_ctx *ctx, const float *raw, const float *target, int32_t n, int32_t h, int32_t w, float *loss,
                float *grad);
/* Network output (linear radiance) for raw inputs: out n x h x w x 3. */
int rc_net_forward(rc_ctx *ctx, const float *raw, int32_t n, int32_t h, int32_t w, float *out);
/* Which denoiser rc_post runs when rc_post_params.atrousEnabled is set. */
int rc_set_denoiser(rc_ctx *ctx, int32_t mode);
/* Network passes per denoise (1, 2, 4 or 8; default 1): the frame as is, then
 * mirrored left-right, top-bottom and both, then (8) those four moved 8 pixels
 * right and down; the denoised image is the mean of the passes' outputs mapped
 * back (an ensemble: lower error per sample for more network time). Applies to
 * rc_post and rc_views_post. */
int rc_set_net_flips(rc_ctx *ctx, int32_t passes);

/* Convolution kernels of the network's forward pass (no counterpart in the
 * reference); every choice computes each output with the same operands in the
 * same order, so the network's outputs are bit-identical:
 *   RC_NET_KERNELS_TILED  (default) the three finest levels stage input tiles in
 *                         LDS, small levels split output channels over waves;
 *   RC_NET_KERNELS_DIRECT every level reads its operands per tap from L1/L2,
 *                         channel groups as in the default;
 *   RC_NET_KERNELS_SINGLE_GROUP  as DIRECT, one channel group per wave. */
enum { RC_NET_KERNELS_TILED = 0, RC_NET_KERNELS_DIRECT = 1, RC_NET_KERNELS_SINGLE_GROUP = 2 };
int rc_set_net_kernels(rc_ctx *ctx, int32_t mode);

/* ---- captured frames (no counterpart in the reference) ------------------- */
/* RenderPipeline.accumulateFrame (src/main/java/com/retrocam/export/RenderPipeline.java:209-239)
 * issues the same chain of dispatches every frame: reset, spp x {photon trace,
 * path trace, gather, radius update}, post chain, display. These calls record
 * such a chain once into a HIP graph and replay it with one launch.
 *
 * Between rc_frame_capture_begin and rc_frame_capture_end, the device work of
 * rc_reset, rc_render_pass(es), rc_sppm_trace / _gather / _update_radius,
 * rc_post, rc_views_render and rc_views_post is recorded instead of run (host
 * state such as the sample count advances as usual). The recording must begin
 * with rc_reset before any single-view rendering, and with RC_VIEWS_RESET in
 * its first rc_views_render, so that every replay recomputes the frame.
 * Nothing recorded may allocate, wait or copy to the host: run the same frame
 * once before capturing it (buffers, network workspace and view tables then
 * exist); calls that would do so return RC_ERR_STATE (or RC_ERR_HIP for a
 * buffer that would have to grow) and end nothing - call rc_frame_capture_end
 * to drop the recording. rc_read, rc_sync, uploads, rc_net_* and the wavefront
 * path tracer are refused while capturing.
 *
 * rc_frame_replay launches the recorded frame on the context's stream and
 * restores the host state to its end (sample count, SPPM radius and
 * iteration, view sample counts). t, when not NULL, replaces the temporal
 * inputs of every recorded post chain (video time, AGC gain, frame index);
 * NULL replays them as recorded. A frame that renders views reads the view
 * tables of the last rc_views_render; replaying while other views' tables are
 * on the device returns RC_ERR_STATE (rendering the recorded views again
 * restores them). One recorded frame per context; capturing again or
 * rc_frame_release drops it. */
int rc_frame_capture_begin(rc_ctx *ctx);
int rc_frame_capture_end(rc_ctx *ctx);
int rc_frame_replay(rc_ctx *ctx, const rc_temporal *t);
/* Camera of the recorded single-view path-trace launches and SPPM gathers for
 * the following replays (a video's moving or refocusing camera, e.g.
 * TemporalState.focalDist changing per frame, TemporalState.java:96-111); its
 * resolution must be the recorded one. NULL returns to the recorded cameras.
 * Views keep the cameras of their tables. */
int rc_frame_set_camera(rc_ctx *ctx, const rc_camera *cam);
int rc_frame_release(rc_ctx *ctx);

/* ---- tuning (no counterpart in the reference) --------------------------- */
/* Launch parameters that change speed, never results. */
#define RC_LANE_VARIANTS 11
typedef struct rc_tuning {
    int32_t laneThreshold;  /* RC_PT_LANES: a wave stops walking and shades once this
                               many of its 64 lanes wait (1..64, default 32)       */
    int32_t laneVariant;    /* RC_PT_LANES kernel variant, 0..RC_LANE_VARIANTS-1:
                               occupancy, LDS stack depth, triangles per step
                               (default 9: 4 waves/SIMD, 20 entries, a triangle
                               and a node per walk step)                           */
    int32_t laneBufferMiB;  /* RC_PT_LANES sample buffer bound (16 B per pixel and
                               sample; larger launches run in chunks), 64..65536,
                               default 24576 (1.6 G pixel-samples per launch)      */
    int32_t laneTileWidth;  /* RC_PT_LANES unit tile: laneTileWidth x (64 /
                               laneTileWidth) pixels; 4, 8 (default), 16 or 32     */
    int32_t ptMode;         /* path-trace form (rc_set_pt_mode); rc_set_tuning
                               sets it too                                         */
} rc_tuning;
int rc_get_tuning(const rc_ctx *ctx, rc_tuning *out);
int rc_set_tuning(rc_ctx *ctx, const rc_tuning *t);
/* Autotuner: times the lane scheduler's variants, wait thresholds and tile
 * shapes, and the megakernel, on the uploaded scene at cam's resolution
 * (samplesPerTrial samples per trial, 1..4096; every candidate renders the same
 * image) and keeps the fastest as the context's tuning and path-trace form. out and msPerSample (may be NULL) receive it and its
 * time per sample. The accumulation is reset on return. RC_ERR_STATE for a scene
 * the lane scheduler does not run (BVH of 30 levels or more). */
int rc_autotune(rc_ctx *ctx, const rc_camera *cam, const rc_pt_params *p, int32_t samplesPerTrial,
                rc_tuning *out, float *msPerSample);

/* ---- instrumentation (no counterpart in the reference) ------------------ */
/* Work counters of the path-trace kernel, for the roofline byte model. Copies
 * the first n uint64 counters gathered since the previous call into out (order:
 * rays, shadowRays, nodePops, childTests, triTests, hitUpdates, shadeHits,
 * shadowPops, shadowTriTests, lightScans, lightReads, paths), clears them, and
 * switches counting on (enable = 1) or off (enable = 0) for later launches.
 * Counting uses a separately compiled kernel variant; results are unchanged. */
int rc_debug_counters(rc_ctx *ctx, int32_t enable, uint64_t *out, int32_t n);
/* Device PCG streams: for each (x, y, frame) triple in xyf[3*i..], writes the
 * seed initSeed(x, y, frame) (pathtrace.comp:94-98) followed by `count` pcg()
 * outputs (pathtrace.comp:80-84) to out[i*(count+1) ..]. n <= 2^20, count <= 1024. */
int rc_debug_pcg(rc_ctx *ctx, const int32_t *xyf, int32_t n, int32_t count, uint32_t *out);
/* Device BVH queries on the uploaded scene. rays: n records of 8 floats
 * {origin.xyz, maxDist, dir.xyz, unused}. Writes the closest-hit triangle
 * (traverseBVH, pathtrace.comp:142-203; -1 = miss) and its t (1e30 on a miss),
 * and the shadow-ray result (shadowRay, pathtrace.comp:211-242; 1 = unoccluded
 * within maxDist). n <= 2^22. */
int rc_debug_rays(rc_ctx *ctx, const float *rays, int32_t n, int32_t *triOut, float *tOut,
                  int32_t *unoccludedOut);
/* The triangle test's reciprocal (hardware reciprocal + one Newton step) against
 * the IEEE division 1.0f / x for every binary32 x with 1e-8 <= |x| <= 2^126:
 * writes the number of results that differ (0 when the fast form is exact). */
int rc_debug_recip_check(rc_ctx *ctx, uint64_t *mismatches);

#ifdef __cplusplus
}
#endif

#endif /* RETROCAM_HIP_H */
