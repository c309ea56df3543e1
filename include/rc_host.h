/*
 * rc_host.h — C API of librc_host.so: host-side scene assembly for tests,
 * benchmarks and C++ callers of libretrocam_hip.so.
 *
 * In the reference these steps are Java and stay Java in production (the Java
 * host keeps the scene, the BVH builder and the struct writers). This library
 * restates them in C++ with Java's binary32 semantics so that tests and the
 * benchmark can produce the same float streams without a JVM:
 *
 *   Box.toTriangles        src/main/java/com/retrocam/scene/Box.java:25-76
 *   Sphere.toTriangles     src/main/java/com/retrocam/scene/Sphere.java:34-102
 *   TransformedPrimitive   src/main/java/com/retrocam/scene/TransformedPrimitive.java:31-97
 *   BVHBuilder.build       src/main/java/com/retrocam/scene/BVHBuilder.java:35-164
 *   uploadLights           src/main/java/com/retrocam/scene/SceneUploader.java:144-216
 *   OrbitCamera basis      src/main/java/com/retrocam/camera/OrbitCamera.java:155-291
 *   tanHalfFov             src/main/java/com/retrocam/camera/ThinLensCamera.java:64-68
 *
 * Output streams use the wire layouts of retrocam_hip.h (32/12/16/20 floats).
 */
#ifndef RC_HOST_H
#define RC_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rcs_scene rcs_scene;

rcs_scene *rcs_new(void);
void rcs_free(rcs_scene *s);

/* Appends a material (16 floats, Material.write layout); returns its index or -1. */
int32_t rcs_add_material(rcs_scene *s, const float *mat16);
/* Box of centre c and half-extents h (Box.java:16-46). Returns triangles added or -1. */
int32_t rcs_add_box(rcs_scene *s, float cx, float cy, float cz,
                    float hx, float hy, float hz, int32_t mat);
/* UV sphere with pole rows skipped (Sphere.java:34-62). Returns triangles added or -1. */
int32_t rcs_add_sphere(rcs_scene *s, float cx, float cy, float cz, float radius,
                       int32_t mat, int32_t stacks, int32_t slices);
/* Two-triangle quad a,b,c,d with flat normal (Box.addQuad, Box.java:49-76). */
int32_t rcs_add_quad(rcs_scene *s, const float *abcd12, int32_t mat);
/* Rotates the triangles added since index `first` about (px,py,pz) by XYZ Euler
 * degrees, R = Rz*Ry*Rx (TransformedPrimitive.java:31-97). Returns 0 or -1. */
int32_t rcs_rotate_since(rcs_scene *s, int32_t first, float px, float py, float pz,
                         float degX, float degY, float degZ);
/* Appends raw triangles (32-float records). Returns 0 or -1. */
int32_t rcs_add_triangles(rcs_scene *s, const float *tris32, int32_t n);
int32_t rcs_triangle_count(const rcs_scene *s);

/* Builds the BVH (reordering triangles) and the light list. Returns 0 or -1;
 * -1 also where BVHBuilder.java:117-119 throws (a bucket index below 0, from a
 * centroid that rounds under a thin node's minimum), with the triangles left in
 * their original order. */
int32_t rcs_build(rcs_scene *s);

typedef struct rcs_info {
    int32_t nTris, nNodes, nMats, nLights;
    float totalLightPower;
    int32_t maxDepth;      /* deepest node level (root = 0) */
    int32_t maxLeafSize;   /* largest primCount of any leaf */
    int32_t nLeaves;
} rcs_info;

int32_t rcs_get_info(const rcs_scene *s, rcs_info *out);
/* Copies the built streams; each dst must hold n*record floats. Return 0 or -1. */
int32_t rcs_copy_tris(const rcs_scene *s, float *dst);
int32_t rcs_copy_nodes(const rcs_scene *s, float *dst);
int32_t rcs_copy_mats(const rcs_scene *s, float *dst);
int32_t rcs_copy_lights(const rcs_scene *s, float *dst);

/* OrbitCamera eye/forward/right/up (each 3 floats) for the given state. */
void rcs_orbit_basis(float theta, float phi, float radius,
                     float tx, float ty, float tz,
                     float *eye, float *fwd, float *right, float *up);
/* Look-at basis with the same construction (forward, right = f x (0,1,0), up = r x f). */
void rcs_lookat_basis(const float *eye, const float *target,
                      float *fwd, float *right, float *up);
/* (float)tan(atan(0.012 / (focalLengthMm / 1000))) */
float rcs_tan_half_fov(float focalLengthMm);

/* ---- scene files (rc_scenejson.cpp) --------------------------------------
 * SceneSerializer.load (src/main/java/com/retrocam/scene/SceneSerializer.java:72-212)
 * with the Java reader's scanning rules, so a file yields the values the Java
 * host reads from it, and SceneEditor.buildScene (SceneEditor.java:101-107).
 * The writer (SceneSerializer.save) is retrocam_amd/scene_file.py. */

enum { RCS_OBJ_BOX = 0, RCS_OBJ_SPHERE = 1 };  /* SceneObject.Type */
enum {
    RCS_JSON_OK = 0,
    RCS_JSON_ERR_ARG = 1,
    RCS_JSON_ERR_OOM = 2,
    RCS_JSON_ERR_STRING = 3  /* a string scan would run past the text (the Java reader throws) */
};

/* SceneObject fields (SceneObject.java:14-30): position, scale (a sphere's
 * radius is scale[0]), XYZ rotation in degrees about the position. */
typedef struct rcs_json_object {
    int32_t type;
    float pos[3];
    float scale[3];
    float rot[3];
    int32_t stacks, slices;    /* 24 unless the file gives > 2 */
    int32_t materialIndex;     /* >= 0 */
} rcs_json_object;

typedef struct rcs_doc rcs_doc;

/* Parses len bytes of scene-file text; NULL on failure with *status (may be NULL) set. */
rcs_doc *rcs_json_parse(const char *text, size_t len, int32_t *status);
void rcs_json_free(rcs_doc *d);
int32_t rcs_json_material_count(const rcs_doc *d);
int32_t rcs_json_object_count(const rcs_doc *d);
/* Material i in the Material.write layout (16 floats, albedoTex -1) and its name
 * (NUL-terminated, truncated to nameCap - 1 bytes). Returns the name's length, or -1. */
int32_t rcs_json_get_material(const rcs_doc *d, int32_t i, float *mat16, char *name, int32_t nameCap);
int32_t rcs_json_get_object(const rcs_doc *d, int32_t i, rcs_json_object *o, char *name, int32_t nameCap);
/* SceneObject.toPrimitive (SceneObject.java:80-88): the box or sphere, rotated
 * when any angle is non-zero. Returns triangles added, or -1. */
int32_t rcs_add_object(rcs_scene *s, const rcs_json_object *o);
/* The document's materials, then its objects, in file order. Returns triangles added, or -1. */
int32_t rcs_json_build(const rcs_doc *d, rcs_scene *s);

/* ---- per-video-frame inputs (rc_temporal.cpp) ---------------------------
 * The values the render path takes per video frame (rc_temporal, the camera's
 * focalDist) come from these Java host objects, restated with binary32
 * semantics:
 *   TemporalState   src/main/java/com/retrocam/core/TemporalState.java:13-111
 *   KeyframeTrack   src/main/java/com/retrocam/keyframe/KeyframeTrack.java:20-178
 *   RenderJob timing src/main/java/com/retrocam/export/RenderJob.java:34-41 and
 *                   RenderPipeline.java:150-151 (t = frame * dt) */

/* RenderSettings fields the filters read (RenderSettings.java:28-43; defaults
 * 2.0, 3.0, 6.0, 1.5). */
typedef struct rct_settings {
    float focusDistM;
    float afSpeed;
    float agcMaxGain;
    float agcSpeed;
} rct_settings;

/* TemporalState fields, including the private IIR targets. */
typedef struct rct_state {
    float agcGain;
    float focalDist;
    float time;
    float agcGainTarget;    /* setAgcGainTarget */
    float focalDistTarget;  /* setFocalDistTarget */
    float whiteBalance[9];  /* row-major 3x3, identity */
} rct_state;

void rct_init(rct_state *s);                                               /* field initialisers :16-29 */
void rct_update(rct_state *s, float dt, const rct_settings *st);           /* update :44-64 */
void rct_reset_for_video(rct_state *s, const rct_settings *st);            /* resetForVideoRendering :77-84 */
void rct_update_for_video_frame(rct_state *s, float canonicalTime, float frameDt,
                                const rct_settings *st);                   /* updateForVideoFrame :96-111 */
/* nFrames video frames from frame 0 (RenderPipeline.tickVideoFrame :147-171):
 * per frame, focalDistTarget := focusTargets[i] (when not NULL), then
 * updateForVideoFrame(i * dt, dt) with dt = rcj_frame_dt(fps); writes the
 * resulting time, agcGain and focalDist per frame (any output may be NULL).
 * Returns nFrames, or -1. */
int32_t rct_run_video(rct_state *s, const rct_settings *st, const float *focusTargets, float fps, int32_t nFrames,
                      float *outTime, float *outAgcGain, float *outFocalDist);

float rcj_frame_dt(float fps);                                  /* RenderJob.getFrameDt */
int32_t rcj_total_frames(float durationSeconds, float fps);     /* RenderJob.getTotalFrames (video) */
float rcj_canonical_time(int32_t frame, float frameDt);         /* frame * dt in binary32 */

/* Keyframe tracks: keys kept in ascending time; interpolation per key. */
typedef struct rck_track rck_track;
enum { RCK_STEP = 0, RCK_LINEAR = 1, RCK_BEZIER = 2 };  /* BEZIER = Catmull-Rom */

rck_track *rck_new(void);
void rck_free(rck_track *t);
/* addKeyframe: overwrites a key within 0.001 s, else sorted insert; interp is
 * clamped to 0..2 (Interpolation.fromIndex). Returns the key's index, or -1. */
int32_t rck_add(rck_track *t, float time, float value, int32_t interp);
/* removeKeyframe: out-of-range indices are ignored. Returns the key count, or -1. */
int32_t rck_remove(rck_track *t, int32_t index);
int32_t rck_count(const rck_track *t);
int32_t rck_get(const rck_track *t, int32_t index, float *time, float *value, int32_t *interp);
/* evaluate(time, defaultValue): clamped outside the keys, STEP / LINEAR /
 * Catmull-Rom inside. */
float rck_evaluate(const rck_track *t, float time, float defaultValue);
int32_t rck_evaluate_many(const rck_track *t, const float *times, int32_t n, float defaultValue, float *out);

#ifdef __cplusplus
}
#endif

#endif /* RC_HOST_H */
