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

/* Builds the BVH (reordering triangles) and the light list. Returns 0 or -1. */
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

#ifdef __cplusplus
}
#endif

#endif /* RC_HOST_H */
