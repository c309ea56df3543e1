/*
 * rc_oracle.h — CPU oracle of the RetroCam render hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is a scalar C restatement of the reference's
 * GLSL programs, used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker. The product library (libretrocam_hip.so) never
 * links, loads or calls it.
 *
 * Each function cites the reference file:line it restates (paths relative to
 * src/main/resources/shaders/ unless noted). Arithmetic is IEEE binary32 with
 * no FMA contraction (-ffp-contract=off) in the reference's operand order;
 * transcendentals come from rc_fmath.h, the routines the device code uses.
 */
#ifndef RC_ORACLE_H
#define RC_ORACLE_H

#include <stdint.h>
#include "retrocam_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rco_scene {
    const float *tris;   /* 32 floats per triangle */
    const float *nodes;  /* 12 floats per node     */
    const float *mats;   /* 16 floats per material */
    const float *lights; /* 20 floats per light    */
    int32_t nTris, nNodes, nMats, nLights;
    float totalLightPower;
} rco_scene;

/* Output images; any pointer may be NULL. Sizes: width*height elements of
 * float4 / float / half4 / half4 / int32 / int32 / double4. */
typedef struct rco_frame {
    float *accum;
    float *variance;
    uint16_t *gbuf;
    uint16_t *galbedo;
    int32_t *primId;
    int32_t *gbufId;
    double *accum64;
} rco_frame;

/* Work counters of a pass (summed over the pixels rendered). */
typedef struct rco_stats {
    uint64_t paths;          /* camera paths traced (3 per pixel with LCA)        */
    uint64_t rays;           /* closest-hit queries (camera + extension + G-buffer) */
    uint64_t shadowRays;     /* any-hit queries                                   */
    uint64_t nodePops;       /* closest-hit node pops (48 B node reads)           */
    uint64_t childTests;     /* closest-hit interior child-pair tests (2 x 32 B)  */
    uint64_t triTests;       /* closest-hit triangle tests (48 B positions)       */
    uint64_t hitUpdates;     /* accepted closer hits (48 B normals + 4 B matIndex) */
    uint64_t shadeHits;      /* material fetches (64 B) at valid hits             */
    uint64_t shadowPops;     /* any-hit node pops (48 B)                          */
    uint64_t shadowTriTests; /* any-hit triangle tests (48 B)                     */
    uint64_t lightScans;     /* CDF entries read during light selection (4 B)     */
    uint64_t lightReads;     /* selected light records read (80 B)                */
} rco_stats;

/* PCG-RXS-M-XS step (pathtrace.comp:80-84). */
uint32_t rco_pcg(uint32_t *state);
/* initSeed (pathtrace.comp:94-98). */
uint32_t rco_init_seed(int32_t x, int32_t y, int32_t frame);
/* rand (pathtrace.comp:86-88). */
float rco_rand(uint32_t *state);

/* Closest hit (pathtrace.comp:142-203); returns the triangle index or -1. */
int32_t rco_closest_hit(const rco_scene *s, const float *ro, const float *rd,
                        float *tOut, float *normalOut);
/* Brute force over all triangles with the same triangle test; smallest t,
 * ties to the lowest index. Returns the triangle index or -1. */
int32_t rco_brute_closest(const rco_scene *s, const float *ro, const float *rd, float *tOut);
/* Shadow ray (pathtrace.comp:211-242): 1 if unoccluded. */
int32_t rco_shadow_ray(const rco_scene *s, const float *ro, const float *rd, float maxDist);

/* One progressive sample per pixel (pathtrace.comp:601-655) for frameIndex.
 * Renders pixels with x % pxStride == 0 and y % pxStride == 0 (1 = all).
 * OpenMP across rows when built with -fopenmp. Returns 0 or -1. */
int32_t rco_render_pass(const rco_scene *s, const rc_camera *cam, const rc_pt_params *p,
                        int32_t frameIndex, int32_t pxStride, rco_frame *out, rco_stats *stats);

/* Element-wise rc_fmath routines: fn 0 sin, 1 cos, 2 exp, 3 log, 4 pow(x, y),
 * 5 binary32 -> binary16 (RNE) -> binary32. Returns 0, or -1 for a bad fn. */
int32_t rco_fmath(int32_t fn, const float *x, const float *y, float *out, int32_t n);

/* Number of OpenMP threads the oracle uses (1 without OpenMP). */
int32_t rco_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* RC_ORACLE_H */
