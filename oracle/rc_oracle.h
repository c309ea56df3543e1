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

/* Post chain of one frame (rc_oracle_post.c): p00 normalize of the accumulation
 * sum, the a-trous passes, the enabled VHS passes p01..p16 (PostProcessStack.
 * runChain, src/main/java/com/retrocam/post/PostProcessStack.java:109-280) and the
 * display encode (display.frag with totalSamples = exposure = 1, then RGBA8).
 * gbuf/galbedo/variance may be NULL (then no a-trous: runOnImage). outHdr
 * (float4) and outRgba8 (uint8 x4) may be NULL. Returns 0 or -1. */
int32_t rco_post_chain(const float *accum, const uint16_t *gbuf, const uint16_t *galbedo, const float *variance,
                       int32_t W, int32_t H, int32_t totalSamples, const rc_post_params *p, const rc_temporal *t,
                       float *outHdr, uint8_t *outRgba8);
/* One pass over a float4 image: 0 = p00, 1..16 = VHS pass pN (5, 6 = copy),
 * 100 + i = a-trous pass i (step 1 << i), 200 = display encode (float, pre-8-bit). */
int32_t rco_post_pass(int32_t pass, const float *src, float *dst, int32_t W, int32_t H, const rc_post_params *p,
                      const rc_temporal *t, const uint16_t *gbuf, const uint16_t *galbedo, const float *variance,
                      int32_t totalSamples);

/* ---- SPPM (rc_oracle_sppm.c) ---- */
/* Photon seed initSeed(idx) (photon_trace.comp:68-72). */
uint32_t rco_photon_seed(uint32_t idx, int32_t seed);
/* cellHash of integer cell coordinates (photon_trace.comp:140-143). */
uint32_t rco_cell_hash(int32_t cx, int32_t cy, int32_t cz, int32_t hashSize);
/* Unordered closest hit of the photon shaders; triangle index or -1. */
int32_t rco_walk_closest(const rco_scene *s, const float *ro, const float *rd, float *tOut);
/* PhotonMap.clear + photon_trace.comp over photonCount photons; photons is
 * photonCount x 12 floats, heads hashSize ints, next photonCount ints. Photons are
 * inserted in photon-index order. *storedOut = number of stored photons. */
int32_t rco_photon_trace(const rco_scene *s, int32_t photonCount, float radius, int32_t seed, int32_t hashSize,
                         float *photons, int32_t *heads, int32_t *next, int32_t *storedOut);
/* photon_gather.comp for every pixel of cam (pinhole, aspect = width/height):
 * accum (float4 per pixel) += (L, 0); accum64 (optional) += L in double. */
int32_t rco_photon_gather(const rco_scene *s, const rc_camera *cam, float radius, int32_t hashSize,
                          const float *photons, const int32_t *heads, const int32_t *next, float *accum,
                          double *accum64);
/* SPPMManager.updateRadius with Java float/double semantics. */
float rco_sppm_next_radius(float r, int32_t iteration, float alpha);

/* VHS integer hashes: kind 0 pcgHash (p07_ccd_noise.frag:50-54), 1 one-round
 * uHash (p12_dropout.frag:32-37), 2 two-round uHash (p13_headswitch.frag:40-45),
 * 3 / 4 the float bits of fHash built on 1 / 2. */
uint32_t rco_hash(int32_t kind, uint32_t a, uint32_t b);

/* Element-wise rc_fmath routines: fn 0 sin, 1 cos, 2 exp, 3 log, 4 pow(x, y),
 * 5 binary32 -> binary16 (RNE) -> binary32. Returns 0, or -1 for a bad fn. */
int32_t rco_fmath(int32_t fn, const float *x, const float *y, float *out, int32_t n);

/* ---- sub-steps of pathtrace.comp, for the oracle's own closed-form pins ---- */
/* generateRay (pathtrace.comp:401-439) for pixel (px, py) from the seed stream
 * *seed (advanced as the shader advances it), focal plane offset focalOffset
 * (the LCA channel offset). Returns 0 or -1. */
int32_t rco_gen_ray(const rc_camera *cam, int32_t px, int32_t py, uint32_t *seed, float focalOffset, float *ro,
                    float *rd);
/* sampleNGon (pathtrace.comp:310-321): lens point xy[2] of radius/blades. */
void rco_sample_ngon(uint32_t *seed, float radius, int32_t blades, float *xy);
/* sampleGGXVNDF (pathtrace.comp:252-276) in the local frame with y = normal. */
void rco_sample_vndf(const float *wo, float alpha, float u1, float u2, float *h);
/* smithG1GGX / smithG2GGX (pathtrace.comp:278-292). */
float rco_smith_g1(float nv, float alpha);
float rco_smith_g2(float nl, float nv, float alpha);
/* sampleCosineHemisphere about n (pathtrace.comp:296-306). */
void rco_sample_cosine(float u1, float u2, const float *n, float *dir);
/* The specular scattering step of traceRadiance at a hit (glass when
 * mat[10] > 0.5, else metal when mat[3] > 0.5; pathtrace.comp:504-566): ray
 * direction rd, the hit's normal before facing, hit position, one material
 * record (16 floats) and the seed stream. Writes the new origin, direction and
 * throughput factor, and for glass the Fresnel term (-1 for metal). Returns 0,
 * or -1 for a diffuse material. */
int32_t rco_scatter(const float *mat, const float *rd, const float *geomNormal, const float *pos, uint32_t *seed,
                    float *roOut, float *rdOut, float *thrOut, float *fresnelOut);

/* Number of OpenMP threads the oracle uses (1 without OpenMP). */
int32_t rco_num_threads(void);

#ifdef __cplusplus
}
#endif

#endif /* RC_ORACLE_H */
