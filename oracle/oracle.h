/*
 * oracle.h - CPU oracle for the path-tracing hot path.  TEST INFRASTRUCTURE:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * liborcpt.so, and only to check or time the product (librtxamd.so); the
 * product never links, loads or calls it.
 *
 * It is a serial, per-path restatement of shaders/raytrace.rgen (main,
 * raytrace, nextEventEstimation, sampleLights, sampleBSDF, evalBsdf,
 * pdfBSDF, ...), random.glsl, transform.glsl and the hit shaders, templated
 * on the arithmetic type: FP64 is the oracle, FP32 the CPU baseline of the
 * same logic.  Traversal is an FP64 BVH whose closest-hit answer is pinned
 * against brute force over all primitives.
 */
#ifndef ORACLE_H
#define ORACLE_H

#include <stddef.h>
#include <stdint.h>

#include "../include/rtx_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

int orc_scene_create(const rtx_scene_desc *desc, orc_scene **out);
void orc_scene_destroy(orc_scene *scene);
const char *orc_last_error(void);

/* Render the listed paths.  Path id = pixel + width*height*sample, where
 * sample counts from params->sampleOffset.  prec = 64 (oracle) or 32 (the
 * FP32 instantiation).  rgb: 3 doubles per path; sig: 2 uint32 per path
 * (final LCG state, hash of hit primitive ids), may be NULL.
 * threads <= 0 uses all hardware threads. */
int orc_render_paths(const orc_scene *scene, const rtx_frame_params *params, const uint64_t *ids,
                     size_t n, int prec, int threads, double *rgb, uint32_t *sig);

/* Closest hit / occlusion with the oracle's FP64 traversal.
 * brute != 0 tests every primitive instead of walking the BVH. */
int orc_trace_closest(const orc_scene *scene, const rtx_ray *rays, size_t n, int brute,
                      double *t, int32_t *prim, double *u, double *v);
/* The same with the traversal arithmetic in FP32 (prec = 32) or FP64 (64). */
int orc_trace_closest_prec(const orc_scene *scene, const rtx_ray *rays, size_t n, int brute, int prec,
                           double *t, int32_t *prim, double *u, double *v);
int orc_trace_occluded(const orc_scene *scene, const rtx_ray *rays, size_t n, int brute,
                       uint8_t *occluded);

/* Building blocks, exposed for known-answer and distribution tests. */
uint32_t orc_tea(uint32_t v0, uint32_t v1);
uint32_t orc_lcg_step(uint32_t state);
uint32_t orc_pcg32_first(uint64_t initstate, uint64_t initseq);
/* Camera ray of random.glsl/raytrace.rgen:1487-1494 for pixel (px,py) of a
 * width x height launch; consumes 2 draws from *state. */
void orc_camera_ray(const rtx_camera *cam, uint32_t width, uint32_t height, uint32_t px, uint32_t py,
                    uint32_t *state, double org[3], double dir[3]);
/* BSDF pieces of raytrace.rgen:144-599 in FP64.  Vectors are world space,
 * normal is the (flipped) shading normal, wi points away from the surface. */
double orc_fresnel(double eta, double cosThetaI);
double orc_fresnel_conductor(double cosThetaI, double eta, double k);
double orc_beckmann_d(const rtx_material *mat, const double n[3], const double m[3]);
double orc_pdf_bsdf(const rtx_material *mat, const double n[3], const double wi[3], const double wo[3]);
void orc_eval_bsdf(const rtx_material *mat, const double n[3], const double wi[3], const double wo[3],
                   int frontFace, double out[3]);
double orc_sample_bsdf(const rtx_material *mat, const double wi[3], const double n[3], int frontFace,
                       uint32_t *state, double wo[3]);
/* Direction samplers of random.glsl with the draw order of the reference. */
void orc_sample_cosine(const double n[3], uint32_t *state, double out[3]);
void orc_sample_cosine_power(const double refl[3], double p, uint32_t *state, double out[3]);
void orc_sample_beckmann_normal(const rtx_material *mat, const double n[3], uint32_t *state,
                                double out[3]);
void orc_sample_unit_sphere(uint32_t *state, double out[3]);
/* sampleLights (raytrace.rgen:358-473): returns the pdf, fills direction,
 * colour and distance. */
double orc_sample_lights(const orc_scene *scene, const double origin[3], const double normal[3],
                         uint32_t *state, double lightDir[3], double lightColor[3],
                         double *lightDistance);

#ifdef __cplusplus
}
#endif

#endif
