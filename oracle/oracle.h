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
 * sample counts from params->sampleOffset; one path is one frame of one
 * pixel (main(), raytrace.rgen:1623-1711): pc.samplesPerPixel camera
 * samples on one random stream, then the frame's queued splits, averaged.  prec = 64 (oracle) or 32 (the
 * FP32 instantiation).  rgb: 3 doubles per path; sig: 2 uint32 per path
 * (final LCG state, hash of hit primitive ids), may be NULL.
 * threads <= 0 uses all hardware threads. */
int orc_render_paths(const orc_scene *scene, const rtx_frame_params *params, const uint64_t *ids,
                     size_t n, int prec, int threads, double *rgb, uint32_t *sig);

/* ADRRS: the per-pixel estimate (RGBA, width*height) that frames with
 * pc.useADRRS read (estimateImage, raytrace.rgen:1632-1636). */
int orc_estimate_set(orc_scene *scene, const float *rgba, uint32_t width, uint32_t height);
/* Counts since the last call (then reset): split records queued, paths
 * ended by the weight window's roulette, split paths traced. */
void orc_adrrs_counters(uint64_t out[3]);
/* applyWeightWindow (raytrace.rgen:830-869) in FP64 with the given pixel
 * estimate, adjoint (cache irradiance) and path throughput; *state is the
 * LCG state (one draw when the window splits).  Returns q, sets *n. */
double orc_weight_window(const double estimate[3], const double adjoint[3], const double throughput[3],
                         double adrrsS, uint32_t *state, int *n);

/* Path-guiding regions (boxes) and their vMF mixtures, read by renders
 * with pc.useGuiding (the reference's bindings 15-17). */
int orc_guiding_set(orc_scene *scene, const rtx_aabb *regions, const rtx_vmm_theta *vmms, uint32_t n);
/* VMM pdf (guiding.glsl:57-65) and sampleVMM (91-104) in FP64. */
double orc_vmm_pdf(const rtx_vmm_theta *m, const double x[3], const double wo[3], int parallax);
void orc_sample_vmm(const rtx_vmm_theta *m, const double x[3], int parallax, uint32_t *state, double out[3]);

/* orc_render_paths with pc.updateGuiding: also returns each path's guiding
 * sample record, RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL DirectionalData per path
 * (the slots of its pixel in the reference's buffer, binding 18). */
int orc_render_paths_record(const orc_scene *scene, const rtx_frame_params *params, const uint64_t *ids,
                            size_t n, int prec, int threads, double *rgb, uint32_t *sig,
                            rtx_directional_data *dd);

/* Closest hit / occlusion with the oracle's FP64 traversal.
 * brute != 0 tests every primitive instead of walking the BVH. */
int orc_trace_closest(const orc_scene *scene, const rtx_ray *rays, size_t n, int brute,
                      double *t, int32_t *prim, double *u, double *v);
/* The same with the traversal arithmetic in FP32 (prec = 32) or FP64 (64). */
int orc_trace_closest_prec(const orc_scene *scene, const rtx_ray *rays, size_t n, int brute, int prec,
                           double *t, int32_t *prim, double *u, double *v);
int orc_trace_occluded(const orc_scene *scene, const rtx_ray *rays, size_t n, int brute,
                       uint8_t *occluded);

/* Irradiance cache (shaders/raytrace.rgen:748-828, 1228-1421).
 * orc_ic_frame renders one frame - every pixel of params at sample
 * params->sampleOffset - against the cache state passed in (frames with
 * pc.useIrradianceCache, or pc.useADRRS, which looks the cache up), then
 * runs the frame's cache updates and creations and writes the new state
 * back (hdr, and entries/spheres of hdr->maxCaches records).  rgb/sig per
 * pixel as in orc_render_paths.  The frame's jobs (updates first, in pixel
 * order, then creations) and their 200 sub-path results each (L rgb, first
 * hit distance; sig) are returned up to jobCap jobs; any of them may be NULL. */
typedef struct orc_ic_job {
    int32_t type;    /* 0 = update of entry `target`, 1 = creation */
    uint32_t target;
    double origin[3];
    double normal[3];
} orc_ic_job;

int orc_ic_frame(const orc_scene *scene, const rtx_frame_params *params, int prec, int threads,
                 rtx_cache_header *hdr, rtx_cache_data *entries, rtx_sphere *spheres, double *rgb,
                 uint32_t *sig, orc_ic_job *jobs, size_t jobCap, size_t *numJobs, double *sub,
                 uint32_t *subSig);
/* The accumulation and commit of a frame's jobs, given their sub-path
 * results (4 doubles each, 200 per job): the state is updated in place. */
int orc_ic_apply(const rtx_frame_params *params, int prec, rtx_cache_header *hdr, rtx_cache_data *entries,
                 rtx_sphere *spheres, const orc_ic_job *jobs, size_t numJobs, const double *sub);
/* calculateCacheData's accumulation over one job's 200 sub-path results
 * (sub: 4 doubles each, stratum order k * 10 + j). */
int orc_ic_reduce(int prec, uint64_t seed, uint32_t sample, uint32_t job, const double normal[3],
                  const double *sub, double color[3], double rot[3], double trans[3], double *harmonicR);
/* calculateCacheData (raytrace.rgen:1231-1316) for one point: the 200
 * sub-paths of job `job` of frame params->sampleOffset against the given
 * cache state, then the accumulation. */
int orc_ic_calculate(const orc_scene *scene, const rtx_frame_params *params, int prec,
                     const rtx_cache_header *hdr, const rtx_cache_data *entries, const rtx_sphere *spheres,
                     uint32_t job, const double origin[3], const double normal[3], double color[3], double rot[3],
                     double trans[3], double *harmonicR);
/* phi, theta and the stream state after the theta draw of stratum (k, j). */
void orc_ic_stratum(uint64_t seed, uint32_t sample, uint32_t job, int k, int j, double *phi, double *theta,
                    uint32_t *state);
/* Cache lookup at x with normal n: 1 and the weighted colour, or 0. */
int orc_ic_query(const orc_scene *scene, const rtx_push_constant *pc, const rtx_cache_header *hdr,
                 const rtx_cache_data *entries, const rtx_sphere *spheres, const double x[3], const double n[3],
                 double color[3]);

/* Path-guiding update (em_oracle.cpp): PathGuiding's regions and per-region
 * vMF mixtures, refit from recorded samples with lightpmm's MAP-EM (fit on
 * the first update, updateFit after), parallax-aware when `parallax`.
 * Component split/merge is not restated. */
typedef struct orc_guide orc_guide;
int orc_guide_create(const rtx_aabb *sceneBox, uint32_t splitCount, int parallax, int prec, orc_guide **out);
void orc_guide_destroy(orc_guide *g);
/* Region boxes and VMM_Theta (as syncPMMsToVMM_Thetas fills them); returns the region count. */
uint32_t orc_guide_regions(const orc_guide *g, rtx_aabb *boxes, rtx_vmm_theta *vmms, uint32_t capacity);
uint32_t orc_guide_regions_wide(const orc_guide *g, rtx_aabb *boxes, rtx_vmm_theta_wide *vmms, uint32_t capacity);
/* PathGuiding::update with these samples (flags = region, 0xFFFFFFFF = unused). */
int orc_guide_update(orc_guide *g, const rtx_directional_data *samples, size_t n);
const char *orc_guide_last_error(void);
float orc_fastexp(float x);
/* postFit's split/merge switches (default: the reference's). */
int orc_guide_set_options(orc_guide *g, const rtx_guiding_options *options);
/* Test hooks in FP64.  A lobe is 8 doubles: weight, mu xyz, kappa, mean
 * cosine r, norm, e^-2kappa.  orc_vmf_combine: VMFKernel::product (divide =
 * 0) or ::division (1) of a by b; orc_vmf_merge: ::mergeComponent;
 * orc_merge_metric: the chi-square merge metric of the pair;
 * orc_eigen2x2: eigenvalues and eigenvector matrix (row-major, columns are
 * the vectors) of the 2x2 matrix a (row-major) as Eigen's EigenSolver
 * computes them. */
void orc_vmf_combine(const double a[8], const double b[8], int divide, double out[8]);
void orc_vmf_merge(const double a[8], const double b[8], double out[8]);
double orc_merge_metric(const double a[8], const double b[8]);
void orc_eigen2x2(const double a[4], double lam[2], double vec[4]);
/* The MAP-EM's steps on one FP64 mixture, for closed-form tests.  A mixture
 * is orc_em_mix_doubles() doubles: K, then 16 x (w, mu xyz, kappa, r), then
 * sampleWeight, numSamples, totalNumSamples.  props: minItr, maxItr,
 * relLogLikelihoodThreshold, maxKappa, vPrior, rPrior, rPriorWeight. */
int orc_em_mix_doubles(void);
/* The order std::partition (libstdc++, bidirectional form) leaves for
 * "keys[i] <= pivot"; out receives the original indices. */
size_t orc_partition_order(const double *keys, size_t n, double pivot, uint32_t *out);
int orc_em_estep(const double *mix, const double *dirs, const double *weights, size_t n, double *rW, double *W,
                 double *sumWeight, double *logLik);
int orc_em_mstep(const double props[7], double *mix, const double *rW, const double *W);
int orc_em_fit(const double props[7], double *mix, const double *dirs, const double *weights, size_t n, int update);

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
/* D and G1(v) of the material's microfacet distribution (rtx_material.distribution) */
double orc_microfacet_d(const rtx_material *mat, const double n[3], const double m[3]);
double orc_microfacet_g1(const rtx_material *mat, const double n[3], const double m[3], const double v[3]);
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
/* Texture reads: a material texture (slot >= 1) at (u, v) as RGBA, and the
 * environment map's colour in direction dir (black without a map). */
int orc_tex_sample(const orc_scene *scene, int slot, double u, double v, double out[4]);
void orc_env_color(const orc_scene *scene, const double dir[3], double out[3]);

#ifdef __cplusplus
}
#endif

#endif
