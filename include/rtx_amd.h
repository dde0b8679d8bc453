/*
 * rtx_amd.h - C ABI of the MI355X path-tracing library (librtxamd.so).
 *
 * The library replaces the device side of FelixFifi/rtx-pathtracer: the
 * Vulkan ray-tracing pipeline launched by RayTracingApp::raytrace
 * (src/RayTracingApp.cpp:1120-1204, traceRaysKHR at 1199-1201) with its
 * raygen/closest-hit/any-hit/miss/intersection shaders
 * (shaders/raytrace.rgen, raytrace.rchit, raytrace.rahit, raytrace.rmiss,
 * raytrace.shadow.rmiss, raytrace.sphere.rint/.rchit) and the driver-built
 * acceleration structures (src/SceneLoader.cpp:1101-1193).
 *
 * Conventions
 *   - every entry point returns int: RTX_OK (0) or a negative RTX_E_* code;
 *     rtx_last_error() gives the message.  Nothing throws across the ABI.
 *   - caller-owned host arrays are copied during the call; device memory is
 *     owned by the library and released by the matching *_destroy.
 *   - one context per host thread; calls on one context are serialised by
 *     the caller (the reference drives one frame at a time from
 *     RayTracingApp::drawCallback, src/RayTracingApp.cpp:79-225).
 *   - the structs below are byte-identical to the reference's device
 *     layouts, so a host that fills the reference's buffers can hand the
 *     same bytes to this library.
 */
#ifndef RTX_AMD_H
#define RTX_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- errors */
#define RTX_OK 0
#define RTX_E_INVALID (-1)     /* bad argument or inconsistent scene data */
#define RTX_E_DEVICE (-2)      /* HIP runtime error                       */
#define RTX_E_MEMORY (-3)      /* device or host allocation failed        */
#define RTX_E_UNSUPPORTED (-4) /* feature outside this build              */
#define RTX_E_STATE (-5)       /* call out of order (e.g. no render yet)  */

/* ------------------------------------------------------- limits / enums */
/* shaders/limits.glsl:1-7 */
#define RTX_SIZE_LIGHT_RANDOM 10000
#define RTX_SIZE_TRI_RANDOM 10000
#define RTX_MAX_DISTRIBUTIONS 16 /* lobes of a VMM_Theta (shaders/limits.glsl:6) */
#define RTX_MAX_LOBES 32         /* lobes a device mixture may hold (rtx_guiding_options.maxComponents) */
#define RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL 16

/* shaders/raytrace.rgen:21-27; src/SceneLoader.h:31-39 */
#define RTX_MAT_DIFFUSE 0
#define RTX_MAT_SPECULAR 1
#define RTX_MAT_DIELECTRIC 2
#define RTX_MAT_LIGHT 3
#define RTX_MAT_PHONG 4
#define RTX_MAT_CONDUCTOR 5
#define RTX_MAT_ROUGH_CONDUCTOR 6

/* rtx_material.distribution of a rough conductor.  BECKMANN is the
 * reference's (shaders/raytrace.rgen:179-230, random.glsl:127-136, with its
 * G1 as written); GGX is new (Walter et al. 2007: GGX D, separable Smith G,
 * sampling of D(m)cos(theta_m)), selected by the Mitsuba roughconductor's
 * "distribution" string, which the reference's loader ignores
 * (src/SceneLoader.cpp:601-605). */
#define RTX_DISTRIBUTION_BECKMANN 0
#define RTX_DISTRIBUTION_GGX 1

/* shaders/raytrace.rgen:29-32; src/SceneLoader.h:41-46 */
#define RTX_LIGHT_AREA 0
#define RTX_LIGHT_POINT 1
#define RTX_LIGHT_SPHERE 2
#define RTX_LIGHT_ENV_MAP 3

/* ------------------------------------------------- device data layouts */
/* Vertex, std140, 48 B: shaders/wavefront.glsl:1-7; src/Model.h:24-33 */
typedef struct rtx_vertex {
    float pos[3];
    float _pad0;
    float normal[3];
    float _pad1;
    float texCoord[2];
    int32_t materialIndex;
    int32_t _pad2;
} rtx_vertex;

/* Material, std430, 96 B: shaders/wavefront.glsl:9-23; src/SceneLoader.h:48-62 */
typedef struct rtx_material {
    float lightColor[3];
    float _pad0;
    float diffuse[3];
    float _pad1;
    float specular[3];
    float specularHighlight;
    float transparency;
    float refractionIndex;
    float refractionIndexInv;
    float eta;
    float k;
    float roughness;
    int32_t textureIdDiffuse;
    int32_t textureIdSpecular;
    int32_t type;
    /* Microfacet distribution of RTX_MAT_ROUGH_CONDUCTOR: RTX_DISTRIBUTION_*.
     * New: the reference has Beckmann only and its std430 padding here is
     * zero, so its bytes select Beckmann. */
    int32_t distribution;
    int32_t _pad2[2];
} rtx_material;

/* InstanceInfo, std140, 144 B: shaders/wavefront.glsl:25-30; src/SceneLoader.h:64-69.
 * Matrices are column-major (GLM / GLSL convention): m[col*4 + row]. */
typedef struct rtx_instance {
    float transform[16];
    float normalTransform[16];
    int32_t modelIndex;
    int32_t iLight;
    int32_t _pad[2];
} rtx_instance;

/* Light, scalar, 40 B: shaders/wavefront.glsl:32-39; src/SceneLoader.h:71-78 */
typedef struct rtx_light {
    float color[3];
    float pos[3];
    uint32_t instanceIndex;
    float sampleProb;
    float area;
    int32_t type;
} rtx_light;

/* FaceSample, 12 B: shaders/wavefront.glsl:41-45; src/SceneLoader.h:80-84 */
typedef struct rtx_face_sample {
    int32_t index;
    float sampleProb;
    float faceArea;
} rtx_face_sample;

/* sphere, scalar, 24 B: shaders/raycommon.glsl:71-76; src/Shapes.h:65-75 */
typedef struct rtx_sphere {
    float center[3];
    float radius;
    int32_t materialIndex;
    int32_t iLight;
} rtx_sphere;

/* aabb, scalar, 24 B: shaders/raycommon.glsl:78-81; src/Shapes.h:19-63 */
typedef struct rtx_aabb {
    float min[3];
    float max[3];
} rtx_aabb;

/* cacheHeader / cacheData, scalar: shaders/raycommon.glsl:83-96;
 * src/IrradianceCache.h:17-30 */
typedef struct rtx_cache_header {
    uint32_t nextCacheSlot;
    uint32_t maxCaches;
    uint32_t nextUpdateSlot;
} rtx_cache_header;

typedef struct rtx_cache_data {
    float color[3];
    float normal[3];
    float rotGrad[3];
    float transGrad[3];
    float harmonicR;
    uint32_t numUpdates;
} rtx_cache_data;

/* VMF_Theta / VMM_Theta, scalar, 40 B / 720 B: shaders/guiding.glsl:14-51;
 * src/PathGuiding.h:36-75 */
typedef struct rtx_vmf_theta {
    float mu[3];
    float k;
    float norm;
    float eMin2K;
    float distance;
    float target[3];
} rtx_vmf_theta;

typedef struct rtx_vmm_theta {
    rtx_vmf_theta thetas[RTX_MAX_DISTRIBUTIONS];
    float pi[RTX_MAX_DISTRIBUTIONS];
    float meanPosition[3];
    int32_t usedDistributions;
} rtx_vmm_theta;

/* DirectionalData, scalar, 40 B: shaders/guiding.glsl:99-113 */
typedef struct rtx_directional_data {
    float position[3];
    float direction[3];
    float weight;
    float pdf;
    float distance;
    uint32_t flags; /* region id, or 0xFFFFFFFF = invalid */
} rtx_directional_data;

/* Push-constant block, 192 B, 48 four-byte fields in the reference's order
 * and with its names (shaders/raycommon.glsl:19-69; src/RayTracingApp.h:81-130).
 * GLSL bools are 4-byte ints.  rtx_push_constant_defaults() fills the
 * reference's defaults.  enableRR was declared but never read by the
 * reference's shaders; here it switches on Russian roulette (see
 * rtx_frame_params.rrStartDepth). */
typedef struct rtx_push_constant {
    uint32_t randomUInt;
    uint32_t previousFrames;
    int32_t maxDepth;
    int32_t maxFollowDiscrete;
    int32_t samplesPerPixel;
    int32_t enableRR;
    int32_t enableNEE;
    int32_t numNEE;
    int32_t enableAverageInsteadOfMix;
    int32_t enableMIS;
    int32_t usePowerHeuristic;
    int32_t storeEstimate;
    int32_t visualizeMode;
    int32_t showIrradianceCacheOnly;
    int32_t showIrradianceGradients;
    int32_t useIrradianceCache;
    int32_t highlightIrradianceCacheColor;
    float irradianceA;
    float irradianceUpdateProb;
    float irradianceCreateProb;
    float irradianceVisualizationScale;
    int32_t useIrradianceGradients;
    int32_t useIrradianceCacheOnGlossy;
    float irradianceGradientsMaxLength;
    int32_t isIrradiancePrepareFrame;
    int32_t irradianceNumNEE;
    float irradianceCacheMinRadius;
    int32_t irradianceCachePerformVisibilityCheck;
    int32_t useVisibleSphereSampling;
    int32_t useADRRS;
    float adrrsS;
    int32_t adrrsSplit;
    int32_t splitOnFirst;
    int32_t useGuiding;
    float guidingProb;
    float guidingVisuScale;
    float guidingVisuMax;
    int32_t guidingVisuIgnoreOcclusioon;
    int32_t updateGuiding;
    int32_t useParallaxCompensation;
    float time;
    int32_t guidingVisuMove;
    float guidingVisuPhiScale;
    float guidingVisuThetaScale;
    int32_t numGuidingRegions;
    int32_t guidingPiPHighlightRegion;
    int32_t guidingPiPShowSpheres;
    float guidingPiPSize;
} rtx_push_constant;

/* Camera uniform block, 256 B: shaders/raytrace.rgen:128-134;
 * src/RayTracingApp.h:65-70.  Column-major 4x4 matrices. */
typedef struct rtx_camera {
    float view[16];
    float proj[16];
    float viewInverse[16];
    float projInverse[16];
} rtx_camera;

/* ------------------------------------------------------ scene upload */
/* One model = one vertex and one index buffer of the reference
 * (descriptor set 1 bindings 1 and 2, src/SceneLoader.cpp:946-984). */
typedef struct rtx_model_desc {
    const rtx_vertex *vertices;
    uint32_t vertexCount;
    const uint32_t *indices; /* 3 per triangle */
    uint32_t indexCount;
} rtx_model_desc;

/* An 8-bit RGBA texture, rows top to bottom as stbi_load returns them
 * (src/SceneLoader.cpp:199): RGB sRGB-encoded, alpha linear, sampled like
 * the reference's R8G8B8A8_SRGB images with a bilinear, repeating sampler
 * (src/SceneLoader.cpp:209-227). */
typedef struct rtx_texture_desc {
    const uint8_t *rgba;
    uint32_t width;
    uint32_t height;
} rtx_texture_desc;

/* Everything descriptor set 1 bindings 1-8 hold for one scene.  The light
 * sampling tables are the reference's LightSamplerBuffer
 * (shaders/raytrace.rgen:107-110): randomLightIndex[10000] followed by
 * randomTriIndex[rows][10000], one row per area light.  The library
 * copies the data and builds the bounding-volume hierarchy that replaces
 * the driver's BLAS/TLAS (src/SceneLoader.cpp:1101-1193). */
typedef struct rtx_scene_desc {
    const rtx_model_desc *models;
    uint32_t modelCount;
    const rtx_instance *instances;
    uint32_t instanceCount;
    const rtx_material *materials;
    uint32_t materialCount;
    const rtx_light *lights;
    uint32_t lightCount;
    const int32_t *randomLightIndex;        /* RTX_SIZE_LIGHT_RANDOM entries */
    const rtx_face_sample *randomTriIndex;  /* randomTriRows * RTX_SIZE_TRI_RANDOM */
    uint32_t randomTriRows;
    const rtx_sphere *spheres;
    uint32_t sphereCount;
    /* Optional lat-long environment map (texture slot 0 of the reference,
     * raytrace.rmiss:15-30), RGBA32F, envWidth*envHeight texels.  NULL means
     * the reference's default black texture (src/SceneLoader.cpp:80-86). */
    const float *envMap;
    uint32_t envWidth;
    uint32_t envHeight;
    /* Material textures by slot, as rtx_material.textureIdDiffuse /
     * textureIdSpecular number them (binding 7, textureSamplers[]): slot 0 is
     * the environment map's, given above, so textures[0] is not read.
     * textureCount = 0 when no material has a texture. */
    const rtx_texture_desc *textures;
    uint32_t textureCount;
} rtx_scene_desc;

/* ------------------------------------------------------ frame params */
/* flags of rtx_frame_params */
#define RTX_FRAME_KEEP_PATHS 0x1u   /* keep per-path radiance + signature     */
#define RTX_FRAME_COUNT_VISITS 0x2u /* count BVH node/primitive fetches       */
#define RTX_FRAME_RESET_ACCUM 0x4u  /* clear the accumulation image first     */
#define RTX_FRAME_TIME_KERNELS 0x8u /* time every kernel with HIP events      */
#define RTX_FRAME_NO_GUIDING_UPDATE 0x10u /* record (pc.updateGuiding) but skip the
                                           * guiding update after the frame     */
#define RTX_FRAME_NRC 0x40u         /* learned radiance cache (rtx_nrc_create): camera
                                     * paths end into the cache after
                                     * terminateBounces bounces, a sparse set of
                                     * training paths continues and trains it */
#define RTX_FRAME_TIME_TRACE 0x20u  /* time only the traversal launches with HIP
                                     * events (msExtend, msShadow; the other
                                     * per-kind times stay 0): each event pair
                                     * costs device time, so this keeps the
                                     * timing of the dominant kernel cheap    */

/* One call to rtx_render renders `spp` frames and adds each frame's result
 * (per pixel: the mean of pc.samplesPerPixel camera samples and their
 * splits, main() of shaders/raytrace.rgen:1623-1711) to the context's
 * accumulation image.
 *
 * Random streams.  The reference seeds one LCG stream per pixel with
 * tea(pixelIndex, pc.randomUInt) (shaders/raytrace.rgen:1625) where
 * randomUInt is drawn per frame on the host (src/RayTracingApp.cpp:1121).
 * Here sample s of a call is rendered exactly as the reference renders one
 * 1-spp frame whose randomUInt is the first output of PCG32 seeded with
 * (initstate = seed, initseq = sampleOffset + s).  Each (pixel, sample)
 * path therefore has its own stream that does not depend on batching. */
typedef struct rtx_frame_params {
    rtx_push_constant pc; /* pc.randomUInt and pc.previousFrames are ignored */
    rtx_camera cam;
    uint32_t width;
    uint32_t height;
    uint32_t spp;
    uint32_t sampleOffset;
    uint64_t seed;
    uint32_t flags;        /* RTX_FRAME_* */
    int32_t rrStartDepth;  /* Russian roulette from this depth on (pc.enableRR) */
    uint32_t batchPaths;   /* paths in flight per wavefront batch, 0 = auto */
    uint32_t _reserved[7];
} rtx_frame_params;

/* Statistics of the last rtx_render (all counts are exact). */
typedef struct rtx_frame_stats {
    uint64_t paths;         /* camera paths traced (= samples)            */
    uint64_t extendRays;    /* closest-hit rays (camera + continuation)   */
    uint64_t shadowRays;    /* any-hit rays: NEE shadow rays, and the
                               occlusion rays of MIS probes that found their
                               emitter first (scenes with at most 64
                               emissive triangles and no environment map) */
    uint64_t probeRays;     /* closest-hit MIS probe rays (other scenes)  */
    uint64_t nodeVisits;    /* BVH node fetches (RTX_FRAME_COUNT_VISITS)  */
    uint64_t primVisits;    /* primitive fetches (RTX_FRAME_COUNT_VISITS) */
    double msTotal;         /* device time of the whole render (HIP events) */
    double msExtend;        /* time in traversal: one launch per bounce
                               (extend rays with the previous bounce's
                               probe and shadow rays)                     */
    double msShadow;        /* unused (0): shadow and probe rays are traced
                               in the extend launch                       */
    double msProbe;         /* time in the probe's BSDF-side resolve        */
    double msShade;         /* time in generate/shade/finish/accumulate    */
    uint32_t batches;
    uint32_t bounces;       /* wavefront iterations of the last batch     */
    uint32_t nodeBytes;     /* bytes per BVH node record fetched          */
    uint32_t primBytes;     /* bytes per triangle record fetched          */
    /* per traversal kind: [0] closest-hit extend, [1] shadow, [2] probe */
    uint64_t nodeVisitsByKind[3];
    uint64_t primVisitsByKind[3];
    uint32_t launchesByKind[4]; /* kernel launches, by the four times above:
                                   extend, shadow, probe, other           */
    uint32_t _pad;
} rtx_frame_stats;

/* BVH statistics of a scene. */
typedef struct rtx_scene_info {
    uint32_t triangles;
    uint32_t spheres;
    uint32_t nodes;
    uint32_t maxDepth;
    uint64_t deviceBytes;   /* nodes + triangle records + shading data */
    double buildMs;         /* host build time */
    /* SceneLoader::calculateSceneSize (src/SceneLoader.cpp:350-367): every
     * instance's model vertices transformed, and the spheres' boxes */
    float boundsMin[3];
    float boundsMax[3];
} rtx_scene_info;

/* Launch-shape knobs of the wavefront (no counterpart in the reference,
 * whose traversal is the driver's).  In rtx_ctx_set_tuning a zero field
 * keeps the context's current value. */
typedef struct rtx_tuning {
    int32_t traversalKernel;  /* 1 = persistent lane-refill while-while (the only traversal) */
    int32_t traceBlocksPerCU; /* 128-thread traversal blocks per CU (default 14:
                                 every block of the persistent kernel is
                                 resident at once, 10 KB of LDS stack each,
                                 7 waves per SIMD) */
    int32_t shadeBlocksPerCU; /* 256-thread shading blocks per CU (default 8) */
    uint32_t batchPaths;      /* paths in flight per batch when rtx_frame_params.batchPaths is 0 (2^22) */
    int32_t ldsStackEntries;  /* traversal stack entries per thread held in LDS: 12, 16, 20 or 24
                                 (20); deeper entries spill to scratch */
    int32_t cacheUpkeep;      /* irradiance-cache frames: 1 = the update jobs' sub-paths run in
                                 the frame's own bounces, beside its camera paths, where the
                                 path buffers hold both (default); 2 = every job after the
                                 frame.  Results are the same either way. */
    int32_t _reserved[2];
} rtx_tuning;

/* Autotuner (no counterpart in the reference).  Renders `probe` (as given,
 * with the context's cache and guiding state; its side effects - the
 * accumulation image, cache upkeep - happen once per trial) under launch
 * shapes chosen by coordinate descent, one knob at a time from the current
 * tuning: traceBlocksPerCU in {8, 10, 12, 14, 16, 20}, ldsStackEntries in
 * {12, 16, 20, 24}, shadeBlocksPerCU in {4, 6, 8, 12, 16}, batchPaths in
 * {2^20, 2^21, 2^22, 2^23}; every candidate is timed `repeats` times (device
 * time of the render, HIP events) and its fastest time kept.  The best shape
 * is set on the context (rtx_ctx_set_tuning) and returned. */
typedef struct rtx_autotune_result {
    rtx_tuning best;
    double bestMs;            /* fastest probe render with `best` */
    double startMs;           /* fastest probe render with the tuning in place at the call */
    uint32_t trials;          /* probe renders run */
    uint32_t reserved;
} rtx_autotune_result;

typedef struct rtx_ctx rtx_ctx;
typedef struct rtx_scene rtx_scene;

/* ------------------------------------------------------- entry points */
/* Fill the reference's push-constant defaults (src/RayTracingApp.h:81-130). */
void rtx_push_constant_defaults(rtx_push_constant *pc);

/* Create a context on HIP device `device` (replaces the Vulkan instance,
 * device and pipeline set-up of src/VulkanWindow.cpp and
 * RayTracingApp::initRayTracing, src/RayTracingApp.cpp:588-604). */
int rtx_ctx_create(int device, rtx_ctx **out);
int rtx_ctx_destroy(rtx_ctx *ctx);
const char *rtx_last_error(const rtx_ctx *ctx);

/* Set / read the launch-shape knobs of a context (new; no counterpart in
 * the reference). */
int rtx_ctx_set_tuning(rtx_ctx *ctx, const rtx_tuning *tuning);
int rtx_ctx_get_tuning(const rtx_ctx *ctx, rtx_tuning *tuning);
int rtx_autotune(rtx_ctx *ctx, const rtx_scene *scene, const rtx_frame_params *probe, uint32_t repeats,
                 rtx_autotune_result *result);

/* Upload a scene and build its BVH (replaces SceneLoader::createVulkanObjects,
 * src/SceneLoader.cpp:67-87, incl. createBottomLevelAS/createTopLevelAS). */
int rtx_scene_create(rtx_ctx *ctx, const rtx_scene_desc *desc, rtx_scene **out);
int rtx_scene_destroy(rtx_ctx *ctx, rtx_scene *scene);
int rtx_scene_get_info(const rtx_scene *scene, rtx_scene_info *info);

/* Render params->spp frames into the accumulation image
 * (replaces RayTracingApp::raytrace + traceRaysKHR, src/RayTracingApp.cpp:
 * 1120-1204, and the per-pixel main() of shaders/raytrace.rgen:1623-1827).
 * Synchronous: returns when the device work is done. stats may be NULL. */
int rtx_render(rtx_ctx *ctx, const rtx_scene *scene, const rtx_frame_params *params,
               rtx_frame_stats *stats);

/* Copy the mean image (accumulated sum / accumulated samples) as RGBA32F,
 * width*height*4 floats, alpha = 1 (PostProcessing::saveOffscreenImage,
 * src/PostProcessing.cpp:310-338). */
int rtx_read_image(rtx_ctx *ctx, float *rgba, size_t floatCount);

/* Number of samples per pixel accumulated since the last reset. */
int rtx_accumulated_samples(const rtx_ctx *ctx, uint32_t *samples);

/* Per-path results of the last render made with RTX_FRAME_KEEP_PATHS:
 * path i = sample (i / (width*height)) of pixel (i % (width*height)).
 * rgb: 3 floats per path.  sig: 2 uint32 per path = (final LCG state,
 * FNV-1a hash of the hit primitive ids along the path); either may be NULL. */
int rtx_read_paths(rtx_ctx *ctx, float *rgb, uint32_t *sig, size_t pathCount);

/* ------------------------------------------------------ irradiance cache */
/* One cache per context (the reference keeps one per scene and recreates it
 * on a scene switch, src/RayTracingApp.cpp:346-347).  rtx_render with
 * pc.useIrradianceCache renders each sample as one 1-spp frame that reads
 * the cache as it was at the frame's start; after the frame the frame's
 * update elections (in pixel order, by the header's nextUpdateSlot rule)
 * and creation candidates (in pixel order; the first maxEntries-count of
 * them) are computed with calculateCacheData's 200 strata and committed in
 * that order.  pc.isIrradiancePrepareFrame frames update the cache only. */

/* Replaces IrradianceCache::IrradianceCache + createBuffers
 * (src/IrradianceCache.cpp:11-17, 46-75): a cache of maxEntries records. */
int rtx_ic_create(rtx_ctx *ctx, uint32_t maxEntries);
/* Empty the cache (as freshly created). */
int rtx_ic_reset(rtx_ctx *ctx);
/* Copy the cache buffers (header, and up to `capacity` records and spheres)
 * out of / into the device (the reference's binding-13 and binding-12
 * buffers, src/IrradianceCache.cpp:106-123).  rtx_ic_write takes
 * hdr->maxCaches records; hdr must match the created cache. */
int rtx_ic_read(rtx_ctx *ctx, rtx_cache_header *hdr, rtx_cache_data *entries, rtx_sphere *spheres,
                uint32_t capacity);
int rtx_ic_write(rtx_ctx *ctx, const rtx_cache_header *hdr, const rtx_cache_data *entries,
                 const rtx_sphere *spheres);

/* The upkeep jobs of the last cache frame of the last rtx_render made with
 * RTX_FRAME_KEEP_PATHS (diagnostics; no counterpart in the reference).
 * sub: 4 floats per sub-path (radiance rgb, first-hit distance), 200 per
 * job in stratum order k * 10 + j; subSig: 2 uint32 per sub-path. */
typedef struct rtx_ic_job {
    int32_t type;      /* 0 = update of record `target`, 1 = creation */
    uint32_t target;
    float origin[3];
    float normal[3];
} rtx_ic_job;
int rtx_ic_read_jobs(rtx_ctx *ctx, rtx_ic_job *jobs, float *sub, uint32_t *subSig, uint32_t capacity,
                     uint32_t *numJobs);

/* ------------------------------------------------------------------- ADRRS */
/* The per-pixel estimate that frames with pc.useADRRS read: the reference's
 * estimateImage (binding 14, shaders/raytrace.rgen:120, 1632-1636), written by
 * a render with pc.storeEstimate (saveEstimate, raytrace.rgen:1602-1605:
 * each pixel's frame result, RGBA with A = 1) or by rtx_estimate_write.
 * rgba holds 4 floats per pixel, width * height pixels; rtx_estimate_read
 * fails with RTX_E_STATE when no estimate exists or its size differs. */
int rtx_estimate_write(rtx_ctx *ctx, const float *rgba, uint32_t width, uint32_t height);
int rtx_estimate_read(rtx_ctx *ctx, float *rgba, uint32_t width, uint32_t height);

/* ------------------------------------------------------------ path guiding */
/* The guiding regions (boxes) and one vMF mixture per region: replaces
 * PathGuiding's AABB and VMM_Theta buffers and its acceleration structure
 * (bindings 15-17, src/PathGuiding.cpp:141-197).  Renders with pc.useGuiding
 * sample non-discrete materials from the mixture of the region holding the
 * vertex with probability pc.guidingProb (getNewDirection,
 * shaders/raytrace.rgen:923-960).  n = 0 removes them. */
int rtx_guiding_set(rtx_ctx *ctx, const rtx_aabb *regions, const rtx_vmm_theta *vmms, uint32_t n);
int rtx_guiding_get(rtx_ctx *ctx, rtx_aabb *regions, rtx_vmm_theta *vmms, uint32_t capacity, uint32_t *n);
/* Replaces PathGuiding's constructor (src/PathGuiding.cpp:11-31): regions
 * from splitting the scene box (grown by Aabb::addEpsilon) splitCount times
 * along the largest axis, each with lightpmm's initial 8-lobe mixture; the
 * context then owns and refits them.  parallax: useParallaxCompensation of
 * the fit (lobe distances and targets). */
int rtx_guiding_init(rtx_ctx *ctx, const rtx_aabb *sceneBox, uint32_t splitCount, int parallax);
/* PathGuiding::update (src/PathGuiding.cpp:276-312) on the samples of the
 * last frame rendered with pc.updateGuiding: per region the samples move to
 * the region's centre (parallax), lightpmm's MAP-EM fits (first update) or
 * refits (later) the mixture, and the lobe distances are updated.  Renders
 * with pc.updateGuiding do this after every frame by themselves when the
 * context owns the mixtures.  With options.splitAndMerge (the default) the
 * fit is followed by postFit's component merge and split. */
int rtx_guiding_update(rtx_ctx *ctx);
/* PathGuiding's public fitting switches (src/PathGuiding.h:118-129), with
 * their defaults: postFit merges pairs whose chi-square merge metric is at
 * most mergeMaxDivergence once more than minSamplesForMerging samples came
 * in since the last merge, and splits components whose weighted chi-square
 * divergence reaches splitMinDivergence (refitting them when the batch has
 * more than minSamplesForPostSplitFitting samples or is the first). */
typedef struct rtx_guiding_options {
    int32_t splitAndMerge;                  /* 1 */
    uint32_t minSamplesForMerging;          /* 8192 */
    uint32_t minSamplesForSplitting;        /* 4096 */
    uint32_t minSamplesForPostSplitFitting; /* 4096 */
    float splitMinDivergence;               /* 0.5 */
    float mergeMaxDivergence;               /* 0.025 */
    /* region splitting (PathGuiding::update, src/PathGuiding.cpp:295-301,
     * splitRegion 328-348): after an update every region whose mixture has
     * seen more than samplesForRegionSplit samples is halved along the largest
     * axis of its box; both halves keep its mixture and statistics, with the
     * sample counts decayed by 0.25, the new half appended as the last region */
    int32_t splitRegions;                   /* 0 */
    uint32_t samplesForRegionSplit;         /* 10000 */
    /* Scale limits (new; the reference fixes MAX_DISTRIBUTIONS = 16 at compile
     * time, shaders/limits.glsl:6, and has no region or sample caps):
     * maxComponents: lobes a mixture may grow to by splits, 1..RTX_MAX_LOBES (16);
     * maxRegions: region splitting stops once there are this many regions (0: 2^20);
     * maxSamplesPerUpdate: an update fits the first this many valid samples in
     * slot order (0: all). */
    uint32_t maxComponents;
    uint32_t maxRegions;
    uint32_t maxSamplesPerUpdate;
    uint32_t reserved;
} rtx_guiding_options;
int rtx_guiding_set_options(rtx_ctx *ctx, const rtx_guiding_options *options);
int rtx_guiding_get_options(rtx_ctx *ctx, rtx_guiding_options *options);
/* A mixture of up to RTX_MAX_LOBES lobes: VMM_Theta's layout with
 * MAX_DISTRIBUTIONS = 32 (shaders/guiding.glsl:49-54).  rtx_guiding_get
 * returns RTX_E_UNSUPPORTED when a region uses more than 16 lobes; this
 * getter returns every mixture. */
typedef struct rtx_vmm_theta_wide {
    rtx_vmf_theta thetas[RTX_MAX_LOBES];
    float pi[RTX_MAX_LOBES];
    float meanPosition[3];
    int32_t usedDistributions;
} rtx_vmm_theta_wide;
int rtx_guiding_get_wide(rtx_ctx *ctx, rtx_aabb *regions, rtx_vmm_theta_wide *vmms, uint32_t capacity, uint32_t *n);
/* The guiding samples of the last frame rendered with pc.updateGuiding: the
 * reference's DirectionalData buffer (binding 18, src/SampleCollector.cpp:
 * 29-51), RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL slots per pixel in pixel order;
 * unused slots have flags 0xFFFFFFFF.  count <= width*height*16. */
int rtx_guiding_read_samples(rtx_ctx *ctx, rtx_directional_data *out, size_t count);

/* ---------------------------------------------------- traversal queries */
/* Rays for the query entry points: origin + tmin, direction + tmax. */
typedef struct rtx_ray {
    float org[3];
    float tmin;
    float dir[3];
    float tmax;
} rtx_ray;

/* Closest-hit result: t, primitive id (triangles first, in instance order,
 * then spheres; -1 = miss) and the hit's barycentrics (u, v) for triangles. */
typedef struct rtx_hit {
    float t;
    int32_t prim;
    float u;
    float v;
} rtx_hit;

/* Closest hit for n rays (the scene traceRayEXT with gl_RayFlagsNoneEXT,
 * shaders/raytrace.rgen:1011-1022). */
int rtx_trace_closest(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays,
                      size_t n, rtx_hit *hits);

/* Occlusion for n rays (the shadow traceRayEXT with TerminateOnFirstHit and
 * SkipClosestHit, shaders/raytrace.rgen:622-638): occluded[i] = 1 if any
 * primitive is hit in [tmin, tmax]. */
int rtx_trace_occluded(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays,
                       size_t n, uint8_t *occluded);

/* Allocation-free closest-hit query: rtx_trace_closest for at most
 * RTX_NOALLOC_MAX_RAYS rays, staged through a fixed device array of the
 * library (3 MiB + 256 B per device, held until the process ends) instead of
 * per-call device allocations; it calls no hipMalloc or hipFree.  Rays past
 * the limit are refused with RTX_E_INVALID.  Calls that overlap on one device
 * share that array and race; each call finishes (stream synchronised) before
 * it returns.  No counterpart in the reference: for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). */
#define RTX_NOALLOC_MAX_RAYS 65536
int rtx_trace_closest_noalloc(rtx_ctx *ctx, const rtx_scene *scene, const rtx_ray *rays,
                              size_t n, rtx_hit *hits);

/* ------------------------------------------------ learned radiance cache */
/* A neural radiance cache (Mueller, Rousselle, Novak and Keller 2021) of the
 * reflected radiance leaving a surface point: a multiresolution hash-grid
 * encoding of the position (Mueller, Evans, Schied and Keller 2022; 2 features
 * per level) plus the direction, normal, roughness and albedo, into a 5 x 64
 * ReLU MLP with an RGB output.  BF16 MFMA kernels with FP32 accumulation run
 * inference and training; FP32 master parameters are updated by Adam on the
 * device.  No counterpart in the reference, whose nearest behaviour is the
 * irradiance cache ending a path at a cache hit
 * (shaders/raytrace.rgen:1117-1123).  One cache per context.
 *
 * Renders with RTX_FRAME_NRC: at the first non-specular vertex reached after
 * terminateBounces bounces a camera path adds throughput x max(cache, 0) and
 * ends (no NEE there; the cache holds everything the tracer would add from
 * that vertex on, emission at the vertex excluded).  About trainingPaths
 * paths per frame (a regular pixel pattern that moves every frame) continue
 * instead, for trainSuffixBounces more non-specular vertices, where they end
 * into the cache themselves (or to their end when it is 0); what they add
 * from the training vertex on, the cache's prediction at their end included,
 * divided by their throughput at the training vertex, is the vertex's
 * training target.  After each frame
 * the cache takes trainSteps Adam steps on trainBatch of these targets each
 * (relative L2 loss, Mueller et al. 2021), so a frame reads the cache as the
 * previous frames left it.  Guided sampling (pc.useGuiding, with a guide
 * trained by earlier recording frames) combines with it; the irradiance
 * cache, ADRRS, guiding updates (pc.updateGuiding), splitting and several
 * samples per frame do not (RTX_E_UNSUPPORTED). */
typedef struct rtx_nrc_config {
    uint32_t levels;           /* hash-grid levels (16; at most 16)             */
    uint32_t log2TableSize;    /* entries per level, log2 (19)                  */
    float baseResolution;      /* cells per axis of the coarsest level (16)     */
    float maxResolution;       /* of the finest (2048)                          */
    float learningRate;        /* Adam (1e-2)                                   */
    float beta1, beta2, epsilon; /* Adam (0.9, 0.99, 1e-15)                     */
    uint32_t trainBatch;       /* rows per Adam step (16384; at most 65536)     */
    uint32_t trainSteps;       /* Adam steps per frame (4)                      */
    uint32_t terminateBounces; /* bounces before a path ends into the cache (2) */
    uint32_t trainingPaths;    /* training paths per frame (65536)              */
    float sceneMin[3];         /* box the positions are encoded in              */
    float sceneMax[3];
    uint64_t seed;             /* parameter initialisation                      */
    uint32_t trainSuffixBounces; /* a training path ends into the cache this many
                                    non-specular vertices past its training vertex,
                                    its target completed there by the cache's own
                                    prediction (self-training, Mueller et al. 2021);
                                    0: it runs to its end (1)                     */
    float emaDecay;            /* inference reads an exponential moving average
                                  of the parameters, e = d e + (1 - d) p after
                                  every Adam step (from e = p at creation and at
                                  rtx_nrc_set_params); 0: the parameters
                                  themselves (0.99; in [0, 1))                    */
} rtx_nrc_config;

/* One cache query: the vertex's position, the direction towards the previous
 * vertex (unit), the shading normal, a roughness (1 Lambert, alpha for rough
 * conductors, sqrt(2 / (exponent + 2)) for Phong, 0 for specular) and an
 * albedo. 64 B. */
typedef struct rtx_nrc_input {
    float pos[3];
    float roughness;
    float dir[3];
    float _pad0;
    float normal[3];
    float _pad1;
    float albedo[3];
    float _pad2;
} rtx_nrc_input;

typedef struct rtx_nrc_info {
    uint32_t levels;
    uint32_t tableSize;         /* entries per level */
    float resolution[16];       /* cells per axis of each level */
    uint32_t dense[16];         /* 1: (resolution + 1)^3 <= tableSize, indexed densely */
    float lo[3];                /* encoded position = clamp((x - lo) * inv, 0, 1) */
    float inv[3];
    uint64_t gridParams;        /* levels * tableSize * 2 */
    uint32_t mlpParams;         /* 5 * 64 * 64 + 16 * 64: W[layer][out][in], output layer 16 x 64
                                   (rows 3..15 stay zero) */
    uint32_t steps;             /* Adam steps taken */
    float lastLoss;             /* mean loss of the last frame's (or call's) steps */
} rtx_nrc_info;

void rtx_nrc_config_defaults(rtx_nrc_config *config);
int rtx_nrc_create(rtx_ctx *ctx, const rtx_nrc_config *config);
int rtx_nrc_destroy(rtx_ctx *ctx);
int rtx_nrc_get_info(rtx_ctx *ctx, rtx_nrc_info *info);
/* FP32 master parameters: the table [level][entry][feature], then the MLP.
 * set_params restarts Adam and the moving average (at the parameters set). */
int rtx_nrc_get_params(rtx_ctx *ctx, float *grid, size_t gridCount, float *mlp, size_t mlpCount);
int rtx_nrc_set_params(rtx_ctx *ctx, const float *grid, size_t gridCount, const float *mlp, size_t mlpCount);
/* The cache's RGB prediction for n queries (rgb: 3 floats each), from the
 * parameters' moving average (rtx_nrc_config.emaDecay). */
int rtx_nrc_infer(rtx_ctx *ctx, const rtx_nrc_input *in, size_t n, float *rgb);
/* The loss gradients of n <= trainBatch rows against targets (3 floats each)
 * without changing the parameters; grid/mlp as in get_params, either may be
 * NULL.  loss: the relative L2 loss, mean over rows and channels. */
int rtx_nrc_gradients(rtx_ctx *ctx, const rtx_nrc_input *in, const float *target, size_t n, float *gridGrad,
                      float *mlpGrad, float *loss);
/* One Adam step on n <= trainBatch rows. */
int rtx_nrc_train(rtx_ctx *ctx, const rtx_nrc_input *in, const float *target, size_t n, float *loss);

/* ------------------------------------------------------------- serving */
/* A persistent request loop on one device (no counterpart in the reference,
 * which renders one scene per process through RayTracingApp::drawFrame,
 * src/RayTracingApp.cpp:74-215).  Render jobs of any resident scene,
 * resolution and sample count are submitted as they arrive; a server thread
 * admits them in arrival order while their images fit the framebuffer
 * budget, cuts each into batches of whole sample planes (or pixel ranges of
 * one plane) of at most slotPaths paths, and runs the batches of all
 * admitted jobs together: the pool's slots hold one batch each, every bounce
 * is one launch of each wavefront kernel over the paths of all slots (rays
 * of different scenes in one persistent traversal launch), and a slot whose
 * paths have all ended takes the next batch at the next bounce (continuous
 * batching).  Each job's batches are added to its image in batch order, and
 * every path draws from its own (pixel, sample) stream, so a job's image is
 * the same bit for bit whatever else runs beside it, and equal to rtx_render
 * of the same frame with batchPaths equal to the job's batch size.
 * Jobs render the plain method with one camera sample per path (NEE, MIS,
 * Russian roulette, textures, environment map); with pc.useIrradianceCache
 * or pc.useGuiding they read the cache or the guiding regions of `state`, a
 * context whose own frames built them (read-only while jobs use them).
 * Splitting, ADRRS, guiding records, samplesPerPixel > 1, cache upkeep and
 * the learned cache are RTX_E_UNSUPPORTED here (rtx_render has them). */
typedef struct rtx_server rtx_server;

typedef struct rtx_server_config {
    uint64_t framebufferBytes; /* budget of the admitted jobs' images: 32 B per pixel (1 GiB) */
    uint32_t slotPaths;        /* paths per slot, a power of two in [2^10, 2^22] (2^18) */
    uint32_t slots;            /* slots in the pool, 1..64 (16) */
    uint32_t maxNumNEE;        /* largest pc.numNEE a job may use, 1..8 (1) */
    uint32_t maxJobs;          /* jobs queued, running or unread at once (4096) */
} rtx_server_config;

typedef struct rtx_job_desc {
    const rtx_scene *scene;
    const rtx_ctx *state;      /* cache / guiding regions read with pc.useIrradianceCache /
                                  pc.useGuiding (NULL: none) */
    rtx_push_constant pc;
    rtx_camera cam;
    uint32_t width, height;
    uint32_t spp;              /* sample planes (one camera path per pixel each) */
    int32_t rrStartDepth;
    uint64_t seed;
    uint32_t sampleOffset;
    uint32_t reserved;
} rtx_job_desc;

#define RTX_JOB_QUEUED 0   /* waiting for framebuffer budget */
#define RTX_JOB_RUNNING 1  /* admitted: batches in the pool */
#define RTX_JOB_DONE 2     /* image complete, not yet read */
#define RTX_JOB_RELEASED 3 /* read (rtx_server_read_image) */

typedef struct rtx_job_status {
    int32_t state;
    uint32_t batches;          /* batches of the job */
    uint32_t batchesDone;      /* added to its image */
    uint32_t reserved;
    uint64_t paths;            /* width * height * spp */
    double queuedMs;           /* submission to admission */
    double latencyMs;          /* submission to the image complete on the device, as the
                                  server loop observed it (resolution: one bounce) */
} rtx_job_status;

typedef struct rtx_server_stats {
    uint64_t jobsDone;
    uint64_t pathsDone;
    double latencyP50Ms, latencyP99Ms, latencyMaxMs, latencyMeanMs; /* jobs done */
    double msamplesPerSec;     /* paths of the jobs done / (last completion - first submission) */
    uint64_t bounces;          /* pool bounces launched */
    uint64_t extendRays, shadowRays, probeRays;
    uint32_t peakSlotsBusy;
    uint32_t peakJobsRunning;
    uint64_t framebufferPeakBytes;
    uint64_t poolBytes;        /* device memory of the path pool */
} rtx_server_stats;

void rtx_server_config_defaults(rtx_server_config *config);
/* The server runs on ctx's device with its own stream and thread; ctx must
 * outlive it. */
int rtx_server_create(rtx_ctx *ctx, const rtx_server_config *config, rtx_server **out);
/* Finishes the jobs already admitted (queued ones are dropped), then stops
 * the thread and frees the pool. */
int rtx_server_destroy(rtx_server *server);
/* Checks the job (RTX_E_INVALID, RTX_E_UNSUPPORTED, RTX_E_MEMORY when its
 * image alone exceeds the budget) and queues it; the scene (and state) must
 * stay alive until the job is done. */
int rtx_server_submit(rtx_server *server, const rtx_job_desc *job, uint64_t *jobId);
int rtx_server_poll(rtx_server *server, uint64_t jobId, rtx_job_status *status);
/* Blocks until the job is done. */
int rtx_server_wait(rtx_server *server, uint64_t jobId, rtx_job_status *status);
/* Copies a done job's image (RGBA32F, width * height * 4 floats, the mean of
 * its samples) and releases its framebuffer. */
int rtx_server_read_image(rtx_server *server, uint64_t jobId, float *rgba, size_t floatCount);
int rtx_server_get_stats(rtx_server *server, rtx_server_stats *stats);
const char *rtx_server_last_error(const rtx_server *server);

/* ------------------------------------------------------ frame schedule */
/* The host loop around the render of the reference's application,
 * RayTracingApp::drawFrame and RayTracingApp::raytrace
 * (src/RayTracingApp.cpp:91-215, 1120-1163, 1206-1217), as a state machine
 * that plans each frame's push constants and accumulation:
 *   - progressive accumulation: previousFrames counts frames since the last
 *     change (rtx_schedule_input_changed, rtx_schedule_set_push_constants);
 *   - irradiance-cache prepare frames: the first prepareFrames frames with
 *     useIrradianceCache or useADRRS are prepare frames (ADRRS runs them as
 *     cache frames with useIrradianceCacheOnGlossy), then ADRRS renders the
 *     estimate frame (setEstimateRTSettings) and returns to the saved
 *     settings with useADRRS;
 *   - guiding cadence: recording frames (pc.updateGuiding) are followed by a
 *     guiding update until guidingOptimizations updates have run; the next
 *     recording frame is rendered without one and recording stops;
 *   - eval modes: evalSamples stops after that many samples per pixel
 *     outside prepare frames and reports the wall time since the first eval
 *     frame; evalSeconds counts the samples rendered in that many seconds.
 * A frame's randomUInt comes from PCG32 (seed, sampleOffset = frames so far)
 * as in rtx_render.  The structs have no counterpart in the reference: their
 * fields are RayTracingApp's (src/RayTracingApp.h:147-190). */
typedef struct rtx_schedule rtx_schedule;

typedef struct rtx_schedule_config {
    int32_t prepareFrames;        /* irradianceCachePrepareFrames (reference: 50) */
    int32_t guidingOptimizations; /* numGuidingOptimizations (reference: 6; -1 = no limit) */
    int32_t accumulate;           /* accumulateResults */
    int32_t evalSamples;          /* evalXSamples with evalNumSamples = this; 0 = off */
    int32_t evalSeconds;          /* evalXSeconds with evalNumSeconds = this; 0 = off */
    int32_t sppToOneAfterEval;    /* setSPPToOneAfterEval */
} rtx_schedule_config;

typedef struct rtx_schedule_state {
    uint64_t frames;                     /* frames rendered so far */
    int32_t currentPrepareFrames;
    int32_t currentGuidingOptimizations; /* -1 when not recording */
    uint32_t previousFrames;             /* of the last planned frame */
    int32_t evalSamplesRunning;          /* evalXSamples */
    int32_t evalSecondsRunning;          /* evalXSeconds */
    int32_t evalCurrentSamples;
    int64_t evalSecondsTotalSamples;     /* evalXSecondsTotalSamples */
    int64_t evalSecondsResult;           /* evalXSecondsLastResult */
    double evalSamplesMs;                /* evalXSamplesTime; -1 until an eval finished */
    double guidingUpdateMs;              /* guidingUpdateTimeMs; -1 until recording stopped */
    rtx_push_constant pc;                /* rtPushConstants for the next frame */
} rtx_schedule_state;

void rtx_schedule_config_defaults(rtx_schedule_config *config);
/* pc: the application's settings (rtPushConstants) at start. */
int rtx_schedule_create(const rtx_schedule_config *config, const rtx_push_constant *pc, rtx_schedule **out);
int rtx_schedule_destroy(rtx_schedule *schedule);
/* A settings change (a UI edit: hasInputChanged); restarts accumulation. */
int rtx_schedule_set_push_constants(rtx_schedule *schedule, const rtx_push_constant *pc);
/* The camera or scene changed: accumulation restarts with the next frame. */
int rtx_schedule_input_changed(rtx_schedule *schedule);
/* Plans the next frame: *frame = *base (camera, size, seed, flags, ...) with
 * the frame's push constants, spp = 1, sampleOffset = frames so far and
 * RTX_FRAME_RESET_ACCUM when accumulation restarts.  No device work. */
int rtx_schedule_begin_frame(rtx_schedule *schedule, const rtx_frame_params *base, rtx_frame_params *frame);
/* The bookkeeping after the planned frame was rendered: guiding cadence and
 * the eval counters. */
int rtx_schedule_end_frame(rtx_schedule *schedule);
/* begin_frame, rtx_render, end_frame. */
int rtx_schedule_frame(rtx_ctx *ctx, const rtx_scene *scene, rtx_schedule *schedule,
                       const rtx_frame_params *base, rtx_frame_stats *stats);
int rtx_schedule_get_state(const rtx_schedule *schedule, rtx_schedule_state *state);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* RTX_AMD_H */
