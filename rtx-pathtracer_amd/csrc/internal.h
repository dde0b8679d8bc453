// Data shared between the host orchestration (api.cpp), the BVH builder
// (bvh_build.cpp) and the kernels (kernels.hip).  Plain structs only.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#include <hip/hip_runtime.h>

#include "bvh.h"
#include "rtx_amd.h"

namespace rtx {

// ----------------------------------------------------------- device scene
struct DevScene {
    const float4 *nodes;      // 8 x float4 per Bvh4Node
    const float4 *tris;       // 3 x float4 per triangle in leaf order: v0 (w = prim id bits), v1, v2
    const int4 *triRef;       // per global triangle id: (instance, global indices of its three vertices)
    const rtx_vertex *vertices;
    const uint32_t *indices;
    const uint32_t *vertexBase; // per model
    const uint32_t *indexBase;  // per model
    const rtx_instance *instances;
    const rtx_material *materials;
    const rtx_light *lights;
    const int32_t *lightIndex;
    const rtx_face_sample *triTable;
    // Area-light sampling without the table -> index -> vertex -> material
    // chain: per face-table entry (light row x RTX_SIZE_TRI_RANDOM) the face's
    // three positions (w: light colour r, g, b) and three normals, model
    // space, and per light row the instance it transforms them by.  Null when
    // the table would exceed kEmitTableBytes (then the chain is walked).
    const float4 *emitTris;
    const rtx_instance *lightInst;
    // The emissive triangles, world space, 3 float4 each as in `tris` (w of
    // the first: the primitive id), for MIS probe rays that look for their
    // light before they trace (FrameConst::probeFirstLight); lightFirstOk is
    // 0 when there are more than kProbeLightMax of them or one is
    // alpha-tested (emissive spheres are found among `spheres`).
    const float4 *lightTris;
    uint32_t numLightTris;
    uint32_t lightFirstOk;
    const rtx_sphere *spheres;
    uint32_t numTris;
    uint32_t numSpheres;
    uint32_t lightCount;
    uint32_t triRows;
    uint32_t numNodes;
    // textures (binding 7 of the reference): slot 0 the environment map
    // (RGBA32F lat-long), slots 1.. 8-bit RGBA material textures
    const float4 *env;          // envW x envH texels, or null (black)
    uint32_t envW, envH;
    const uint32_t *texels;     // all material textures, RGBA8 packed little-endian
    const int4 *texInfo;        // per slot: (first texel, width, height, 0)
    uint32_t numTextures;
    uint32_t alphaTris;         // triangles whose material has a diffuse texture (any-hit alpha test)
    const float *srgbToLinear;  // 256 entries
};

constexpr size_t kEmitTableBytes = size_t(256) << 20;
// Emissive triangles up to which probe rays look for their light first.
constexpr uint32_t kProbeLightMax = 64;
// Bit of a shadow-queue id marking a probe's occlusion ray (the rest: its
// NEE slot); its visibility goes to pendProbe instead of pendShadow.
constexpr int32_t kProbeOccBit = 1 << 30;

// Bit of a triangle's id (tris[3k].w) marking it for the any-hit alpha test
// (raytrace.rahit:21-47); ids stay below 2^27.
constexpr int32_t kAlphaBit = 1 << 30;

// ----------------------------------------------------------- path state
// Queues and per-path state of one wavefront batch (structure of arrays).
struct Queue {
    int32_t *id;    // path id (extend queue, done queue) or NEE slot (shadow/probe)
    float4 *org;    // origin.xyz, tmin
    float4 *dir;    // direction.xyz, tmax
    float4 *hit;    // closest hit (t, prim bits, u, v): extend and probe queues
    float4 *aux;    // probe: (premultiplied BSDF weight rgb, pdfMat)
    // the frame's randomUInt of entry i's path is pathSeed[(id[i] & 0x7fffffff) / seedDiv]
    // (read by the any-hit alpha test only)
    const uint32_t *pathSeed;
    uint32_t seedDiv;
};

struct Counters {
    uint32_t extend[2]; // ping-pong extend queue sizes
    uint32_t shadow;
    uint32_t probe;
    uint32_t done[2];   // done lists, by the bounce's extend queue
    uint32_t pad[2];
    uint32_t fetch[3];  // persistent-kernel work counters ([1]: the per-bounce trace)
    uint32_t pad2;
    unsigned long long extendRays, shadowRays, probeRays;
    unsigned long long nodeVisits[3], primVisits[3]; // extend, shadow, probe
};

struct FrameConst {
    rtx_push_constant pc;
    rtx_camera cam;
    uint32_t width, height;
    uint64_t seed;
    uint32_t sampleOffset;   // first sample of this render call
    int32_t rrStartDepth;
    uint32_t numNEE;         // NEE samples per vertex (>= 1)
    uint32_t neeStride;      // NEE slot j of path pid is pid * neeStride + j (>= numNEE)
    uint32_t countVisits;
    uint64_t debugPath;      // path whose bounces are printed (RTX_DEBUG_PATH), ~0 = none
    // The arguments of raytrace() (raytrace.rgen:992) for this class of paths:
    // camera paths (main, raytrace.rgen:1671) or the hemisphere strata of
    // the irradiance cache (calculateCacheData, raytrace.rgen:1260).
    int32_t maxDepth, maxFollow;
    uint32_t addDirect, useNEE, useIC, createIC, allowRR;
    uint32_t record;          // saveSamples: write guiding sample records (pc.updateGuiding)
    // main()'s frame structure (raytrace.rgen:1662-1708): camera samples per
    // pixel and frame, and whether a finished path continues with the
    // frame's next camera sample or queued split (segments)
    uint32_t samplesPerFrame, segments;
    uint32_t useADRRS, splitOnFirst; // pc's, for camera paths only
    // learned radiance cache (RTX_FRAME_NRC): paths end into the cache at the
    // first non-specular vertex of depth >= nrcDepth; pixels with
    // (pixel + frame offset) % nrcTrainEvery == 0 are training paths
    uint32_t nrc, nrcDepth, nrcTrainEvery, nrcTrainSlots;
    uint32_t nrcSuffix; // a training path ends into the cache this many non-specular vertices past its training vertex (0: never)
    uint32_t probeFirstLight; // MIS probe rays find their light first and trace only an occlusion ray to it
};

// ------------------------------------------------------------ path guiding
// Box hierarchy over the guiding regions, depth-first: an inner node's left
// child follows it, `right` names the other; a leaf holds one region.
struct GuideNode {
    float lo[3];
    int32_t right;
    float hi[3];
    int32_t region; // >= 0: leaf
};

struct DevGuiding {
    const rtx_aabb *boxes;
    const rtx_vmm_theta_wide *vmms;
    const GuideNode *nodes;
    uint32_t numRegions, numNodes;
};

// Builds the hierarchy over region boxes (median splits on the widest axis
// of the centres); api.cpp.
std::vector<GuideNode> buildGuideNodes(const rtx_aabb *boxes, uint32_t n);

// One guiding region's mixture as lightpmm keeps it (ParametricMixtureModel
// of VMFKernels, external/lightpmm/include/pmm/ParametricMixtureModel.h:56-73)
// plus PathGuiding's per-region extras (parallax mean, harmonic-mean lobe
// distances; src/PathGuiding.h:77-84).
struct GuideRegion {
    int32_t K;
    uint32_t numEMIterations;
    float sampleWeight, numSamples;
    uint64_t totalNumSamples;
    float w[RTX_MAX_LOBES], kappa[RTX_MAX_LOBES], r[RTX_MAX_LOBES];
    float norm[RTX_MAX_LOBES], e2k[RTX_MAX_LOBES];
    float mu[RTX_MAX_LOBES][3];
    float dist[RTX_MAX_LOBES], distW[RTX_MAX_LOBES];
    float mean[3], lastMean[3];
    // postFit's statistics (external/guiding/incrementalpearsonchisquared.h,
    // incrementalcovariance2d.h): divergence + 1 times the squared integral and
    // its sample count; tangent-plane (var x, var y, cov xy) and weight sum
    float chi[RTX_MAX_LOBES], chiN[RTX_MAX_LOBES];
    float cov[RTX_MAX_LOBES][3], covW[RTX_MAX_LOBES];
    uint32_t samplesSinceLastMerge;
};

// lightpmm factory properties (VMMFactory.h:45-61 as PathGuiding sets them)
struct EmProps {
    uint32_t minItr, maxItr;
    float relLogLikelihoodThreshold;
    float maxKappa, vPrior, rPrior, rPriorWeight;
};

// Work buffers of one guiding update (device).
struct GuideFitBuffers {
    GuideRegion *regions;
    const rtx_aabb *boxes;
    uint32_t numRegions;
    uint32_t *keysA, *keysB, *valsA, *valsB; // region / slot pairs while sorting
    uint32_t *blockSums, *hist;
    uint32_t *regionStart;                   // numRegions + 1
    uint32_t *counters;                      // [0] valid samples
    float4 *samples;                         // 2 per sample: (direction, weight), (distance, -, -, -)
    uint32_t slotCap, blockCap;
};

void launchGuideInit(hipStream_t s, GuideFitBuffers g, uint32_t numInitialComponents, float initKappa);
// PathGuiding::update on the recorded samples (slots = pixels * 16)
void launchGuideUpdate(hipStream_t s, GuideFitBuffers g, const rtx_directional_data *dd, uint32_t slots,
                       EmProps props, const rtx_guiding_options &sm, bool parallax, bool firstFit,
                       rtx_vmm_theta_wide *vmmsOut);
void launchGuideToVmm(hipStream_t s, GuideFitBuffers g, bool parallax, rtx_vmm_theta_wide *vmmsOut);

// ------------------------------------------------------ irradiance cache
// One (record, grid cell) pair of the lookup index, 64 bytes.  A lookup's
// scan reads the first 16 (the sphere); a record whose sphere contains the
// point then reads the next 32 (normal, harmonic radius, colour) from the
// same cache line, and its gradients, when they are used, from its entry.
// The item is the same for all of a record's cells.
struct alignas(16) IcItem {
    float center[3];
    float radius;
    float normal[3];
    float harmonicR;
    float color[3];
    uint32_t entry;
    uint32_t pad[4];
};
static_assert(sizeof(IcItem) == 64, "index item layout");

// What lookups read (dev_icache.h).
struct DevIC {
    const rtx_cache_data *entries;
    const rtx_sphere *spheres;
    const uint32_t *bucketStart; // nb + 1
    const IcItem *items;
    const uint32_t *meta;        // [0] occupied levels, [1] items, [2] level-0 cell size (float bits)
                                 // [3] slots of the records in use
    uint32_t nb;
};

// The cache, its index and the per-frame work lists (device memory).
struct IcBuffers {
    rtx_cache_header *hdr;
    rtx_cache_data *entries;
    rtx_sphere *spheres;
    uint32_t maxEntries;
    uint32_t nb, itemCap;
    uint32_t *bucketCount, *bucketStart, *cursor, *meta;
    IcItem *items;
    // the index build's sort: 27 fixed slots per record (bucket keys,
    // unused slots keyed nb); (key, slot) pairs are sorted stably by bucket
    // (keys, slot ids and their ping-pong copies, digit histograms), then
    // each sorted slot's item is written from its record (slot / 27)
    uint32_t *slotKeys, *slotKeysB, *slotIdx, *slotIdxB, *sortHist;
    uint32_t bucketBits; // log2(nb)
    // frame: [0] update elections, [1] creation candidates, [2] free slots at
    // frame start, [3] update jobs, [4] jobs
    uint32_t *frame;
    float4 *cand;                // 2 per candidate: origin, normal (first maxEntries kept)
    float4 *jobs;                // 2 per job: (origin, type), (normal, target)
    uint32_t jobCap;
    float *jobRes;               // 10 per job: color, rot, trans, harmonicR
    uint32_t *blockSums;
    uint32_t blockCap;
    // pinned host word: 1 once a commit has left every record slot in use
    // (the cache then takes no creation job again until it is reset or
    // written); 0 before, or not yet known
    uint32_t *fullHost;
};
constexpr int kMaxNewCacheEntries = 5; // MAX_NEW_IRRADIANCE_ENTRIES, raytrace.rgen:65
constexpr int kIcStrataPhi = 20;       // N, raytrace.rgen:1232
constexpr int kIcStrataTheta = 10;     // M, raytrace.rgen:1233
constexpr int kIcStrata = kIcStrataPhi * kIcStrataTheta;
constexpr uint32_t kIcPhiKey = 0x9e3779b9u;     // keys of the stratum streams
constexpr uint32_t kIcStratumKey = 0x7f4a7c15u;

struct BatchDesc {
    uint64_t firstPath;   // global path index (within the call) of local path 0
    uint32_t numPaths;    // paths in this batch
    uint32_t npix;        // width * height
    uint32_t pathBase;    // buffer index of local path 0 (0, or the serving slot's first path)
    uint32_t pad;
};

struct PathBuffers {
    uint32_t *rng;
    uint32_t *hash;
    float4 *thr;      // throughput xyz, w = flags (uint bits)
    float4 *result;   // radiance xyz
    float4 *pendShadow; // [path * numNEE + j]: premultiplied shadow contribution, w = 1 if to add
    float4 *pendProbe;  // [path * numNEE + j]: premultiplied probe contribution
    float4 *out;      // per local path: radiance xyz, w = first hit distance
    uint2 *outSig;    // per local path: (final LCG state, primitive hash)
    float4 *cand;     // 2 per candidate, kMaxNewCacheEntries per path: origin, normal
    // guiding sample record (pc.updateGuiding)
    int4 *grec;       // per path: (current slot, slot whose distance grows, distanceFactor bits, NEE slot base)
    float4 *grecThr;  // per path: throughput at the pending NEE
    float4 *lightSum; // kMaxRecords per path
    float4 *sampleThr;// kMaxRecords per path: throughput change of the slot's vertex
    rtx_directional_data *dirData; // per pixel of the frame: kMaxRecords slots
    uint32_t dirFirstPixel;        // pixel of local path 0 (frames are batched by pixel ranges)
    // frames of several segments (FrameConst::segments): one path slot is one
    // pixel's frame, main() of the reference
    int4 *task;       // per path: (camera samples left, splits queued, next split, record sampleOffset)
    float4 *taskRes;  // per path: main()'s result so far (the segment's own result is in `result`)
    float4 *splitRec; // kMaxSplits x 4 per path: (origin, depth bits), (normal, material | frontFace << 31),
                      // (wi, texture u), (throughput, texture v)  (splitInfo, raytrace.rgen:70-80)
    const float4 *estimate; // per pixel: the ADRRS estimate (estimateImage)
    uint32_t *frameSeed;    // per path: the frame's randomUInt (any-hit alpha test seeds)
    // learned radiance cache: the queries of paths that end into it (4 float4
    // of rtx_nrc_input each; aux = (throughput, path id bits)), training
    // vertices by slot (input, and target rgb with w = 1 once the path ended),
    // and per path (throughput at the training vertex, slot bits), (radiance so far)
    float4 *nrcQuery, *nrcQueryAux;
    uint32_t *nrcQueryCount;
    float4 *nrcTrainIn, *nrcTrainTarget;
    float4 *nrcPath;
};
constexpr int kMaxSplits = 10; // MAX_SPLITS, raytrace.rgen:82
constexpr int kMaxRecords = RTX_MAX_DIRECTIONAL_DATA_PER_PIXEL;
constexpr uint32_t kDirInvalid = 0xFFFFFFFFu; // DirectionalData INVALID, guiding.glsl:112

// Path flags packed in thr.w
constexpr uint32_t kFlagDepthMask = 0xFFu;        // bits 0-7 depth
constexpr uint32_t kFlagFollowCountShift = 8;     // bits 8-15 followCount
constexpr uint32_t kFlagFollow = 1u << 16;
constexpr uint32_t kFlagAddNext = 1u << 17;
constexpr uint32_t kFlagPending = 1u << 18;       // NEE contributions waiting to be added
constexpr uint32_t kFlagCandShift = 19;           // bits 19-21 irradiance-cache candidates
constexpr uint32_t kFlagCandMask = 7u << kFlagCandShift;
constexpr uint32_t kFlagPendTask = 1u << 22;      // the pending NEE terms belong to main()'s result (a split vertex)
constexpr uint32_t kFlagNrcTrain = 1u << 23;      // a training path past its training vertex
constexpr uint32_t kFlagNrcTrainEnd = 1u << 24;   // ... that ended into the cache (its target adds the prediction)
// Extend-queue id bit: the entry is a queued split to resume (its ray is
// empty); kShade takes the vertex from the path's split record.
constexpr uint32_t kSplitVertex = 1u << 31;

// ------------------------------------------------------ serving pool
// One slot of the serving pool: a batch of one job.  Paths of slot k have
// buffer ids k * 2^slotShift + local; the bounce kernels' multi-job builds
// read each path's scene, cache, guide, frame constants and batch here.
struct SlotDesc {
    DevScene S;
    DevIC ic;
    DevGuiding g;
    FrameConst fc;
    BatchDesc b;
};
static_assert(sizeof(SlotDesc) % 4 == 0, "slot descriptors are copied as words");

struct MultiArgs {
    const SlotDesc *slots;
    uint32_t *slotAlive;   // per slot: paths not yet ended
    uint32_t slotShift;    // log2 of paths per slot
    uint32_t neeStride;    // NEE slots per path in the pool
    uint32_t numSlots;
};

// ----------------------------------------------------------- launchers

void launchGenerate(hipStream_t s, const FrameConst *fc, BatchDesc b, PathBuffers pb, Queue q0,
                    Counters *cnt, int qsel, uint32_t *electCounter);
// persistent lane-refill traversal of one bounce: extend queue qe (count
// extend[esel]), then the probe and shadow queues; fetch[1] must be zero.
// poll (may be NULL): host-visible slot for the bounce's starting queue
// lengths (extend, probe, shadow, done), written by the first thread
// ldsEntries: stack entries per thread in LDS, 12, 16, 20 or 24
void launchTraceP(hipStream_t s, const DevScene &S, Queue qe, int esel, Queue qs, Queue qp, PathBuffers pb,
                  Counters *cnt, uint32_t *poll, bool countVisits, int blocks, int ldsEntries);
// the two ray queries (closest hit / occlusion); *fetch must be zero
void launchQueryP(hipStream_t s, const DevScene &S, const rtx_ray *rays, uint32_t n, void *out, bool closest,
                  uint32_t *fetch, int blocks);
void launchProbeResolve(hipStream_t s, const DevScene &S, const FrameConst *fc, Queue q, PathBuffers pb,
                        Counters *cnt, int qsel, int blocks);
// A second class of paths in the same wavefront: ids >= split use (fc, b).
struct PathClass2 {
    const FrameConst *fc;
    BatchDesc b;
    uint32_t split;
};
void launchShade(hipStream_t s, const DevScene &S, const DevIC &ic, const DevGuiding &g, const FrameConst *fc,
                 BatchDesc b, PathBuffers pb, Queue qin, Queue qout, Queue qs, Queue qp, int32_t *done, Counters *cnt,
                 int qsel, int blocks, const PathClass2 *second = nullptr);
// Which jobs of the frame's list a sub-path launch or a reduction covers:
// [firstJob, firstJob + count) as the host says, or from device memory (the
// plan's counts, ic.frame[3] updates and ic.frame[4] jobs in all), so that no
// host round trip sizes the launch; then the host's count is an upper bound.
enum IcJobRange { kIcJobsHost = 0, kIcJobsFromFirst = 1, kIcJobsUpdates = 2, kIcJobsCreations = 3 };
// (icJobRange, dev_icache.h, reads the range on the device.)
// pidBase / qBase: the sub-paths' first path id and extend-queue entry (after
// other paths already queued, which keep their entries)
void launchGenerateIC(hipStream_t s, IcBuffers ic, uint32_t frameSeed, uint32_t firstJob, uint32_t numPaths,
                      PathBuffers pb, Queue q0, Counters *cnt, int qsel, int jobRange, uint32_t pidBase = 0,
                      uint32_t qBase = 0);
// irradiance-cache frame bookkeeping (icache.hip)
void launchIcFrameBegin(hipStream_t s, IcBuffers ic);
void launchIcCollect(hipStream_t s, IcBuffers ic, PathBuffers pb, uint32_t numPaths);
// The frame's job list (kIcPlan): update elections, then kept candidates.
// Planned in one go after the frame, or in two parts: the updates (known
// once the camera paths are generated) and then the creations.
enum IcPlanPart { kIcPlanAll = 0, kIcPlanUpdates = 1, kIcPlanCreations = 2 };
void launchIcPlan(hipStream_t s, IcBuffers ic, int part = kIcPlanAll);
void launchIcReduce(hipStream_t s, IcBuffers ic, uint32_t frameSeed, uint32_t firstJob, uint32_t numJobs,
                    PathBuffers pb, int jobRange, uint32_t outBase = 0);
void launchIcCommit(hipStream_t s, IcBuffers ic, const FrameConst *fc);
void launchIcIndex(hipStream_t s, IcBuffers ic, float c0);
void launchIcInit(hipStream_t s, IcBuffers ic);
// Ends the segments of the previous bounce's done list (done[qsel ^ 1]);
// with FrameConst::segments a path with camera samples left or splits queued
// goes on in qnext (count extend[qsel ^ 1]) instead.
void launchFinish(hipStream_t s, const FrameConst *fc, BatchDesc b, PathBuffers pb, const int32_t *done,
                  Counters *cnt, Queue qnext, int qsel, int blocks, const PathClass2 *second = nullptr);
// accum may be NULL (irradiance-cache prepare frames): then only the kept paths are written
// estimate != NULL (pc.storeEstimate): each pixel's last frame result is stored there too
void launchAccumulate(hipStream_t s, BatchDesc b, PathBuffers pb, float4 *accum, float4 *keepRgb, uint2 *keepSig,
                      bool keep, float4 *estimate);
void launchResetCounters(hipStream_t s, Counters *cnt, int which);
// p[0..n) = v (a kernel, so that captured streams hold no memset nodes)
void launchFillU32(hipStream_t s, uint32_t *p, uint32_t n, uint32_t v);
// serving pool (multi-job builds of the bounce kernels)
void launchInstallSlot(hipStream_t s, const SlotDesc &sd, SlotDesc *slots, uint32_t slot, uint32_t *alive);
void launchGenerateM(hipStream_t s, const SlotDesc *slots, uint32_t slot, uint32_t numPaths, PathBuffers pb, Queue q0,
                     Counters *cnt, int qsel);
void launchTraceM(hipStream_t s, const MultiArgs &ma, bool alpha, Queue qe, int esel, Queue qs, Queue qp,
                  PathBuffers pb, Counters *cnt, uint32_t *poll, int blocks);
void launchProbeResolveM(hipStream_t s, const MultiArgs &ma, Queue q, PathBuffers pb, Counters *cnt, int qsel,
                         int blocks);
void launchFinishM(hipStream_t s, const MultiArgs &ma, PathBuffers pb, const int32_t *done, Counters *cnt,
                   Queue qnext, int qsel, int blocks);
void launchShadeM(hipStream_t s, const MultiArgs &ma, PathBuffers pb, Queue qin, Queue qout, Queue qs, Queue qp,
                  int32_t *done, Counters *cnt, int qsel, int blocks);
void launchResolveImage(hipStream_t s, const float4 *accum, uint32_t npix, float invSamples, float4 *out);

} // namespace rtx
