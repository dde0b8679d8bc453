"""ctypes mirror of include/rtx_amd.h and include/rtx_host.h.

Every structure asserts the byte size the C header (and the reference's
device layouts, shaders/raycommon.glsl, shaders/wavefront.glsl) fixes.
"""
import ctypes as C

import numpy as np

RTX_OK = 0
RTX_E_INVALID = -1
RTX_E_DEVICE = -2
RTX_E_MEMORY = -3
RTX_E_UNSUPPORTED = -4
RTX_E_STATE = -5

RTX_FRAME_KEEP_PATHS = 0x1
RTX_FRAME_COUNT_VISITS = 0x2
RTX_FRAME_RESET_ACCUM = 0x4
RTX_FRAME_TIME_KERNELS = 0x8
RTX_FRAME_NO_GUIDING_UPDATE = 0x10
RTX_FRAME_TIME_TRACE = 0x20
RTX_FRAME_NRC = 0x40

RTX_NOALLOC_MAX_RAYS = 65536  # rays per rtx_trace_closest_noalloc call

MAT_DIFFUSE, MAT_SPECULAR, MAT_DIELECTRIC, MAT_LIGHT, MAT_PHONG, MAT_CONDUCTOR, MAT_ROUGH_CONDUCTOR = range(7)
LIGHT_AREA, LIGHT_POINT, LIGHT_SPHERE, LIGHT_ENV_MAP = range(4)
DISTRIBUTION_BECKMANN, DISTRIBUTION_GGX = 0, 1

SCENE_CORNELL = 1
SCENE_ROOM = 2
SCENE_PLATES = 3
SCENE_SPOT = 4

F3 = C.c_float * 3
M4 = C.c_float * 16


class Vertex(C.Structure):
    _fields_ = [("pos", F3), ("_pad0", C.c_float), ("normal", F3), ("_pad1", C.c_float),
                ("texCoord", C.c_float * 2), ("materialIndex", C.c_int32), ("_pad2", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("lightColor", F3), ("_pad0", C.c_float), ("diffuse", F3), ("_pad1", C.c_float),
                ("specular", F3), ("specularHighlight", C.c_float), ("transparency", C.c_float),
                ("refractionIndex", C.c_float), ("refractionIndexInv", C.c_float), ("eta", C.c_float),
                ("k", C.c_float), ("roughness", C.c_float), ("textureIdDiffuse", C.c_int32),
                ("textureIdSpecular", C.c_int32), ("type", C.c_int32), ("distribution", C.c_int32),
                ("_pad2", C.c_int32 * 2)]


class Instance(C.Structure):
    _fields_ = [("transform", M4), ("normalTransform", M4), ("modelIndex", C.c_int32),
                ("iLight", C.c_int32), ("_pad", C.c_int32 * 2)]


class Light(C.Structure):
    _fields_ = [("color", F3), ("pos", F3), ("instanceIndex", C.c_uint32), ("sampleProb", C.c_float),
                ("area", C.c_float), ("type", C.c_int32)]


class FaceSample(C.Structure):
    _fields_ = [("index", C.c_int32), ("sampleProb", C.c_float), ("faceArea", C.c_float)]


class Sphere(C.Structure):
    _fields_ = [("center", F3), ("radius", C.c_float), ("materialIndex", C.c_int32), ("iLight", C.c_int32)]


class Aabb(C.Structure):
    _fields_ = [("min", F3), ("max", F3)]


MAX_DISTRIBUTIONS = 16
RTX_MAX_LOBES = 32


class VmfTheta(C.Structure):
    _fields_ = [("mu", F3), ("k", C.c_float), ("norm", C.c_float), ("eMin2K", C.c_float), ("distance", C.c_float),
                ("target", F3)]


class VmmTheta(C.Structure):
    _fields_ = [("thetas", VmfTheta * MAX_DISTRIBUTIONS), ("pi", C.c_float * MAX_DISTRIBUTIONS),
                ("meanPosition", F3), ("usedDistributions", C.c_int32)]


class VmmThetaWide(C.Structure):
    _fields_ = [("thetas", VmfTheta * RTX_MAX_LOBES), ("pi", C.c_float * RTX_MAX_LOBES),
                ("meanPosition", F3), ("usedDistributions", C.c_int32)]


DIRECTIONAL_DTYPE = np.dtype([("position", "<f4", 3), ("direction", "<f4", 3), ("weight", "<f4"), ("pdf", "<f4"),
                              ("distance", "<f4"), ("flags", "<u4")])
DIR_INVALID = 0xFFFFFFFF
MAX_DIRECTIONAL_DATA_PER_PIXEL = 16


class CacheHeader(C.Structure):
    _fields_ = [("nextCacheSlot", C.c_uint32), ("maxCaches", C.c_uint32), ("nextUpdateSlot", C.c_uint32)]


class CacheData(C.Structure):
    _fields_ = [("color", F3), ("normal", F3), ("rotGrad", F3), ("transGrad", F3), ("harmonicR", C.c_float),
                ("numUpdates", C.c_uint32)]


class IcJob(C.Structure):
    _fields_ = [("type", C.c_int32), ("target", C.c_uint32), ("origin", F3), ("normal", F3)]


# numpy views of the cache buffers (same bytes as CacheData / Sphere arrays)
CACHE_DTYPE = np.dtype([("color", "<f4", 3), ("normal", "<f4", 3), ("rotGrad", "<f4", 3), ("transGrad", "<f4", 3),
                        ("harmonicR", "<f4"), ("numUpdates", "<u4")])
SPHERE_DTYPE = np.dtype([("center", "<f4", 3), ("radius", "<f4"), ("materialIndex", "<i4"), ("iLight", "<i4")])


PUSH_CONSTANT_FIELDS = [
    ("randomUInt", C.c_uint32), ("previousFrames", C.c_uint32), ("maxDepth", C.c_int32),
    ("maxFollowDiscrete", C.c_int32), ("samplesPerPixel", C.c_int32), ("enableRR", C.c_int32),
    ("enableNEE", C.c_int32), ("numNEE", C.c_int32), ("enableAverageInsteadOfMix", C.c_int32),
    ("enableMIS", C.c_int32), ("usePowerHeuristic", C.c_int32), ("storeEstimate", C.c_int32),
    ("visualizeMode", C.c_int32), ("showIrradianceCacheOnly", C.c_int32),
    ("showIrradianceGradients", C.c_int32), ("useIrradianceCache", C.c_int32),
    ("highlightIrradianceCacheColor", C.c_int32), ("irradianceA", C.c_float),
    ("irradianceUpdateProb", C.c_float), ("irradianceCreateProb", C.c_float),
    ("irradianceVisualizationScale", C.c_float), ("useIrradianceGradients", C.c_int32),
    ("useIrradianceCacheOnGlossy", C.c_int32), ("irradianceGradientsMaxLength", C.c_float),
    ("isIrradiancePrepareFrame", C.c_int32), ("irradianceNumNEE", C.c_int32),
    ("irradianceCacheMinRadius", C.c_float), ("irradianceCachePerformVisibilityCheck", C.c_int32),
    ("useVisibleSphereSampling", C.c_int32), ("useADRRS", C.c_int32), ("adrrsS", C.c_float),
    ("adrrsSplit", C.c_int32), ("splitOnFirst", C.c_int32), ("useGuiding", C.c_int32),
    ("guidingProb", C.c_float), ("guidingVisuScale", C.c_float), ("guidingVisuMax", C.c_float),
    ("guidingVisuIgnoreOcclusioon", C.c_int32), ("updateGuiding", C.c_int32),
    ("useParallaxCompensation", C.c_int32), ("time", C.c_float), ("guidingVisuMove", C.c_int32),
    ("guidingVisuPhiScale", C.c_float), ("guidingVisuThetaScale", C.c_float),
    ("numGuidingRegions", C.c_int32), ("guidingPiPHighlightRegion", C.c_int32),
    ("guidingPiPShowSpheres", C.c_int32), ("guidingPiPSize", C.c_float),
]


class PushConstant(C.Structure):
    _fields_ = PUSH_CONSTANT_FIELDS


class Camera(C.Structure):
    _fields_ = [("view", M4), ("proj", M4), ("viewInverse", M4), ("projInverse", M4)]


class ModelDesc(C.Structure):
    _fields_ = [("vertices", C.POINTER(Vertex)), ("vertexCount", C.c_uint32),
                ("indices", C.POINTER(C.c_uint32)), ("indexCount", C.c_uint32)]


class TextureDesc(C.Structure):
    _fields_ = [("rgba", C.POINTER(C.c_uint8)), ("width", C.c_uint32), ("height", C.c_uint32)]


class SceneDesc(C.Structure):
    _fields_ = [("models", C.POINTER(ModelDesc)), ("modelCount", C.c_uint32),
                ("instances", C.POINTER(Instance)), ("instanceCount", C.c_uint32),
                ("materials", C.POINTER(Material)), ("materialCount", C.c_uint32),
                ("lights", C.POINTER(Light)), ("lightCount", C.c_uint32),
                ("randomLightIndex", C.POINTER(C.c_int32)),
                ("randomTriIndex", C.POINTER(FaceSample)), ("randomTriRows", C.c_uint32),
                ("spheres", C.POINTER(Sphere)), ("sphereCount", C.c_uint32),
                ("envMap", C.POINTER(C.c_float)), ("envWidth", C.c_uint32), ("envHeight", C.c_uint32),
                ("textures", C.POINTER(TextureDesc)), ("textureCount", C.c_uint32)]


class FrameParams(C.Structure):
    _fields_ = [("pc", PushConstant), ("cam", Camera), ("width", C.c_uint32), ("height", C.c_uint32),
                ("spp", C.c_uint32), ("sampleOffset", C.c_uint32), ("seed", C.c_uint64),
                ("flags", C.c_uint32), ("rrStartDepth", C.c_int32), ("batchPaths", C.c_uint32),
                ("_reserved", C.c_uint32 * 7)]


class FrameStats(C.Structure):
    _fields_ = [("paths", C.c_uint64), ("extendRays", C.c_uint64), ("shadowRays", C.c_uint64),
                ("probeRays", C.c_uint64), ("nodeVisits", C.c_uint64), ("primVisits", C.c_uint64),
                ("msTotal", C.c_double), ("msExtend", C.c_double), ("msShadow", C.c_double),
                ("msProbe", C.c_double), ("msShade", C.c_double), ("batches", C.c_uint32),
                ("bounces", C.c_uint32), ("nodeBytes", C.c_uint32), ("primBytes", C.c_uint32),
                ("nodeVisitsByKind", C.c_uint64 * 3), ("primVisitsByKind", C.c_uint64 * 3),
                ("launchesByKind", C.c_uint32 * 4), ("_pad", C.c_uint32)]


class Tuning(C.Structure):
    _fields_ = [("traversalKernel", C.c_int32), ("traceBlocksPerCU", C.c_int32),
                ("shadeBlocksPerCU", C.c_int32), ("batchPaths", C.c_uint32), ("ldsStackEntries", C.c_int32),
                ("cacheUpkeep", C.c_int32), ("_reserved", C.c_int32 * 2)]


class AutotuneResult(C.Structure):
    _fields_ = [("best", Tuning), ("bestMs", C.c_double), ("startMs", C.c_double), ("trials", C.c_uint32),
                ("reserved", C.c_uint32)]


class ServerConfig(C.Structure):
    _fields_ = [("framebufferBytes", C.c_uint64), ("slotPaths", C.c_uint32), ("slots", C.c_uint32),
                ("maxNumNEE", C.c_uint32), ("maxJobs", C.c_uint32)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("state", C.c_void_p), ("pc", PushConstant), ("cam", Camera),
                ("width", C.c_uint32), ("height", C.c_uint32), ("spp", C.c_uint32), ("rrStartDepth", C.c_int32),
                ("seed", C.c_uint64), ("sampleOffset", C.c_uint32), ("reserved", C.c_uint32)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("batches", C.c_uint32), ("batchesDone", C.c_uint32),
                ("reserved", C.c_uint32), ("paths", C.c_uint64), ("queuedMs", C.c_double),
                ("latencyMs", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("jobsDone", C.c_uint64), ("pathsDone", C.c_uint64), ("latencyP50Ms", C.c_double),
                ("latencyP99Ms", C.c_double), ("latencyMaxMs", C.c_double), ("latencyMeanMs", C.c_double),
                ("msamplesPerSec", C.c_double), ("bounces", C.c_uint64), ("extendRays", C.c_uint64),
                ("shadowRays", C.c_uint64), ("probeRays", C.c_uint64), ("peakSlotsBusy", C.c_uint32),
                ("peakJobsRunning", C.c_uint32), ("framebufferPeakBytes", C.c_uint64), ("poolBytes", C.c_uint64)]


RTX_JOB_QUEUED, RTX_JOB_RUNNING, RTX_JOB_DONE, RTX_JOB_RELEASED = range(4)


class ScheduleConfig(C.Structure):
    _fields_ = [("prepareFrames", C.c_int32), ("guidingOptimizations", C.c_int32), ("accumulate", C.c_int32),
                ("evalSamples", C.c_int32), ("evalSeconds", C.c_int32), ("sppToOneAfterEval", C.c_int32)]


class ScheduleState(C.Structure):
    _fields_ = [("frames", C.c_uint64), ("currentPrepareFrames", C.c_int32),
                ("currentGuidingOptimizations", C.c_int32), ("previousFrames", C.c_uint32),
                ("evalSamplesRunning", C.c_int32), ("evalSecondsRunning", C.c_int32),
                ("evalCurrentSamples", C.c_int32), ("evalSecondsTotalSamples", C.c_int64),
                ("evalSecondsResult", C.c_int64), ("evalSamplesMs", C.c_double), ("guidingUpdateMs", C.c_double),
                ("pc", PushConstant)]


class NrcConfig(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("log2TableSize", C.c_uint32), ("baseResolution", C.c_float),
                ("maxResolution", C.c_float), ("learningRate", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("epsilon", C.c_float), ("trainBatch", C.c_uint32), ("trainSteps", C.c_uint32),
                ("terminateBounces", C.c_uint32), ("trainingPaths", C.c_uint32), ("sceneMin", F3),
                ("sceneMax", F3), ("seed", C.c_uint64), ("trainSuffixBounces", C.c_uint32), ("emaDecay", C.c_float)]


class NrcInput(C.Structure):
    _fields_ = [("pos", F3), ("roughness", C.c_float), ("dir", F3), ("_pad0", C.c_float), ("normal", F3),
                ("_pad1", C.c_float), ("albedo", F3), ("_pad2", C.c_float)]


class NrcInfo(C.Structure):
    _fields_ = [("levels", C.c_uint32), ("tableSize", C.c_uint32), ("resolution", C.c_float * 16),
                ("dense", C.c_uint32 * 16), ("lo", F3), ("inv", F3), ("gridParams", C.c_uint64),
                ("mlpParams", C.c_uint32), ("steps", C.c_uint32), ("lastLoss", C.c_float)]


class GuidingOptions(C.Structure):
    _fields_ = [("splitAndMerge", C.c_int32), ("minSamplesForMerging", C.c_uint32),
                ("minSamplesForSplitting", C.c_uint32), ("minSamplesForPostSplitFitting", C.c_uint32),
                ("splitMinDivergence", C.c_float), ("mergeMaxDivergence", C.c_float),
                ("splitRegions", C.c_int32), ("samplesForRegionSplit", C.c_uint32),
                ("maxComponents", C.c_uint32), ("maxRegions", C.c_uint32), ("maxSamplesPerUpdate", C.c_uint32),
                ("reserved", C.c_uint32)]


def guiding_options(**kw):
    """PathGuiding's defaults (src/PathGuiding.h:118-129; 16 lobes, no region
    or sample cap), overridden by kw."""
    o = GuidingOptions(1, 8192, 4096, 4096, 0.5, 0.025, 0, 10000, MAX_DISTRIBUTIONS, 0, 0, 0)
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class SceneInfo(C.Structure):
    _fields_ = [("triangles", C.c_uint32), ("spheres", C.c_uint32), ("nodes", C.c_uint32),
                ("maxDepth", C.c_uint32), ("deviceBytes", C.c_uint64), ("buildMs", C.c_double),
                ("boundsMin", F3), ("boundsMax", F3)]


class Ray(C.Structure):
    _fields_ = [("org", F3), ("tmin", C.c_float), ("dir", F3), ("tmax", C.c_float)]


class Hit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim", C.c_int32), ("u", C.c_float), ("v", C.c_float)]


class View(C.Structure):
    _fields_ = [("origin", F3), ("target", F3), ("up", F3), ("vfov", C.c_float)]


EXPECTED_SIZES = {
    Vertex: 48, Material: 96, Instance: 144, Light: 40, FaceSample: 12, Sphere: 24,
    PushConstant: 192, Camera: 256, Ray: 32, Hit: 16, Tuning: 32, CacheHeader: 12, CacheData: 56, IcJob: 32,
    Aabb: 24, VmfTheta: 40, VmmTheta: 720, VmmThetaWide: 1424, GuidingOptions: 48,
}
for _t, _n in EXPECTED_SIZES.items():
    assert C.sizeof(_t) == _n, (_t.__name__, C.sizeof(_t), _n)


def push_constant_defaults():
    """The reference's push-constant defaults, src/RayTracingApp.h:81-130."""
    pc = PushConstant()
    pc.previousFrames = 0xFFFFFFFF
    pc.maxDepth = 30
    pc.maxFollowDiscrete = 3
    pc.samplesPerPixel = 1
    pc.enableNEE = 1
    pc.numNEE = 1
    pc.usePowerHeuristic = 1
    pc.irradianceA = 0.2
    pc.irradianceUpdateProb = 0.00001
    pc.irradianceCreateProb = 0.0005
    pc.irradianceVisualizationScale = 1.0
    pc.irradianceGradientsMaxLength = 5
    pc.irradianceNumNEE = 1
    pc.irradianceCacheMinRadius = 0.1
    pc.adrrsS = 5
    pc.adrrsSplit = 1
    pc.guidingProb = 0.5
    pc.guidingVisuScale = 0.5
    pc.guidingVisuMax = 1.0
    pc.useParallaxCompensation = 1
    pc.guidingVisuPhiScale = 0.004
    pc.guidingVisuThetaScale = 0.01
    pc.guidingPiPShowSpheres = 1
    pc.guidingPiPSize = 0.3
    return pc
