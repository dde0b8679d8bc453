// Serving pool: a persistent request loop that runs the batches of many
// render jobs together in one wavefront (include/rtx_amd.h, "serving").
//
// Device side: a pool of `slots` x `slotPaths` path states and the queues of
// one wavefront.  Slot k holds one batch of one job (SlotDesc: the job's
// scene, cache, guide, frame constants and the batch); its paths have buffer
// ids k * slotPaths + local.  Each bounce is one launch of kTraceM,
// kProbeResolveM, kFinishM and kShadeM over every slot's paths; kFinishM
// counts each slot's paths down as they end, and kTraceM copies the counts
// into a pinned ring that the loop reads kPollLead bounces behind the
// device, so the stream never drains while jobs run.
//
// Host side: one thread.  Per bounce it admits queued jobs whose images fit
// the framebuffer arena (in arrival order), gives each free slot the next
// batch of the admitted job holding the fewest slots (install + generate,
// appended to the next bounce's extend queue), launches the bounce, and reads the
// ring of an earlier bounce: a slot whose count reached zero holds a
// finished batch.  Finished batches are added to their job's image in batch
// order (then the slot is free), and a job whose batches are all added gets
// its mean image and a completion event.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "abi_guard.h"
#include "context.h"

using Clock = std::chrono::steady_clock;

namespace {

double msBetween(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

// First-fit allocator over one device allocation (the framebuffer budget).
class Arena {
  public:
    void init(size_t bytes) {
        free_.clear();
        free_[0] = bytes;
        used_ = peak_ = 0;
    }
    // offset of a block of `bytes` (multiple of 256), or SIZE_MAX
    size_t alloc(size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const size_t off = it->first, rest = it->second - bytes;
                free_.erase(it);
                if (rest) free_[off + bytes] = rest;
                used_ += bytes;
                peak_ = std::max(peak_, used_);
                return off;
            }
        }
        return SIZE_MAX;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + 255) & ~(size_t)255;
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto next = std::next(it);
        if (next != free_.end() && it->first + it->second == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t peak() const { return peak_; }

  private:
    std::map<size_t, size_t> free_;
    size_t used_ = 0, peak_ = 0;
};

struct Job {
    uint64_t id = 0;
    rtx_job_desc desc{};
    SlotDesc proto{};          // the job's slot descriptor without the batch
    uint64_t npix = 0, total = 0;
    uint32_t batchPaths = 0, numBatches = 0;
    uint32_t nextBatch = 0;    // next batch to give a slot
    uint32_t nextAdd = 0;      // next batch to add to the image
    std::vector<int> finishedSlot; // per batch: slot holding it once finished (-1: not yet)
    size_t fbOffset = SIZE_MAX, fbBytes = 0;
    float4 *accum = nullptr, *image = nullptr;
    int32_t state = RTX_JOB_QUEUED;
    bool resolving = false;
    hipEvent_t doneEv = nullptr;
    Clock::time_point tSubmit, tAdmit, tDone;
};

struct Slot {
    Job *job = nullptr;
    uint32_t batch = 0;
    uint64_t installedAt = 0;  // bounce before which it was installed
    bool finished = false;
};

} // namespace

struct rtx_server {
    rtx_ctx *ctx = nullptr;
    rtx_server_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t slotShift = 0;
    size_t poolPaths = 0;
    std::vector<void *> allocs;
    PathBuffers pb{};
    Queue q[2]{}, qs{}, qp{};
    int32_t *done = nullptr;
    Counters *cnt = nullptr;
    SlotDesc *slotTable = nullptr;
    uint32_t *slotAlive = nullptr;
    uint32_t *ring = nullptr;  // pinned: kPollSlots x (4 + slots) words
    uint32_t ringStride = 0;
    hipEvent_t ringEv[kPollSlots] = {};
    char *fbBase = nullptr;
    Arena arena;
    size_t poolBytes = 0;
    int traceBlocks = 0, shadeBlocks = 0;
    // jobs and the loop
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::map<uint64_t, std::unique_ptr<Job>> jobs;
    std::deque<Job *> waiting;  // submitted, not admitted (arrival order)
    std::vector<Job *> running; // admitted, image not yet complete (admission order)
    uint64_t nextId = 1;
    bool stop = false;
    std::thread thread;
    std::vector<Slot> slots;
    // statistics (under mu)
    std::vector<double> latencies;
    uint64_t pathsDone = 0;
    bool anySubmit = false;
    Clock::time_point firstSubmit, lastDone;
    uint64_t bounces = 0, extendRays = 0, shadowRays = 0, probeRays = 0;
    uint32_t peakSlots = 0, peakRunning = 0;
    std::string err;           // last API error
    std::string loopErr;       // a device error that stopped the loop
};

namespace {

template <class T> hipError_t poolAlloc(rtx_server *sv, T **p, size_t count) {
    void *q = nullptr;
    hipError_t e = devMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) {
        sv->allocs.push_back(q);
        sv->poolBytes += std::max<size_t>(count, 1) * sizeof(T);
        *p = (T *)q;
    }
    return e;
}

void freePool(rtx_server *sv) {
    for (void *p : sv->allocs) (void)hipFree(p);
    sv->allocs.clear();
    for (hipEvent_t &e : sv->ringEv)
        if (e) (void)hipEventDestroy(e), e = nullptr;
    if (sv->ring) (void)hipHostFree(sv->ring), sv->ring = nullptr;
    if (sv->stream) (void)hipStreamDestroy(sv->stream), sv->stream = nullptr;
}

int setErr(rtx_server *sv, int code, const std::string &msg) {
    sv->err = msg;
    return code;
}

// What rtx_render would not run here (the serving pool has no per-frame
// upkeep, records or multi-sample frames).
const char *unsupportedInPool(const rtx_push_constant &pc) {
    if (pc.visualizeMode != 0 || pc.showIrradianceCacheOnly) return "visualisation modes";
    if (pc.useADRRS || pc.splitOnFirst) return "ADRRS and splitting";
    if (pc.updateGuiding) return "guiding records (updateGuiding)";
    if (pc.storeEstimate) return "storeEstimate";
    if (pc.isIrradiancePrepareFrame) return "irradiance-cache prepare frames";
    if (pc.samplesPerPixel > 1) return "samplesPerPixel > 1";
    return nullptr;
}

// Frame constants of a job's camera paths, as rtx_render sets them for a
// frame of one camera sample (raytrace(..., maxDepth, maxFollowDiscrete,
// !enableNEE, true, enableNEE, numNEE, useIrradianceCache, ...),
// raytrace.rgen:1671), with the pool's NEE stride and no cache upkeep.
FrameConst jobFrame(const rtx_job_desc &d, uint32_t neeStride) {
    FrameConst fc;
    std::memset(&fc, 0, sizeof(fc));
    fc.pc = d.pc;
    fc.cam = d.cam;
    fc.width = d.width;
    fc.height = d.height;
    fc.seed = d.seed;
    fc.sampleOffset = d.sampleOffset;
    fc.rrStartDepth = d.rrStartDepth;
    const bool useIC = d.pc.useIrradianceCache != 0;
    const bool needNEE = d.pc.enableNEE || useIC;
    fc.numNEE = needNEE ? (uint32_t)d.pc.numNEE : 1u;
    fc.neeStride = neeStride;
    fc.countVisits = 0;
    fc.debugPath = ~0ull;
    fc.maxDepth = d.pc.maxDepth;
    fc.maxFollow = d.pc.maxFollowDiscrete;
    fc.addDirect = d.pc.enableNEE ? 0u : 1u;
    fc.useNEE = d.pc.enableNEE ? 1u : 0u;
    fc.useIC = useIC ? 1u : 0u;
    fc.createIC = 0;
    fc.allowRR = 1;
    fc.record = 0;
    fc.samplesPerFrame = 1;
    fc.segments = 0;
    fc.useADRRS = 0;
    fc.splitOnFirst = 0;
    fc.nrc = 0;
    fc.probeFirstLight = 1; // taken for slots whose scene has few emitters and no environment map
    return fc;
}

// Batch size of a job: whole sample planes when one fits the slot, else
// pixel ranges of one plane (rtx_render's rule with batchPaths = slotPaths).
uint32_t jobBatchPaths(uint64_t npix, uint64_t total, uint32_t slotPaths) {
    uint64_t batch = std::min<uint64_t>(slotPaths, total);
    if (batch >= npix) batch -= batch % npix;
    return (uint32_t)batch;
}

// --------------------------------------------------------------- the loop
// All device work of the server runs here, on sv->stream.
class Loop {
  public:
    explicit Loop(rtx_server *sv) : sv_(sv) {}
    void run();

  private:
    bool check(hipError_t e, const char *what);
    void admit();                      // under mu
    void fillSlots();
    void bounce();
    void readBack(uint64_t j);
    void addFinished();
    void pollDone();                   // under mu
    uint32_t busySlots() const;

    rtx_server *sv_;
    int qsel_ = 0;
    uint64_t it_ = 0;                  // bounces launched
    uint64_t read_ = 0;                // bounces whose ring entry was read
    bool fresh_ = true;                // counters need a reset before the next install
    bool alpha_ = false;
};

bool Loop::check(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    std::lock_guard<std::mutex> lock(sv_->mu);
    if (sv_->loopErr.empty()) sv_->loopErr = std::string(what) + ": " + hipGetErrorString(e);
    sv_->cvDone.notify_all();
    return false;
}

uint32_t Loop::busySlots() const {
    uint32_t n = 0;
    for (const Slot &s : sv_->slots) n += s.job ? 1u : 0u;
    return n;
}

void Loop::admit() {
    while (!sv_->waiting.empty()) {
        Job *j = sv_->waiting.front();
        const size_t off = sv_->arena.alloc(j->fbBytes);
        if (off == SIZE_MAX) break; // arrival order: later jobs wait behind it
        sv_->waiting.pop_front();
        j->fbOffset = off;
        j->accum = reinterpret_cast<float4 *>(sv_->fbBase + off);
        j->image = j->accum + j->npix;
        j->state = RTX_JOB_RUNNING;
        j->tAdmit = Clock::now();
        sv_->running.push_back(j);
        sv_->peakRunning = std::max<uint32_t>(sv_->peakRunning, (uint32_t)sv_->running.size());
        // a kernel clears the accumulation image (no memset node in the stream)
        launchFillU32(sv_->stream, reinterpret_cast<uint32_t *>(j->accum), (uint32_t)(j->npix * 4), 0u);
    }
}

void Loop::fillSlots() {
    std::vector<Job *> order;
    {
        std::lock_guard<std::mutex> lock(sv_->mu);
        admit();
        order = sv_->running;
    }
    // A free slot goes to the admitted job holding the fewest slots (the
    // earliest of those): every admitted job keeps batches in flight, so a
    // small job is not held behind a large one that arrived first.
    std::unordered_map<const Job *, uint32_t> held;
    for (const Slot &s : sv_->slots)
        if (s.job) held[s.job]++;
    for (uint32_t k = 0; k < sv_->slots.size(); ++k) {
        Slot &s = sv_->slots[k];
        if (s.job) continue;
        Job *j = nullptr;
        uint32_t best = UINT32_MAX;
        for (Job *c : order) {
            if (c->nextBatch >= c->numBatches) continue;
            const uint32_t h = held[c];
            if (h < best) {
                best = h;
                j = c;
            }
        }
        if (!j) break;
        held[j]++;
        if (fresh_) { // the pool was idle: every counter starts at zero
            launchFillU32(sv_->stream, reinterpret_cast<uint32_t *>(sv_->cnt), sizeof(Counters) / 4, 0u);
            qsel_ = 0;
            fresh_ = false;
        }
        const uint32_t b = j->nextBatch++;
        SlotDesc sd = j->proto;
        sd.b.firstPath = (uint64_t)b * j->batchPaths;
        sd.b.numPaths = (uint32_t)std::min<uint64_t>(j->batchPaths, j->total - sd.b.firstPath);
        sd.b.npix = (uint32_t)j->npix;
        sd.b.pathBase = k << sv_->slotShift;
        launchInstallSlot(sv_->stream, sd, sv_->slotTable, k, sv_->slotAlive);
        launchGenerateM(sv_->stream, sv_->slotTable, k, sd.b.numPaths, sv_->pb, sv_->q[qsel_], sv_->cnt, qsel_);
        s.job = j;
        s.batch = b;
        s.installedAt = it_;
        s.finished = false;
        if (j->proto.S.alphaTris) alpha_ = true;
    }
    std::lock_guard<std::mutex> lock(sv_->mu);
    sv_->peakSlots = std::max(sv_->peakSlots, busySlots());
}

void Loop::bounce() {
    MultiArgs ma;
    ma.slots = sv_->slotTable;
    ma.slotAlive = sv_->slotAlive;
    ma.slotShift = sv_->slotShift;
    ma.neeStride = sv_->cfg.maxNumNEE;
    ma.numSlots = (uint32_t)sv_->slots.size();
    hipStream_t s = sv_->stream;
    const int r = (int)(it_ % kPollSlots);
    uint32_t *poll = sv_->ring + (size_t)r * sv_->ringStride;
    launchTraceM(s, ma, alpha_, sv_->q[qsel_], qsel_, sv_->qs, sv_->qp, sv_->pb, sv_->cnt, poll, sv_->traceBlocks);
    (void)hipEventRecord(sv_->ringEv[r], s);
    launchProbeResolveM(s, ma, sv_->qp, sv_->pb, sv_->cnt, qsel_, sv_->shadeBlocks);
    launchFinishM(s, ma, sv_->pb, sv_->done, sv_->cnt, sv_->q[qsel_ ^ 1], qsel_, sv_->shadeBlocks);
    launchShadeM(s, ma, sv_->pb, sv_->q[qsel_], sv_->q[qsel_ ^ 1], sv_->qs, sv_->qp, sv_->done, sv_->cnt, qsel_,
                 sv_->shadeBlocks);
    qsel_ ^= 1;
    it_++;
}

// Reads the ring entry of bounce j: its ray counts and every slot's live
// paths when it was traced (finish of bounce j - 1 done).
void Loop::readBack(uint64_t j) {
    const int r = (int)(j % kPollSlots);
    if (!check(hipEventSynchronize(sv_->ringEv[r]), "hipEventSynchronize")) {
        read_ = it_; // the loop stops on the error
        return;
    }
    const uint32_t *c = sv_->ring + (size_t)r * sv_->ringStride;
    {
        std::lock_guard<std::mutex> lock(sv_->mu);
        sv_->extendRays += c[0];
        sv_->probeRays += c[1];
        sv_->shadowRays += c[2];
    }
    for (uint32_t k = 0; k < sv_->slots.size(); ++k) {
        Slot &s = sv_->slots[k];
        // counts of bounces traced after the batch was installed
        if (s.job && !s.finished && j >= s.installedAt && c[4 + k] == 0) {
            s.finished = true;
            s.job->finishedSlot[s.batch] = (int)k;
        }
    }
    read_ = j + 1;
}

// Adds finished batches to their images in batch order; frees their slots.
void Loop::addFinished() {
    for (Job *j : std::vector<Job *>(sv_->running)) {
        while (j->nextAdd < j->numBatches && j->finishedSlot[j->nextAdd] >= 0) {
            const int k = j->finishedSlot[j->nextAdd];
            Slot &s = sv_->slots[(size_t)k];
            BatchDesc b;
            b.firstPath = (uint64_t)j->nextAdd * j->batchPaths;
            b.numPaths = (uint32_t)std::min<uint64_t>(j->batchPaths, j->total - b.firstPath);
            b.npix = (uint32_t)j->npix;
            b.pathBase = 0;
            PathBuffers pb = sv_->pb;
            pb.out = sv_->pb.out + ((size_t)k << sv_->slotShift);
            pb.outSig = sv_->pb.outSig + ((size_t)k << sv_->slotShift);
            launchAccumulate(sv_->stream, b, pb, j->accum, nullptr, nullptr, false, nullptr);
            s.job = nullptr;
            s.finished = false;
            std::lock_guard<std::mutex> lock(sv_->mu);
            j->nextAdd++;
        }
        if (j->nextAdd == j->numBatches && !j->resolving) {
            launchResolveImage(sv_->stream, j->accum, (uint32_t)j->npix, 1.0f / (float)j->desc.spp, j->image);
            if (!check(hipEventRecord(j->doneEv, sv_->stream), "hipEventRecord")) return;
            std::lock_guard<std::mutex> lock(sv_->mu);
            j->resolving = true;
        }
    }
}

void Loop::pollDone() {
    const auto now = Clock::now();
    for (size_t i = 0; i < sv_->running.size();) {
        Job *j = sv_->running[i];
        if (j->resolving && hipEventQuery(j->doneEv) == hipSuccess) {
            j->tDone = now;
            j->state = RTX_JOB_DONE;
            sv_->latencies.push_back(msBetween(j->tSubmit, now));
            sv_->pathsDone += j->total;
            sv_->lastDone = now;
            sv_->running.erase(sv_->running.begin() + (long)i);
            sv_->cvDone.notify_all();
        } else {
            ++i;
        }
    }
}

void Loop::run() {
    if (!check(hipSetDevice(sv_->device), "hipSetDevice")) return;
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(sv_->mu);
            if (!sv_->loopErr.empty()) break;
            pollDone();
            const bool busy = busySlots() > 0;
            bool resolving = false;
            for (Job *j : sv_->running) resolving = resolving || j->resolving;
            // stopping: the admitted jobs finish, queued ones are dropped
            if (!busy && sv_->running.empty() && sv_->stop) break;
            if (!busy) {
                // nothing in flight: every ring entry launched is read, the
                // next install starts from fresh counters
                if (it_ > read_) {
                    lock.unlock();
                    while (read_ < it_) readBack(read_);
                    continue;
                }
                fresh_ = true;
                alpha_ = false;
                bool admissible = false;
                for (Job *j : sv_->running) admissible = admissible || j->nextBatch < j->numBatches;
                if (!admissible && (sv_->waiting.empty() || sv_->stop)) {
                    if (resolving) sv_->cvWork.wait_for(lock, std::chrono::microseconds(100));
                    else if (!sv_->stop) sv_->cvWork.wait(lock, [&] { return sv_->stop || !sv_->waiting.empty(); });
                    continue;
                }
                if (!admissible) { // queued jobs wait for images to be read
                    admit();
                    bool any = false;
                    for (Job *j : sv_->running) any = any || j->nextBatch < j->numBatches;
                    if (!any) {
                        sv_->cvWork.wait_for(lock, std::chrono::microseconds(200));
                        continue;
                    }
                }
            }
        }
        fillSlots();
        if (busySlots() == 0) continue;
        bounce();
        {
            std::lock_guard<std::mutex> lock(sv_->mu);
            sv_->bounces = it_;
        }
        if (it_ > (uint64_t)kPollLead) readBack(it_ - 1 - kPollLead);
        addFinished();
        if (!check(hipGetLastError(), "kernel launch")) break;
    }
    (void)hipStreamSynchronize(sv_->stream);
}

} // namespace

extern "C" {

void rtx_server_config_defaults(rtx_server_config *c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->framebufferBytes = 1ull << 30;
    c->slotPaths = 1u << 18;
    c->slots = 16;
    c->maxNumNEE = 1;
    c->maxJobs = 4096;
}

int rtx_server_create(rtx_ctx *ctx, const rtx_server_config *c, rtx_server **out) try {
    if (!ctx || !out) return RTX_E_INVALID;
    *out = nullptr;
    rtx_server_config cfg;
    if (c) cfg = *c;
    else rtx_server_config_defaults(&cfg);
    const bool pow2 = cfg.slotPaths && (cfg.slotPaths & (cfg.slotPaths - 1)) == 0;
    if (!pow2 || cfg.slotPaths < (1u << 10) || cfg.slotPaths > (1u << 22) || cfg.slots < 1 || cfg.slots > 64 ||
        cfg.maxNumNEE < 1 || cfg.maxNumNEE > 8 || cfg.maxJobs < 1 || cfg.framebufferBytes < 4096) {
        ctx->err = "rtx_server_config: slotPaths a power of two in [2^10, 2^22], slots in [1, 64], maxNumNEE in "
                   "[1, 8], maxJobs >= 1, framebufferBytes >= 4096";
        return RTX_E_INVALID;
    }
    std::unique_ptr<rtx_server> sv(new (std::nothrow) rtx_server());
    if (!sv) return RTX_E_MEMORY;
    sv->ctx = ctx;
    sv->cfg = cfg;
    sv->device = ctx->device;
    sv->slotShift = (uint32_t)__builtin_ctz(cfg.slotPaths);
    sv->poolPaths = (size_t)cfg.slots * cfg.slotPaths;
    sv->slots.resize(cfg.slots);
    sv->traceBlocks = ctx->cuCount * ctx->tuning.traceBlocksPerCU;
    sv->shadeBlocks = ctx->cuCount * ctx->tuning.shadeBlocksPerCU;
    sv->ringStride = 4 + cfg.slots;
    auto fail = [&](hipError_t e, const char *what) {
        ctx->err = std::string("rtx_server_create: ") + what + ": " + hipGetErrorString(e);
        freePool(sv.get());
        return e == hipErrorOutOfMemory ? RTX_E_MEMORY : RTX_E_DEVICE;
    };
    hipError_t e = hipSetDevice(sv->device);
    if (e != hipSuccess) return fail(e, "hipSetDevice");
    if ((e = hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking)) != hipSuccess)
        return fail(e, "hipStreamCreateWithFlags");
    const size_t P = sv->poolPaths, PN = P * cfg.maxNumNEE;
    PathBuffers &pb = sv->pb;
    struct A {
        hipError_t e;
    } a{hipSuccess};
    auto al = [&](auto **p, size_t n) {
        if (a.e == hipSuccess) a.e = poolAlloc(sv.get(), p, n);
    };
    al(&pb.rng, P);
    al(&pb.hash, P);
    al(&pb.thr, P);
    al(&pb.result, P);
    al(&pb.pendShadow, PN);
    al(&pb.pendProbe, PN);
    al(&pb.out, P);
    al(&pb.outSig, P);
    al(&pb.frameSeed, P);
    for (Queue &q : sv->q) {
        al(&q.id, P);
        al(&q.org, P);
        al(&q.dir, P);
        al(&q.hit, P);
    }
    al(&sv->qs.id, 2 * PN); // a shadow ray and a probe's occlusion ray per NEE slot
    al(&sv->qs.org, 2 * PN);
    al(&sv->qs.dir, 2 * PN);
    al(&sv->qp.id, PN);
    al(&sv->qp.org, PN);
    al(&sv->qp.dir, PN);
    al(&sv->qp.hit, PN);
    al(&sv->qp.aux, PN);
    al(&sv->done, P);
    al(&sv->cnt, 1);
    al(&sv->slotTable, cfg.slots);
    al(&sv->slotAlive, cfg.slots);
    if (a.e != hipSuccess) return fail(a.e, "hipMalloc (path pool)");
    for (Queue *q : {&sv->q[0], &sv->q[1], &sv->qs, &sv->qp}) {
        q->pathSeed = pb.frameSeed;
        q->seedDiv = q == &sv->qs || q == &sv->qp ? cfg.maxNumNEE : 1u;
    }
    void *fb = nullptr;
    if ((e = devMalloc(&fb, cfg.framebufferBytes)) != hipSuccess) return fail(e, "hipMalloc (framebuffers)");
    sv->allocs.push_back(fb);
    sv->fbBase = static_cast<char *>(fb);
    sv->arena.init(cfg.framebufferBytes);
    if ((e = hipHostMalloc((void **)&sv->ring, kPollSlots * sv->ringStride * sizeof(uint32_t),
                           hipHostMallocDefault)) != hipSuccess)
        return fail(e, "hipHostMalloc (ring)");
    for (hipEvent_t &ev : sv->ringEv)
        if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return fail(e, "hipEventCreate");
    launchFillU32(sv->stream, sv->slotAlive, cfg.slots, 0u);
    if ((e = hipStreamSynchronize(sv->stream)) != hipSuccess) return fail(e, "hipStreamSynchronize");
    rtx_server *raw = sv.release();
    raw->thread = std::thread([raw] { Loop(raw).run(); });
    *out = raw;
    return RTX_OK;
} catch (...) {
    return rtx::abiFail(ctx);
}

int rtx_server_destroy(rtx_server *sv) try {
    if (!sv) return RTX_E_INVALID;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        sv->stop = true;
        sv->cvWork.notify_all();
    }
    if (sv->thread.joinable()) sv->thread.join();
    (void)hipSetDevice(sv->device);
    for (auto &kv : sv->jobs)
        if (kv.second->doneEv) (void)hipEventDestroy(kv.second->doneEv);
    freePool(sv);
    delete sv;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_server_submit(rtx_server *sv, const rtx_job_desc *d, uint64_t *jobId) try {
    if (!sv) return RTX_E_INVALID;
    if (!d || !jobId || !d->scene) return setErr(sv, RTX_E_INVALID, "NULL argument");
    if (d->width == 0 || d->height == 0 || d->spp == 0) return setErr(sv, RTX_E_INVALID, "empty job");
    if (const char *f = unsupportedInPool(d->pc))
        return setErr(sv, RTX_E_UNSUPPORTED, std::string("not in the serving pool: ") + f);
    if (d->pc.maxDepth < 0 || d->pc.maxDepth > 200 || d->pc.maxFollowDiscrete < 0 || d->pc.maxFollowDiscrete > 50)
        return setErr(sv, RTX_E_INVALID, "maxDepth must be in [0, 200] and maxFollowDiscrete in [0, 50]");
    const bool useIC = d->pc.useIrradianceCache != 0;
    if ((d->pc.enableNEE || useIC) && (d->pc.numNEE < 1 || (uint32_t)d->pc.numNEE > sv->cfg.maxNumNEE))
        return setErr(sv, RTX_E_INVALID, "numNEE must be in [1, maxNumNEE of the server]");
    if (d->pc.useGuiding && !(d->pc.guidingProb >= 0.0f && d->pc.guidingProb <= 1.0f))
        return setErr(sv, RTX_E_INVALID, "guidingProb must be in [0, 1]");
    if (d->scene->device != sv->device) return setErr(sv, RTX_E_INVALID, "scene on another device");
    auto job = std::make_unique<Job>();
    job->desc = *d;
    job->npix = (uint64_t)d->width * d->height;
    job->total = job->npix * d->spp;
    if (job->npix > (1u << 30) || job->total > (1ull << 40)) return setErr(sv, RTX_E_INVALID, "job too large");
    job->fbBytes = (size_t)job->npix * 2 * sizeof(float4);
    if (job->fbBytes > sv->cfg.framebufferBytes)
        return setErr(sv, RTX_E_MEMORY, "the job's image alone exceeds framebufferBytes");
    SlotDesc &sd = job->proto;
    std::memset(&sd, 0, sizeof(sd));
    sd.S = d->scene->view;
    sd.fc = jobFrame(*d, sv->cfg.maxNumNEE);
    if (useIC) {
        if (!d->state || !d->state->icValid)
            return setErr(sv, RTX_E_STATE, "useIrradianceCache needs a state context with a cache (rtx_ic_create)");
        sd.ic = icView(d->state->ic);
    }
    if (d->pc.useGuiding) {
        if (!d->state || d->state->guide.numRegions == 0)
            return setErr(sv, RTX_E_STATE, "useGuiding needs a state context with guiding regions");
        sd.g = d->state->guide;
    }
    job->batchPaths = jobBatchPaths(job->npix, job->total, sv->cfg.slotPaths);
    job->numBatches = (uint32_t)((job->total + job->batchPaths - 1) / job->batchPaths);
    job->finishedSlot.assign(job->numBatches, -1);
    hipError_t e = hipSetDevice(sv->device);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&job->doneEv, hipEventDisableTiming);
    if (e != hipSuccess) return setErr(sv, RTX_E_DEVICE, std::string("hipEventCreate: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(sv->mu);
    if (!sv->loopErr.empty()) {
        (void)hipEventDestroy(job->doneEv);
        return setErr(sv, RTX_E_DEVICE, "server stopped: " + sv->loopErr);
    }
    if (sv->stop) {
        (void)hipEventDestroy(job->doneEv);
        return setErr(sv, RTX_E_STATE, "server is stopping");
    }
    if (sv->jobs.size() >= sv->cfg.maxJobs) {
        (void)hipEventDestroy(job->doneEv);
        return setErr(sv, RTX_E_MEMORY, "maxJobs jobs queued, running or unread");
    }
    job->id = sv->nextId++;
    job->tSubmit = Clock::now();
    if (!sv->anySubmit) {
        sv->anySubmit = true;
        sv->firstSubmit = job->tSubmit;
    }
    *jobId = job->id;
    sv->waiting.push_back(job.get());
    sv->jobs.emplace(job->id, std::move(job));
    sv->cvWork.notify_all();
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

static void fillStatus(const Job &j, rtx_job_status *st) {
    std::memset(st, 0, sizeof(*st));
    st->state = j.state;
    st->batches = j.numBatches;
    st->batchesDone = j.nextAdd;
    st->paths = j.total;
    const auto now = Clock::now();
    st->queuedMs = j.state == RTX_JOB_QUEUED ? msBetween(j.tSubmit, now) : msBetween(j.tSubmit, j.tAdmit);
    st->latencyMs = j.state >= RTX_JOB_DONE ? msBetween(j.tSubmit, j.tDone) : -1.0;
}

int rtx_server_poll(rtx_server *sv, uint64_t id, rtx_job_status *st) try {
    if (!sv) return RTX_E_INVALID;
    if (!st) return setErr(sv, RTX_E_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return setErr(sv, RTX_E_INVALID, "unknown job");
    fillStatus(*it->second, st);
    if (!sv->loopErr.empty() && it->second->state < RTX_JOB_DONE)
        return setErr(sv, RTX_E_DEVICE, "server stopped: " + sv->loopErr);
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_server_wait(rtx_server *sv, uint64_t id, rtx_job_status *st) try {
    if (!sv) return RTX_E_INVALID;
    std::unique_lock<std::mutex> lock(sv->mu);
    auto it = sv->jobs.find(id);
    if (it == sv->jobs.end()) return setErr(sv, RTX_E_INVALID, "unknown job");
    Job &j = *it->second;
    sv->cvDone.wait(lock, [&] { return j.state >= RTX_JOB_DONE || !sv->loopErr.empty(); });
    if (st) fillStatus(j, st);
    if (j.state < RTX_JOB_DONE) return setErr(sv, RTX_E_DEVICE, "server stopped: " + sv->loopErr);
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_server_read_image(rtx_server *sv, uint64_t id, float *rgba, size_t floatCount) try {
    if (!sv) return RTX_E_INVALID;
    if (!rgba) return setErr(sv, RTX_E_INVALID, "NULL argument");
    Job *j = nullptr;
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        auto it = sv->jobs.find(id);
        if (it == sv->jobs.end()) return setErr(sv, RTX_E_INVALID, "unknown job");
        j = it->second.get();
        if (j->state != RTX_JOB_DONE) return setErr(sv, RTX_E_STATE, "the job's image is not complete (or was read)");
        if (floatCount < j->npix * 4) return setErr(sv, RTX_E_INVALID, "buffer smaller than width * height * 4");
    }
    hipError_t e = hipSetDevice(sv->device);
    if (e == hipSuccess) e = hipMemcpy(rgba, j->image, j->npix * sizeof(float4), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return setErr(sv, RTX_E_DEVICE, std::string("hipMemcpy: ") + hipGetErrorString(e));
    std::lock_guard<std::mutex> lock(sv->mu);
    sv->arena.release(j->fbOffset, j->fbBytes);
    j->state = RTX_JOB_RELEASED;
    j->accum = j->image = nullptr;
    (void)hipEventDestroy(j->doneEv);
    j->doneEv = nullptr;
    sv->jobs.erase(id);
    sv->cvWork.notify_all(); // queued jobs may fit now
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

int rtx_server_get_stats(rtx_server *sv, rtx_server_stats *st) try {
    if (!sv) return RTX_E_INVALID;
    if (!st) return setErr(sv, RTX_E_INVALID, "NULL argument");
    std::lock_guard<std::mutex> lock(sv->mu);
    std::memset(st, 0, sizeof(*st));
    st->jobsDone = sv->latencies.size();
    st->pathsDone = sv->pathsDone;
    if (!sv->latencies.empty()) {
        std::vector<double> l = sv->latencies;
        std::sort(l.begin(), l.end());
        auto pct = [&](double p) { // nearest rank
            size_t k = (size_t)std::ceil(p * (double)l.size());
            return l[std::min(l.size(), std::max<size_t>(k, 1)) - 1];
        };
        st->latencyP50Ms = pct(0.50);
        st->latencyP99Ms = pct(0.99);
        st->latencyMaxMs = l.back();
        double sum = 0;
        for (double x : l) sum += x;
        st->latencyMeanMs = sum / (double)l.size();
        const double span = msBetween(sv->firstSubmit, sv->lastDone);
        st->msamplesPerSec = span > 0 ? (double)sv->pathsDone / (span * 1e3) : 0.0;
    }
    st->bounces = sv->bounces;
    st->extendRays = sv->extendRays;
    st->shadowRays = sv->shadowRays;
    st->probeRays = sv->probeRays;
    st->peakSlotsBusy = sv->peakSlots;
    st->peakJobsRunning = sv->peakRunning;
    st->framebufferPeakBytes = sv->arena.peak();
    st->poolBytes = sv->poolBytes;
    return RTX_OK;
} catch (...) {
    return rtx::abiStatus();
}

const char *rtx_server_last_error(const rtx_server *sv) {
    if (!sv) return "NULL server";
    return sv->err.c_str();
}

} // extern "C"
