// Command line of the path tracer, after src/main.cpp:8-35:
//   rtx_pathtracer WIDTH HEIGHT IC_SIZE GUIDING_SPLITS scenes... [options]
// As in the reference (src/RayTracingApp.cpp:44-46, 346-351), every scene
// gets an irradiance cache of IC_SIZE records and 2^GUIDING_SPLITS guiding
// regions over the scene box; the options switch their use on (the
// reference's UI toggles, off by default as in src/RayTracingApp.h:97-120).
// Scenes are Mitsuba XML, JSON or OBJ files (rtxh_scene_load) or the
// procedural scenes gen:cornell, gen:room[:TRIANGLES], gen:plates and gen:spot.
// Frames run through the reference's per-frame host logic (rtx_schedule_*):
// cache prepare frames, the guiding cadence, progressive accumulation, and
// eval-by-samples, which stops after --spp samples per pixel.  Each scene is
// written to an EXR named after the scene, the sample count and the eval time
// (src/RayTracingApp.cpp:268-286).
// Options: --spp N  --depth N  --follow N  --seed N  --mis 0|1  --nee 0|1
//          --rr 0|1  --ic 0|1  --guiding 0|1  --adrrs 0|1  --prepare N
//          --guiding-updates N  --out DIR
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rtx_amd.h"
#include "rtx_host.h"

static int usage() {
    std::fprintf(stderr,
                 "usage: rtx_pathtracer WIDTH HEIGHT IC_SIZE GUIDING_SPLITS scenes... [--spp N] [--depth N]\n"
                 "       [--follow N] [--seed N] [--mis 0|1] [--nee 0|1] [--rr 0|1] [--ic 0|1] [--guiding 0|1]\n"
                 "       [--adrrs 0|1] [--prepare N] [--guiding-updates N] [--out DIR]\n");
    return 2;
}

static bool parseUint(const char *s, uint32_t &out) {
    char *end = nullptr;
    unsigned long v = std::strtoul(s, &end, 10);
    if (!s[0] || *end || v > 0xFFFFFFFFul) return false;
    out = (uint32_t)v;
    return true;
}

struct Options {
    uint32_t width = 0, height = 0, icSize = 0, guidingSplits = 0;
    uint32_t spp = 16, seed = 1;
    int depth = 8, follow = 3, mis = 1, nee = 1, rr = 0, ic = 0, guiding = 0, adrrs = 0;
    int prepare = 50, guidingUpdates = 6;
    std::string outDir = ".";
    std::vector<std::string> scenes;
};

static bool parseArgs(int argc, char **argv, Options &o) {
    if (argc < 6) return false;
    if (!parseUint(argv[1], o.width) || !parseUint(argv[2], o.height) || !parseUint(argv[3], o.icSize) ||
        !parseUint(argv[4], o.guidingSplits))
        return false;
    for (int i = 5; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](int &v) {
            if (i + 1 >= argc) return false;
            v = std::atoi(argv[++i]);
            return true;
        };
        int v = 0;
        if (a == "--spp" && next(v)) o.spp = (uint32_t)v;
        else if (a == "--depth" && next(v)) o.depth = v;
        else if (a == "--follow" && next(v)) o.follow = v;
        else if (a == "--seed" && next(v)) o.seed = (uint32_t)v;
        else if (a == "--mis" && next(v)) o.mis = v;
        else if (a == "--nee" && next(v)) o.nee = v;
        else if (a == "--rr" && next(v)) o.rr = v;
        else if (a == "--ic" && next(v)) o.ic = v;
        else if (a == "--guiding" && next(v)) o.guiding = v;
        else if (a == "--adrrs" && next(v)) o.adrrs = v;
        else if (a == "--prepare" && next(v)) o.prepare = v;
        else if (a == "--guiding-updates" && next(v)) o.guidingUpdates = v;
        else if (a == "--out" && i + 1 < argc) o.outDir = argv[++i];
        else if (a.rfind("--", 0) == 0) return false;
        else o.scenes.push_back(a);
    }
    if (o.scenes.empty() || o.width == 0 || o.height == 0 || o.spp == 0) return false;
    // the regions are 2^GUIDING_SPLITS boxes (src/RayTracingApp.cpp:45)
    if (o.guidingSplits > 12) return false;
    if ((o.ic || o.adrrs) && o.icSize == 0) return false;
    return true;
}

static int loadScene(const std::string &name, uint32_t seed, rtxh_scene **hs) {
    if (name == "gen:cornell") return rtxh_scene_generate(RTXH_SCENE_CORNELL, seed, 0, hs);
    if (name == "gen:plates") return rtxh_scene_generate(RTXH_SCENE_PLATES, seed, 0, hs);
    if (name == "gen:spot") return rtxh_scene_generate(RTXH_SCENE_SPOT, seed, 0, hs);
    if (name.rfind("gen:room", 0) == 0) {
        uint32_t tris = name.size() > 9 ? (uint32_t)std::strtoul(name.c_str() + 9, nullptr, 10) : 0;
        return rtxh_scene_generate(RTXH_SCENE_ROOM, seed, tris, hs);
    }
    return rtxh_scene_load(name.c_str(), hs);
}

// Renders one scene; returns 0 on success.
static int renderScene(rtx_ctx *ctx, const Options &o, const std::string &name) {
    rtxh_scene *hs = nullptr;
    if (loadScene(name, o.seed, &hs) != RTX_OK) {
        std::fprintf(stderr, "%s: %s\n", name.c_str(), rtxh_last_error());
        return 1;
    }
    rtx_scene_desc desc;
    rtxh_scene_desc(hs, &desc);
    rtxh_view view;
    rtxh_scene_view(hs, &view);
    rtx_scene *sc = nullptr;
    rtx_schedule *sched = nullptr;
    int status = 1;
    auto t0 = std::chrono::steady_clock::now();
    do {
        if (rtx_scene_create(ctx, &desc, &sc) != RTX_OK) break;
        // the per-scene cache and guiding regions (src/RayTracingApp.cpp:346-351)
        if (o.icSize > 0 && rtx_ic_create(ctx, o.icSize) != RTX_OK) break;
        rtx_scene_info info;
        if (rtx_scene_get_info(sc, &info) != RTX_OK) break;
        rtx_aabb box;
        for (int a = 0; a < 3; ++a) {
            box.min[a] = info.boundsMin[a];
            box.max[a] = info.boundsMax[a];
        }
        if (rtx_guiding_init(ctx, &box, o.guidingSplits, 1) != RTX_OK) break;
        auto t1 = std::chrono::steady_clock::now();

        rtx_frame_params base;
        std::memset(&base, 0, sizeof(base));
        rtx_push_constant pc;
        rtx_push_constant_defaults(&pc);
        pc.maxDepth = o.depth;
        pc.maxFollowDiscrete = o.follow;
        pc.enableMIS = o.mis;
        pc.enableNEE = o.nee;
        pc.enableRR = o.rr;
        pc.useIrradianceCache = o.ic;
        pc.useADRRS = o.adrrs;
        pc.useGuiding = o.guiding;
        pc.updateGuiding = o.guiding;
        rtxh_camera(&view, o.width, o.height, &base.cam);
        base.width = o.width;
        base.height = o.height;
        base.seed = o.seed;
        base.rrStartDepth = 3;
        rtx_schedule_config cfg;
        rtx_schedule_config_defaults(&cfg);
        cfg.prepareFrames = o.prepare;
        cfg.guidingOptimizations = o.guidingUpdates;
        cfg.evalSamples = (int32_t)o.spp;
        if (rtx_schedule_create(&cfg, &pc, &sched) != RTX_OK) break;
        rtx_schedule_state state;
        uint64_t rays = 0, paths = 0;
        double deviceMs = 0;
        // eval-by-samples ends after spp samples outside prepare frames; the
        // bound covers the prepare frames and an ADRRS estimate frame
        const uint64_t maxFrames = (uint64_t)o.spp + (uint64_t)o.prepare + 2;
        bool ok = true;
        do {
            rtx_frame_stats st;
            if (rtx_schedule_frame(ctx, sc, sched, &base, &st) != RTX_OK) {
                ok = false;
                break;
            }
            rays += st.extendRays + st.shadowRays + st.probeRays;
            paths += st.paths;
            deviceMs += st.msTotal;
            rtx_schedule_get_state(sched, &state);
        } while (state.evalSamplesRunning && state.frames < maxFrames);
        if (!ok) break;
        uint32_t samples = 0;
        rtx_accumulated_samples(ctx, &samples);
        std::vector<float> img((size_t)o.width * o.height * 4);
        if (rtx_read_image(ctx, img.data(), img.size()) != RTX_OK) break;
        std::string stem = name;
        for (char &c : stem)
            if (c == '/' || c == ':') c = '_';
        char file[1024];
        std::snprintf(file, sizeof(file), "%s/%s_%uspp_%.0fms.exr", o.outDir.c_str(), stem.c_str(), samples,
                      state.evalSamplesMs);
        if (rtxh_write_exr(file, img.data(), o.width, o.height) != RTX_OK) {
            std::fprintf(stderr, "%s\n", rtxh_last_error());
            rtx_schedule_destroy(sched);
            rtx_scene_destroy(ctx, sc);
            rtxh_scene_destroy(hs);
            return 1;
        }
        std::printf("%s: startup %.0f ms, %llu frames, %u spp accumulated, eval %.1f ms, device %.1f ms, "
                    "%.2f Msamples/s, %.1f Mrays/s -> %s\n",
                    name.c_str(), std::chrono::duration<double, std::milli>(t1 - t0).count(),
                    (unsigned long long)state.frames, samples, state.evalSamplesMs, deviceMs,
                    (double)paths / (deviceMs * 1e3), (double)rays / (deviceMs * 1e3), file);
        status = 0;
    } while (false);
    if (status != 0) std::fprintf(stderr, "%s: %s\n", name.c_str(), rtx_last_error(ctx));
    if (sched) rtx_schedule_destroy(sched);
    if (sc) rtx_scene_destroy(ctx, sc);
    rtxh_scene_destroy(hs);
    return status;
}

int main(int argc, char **argv) {
    Options o;
    if (!parseArgs(argc, argv, o)) return usage();
    rtx_ctx *ctx = nullptr;
    if (rtx_ctx_create(0, &ctx) != RTX_OK) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    int status = 0;
    for (const std::string &name : o.scenes) status |= renderScene(ctx, o, name);
    rtx_ctx_destroy(ctx);
    return status;
}
