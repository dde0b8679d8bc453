// Command line of the path tracer, after src/main.cpp:8-35:
//   rtx_pathtracer WIDTH HEIGHT IC_SIZE GUIDING_SPLITS scenes... [options]
// Scenes are Mitsuba XML files or the procedural scenes gen:cornell,
// gen:room[:TRIANGLES] and gen:plates.  Each scene is rendered to an EXR
// named after the scene and the sample count (src/RayTracingApp.cpp:268-286).
// Options: --spp N  --depth N  --follow N  --seed N  --mis 0|1  --nee 0|1
//          --rr 0|1  --out DIR
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
                 "       [--follow N] [--seed N] [--mis 0|1] [--nee 0|1] [--rr 0|1] [--out DIR]\n");
    return 2;
}

int main(int argc, char **argv) {
    if (argc < 6) return usage();
    uint32_t width = (uint32_t)std::strtoul(argv[1], nullptr, 10);
    uint32_t height = (uint32_t)std::strtoul(argv[2], nullptr, 10);
    uint32_t icSize = (uint32_t)std::strtoul(argv[3], nullptr, 10);
    uint32_t guidingSplits = (uint32_t)std::strtoul(argv[4], nullptr, 10);
    (void)icSize;
    (void)guidingSplits;
    std::vector<std::string> scenes;
    uint32_t spp = 16, seed = 1;
    int depth = 8, follow = 3, mis = 1, nee = 1, rr = 0;
    std::string outDir = ".";
    for (int i = 5; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](int &v) {
            if (i + 1 >= argc) return false;
            v = std::atoi(argv[++i]);
            return true;
        };
        int v = 0;
        if (a == "--spp" && next(v)) spp = (uint32_t)v;
        else if (a == "--depth" && next(v)) depth = v;
        else if (a == "--follow" && next(v)) follow = v;
        else if (a == "--seed" && next(v)) seed = (uint32_t)v;
        else if (a == "--mis" && next(v)) mis = v;
        else if (a == "--nee" && next(v)) nee = v;
        else if (a == "--rr" && next(v)) rr = v;
        else if (a == "--out" && i + 1 < argc) outDir = argv[++i];
        else if (a.rfind("--", 0) == 0) return usage();
        else scenes.push_back(a);
    }
    if (scenes.empty() || width == 0 || height == 0) return usage();

    rtx_ctx *ctx = nullptr;
    if (rtx_ctx_create(0, &ctx) != RTX_OK) {
        std::fprintf(stderr, "no HIP device\n");
        return 1;
    }
    int status = 0;
    for (const std::string &name : scenes) {
        rtxh_scene *hs = nullptr;
        int rc;
        if (name == "gen:cornell") rc = rtxh_scene_generate(RTXH_SCENE_CORNELL, seed, 0, &hs);
        else if (name == "gen:plates") rc = rtxh_scene_generate(RTXH_SCENE_PLATES, seed, 0, &hs);
        else if (name.rfind("gen:room", 0) == 0) {
            uint32_t tris = name.size() > 9 ? (uint32_t)std::strtoul(name.c_str() + 9, nullptr, 10) : 0;
            rc = rtxh_scene_generate(RTXH_SCENE_ROOM, seed, tris, &hs);
        } else rc = rtxh_scene_load_xml(name.c_str(), &hs);
        if (rc != RTX_OK) {
            std::fprintf(stderr, "%s: %s\n", name.c_str(), rtxh_last_error());
            status = 1;
            continue;
        }
        rtx_scene_desc desc;
        rtxh_scene_desc(hs, &desc);
        rtxh_view view;
        rtxh_scene_view(hs, &view);
        rtx_scene *sc = nullptr;
        auto t0 = std::chrono::steady_clock::now();
        if (rtx_scene_create(ctx, &desc, &sc) != RTX_OK) {
            std::fprintf(stderr, "%s: %s\n", name.c_str(), rtx_last_error(ctx));
            rtxh_scene_destroy(hs);
            status = 1;
            continue;
        }
        auto t1 = std::chrono::steady_clock::now();
        rtx_frame_params fp;
        std::memset(&fp, 0, sizeof(fp));
        rtx_push_constant_defaults(&fp.pc);
        fp.pc.maxDepth = depth;
        fp.pc.maxFollowDiscrete = follow;
        fp.pc.enableMIS = mis;
        fp.pc.enableNEE = nee;
        fp.pc.enableRR = rr;
        rtxh_camera(&view, width, height, &fp.cam);
        fp.width = width;
        fp.height = height;
        fp.spp = spp;
        fp.seed = seed;
        fp.flags = RTX_FRAME_RESET_ACCUM;
        fp.rrStartDepth = 3;
        rtx_frame_stats st;
        if (rtx_render(ctx, sc, &fp, &st) != RTX_OK) {
            std::fprintf(stderr, "%s: %s\n", name.c_str(), rtx_last_error(ctx));
            status = 1;
        } else {
            std::vector<float> img((size_t)width * height * 4);
            rtx_read_image(ctx, img.data(), img.size());
            std::string base = name;
            for (char &c : base)
                if (c == '/' || c == ':') c = '_';
            char file[512];
            std::snprintf(file, sizeof(file), "%s/%s_%uspp_%.0fms.exr", outDir.c_str(), base.c_str(), spp, st.msTotal);
            if (rtxh_write_exr(file, img.data(), width, height) != RTX_OK) {
                std::fprintf(stderr, "%s\n", rtxh_last_error());
                status = 1;
            }
            double msamples = (double)st.paths / (st.msTotal * 1e3);
            std::printf("%s: startup %.0f ms, render %.1f ms, %.2f Msamples/s, %.1f Mrays/s -> %s\n", name.c_str(),
                        std::chrono::duration<double, std::milli>(t1 - t0).count(), st.msTotal, msamples,
                        (double)(st.extendRays + st.shadowRays + st.probeRays) / (st.msTotal * 1e3), file);
        }
        rtx_scene_destroy(ctx, sc);
        rtxh_scene_destroy(hs);
    }
    rtx_ctx_destroy(ctx);
    return status;
}
