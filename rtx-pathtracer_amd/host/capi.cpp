// C entry points of librtxhost.so (include/rtx_host.h).
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>

#include "scene.h"

namespace {
thread_local std::string g_error;
int fail(const std::string &msg) {
    g_error = msg;
    return RTX_E_INVALID;
}
} // namespace

extern "C" {

const char *rtxh_last_error(void) { return g_error.c_str(); }

int rtxh_scene_generate(int kind, uint64_t seed, uint32_t triangles, rtxh_scene **out) {
    if (!out) return fail("out is NULL");
    *out = nullptr;
    auto *s = new (std::nothrow) rtxh_scene();
    if (!s) return fail("out of memory");
    std::string err;
    bool ok = false;
    try {
        switch (kind) {
        case RTXH_SCENE_CORNELL: ok = rtxh::generateCornell(seed, s->scene, err); break;
        case RTXH_SCENE_ROOM: ok = rtxh::generateRoom(seed, triangles, s->scene, err); break;
        case RTXH_SCENE_PLATES: ok = rtxh::generatePlates(seed, s->scene, err); break;
        default: err = "unknown scene kind " + std::to_string(kind);
        }
    } catch (const std::exception &e) {
        err = e.what();
        ok = false;
    }
    if (!ok) {
        delete s;
        return fail(err);
    }
    *out = s;
    return RTX_OK;
}

int rtxh_scene_load_xml(const char *path, rtxh_scene **out) {
    if (!path || !out) return fail("NULL argument");
    *out = nullptr;
    auto *s = new (std::nothrow) rtxh_scene();
    if (!s) return fail("out of memory");
    std::string err;
    bool ok = false;
    try {
        ok = rtxh::loadMitsubaXml(path, s->scene, err);
    } catch (const std::exception &e) {
        err = e.what();
    }
    if (!ok) {
        delete s;
        return fail(err);
    }
    *out = s;
    return RTX_OK;
}

void rtxh_scene_destroy(rtxh_scene *scene) { delete scene; }

int rtxh_scene_desc(const rtxh_scene *scene, rtx_scene_desc *desc) {
    if (!scene || !desc) return fail("NULL argument");
    const_cast<rtxh_scene *>(scene)->scene.fillDesc(desc);
    return RTX_OK;
}

int rtxh_scene_view(const rtxh_scene *scene, rtxh_view *view) {
    if (!scene || !view) return fail("NULL argument");
    *view = scene->scene.view;
    return RTX_OK;
}

int rtxh_camera(const rtxh_view *view, uint32_t width, uint32_t height, rtx_camera *cam) {
    if (!view || !cam) return fail("NULL argument");
    if (!rtxh::cameraMatrices(*view, width, height, cam)) return fail("degenerate camera");
    return RTX_OK;
}

int rtxh_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height) {
    if (!path || !rgba) return fail("NULL argument");
    std::string err;
    if (!rtxh::writeExr(path, rgba, width, height, err)) return fail(err);
    return RTX_OK;
}

int rtxh_read_exr(const char *path, float **rgba, uint32_t *width, uint32_t *height) {
    if (!path || !rgba || !width || !height) return fail("NULL argument");
    std::vector<float> data;
    std::string err;
    if (!rtxh::readExr(path, data, *width, *height, err)) return fail(err);
    *rgba = (float *)std::malloc(data.size() * sizeof(float));
    if (!*rgba) return fail("out of memory");
    std::memcpy(*rgba, data.data(), data.size() * sizeof(float));
    return RTX_OK;
}

void rtxh_free(void *p) { std::free(p); }

uint32_t rtxh_pcg32_first(uint64_t initstate, uint64_t initseq) {
    rtxh::Pcg32 r(initstate, initseq);
    return r.next();
}

} // extern "C"
