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
// Handler of every entry point's function-try-block: no exception crosses
// the C ABI.  std::bad_alloc becomes RTX_E_MEMORY, anything else
// RTX_E_INVALID (the host layer reports every failure as invalid input).
int abiFail() noexcept {
    try {
        throw;
    } catch (const std::bad_alloc &) {
        try {
            g_error = "out of memory";
        } catch (...) {
        }
        return RTX_E_MEMORY;
    } catch (const std::exception &e) {
        try {
            g_error = e.what();
        } catch (...) {
        }
    } catch (...) {
        try {
            g_error = "internal error";
        } catch (...) {
        }
    }
    return RTX_E_INVALID;
}
} // namespace

extern "C" {

const char *rtxh_last_error(void) { return g_error.c_str(); }

int rtxh_scene_generate(int kind, uint64_t seed, uint32_t triangles, rtxh_scene **out) try {
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
        case RTXH_SCENE_SPOT: ok = rtxh::generateSpot(seed, s->scene, err); break;
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
} catch (...) {
    return abiFail();
}

int rtxh_scene_load_xml(const char *path, rtxh_scene **out) try {
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
} catch (...) {
    return abiFail();
}

int rtxh_scene_load(const char *path, rtxh_scene **out) try {
    if (!path || !out) return fail("NULL argument");
    *out = nullptr;
    auto *s = new (std::nothrow) rtxh_scene();
    if (!s) return fail("out of memory");
    std::string err;
    bool ok = false;
    try {
        ok = rtxh::loadSceneFile(path, s->scene, err);
    } catch (const std::exception &e) {
        err = e.what();
    }
    if (!ok) {
        delete s;
        return fail(err);
    }
    *out = s;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

void rtxh_scene_destroy(rtxh_scene *scene) { delete scene; }

int rtxh_scene_new(rtxh_scene **out) try {
    if (!out) return fail("out is NULL");
    *out = new (std::nothrow) rtxh_scene();
    return *out ? RTX_OK : fail("out of memory");
} catch (...) {
    return abiFail();
}

int rtxh_scene_add_material(rtxh_scene *scene, const rtx_material *m, int32_t *index) try {
    if (!scene || !m) return fail("NULL argument");
    int i = scene->scene.addMaterial(*m);
    if (index) *index = i;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_add_mesh(rtxh_scene *scene, const rtx_vertex *v, uint32_t nv, const uint32_t *idx, uint32_t ni,
                        const float *transform, const float *normalTransform, int32_t *instance) try {
    if (!scene || (nv && !v) || (ni && !idx)) return fail("NULL argument");
    if (ni % 3) return fail("index count must be a multiple of 3");
    rtxh::Model m;
    m.vertices.assign(v, v + nv);
    m.indices.assign(idx, idx + ni);
    for (uint32_t i = 0; i < ni; ++i)
        if (idx[i] >= nv) return fail("vertex index out of range");
    for (uint32_t i = 0; i < nv; ++i)
        if (v[i].materialIndex < 0 || v[i].materialIndex >= (int)scene->scene.materials.size())
            return fail("vertex material index out of range");
    int inst = scene->scene.addModelInstance(std::move(m), transform, normalTransform);
    if (instance) *instance = inst;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_add_sphere(rtxh_scene *scene, const float c[3], float radius, int32_t material, int32_t *index) try {
    if (!scene || !c) return fail("NULL argument");
    if (material < 0 || material >= (int)scene->scene.materials.size()) return fail("material index out of range");
    int i = scene->scene.addSphere(rtxh::Vec3(c[0], c[1], c[2]), radius, material);
    if (index) *index = i;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_add_point_light(rtxh_scene *scene, const float p[3], const float col[3]) try {
    if (!scene || !p || !col) return fail("NULL argument");
    scene->scene.addPointLight(rtxh::Vec3(p[0], p[1], p[2]), rtxh::Vec3(col[0], col[1], col[2]));
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_set_view(rtxh_scene *scene, const rtxh_view *view) try {
    if (!scene || !view) return fail("NULL argument");
    scene->scene.view = *view;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_add_texture(rtxh_scene *scene, const uint8_t *rgba, uint32_t width, uint32_t height, int32_t *slot) try {
    if (!scene || !rgba) return fail("NULL argument");
    if (width == 0 || height == 0 || (uint64_t)width * height >= (1ull << 31)) return fail("bad texture size");
    rtxh::Texture t;
    t.width = width;
    t.height = height;
    t.rgba.assign(rgba, rgba + (size_t)width * height * 4);
    int s = scene->scene.addTexture(std::move(t), std::string());
    if (slot) *slot = s;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_set_environment(rtxh_scene *scene, const float *rgba, uint32_t width, uint32_t height, int addLight) try {
    if (!scene || !rgba) return fail("NULL argument");
    if (width == 0 || height == 0) return fail("bad environment map size");
    scene->scene.envMap.assign(rgba, rgba + (size_t)width * height * 4);
    scene->scene.envWidth = width;
    scene->scene.envHeight = height;
    if (addLight) {
        rtx_light light;
        std::memset(&light, 0, sizeof(light));
        light.type = RTX_LIGHT_ENV_MAP;
        scene->scene.lights.push_back(light);
    }
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_finalize(rtxh_scene *scene) try {
    if (!scene) return fail("NULL argument");
    scene->scene.buildLightTables();
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_desc(const rtxh_scene *scene, rtx_scene_desc *desc) try {
    if (!scene || !desc) return fail("NULL argument");
    const_cast<rtxh_scene *>(scene)->scene.fillDesc(desc);
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_scene_view(const rtxh_scene *scene, rtxh_view *view) try {
    if (!scene || !view) return fail("NULL argument");
    *view = scene->scene.view;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_camera(const rtxh_view *view, uint32_t width, uint32_t height, rtx_camera *cam) try {
    if (!view || !cam) return fail("NULL argument");
    if (!rtxh::cameraMatrices(*view, width, height, cam)) return fail("degenerate camera");
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height) try {
    if (!path || !rgba) return fail("NULL argument");
    std::string err;
    if (!rtxh::writeExr(path, rgba, width, height, err)) return fail(err);
    return RTX_OK;
} catch (...) {
    return abiFail();
}

int rtxh_read_exr(const char *path, float **rgba, uint32_t *width, uint32_t *height) try {
    if (!path || !rgba || !width || !height) return fail("NULL argument");
    std::vector<float> data;
    std::string err;
    if (!rtxh::readExr(path, data, *width, *height, err)) return fail(err);
    *rgba = (float *)std::malloc(data.size() * sizeof(float));
    if (!*rgba) return fail("out of memory");
    std::memcpy(*rgba, data.data(), data.size() * sizeof(float));
    return RTX_OK;
} catch (...) {
    return abiFail();
}

void rtxh_free(void *p) { std::free(p); }

int rtxh_decode_image(const char *path, uint8_t **rgba, uint32_t *width, uint32_t *height) try {
    if (!path || !rgba || !width || !height) return fail("NULL argument");
    rtxh::Texture t;
    std::string err;
    if (!rtxh::loadImageRgba8(path, t, err)) return fail(err);
    *rgba = (uint8_t *)std::malloc(t.rgba.size() ? t.rgba.size() : 1);
    if (!*rgba) return fail("out of memory");
    std::memcpy(*rgba, t.rgba.data(), t.rgba.size());
    *width = t.width;
    *height = t.height;
    return RTX_OK;
} catch (...) {
    return abiFail();
}

uint32_t rtxh_pcg32_first(uint64_t initstate, uint64_t initseq) {
    rtxh::Pcg32 r(initstate, initseq);
    return r.next();
}

} // extern "C"
