/*
 * rtx_host.h - host side of the path tracer (librtxhost.so).
 *
 * This is the C++ host that stays on the CPU in front of librtxamd.so:
 * the scene layer (SceneLoader / Model / Shapes / WeightedSampler of the
 * reference), camera matrices (CameraController) and EXR output
 * (CommonOps::writeEXR).  It produces rtx_scene_desc / rtx_camera values
 * whose bytes are laid out exactly as the reference's device buffers.
 *
 * The seeded procedural scenes of the benchmark workloads live here too,
 * since the reference only loads scenes from files.
 */
#ifndef RTX_HOST_H
#define RTX_HOST_H

#include <stdint.h>
#include "rtx_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rtxh_scene rtxh_scene;

/* Procedural scenes. */
#define RTXH_SCENE_CORNELL 1 /* 5 walls + 2 boxes (34 tris) + quad light + dielectric sphere */
#define RTXH_SCENE_ROOM 2    /* 20x10x20 m room, exactly `triangles` triangles, 4 area lights */
#define RTXH_SCENE_PLATES 3  /* plates-and-sphere-lights test (veachMIS-like) */
#define RTXH_SCENE_SPOT 4    /* closed diffuse room lit only through a spot on its ceiling (a
                                lamp on a black box under it): the case path guiding is for */

/* Camera placement of a scene (SceneLoader origin/target/upDir/vfov,
 * src/SceneLoader.h:147-150). */
typedef struct rtxh_view {
    float origin[3];
    float target[3];
    float up[3];
    float vfov; /* degrees */
} rtxh_view;

/* Build procedural scene `kind` from PCG32 stream `seed`.  `triangles` is
 * the exact triangle count for RTXH_SCENE_ROOM (0 = 262144) and ignored
 * otherwise. */
int rtxh_scene_generate(int kind, uint64_t seed, uint32_t triangles, rtxh_scene **out);

/* Load a Mitsuba-0.6 XML scene (subset read by SceneLoader::parseMitsubaSceneFile,
 * src/SceneLoader.cpp:329-348) with its OBJ shapes. */
int rtxh_scene_load_xml(const char *path, rtxh_scene **out);

/* Load a scene file the way the reference's SceneLoader does
 * (src/SceneLoader.cpp:29-63): ".xml" is read as Mitsuba 0.6 with its OBJ,
 * MTL, texture and EXR environment-map files beside it; ".json" as the
 * reference's own format ({"models": [{name: obj}], "instances":
 * [{name: {translate, rotate, scale}}], "lights": [{position, color}],
 * "camera": {origin, target, up}}) with models/, materials/ and textures/
 * below the scene's folder.  Light tables are built. */
int rtxh_scene_load(const char *path, rtxh_scene **out);

void rtxh_scene_destroy(rtxh_scene *scene);

/* Programmatic scene building, in the order SceneLoader adds objects while
 * parsing (src/SceneLoader.cpp:405-514, 700-744): a mesh becomes a model
 * plus an instance and, if any face uses a light material, an area light;
 * a sphere with a light material becomes a sphere light; point lights are
 * appended (the reference's JSON loader adds them last, after all area
 * lights, src/SceneLoader.cpp:686-702).  rtxh_scene_finalize builds the
 * light-sampling tables.  transform / normalTransform may be NULL
 * (identity). */
int rtxh_scene_new(rtxh_scene **out);
int rtxh_scene_add_material(rtxh_scene *scene, const rtx_material *material, int32_t *index);
int rtxh_scene_add_mesh(rtxh_scene *scene, const rtx_vertex *vertices, uint32_t vertexCount,
                        const uint32_t *indices, uint32_t indexCount, const float *transform,
                        const float *normalTransform, int32_t *instance);
int rtxh_scene_add_sphere(rtxh_scene *scene, const float center[3], float radius, int32_t material,
                          int32_t *index);
int rtxh_scene_add_point_light(rtxh_scene *scene, const float pos[3], const float color[3]);
int rtxh_scene_set_view(rtxh_scene *scene, const rtxh_view *view);
/* An 8-bit RGBA texture (rows top to bottom, RGB sRGB-encoded) for
 * rtx_material.textureIdDiffuse / textureIdSpecular; *slot receives its id
 * (1, 2, ...; slot 0 is the environment map's). */
int rtxh_scene_add_texture(rtxh_scene *scene, const uint8_t *rgba, uint32_t width, uint32_t height, int32_t *slot);
/* The lat-long environment map (RGBA32F, width*height texels, rows from
 * +y down) seen by rays that miss; addLight != 0 also appends the
 * environment light of SceneLoader::parseEnvMap (src/SceneLoader.cpp:370-403),
 * which must come after all area lights. */
int rtxh_scene_set_environment(rtxh_scene *scene, const float *rgba, uint32_t width, uint32_t height, int addLight);
int rtxh_scene_finalize(rtxh_scene *scene);

/* Pointers stay valid while the scene lives. */
int rtxh_scene_desc(const rtxh_scene *scene, rtx_scene_desc *desc);
int rtxh_scene_view(const rtxh_scene *scene, rtxh_view *view);

/* View/projection matrices and their inverses as RayTracingApp::
 * updateUniformBuffer computes them (src/RayTracingApp.cpp:559-585,
 * src/CameraController.cpp:76-107). */
int rtxh_camera(const rtxh_view *view, uint32_t width, uint32_t height, rtx_camera *cam);

/* Write width*height RGBA32F pixels as an uncompressed RGB FLOAT OpenEXR
 * file (CommonOps::writeEXR, src/CommonOps.cpp:12-38). */
int rtxh_write_exr(const char *path, const float *rgba, uint32_t width, uint32_t height);

/* Read an uncompressed or ZIP-compressed scanline EXR with HALF or FLOAT
 * R,G,B[,A] channels into RGBA32F (CommonOps::readEXR, src/CommonOps.cpp:40-65).
 * *rgba is malloc'ed; free with rtxh_free. */
int rtxh_read_exr(const char *path, float **rgba, uint32_t *width, uint32_t *height);
void rtxh_free(void *p);

/* Decode a PNG or baseline JPEG file to width*height 8-bit RGBA texels
 * (stbi_load(path, ..., 4) of the reference, src/SceneLoader.cpp:199).
 * *rgba is malloc'ed; free with rtxh_free. */
int rtxh_decode_image(const char *path, uint8_t **rgba, uint32_t *width, uint32_t *height);

/* PCG32 (O'Neill 2014): first output after seeding with (initstate, initseq).
 * This is how rtx_render derives the per-sample frame seed. */
uint32_t rtxh_pcg32_first(uint64_t initstate, uint64_t initseq);

const char *rtxh_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* RTX_HOST_H */
