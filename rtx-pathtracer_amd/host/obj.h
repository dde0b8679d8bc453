// Wavefront OBJ/MTL reading and the conversion of an OBJ into one model.
//
// The reference reads OBJ files with the vendored tinyobjloader 2.0
// (external/tiny_obj_loader.h, LoadObj with triangulate = true,
// mtl_basedir = the scene's material folder) and turns the result into one
// vertex/index buffer with SceneLoader::converteObjData
// (src/SceneLoader.cpp:245-327).  readObj restates the parts of LoadObj that
// reach that buffer: v/vn/vt records, faces with relative indices, the
// ear-clipping triangulation, usemtl/mtllib and the MTL fields that
// SceneLoader::addMaterials reads (src/SceneLoader.cpp:138-182).
#pragma once

#include <string>
#include <vector>

#include "scene.h"

namespace rtxh {

struct ObjMaterial {
    std::string name;
    float diffuse[3] = {0, 0, 0};
    float specular[3] = {0, 0, 0};
    float emission[3] = {0, 0, 0};
    float shininess = 1.0f;
    float ior = 1.0f;
    int illum = 0;
    std::string diffuseTex;  // map_Kd, as written (options stripped)
    std::string specularTex; // map_Ks
};

// One triangle corner: 0-based indices into v/vt/vn, -1 when absent.
struct ObjCorner {
    int v = -1, vt = -1, vn = -1;
};

struct ObjData {
    std::vector<float> v;  // x y z per position
    std::vector<float> vn; // x y z per normal
    std::vector<float> vt; // u v per texture coordinate
    std::vector<ObjCorner> corners; // 3 per triangle, in file order
    std::vector<int> triMaterial;   // per triangle: index into materials, or -1
    std::vector<ObjMaterial> materials;
    std::string warnings;
};

// tinyobjloader's number reader (tryParseDouble, tiny_obj_loader.h:814-938):
// digits accumulated in double, decimal digits scaled by 10^-k, the decimal
// exponent applied as ldexp(m * 5^e, e).  Returns false for text that is not
// a number (the caller then keeps its default).
bool objParseDouble(const char *s, const char *end, double *out);

// Reads an OBJ file.  mtlDir is prepended to mtllib file names.
bool readObj(const std::string &path, const std::string &mtlDir, ObjData &out, std::string &err);

// Reads one MTL file and appends its materials (LoadMtl,
// tiny_obj_loader.h:1666-2051; the last material is always appended, even
// one without a name).  names receives (name, index) for usemtl look-ups.
bool readMtl(const std::string &path, std::vector<ObjMaterial> &materials,
             std::vector<std::pair<std::string, int>> &names, std::string &err);

// SceneLoader::addMaterials (src/SceneLoader.cpp:138-182): the illum model
// picks the type (0,1 diffuse; 2 phong; 3 specular; 4,7 dielectric;
// 11 light).  textureId receives the map_Kd / map_Ks texture index
// (map_Ks overrides map_Kd, as in the reference) or -1.
bool objMaterialToRtx(const ObjMaterial &in, int textureId, rtx_material &out, std::string &err);

// SceneLoader::converteObjData: one vertex per distinct (position, normal,
// texture coordinate, material) in first-use order, texture v flipped,
// per-face normals when the file has none.  materialOverride >= 0 gives every
// face that material; otherwise a face's material is materialOffset plus its
// usemtl index.  materials is the scene's list so far (to find lights).
bool objToModel(const ObjData &obj, int materialOffset, int materialOverride,
                const std::vector<rtx_material> &materials, Model &model, std::string &err);

} // namespace rtxh
