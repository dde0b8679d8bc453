// Mitsuba-0.6 XML scene loading (SceneLoader::parseMitsubaSceneFile,
// src/SceneLoader.cpp:329-348).
#include "scene.h"

namespace rtxh {

bool loadMitsubaXml(const std::string &path, Scene &s, std::string &err) {
    (void)s;
    err = "XML scene loading is not available in this build: " + path;
    return false;
}

} // namespace rtxh
