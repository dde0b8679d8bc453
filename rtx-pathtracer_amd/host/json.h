// A small JSON reader for scene files (RFC 8259 values: objects, arrays,
// strings with escapes, numbers, true/false/null).  Objects keep their keys
// sorted and the last of duplicate keys, as the nlohmann::json objects the
// reference parses its scenes into do (src/SceneLoader.cpp:617-630).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace rtxh {

struct JsonValue {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool boolean = false;
    double number = 0.0;
    std::string string;
    std::vector<JsonValue> array;
    std::map<std::string, JsonValue> object;

    bool has(const std::string &key) const { return type == Object && object.count(key) != 0; }
    const JsonValue *get(const std::string &key) const {
        if (type != Object) return nullptr;
        auto it = object.find(key);
        return it == object.end() ? nullptr : &it->second;
    }
};

bool parseJson(const std::string &text, JsonValue &out, std::string &err);

} // namespace rtxh
