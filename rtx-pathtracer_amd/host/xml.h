// A small XML reader for scene files: elements, attributes, nesting.  Text
// content, comments, processing instructions and DOCTYPE are skipped; the five
// predefined entities and numeric character references are decoded in
// attribute values.  (The reference reads its Mitsuba scenes with tinyxml2;
// only the element tree is used, src/SceneLoader.cpp:329-648.)
#pragma once

#include <string>
#include <utility>
#include <vector>

namespace rtxh {

struct XmlNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XmlNode> children;

    // The attribute's value, or nullptr.
    const char *attr(const char *key) const;
    // Children named `name` (all children for nullptr), in document order.
    std::vector<const XmlNode *> all(const char *name) const;
    // The first child named `name`, or nullptr.
    const XmlNode *first(const char *name) const;
};

// Parses the document; root receives the document element.
bool parseXml(const std::string &text, XmlNode &root, std::string &err);

} // namespace rtxh
