// Recursive-descent XML reader (see xml.h).
#include "xml.h"

#include <cstdlib>
#include <cstring>

namespace rtxh {

const char *XmlNode::attr(const char *key) const {
    for (const auto &a : attrs)
        if (a.first == key) return a.second.c_str();
    return nullptr;
}

std::vector<const XmlNode *> XmlNode::all(const char *n) const {
    std::vector<const XmlNode *> out;
    for (const auto &c : children)
        if (!n || c.name == n) out.push_back(&c);
    return out;
}

const XmlNode *XmlNode::first(const char *n) const {
    for (const auto &c : children)
        if (!n || c.name == n) return &c;
    return nullptr;
}

namespace {

struct Reader {
    const std::string &t;
    size_t p = 0;
    std::string err;

    explicit Reader(const std::string &text) : t(text) {}

    bool fail(const std::string &m) {
        if (err.empty()) {
            size_t line = 1;
            for (size_t i = 0; i < p && i < t.size(); ++i) line += t[i] == '\n';
            err = "XML line " + std::to_string(line) + ": " + m;
        }
        return false;
    }
    bool eof() const { return p >= t.size(); }
    bool starts(const char *s) const { return t.compare(p, std::strlen(s), s) == 0; }
    void skipSpace() {
        while (!eof() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) ++p;
    }
    bool skipPast(const char *end) {
        size_t q = t.find(end, p);
        if (q == std::string::npos) return fail(std::string("missing ") + end);
        p = q + std::strlen(end);
        return true;
    }
    // Comments, processing instructions and DOCTYPE between elements.
    bool skipMisc() {
        for (;;) {
            skipSpace();
            if (starts("<!--")) {
                if (!skipPast("-->")) return false;
            } else if (starts("<?")) {
                if (!skipPast("?>")) return false;
            } else if (starts("<!DOCTYPE") || starts("<!doctype")) {
                if (!skipPast(">")) return false;
            } else {
                return true;
            }
        }
    }
    static bool nameChar(char c) {
        return (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') || c == '_' || c == '-' ||
               c == ':' || c == '.';
    }
    bool readName(std::string &out) {
        size_t s = p;
        while (!eof() && nameChar(t[p])) ++p;
        if (p == s) return fail("expected a name");
        out.assign(t, s, p - s);
        return true;
    }
    static void appendUtf8(std::string &o, unsigned long cp) {
        if (cp < 0x80) {
            o += (char)cp;
        } else if (cp < 0x800) {
            o += (char)(0xC0 | (cp >> 6));
            o += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            o += (char)(0xE0 | (cp >> 12));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        } else {
            o += (char)(0xF0 | (cp >> 18));
            o += (char)(0x80 | ((cp >> 12) & 0x3F));
            o += (char)(0x80 | ((cp >> 6) & 0x3F));
            o += (char)(0x80 | (cp & 0x3F));
        }
    }
    static std::string decode(const std::string &v) {
        std::string o;
        for (size_t i = 0; i < v.size(); ++i) {
            if (v[i] != '&') {
                o += v[i];
                continue;
            }
            size_t semi = v.find(';', i);
            if (semi == std::string::npos) {
                o += v[i];
                continue;
            }
            std::string ent = v.substr(i + 1, semi - i - 1);
            if (ent == "lt") o += '<';
            else if (ent == "gt") o += '>';
            else if (ent == "amp") o += '&';
            else if (ent == "quot") o += '"';
            else if (ent == "apos") o += '\'';
            else if (!ent.empty() && ent[0] == '#')
                appendUtf8(o, std::strtoul(ent.c_str() + (ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? 2 : 1),
                                           nullptr, ent.size() > 1 && (ent[1] == 'x' || ent[1] == 'X') ? 16 : 10));
            else {
                o += v.substr(i, semi - i + 1);
            }
            i = semi;
        }
        return o;
    }
    bool element(XmlNode &n, int depth) {
        if (depth > 256) return fail("elements nested too deeply");
        if (eof() || t[p] != '<') return fail("expected '<'");
        ++p;
        if (!readName(n.name)) return false;
        for (;;) {
            skipSpace();
            if (eof()) return fail("unterminated tag <" + n.name);
            if (starts("/>")) {
                p += 2;
                return true;
            }
            if (t[p] == '>') {
                ++p;
                break;
            }
            std::string key;
            if (!readName(key)) return false;
            skipSpace();
            if (eof() || t[p] != '=') return fail("expected '=' after attribute " + key);
            ++p;
            skipSpace();
            if (eof() || (t[p] != '"' && t[p] != '\'')) return fail("expected a quoted value for " + key);
            const char q = t[p++];
            size_t e = t.find(q, p);
            if (e == std::string::npos) return fail("unterminated value of " + key);
            n.attrs.emplace_back(key, decode(t.substr(p, e - p)));
            p = e + 1;
        }
        // content: children, text (skipped), comments, CDATA
        for (;;) {
            if (eof()) return fail("missing </" + n.name + ">");
            if (starts("</")) {
                p += 2;
                std::string close;
                if (!readName(close)) return false;
                if (close != n.name) return fail("</" + close + "> closes <" + n.name + ">");
                skipSpace();
                if (eof() || t[p] != '>') return fail("expected '>'");
                ++p;
                return true;
            }
            if (starts("<!--")) {
                if (!skipPast("-->")) return false;
            } else if (starts("<![CDATA[")) {
                if (!skipPast("]]>")) return false;
            } else if (starts("<?")) {
                if (!skipPast("?>")) return false;
            } else if (t[p] == '<') {
                n.children.emplace_back();
                if (!element(n.children.back(), depth + 1)) return false;
            } else {
                ++p; // text
            }
        }
    }
};

} // namespace

bool parseXml(const std::string &text, XmlNode &root, std::string &err) {
    Reader r(text);
    if (r.starts("\xEF\xBB\xBF")) r.p = 3; // UTF-8 byte-order mark
    if (!r.skipMisc() || !r.element(root, 0)) {
        err = r.err;
        return false;
    }
    if (!r.skipMisc()) {
        err = r.err;
        return false;
    }
    if (!r.eof()) {
        err = "content after the document element";
        return false;
    }
    return true;
}

} // namespace rtxh
