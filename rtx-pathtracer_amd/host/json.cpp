// Recursive-descent JSON reader (see json.h).
#include "json.h"

#include <cstdlib>
#include <cstring>

namespace rtxh {

namespace {

struct Parser {
    const std::string &t;
    size_t p = 0;
    std::string err;

    explicit Parser(const std::string &text) : t(text) {}

    bool fail(const std::string &m) {
        if (err.empty()) err = "JSON offset " + std::to_string(p) + ": " + m;
        return false;
    }
    void ws() {
        while (p < t.size() && (t[p] == ' ' || t[p] == '\t' || t[p] == '\n' || t[p] == '\r')) ++p;
    }
    bool lit(const char *s) {
        size_t n = std::strlen(s);
        if (t.compare(p, n, s) != 0) return false;
        p += n;
        return true;
    }
    static void utf8(std::string &o, unsigned cp) {
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
    bool hex4(unsigned &v) {
        if (p + 4 > t.size()) return fail("short \\u escape");
        v = 0;
        for (int i = 0; i < 4; ++i) {
            char c = t[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (unsigned)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (unsigned)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (unsigned)(c - 'A' + 10);
            else return fail("bad \\u escape");
        }
        return true;
    }
    bool str(std::string &o) {
        if (p >= t.size() || t[p] != '"') return fail("expected a string");
        ++p;
        o.clear();
        for (;;) {
            if (p >= t.size()) return fail("unterminated string");
            char c = t[p++];
            if (c == '"') return true;
            if ((unsigned char)c < 0x20) return fail("control character in string");
            if (c != '\\') {
                o += c;
                continue;
            }
            if (p >= t.size()) return fail("unterminated escape");
            char e = t[p++];
            switch (e) {
            case '"': o += '"'; break;
            case '\\': o += '\\'; break;
            case '/': o += '/'; break;
            case 'b': o += '\b'; break;
            case 'f': o += '\f'; break;
            case 'n': o += '\n'; break;
            case 'r': o += '\r'; break;
            case 't': o += '\t'; break;
            case 'u': {
                unsigned cp;
                if (!hex4(cp)) return false;
                if (cp >= 0xD800 && cp < 0xDC00) { // surrogate pair
                    unsigned lo;
                    if (!lit("\\u") || !hex4(lo) || lo < 0xDC00 || lo >= 0xE000) return fail("bad surrogate pair");
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                } else if (cp >= 0xDC00 && cp < 0xE000) {
                    return fail("lone low surrogate");
                }
                utf8(o, cp);
                break;
            }
            default: return fail("bad escape");
            }
        }
    }
    bool number(double &v) {
        size_t s = p;
        if (p < t.size() && t[p] == '-') ++p;
        if (p >= t.size()) return fail("expected a number");
        if (t[p] == '0') {
            ++p;
        } else if (t[p] >= '1' && t[p] <= '9') {
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        } else {
            return fail("expected a value");
        }
        if (p < t.size() && t[p] == '.') {
            ++p;
            if (p >= t.size() || t[p] < '0' || t[p] > '9') return fail("expected a digit after '.'");
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        if (p < t.size() && (t[p] == 'e' || t[p] == 'E')) {
            ++p;
            if (p < t.size() && (t[p] == '+' || t[p] == '-')) ++p;
            if (p >= t.size() || t[p] < '0' || t[p] > '9') return fail("expected an exponent");
            while (p < t.size() && t[p] >= '0' && t[p] <= '9') ++p;
        }
        v = std::strtod(t.substr(s, p - s).c_str(), nullptr);
        return true;
    }
    bool value(JsonValue &v, int depth) {
        if (depth > 512) return fail("nested too deeply");
        ws();
        if (p >= t.size()) return fail("unexpected end");
        char c = t[p];
        if (c == '{') {
            ++p;
            v.type = JsonValue::Object;
            ws();
            if (p < t.size() && t[p] == '}') {
                ++p;
                return true;
            }
            for (;;) {
                ws();
                std::string key;
                if (!str(key)) return false;
                ws();
                if (p >= t.size() || t[p] != ':') return fail("expected ':'");
                ++p;
                JsonValue child;
                if (!value(child, depth + 1)) return false;
                v.object[key] = std::move(child);
                ws();
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < t.size() && t[p] == '}') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            ++p;
            v.type = JsonValue::Array;
            ws();
            if (p < t.size() && t[p] == ']') {
                ++p;
                return true;
            }
            for (;;) {
                v.array.emplace_back();
                if (!value(v.array.back(), depth + 1)) return false;
                ws();
                if (p < t.size() && t[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < t.size() && t[p] == ']') {
                    ++p;
                    return true;
                }
                return fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.type = JsonValue::String;
            return str(v.string);
        }
        if (lit("true")) {
            v.type = JsonValue::Bool;
            v.boolean = true;
            return true;
        }
        if (lit("false")) {
            v.type = JsonValue::Bool;
            return true;
        }
        if (lit("null")) return true;
        v.type = JsonValue::Number;
        return number(v.number);
    }
};

} // namespace

bool parseJson(const std::string &text, JsonValue &out, std::string &err) {
    Parser ps(text);
    if (text.compare(0, 3, "\xEF\xBB\xBF") == 0) ps.p = 3;
    out = JsonValue();
    if (!ps.value(out, 0)) {
        err = ps.err;
        return false;
    }
    ps.ws();
    if (ps.p != text.size()) {
        err = "JSON: content after the value";
        return false;
    }
    return true;
}

} // namespace rtxh
