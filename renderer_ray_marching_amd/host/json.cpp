// Recursive-descent JSON parser (RFC 8259 grammar) for the matrices file.
#include <cstdlib>
#include <stdexcept>

#include "rm_io.h"

namespace rmh {

namespace {

struct Parser {
    const std::string& s;
    size_t i = 0;
    int depth = 0;

    explicit Parser(const std::string& text) : s(text) {}

    [[noreturn]] void fail(const std::string& what) const {
        throw std::runtime_error("JSON parse error at byte " + std::to_string(i) + ": " + what);
    }
    void ws() {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\n' || s[i] == '\r')) ++i;
    }
    bool eat(const char* lit) {
        size_t k = 0;
        while (lit[k]) {
            if (i + k >= s.size() || s[i + k] != lit[k]) return false;
            ++k;
        }
        i += k;
        return true;
    }
    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += (char)cp;
        } else if (cp < 0x800) {
            out += (char)(0xC0 | (cp >> 6));
            out += (char)(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += (char)(0xE0 | (cp >> 12));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        } else {
            out += (char)(0xF0 | (cp >> 18));
            out += (char)(0x80 | ((cp >> 12) & 0x3F));
            out += (char)(0x80 | ((cp >> 6) & 0x3F));
            out += (char)(0x80 | (cp & 0x3F));
        }
    }
    uint32_t hex4() {
        if (i + 4 > s.size()) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            char c = s[i++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("bad hex digit in \\u escape");
        }
        return v;
    }
    std::string string() {
        if (i >= s.size() || s[i] != '"') fail("expected '\"'");
        ++i;
        std::string out;
        for (;;) {
            if (i >= s.size()) fail("unterminated string");
            unsigned char c = (unsigned char)s[i++];
            if (c == '"') break;
            if (c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += (char)c;
                continue;
            }
            if (i >= s.size()) fail("unterminated escape");
            char e = s[i++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = hex4();
                    if (cp >= 0xD800 && cp <= 0xDBFF) {
                        if (!eat("\\u")) fail("unpaired surrogate");
                        uint32_t lo = hex4();
                        if (lo < 0xDC00 || lo > 0xDFFF) fail("bad low surrogate");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp <= 0xDFFF) {
                        fail("unpaired surrogate");
                    }
                    put_utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    double number() {
        const size_t b = i;
        if (i < s.size() && s[i] == '-') ++i;
        if (i >= s.size()) fail("truncated number");
        if (s[i] == '0') {
            ++i;
        } else if (s[i] >= '1' && s[i] <= '9') {
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        } else {
            fail("bad number");
        }
        if (i < s.size() && s[i] == '.') {
            ++i;
            if (i >= s.size() || s[i] < '0' || s[i] > '9') fail("bad fraction");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        }
        if (i < s.size() && (s[i] == 'e' || s[i] == 'E')) {
            ++i;
            if (i < s.size() && (s[i] == '+' || s[i] == '-')) ++i;
            if (i >= s.size() || s[i] < '0' || s[i] > '9') fail("bad exponent");
            while (i < s.size() && s[i] >= '0' && s[i] <= '9') ++i;
        }
        const std::string tok = s.substr(b, i - b);
        return std::strtod(tok.c_str(), nullptr);   // correctly rounded to double
    }
    Json value() {
        if (++depth > 512) fail("nesting deeper than 512");
        ws();
        if (i >= s.size()) fail("unexpected end of input");
        Json j;
        const char c = s[i];
        if (c == '{') {
            ++i;
            j.kind = Json::Object;
            ws();
            if (i < s.size() && s[i] == '}') {
                ++i;
            } else {
                for (;;) {
                    ws();
                    std::string key = string();
                    ws();
                    if (i >= s.size() || s[i] != ':') fail("expected ':'");
                    ++i;
                    j.obj[key] = value();
                    ws();
                    if (i < s.size() && s[i] == ',') { ++i; continue; }
                    if (i < s.size() && s[i] == '}') { ++i; break; }
                    fail("expected ',' or '}'");
                }
            }
        } else if (c == '[') {
            ++i;
            j.kind = Json::Array;
            ws();
            if (i < s.size() && s[i] == ']') {
                ++i;
            } else {
                for (;;) {
                    j.arr.push_back(value());
                    ws();
                    if (i < s.size() && s[i] == ',') { ++i; continue; }
                    if (i < s.size() && s[i] == ']') { ++i; break; }
                    fail("expected ',' or ']'");
                }
            }
        } else if (c == '"') {
            j.kind = Json::String;
            j.str = string();
        } else if (c == 't') {
            if (!eat("true")) fail("bad literal");
            j.kind = Json::Bool;
            j.boolean = true;
        } else if (c == 'f') {
            if (!eat("false")) fail("bad literal");
            j.kind = Json::Bool;
        } else if (c == 'n') {
            if (!eat("null")) fail("bad literal");
        } else {
            j.kind = Json::Number;
            j.number = number();
        }
        --depth;
        return j;
    }
};

}  // namespace

const Json& Json::at(const std::string& key) const {
    if (kind != Object) throw std::runtime_error("JSON: cannot use at(\"" + key + "\") on a non-object");
    auto it = obj.find(key);
    if (it == obj.end()) throw std::runtime_error("JSON: key '" + key + "' not found");
    return it->second;
}

const Json* Json::find(const std::string& key) const {
    if (kind != Object) return nullptr;
    auto it = obj.find(key);
    return it == obj.end() ? nullptr : &it->second;
}

// Parses the first JSON value of `text`; what follows it is not read (the reference reads the
// file with `stream >> json`, which stops after one value as well).
Json parse_json(const std::string& text) {
    Parser p(text);
    return p.value();
}

}  // namespace rmh
