/*
 * lx_json.h -- a small JSON document model for the glTF loader.
 *
 * Parses RFC 8259 text into a tree of values (null, bool, number, string,
 * array, object).  Strings decode \uXXXX escapes (with surrogate pairs) to
 * UTF-8.  Errors throw LxError with LX_ERR_FORMAT and the byte offset.
 */
#ifndef LX_JSON_H
#define LX_JSON_H

#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "lx_internal.h"

struct LxJson {
    enum Kind { NUL, BOOL, NUM, STR, ARR, OBJ } kind = NUL;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<LxJson> arr;
    std::vector<std::pair<std::string, LxJson>> obj; /* document order */

    const LxJson* get(const char* key) const {
        if (kind != OBJ) return nullptr;
        for (auto& kv : obj)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    bool has(const char* key) const { return get(key) != nullptr; }
    double number(const char* key, double dflt) const {
        const LxJson* v = get(key);
        return (v && v->kind == NUM) ? v->num : dflt;
    }
    int integer(const char* key, int dflt) const {
        const LxJson* v = get(key);
        return (v && v->kind == NUM) ? (int)v->num : dflt;
    }
    std::string string(const char* key, const std::string& dflt = "") const {
        const LxJson* v = get(key);
        return (v && v->kind == STR) ? v->str : dflt;
    }
    const std::vector<LxJson>& array(const char* key) const {
        static const std::vector<LxJson> empty;
        const LxJson* v = get(key);
        return (v && v->kind == ARR) ? v->arr : empty;
    }
    size_t size() const { return kind == ARR ? arr.size() : (kind == OBJ ? obj.size() : 0); }
};

class LxJsonParser {
  public:
    LxJsonParser(const char* p, size_t n) : s_(p), n_(n) {}
    LxJson parse() {
        LxJson v = value(0);
        ws();
        if (i_ != n_) fail("trailing characters");
        return v;
    }

  private:
    const char* s_;
    size_t n_, i_ = 0;

    [[noreturn]] void fail(const char* what) {
        throw LxError{LX_ERR_FORMAT, std::string("glTF JSON: ") + what + " at byte " + std::to_string(i_)};
    }
    void ws() {
        while (i_ < n_ && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
    }
    bool lit(const char* w) {
        size_t k = 0;
        while (w[k] && i_ + k < n_ && s_[i_ + k] == w[k]) k++;
        if (w[k]) return false;
        i_ += k;
        return true;
    }
    static void utf8(std::string& out, uint32_t cp) {
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
        if (i_ + 4 > n_) fail("truncated \\u escape");
        uint32_t v = 0;
        for (int k = 0; k < 4; k++) {
            char c = s_[i_++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= (uint32_t)(c - '0');
            else if (c >= 'a' && c <= 'f') v |= (uint32_t)(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= (uint32_t)(c - 'A' + 10);
            else fail("bad \\u escape");
        }
        return v;
    }
    std::string string_body() {
        std::string out;
        i_++; /* opening quote */
        while (true) {
            if (i_ >= n_) fail("unterminated string");
            char c = s_[i_++];
            if (c == '"') break;
            if ((unsigned char)c < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= n_) fail("unterminated escape");
            char e = s_[i_++];
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
                if (cp >= 0xD800 && cp < 0xDC00 && i_ + 1 < n_ && s_[i_] == '\\' && s_[i_ + 1] == 'u') {
                    i_ += 2;
                    uint32_t lo = hex4();
                    if (lo >= 0xDC00 && lo < 0xE000) cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    else fail("bad surrogate pair");
                }
                utf8(out, cp);
                break;
            }
            default: fail("bad escape");
            }
        }
        return out;
    }
    LxJson value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (i_ >= n_) fail("unexpected end");
        LxJson v;
        char c = s_[i_];
        if (c == '{') {
            v.kind = LxJson::OBJ;
            i_++;
            ws();
            if (i_ < n_ && s_[i_] == '}') {
                i_++;
                return v;
            }
            while (true) {
                ws();
                if (i_ >= n_ || s_[i_] != '"') fail("expected a key");
                std::string k = string_body();
                ws();
                if (i_ >= n_ || s_[i_] != ':') fail("expected ':'");
                i_++;
                v.obj.emplace_back(std::move(k), value(depth + 1));
                ws();
                if (i_ < n_ && s_[i_] == ',') {
                    i_++;
                    continue;
                }
                if (i_ < n_ && s_[i_] == '}') {
                    i_++;
                    break;
                }
                fail("expected ',' or '}'");
            }
        } else if (c == '[') {
            v.kind = LxJson::ARR;
            i_++;
            ws();
            if (i_ < n_ && s_[i_] == ']') {
                i_++;
                return v;
            }
            while (true) {
                v.arr.push_back(value(depth + 1));
                ws();
                if (i_ < n_ && s_[i_] == ',') {
                    i_++;
                    continue;
                }
                if (i_ < n_ && s_[i_] == ']') {
                    i_++;
                    break;
                }
                fail("expected ',' or ']'");
            }
        } else if (c == '"') {
            v.kind = LxJson::STR;
            v.str = string_body();
        } else if (lit("true")) {
            v.kind = LxJson::BOOL;
            v.b = true;
        } else if (lit("false")) {
            v.kind = LxJson::BOOL;
        } else if (lit("null")) {
            v.kind = LxJson::NUL;
        } else if (c == '-' || (c >= '0' && c <= '9')) {
            size_t b = i_;
            if (s_[i_] == '-') i_++;
            while (i_ < n_ && ((s_[i_] >= '0' && s_[i_] <= '9') || s_[i_] == '.' || s_[i_] == 'e' || s_[i_] == 'E' ||
                               s_[i_] == '+' || s_[i_] == '-'))
                i_++;
            std::string t(s_ + b, i_ - b);
            char* end = nullptr;
            v.kind = LxJson::NUM;
            v.num = std::strtod(t.c_str(), &end);
            if (!end || *end) fail("bad number");
        } else {
            fail("unexpected character");
        }
        return v;
    }
};

#endif
