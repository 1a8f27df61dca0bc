// Minimal JSON reader for the scene files (RFC 8259: objects, arrays, strings with escapes and
// \uXXXX / surrogate pairs as UTF-8, numbers, true / false / null). Throws std::invalid_argument
// with the byte offset on malformed input. Objects keep member order; lookups are linear, which
// is fine for scene documents.
#pragma once

#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace mrg_host {

class Json {
  public:
    enum class Type { null, boolean, number, string, array, object };

    Type type() const { return type_; }
    bool is_number() const { return type_ == Type::number; }
    bool is_string() const { return type_ == Type::string; }
    bool is_array() const { return type_ == Type::array; }
    bool is_object() const { return type_ == Type::object; }
    bool is_boolean() const { return type_ == Type::boolean; }
    // An integer-valued number written without fraction or exponent (nlohmann's is_number_integer).
    bool is_integer() const { return type_ == Type::number && integer_; }

    double number() const {
        require(Type::number, "number");
        return number_;
    }
    bool boolean() const {
        require(Type::boolean, "boolean");
        return boolean_;
    }
    const std::string& string() const {
        require(Type::string, "string");
        return string_;
    }
    const std::vector<Json>& array() const {
        require(Type::array, "array");
        return array_;
    }
    size_t size() const { return type_ == Type::array ? array_.size() : type_ == Type::object ? object_.size() : 0; }
    const Json& operator[](size_t i) const {
        require(Type::array, "array");
        if (i >= array_.size()) throw std::invalid_argument("JSON array index out of range");
        return array_[i];
    }
    bool contains(const std::string& key) const { return find(key) != nullptr; }
    // Member access that throws for a missing key (nlohmann's at()).
    const Json& at(const std::string& key) const {
        const Json* v = find(key);
        if (!v) throw std::invalid_argument("Missing JSON key: " + key);
        return *v;
    }
    const Json* find(const std::string& key) const {
        require(Type::object, "object");
        for (const auto& kv : object_)
            if (kv.first == key) return &kv.second;
        return nullptr;
    }
    const std::vector<std::pair<std::string, Json>>& members() const {
        require(Type::object, "object");
        return object_;
    }

    static Json parse(const std::string& text) {
        size_t pos = 0;
        Json   v   = parse_value(text, pos, 0);
        skip_ws(text, pos);
        if (pos != text.size()) fail("trailing characters", pos);
        return v;
    }

  private:
    Type                                      type_    = Type::null;
    bool                                      boolean_ = false;
    bool                                      integer_ = false;
    double                                    number_  = 0;
    std::string                               string_;
    std::vector<Json>                         array_;
    std::vector<std::pair<std::string, Json>> object_;

    void require(Type t, const char* what) const {
        if (type_ != t) throw std::invalid_argument(std::string("JSON value is not a ") + what);
    }
    [[noreturn]] static void fail(const char* what, size_t pos) {
        throw std::invalid_argument(std::string("JSON parse error (") + what + ") at byte " + std::to_string(pos));
    }
    static void skip_ws(const std::string& s, size_t& p) {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }
    static void expect(const std::string& s, size_t& p, const char* word) {
        for (const char* c = word; *c; ++c, ++p)
            if (p >= s.size() || s[p] != *c) fail("unexpected literal", p);
    }
    static void put_utf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += char(cp);
        } else if (cp < 0x800) {
            out += char(0xC0 | (cp >> 6));
            out += char(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += char(0xE0 | (cp >> 12));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        } else {
            out += char(0xF0 | (cp >> 18));
            out += char(0x80 | ((cp >> 12) & 0x3F));
            out += char(0x80 | ((cp >> 6) & 0x3F));
            out += char(0x80 | (cp & 0x3F));
        }
    }
    static uint32_t hex4(const std::string& s, size_t& p) {
        if (p + 4 > s.size()) fail("short \\u escape", p);
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i, ++p) {
            const char c = s[p];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= uint32_t(c - '0');
            else if (c >= 'a' && c <= 'f') v |= uint32_t(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= uint32_t(c - 'A' + 10);
            else fail("bad hex digit", p);
        }
        return v;
    }
    static std::string parse_string(const std::string& s, size_t& p) {
        if (p >= s.size() || s[p] != '"') fail("expected string", p);
        ++p;
        std::string out;
        for (;;) {
            if (p >= s.size()) fail("unterminated string", p);
            const char c = s[p++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string", p - 1);
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size()) fail("unterminated escape", p);
            const char e = s[p++];
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
                uint32_t cp = hex4(s, p);
                if (cp >= 0xD800 && cp < 0xDC00) {
                    if (p + 2 > s.size() || s[p] != '\\' || s[p + 1] != 'u') fail("lone high surrogate", p);
                    p += 2;
                    const uint32_t lo = hex4(s, p);
                    if (lo < 0xDC00 || lo >= 0xE000) fail("bad low surrogate", p);
                    cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                } else if (cp >= 0xDC00 && cp < 0xE000) {
                    fail("lone low surrogate", p);
                }
                put_utf8(out, cp);
                break;
            }
            default: fail("bad escape", p - 1);
            }
        }
    }
    static Json parse_number(const std::string& s, size_t& p) {
        const size_t start = p;
        bool         integer = true;
        if (p < s.size() && s[p] == '-') ++p;
        if (p >= s.size() || !(s[p] >= '0' && s[p] <= '9')) fail("bad number", p);
        if (s[p] == '0') {
            ++p;
        } else {
            while (p < s.size() && s[p] >= '0' && s[p] <= '9') ++p;
        }
        if (p < s.size() && s[p] == '.') {
            integer = false;
            ++p;
            if (p >= s.size() || !(s[p] >= '0' && s[p] <= '9')) fail("bad fraction", p);
            while (p < s.size() && s[p] >= '0' && s[p] <= '9') ++p;
        }
        if (p < s.size() && (s[p] == 'e' || s[p] == 'E')) {
            integer = false;
            ++p;
            if (p < s.size() && (s[p] == '+' || s[p] == '-')) ++p;
            if (p >= s.size() || !(s[p] >= '0' && s[p] <= '9')) fail("bad exponent", p);
            while (p < s.size() && s[p] >= '0' && s[p] <= '9') ++p;
        }
        Json v;
        v.type_    = Type::number;
        v.integer_ = integer;
        v.number_  = std::strtod(s.substr(start, p - start).c_str(), nullptr);  // C locale: '.' decimal point
        return v;
    }
    static Json parse_value(const std::string& s, size_t& p, int depth) {
        if (depth > 256) fail("nesting too deep", p);
        skip_ws(s, p);
        if (p >= s.size()) fail("unexpected end", p);
        Json v;
        const char c = s[p];
        if (c == '{') {
            v.type_ = Type::object;
            ++p;
            skip_ws(s, p);
            if (p < s.size() && s[p] == '}') {
                ++p;
                return v;
            }
            for (;;) {
                skip_ws(s, p);
                std::string key = parse_string(s, p);
                skip_ws(s, p);
                if (p >= s.size() || s[p] != ':') fail("expected ':'", p);
                ++p;
                v.object_.emplace_back(std::move(key), parse_value(s, p, depth + 1));
                skip_ws(s, p);
                if (p < s.size() && s[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < s.size() && s[p] == '}') {
                    ++p;
                    return v;
                }
                fail("expected ',' or '}'", p);
            }
        }
        if (c == '[') {
            v.type_ = Type::array;
            ++p;
            skip_ws(s, p);
            if (p < s.size() && s[p] == ']') {
                ++p;
                return v;
            }
            for (;;) {
                v.array_.push_back(parse_value(s, p, depth + 1));
                skip_ws(s, p);
                if (p < s.size() && s[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < s.size() && s[p] == ']') {
                    ++p;
                    return v;
                }
                fail("expected ',' or ']'", p);
            }
        }
        if (c == '"') {
            v.type_   = Type::string;
            v.string_ = parse_string(s, p);
            return v;
        }
        if (c == 't') {
            expect(s, p, "true");
            v.type_    = Type::boolean;
            v.boolean_ = true;
            return v;
        }
        if (c == 'f') {
            expect(s, p, "false");
            v.type_ = Type::boolean;
            return v;
        }
        if (c == 'n') {
            expect(s, p, "null");
            return v;
        }
        return parse_number(s, p);
    }
};

}  // namespace mrg_host
