// json_lite.h - a small JSON reader for inp.json parameter files.
//
// Accepts what the reference's parameter files use (jsoncpp 0.7 with comments
// allowed, the mode read_json_params relies on, GPUbounce.cu:1914-1933):
// objects, arrays, numbers, strings, true/false/null, and // line and /* */
// block comments.  The accessors follow jsoncpp's conversions for a missing
// key (0, 0.0, false, "") and between numbers and booleans, so that parameters
// absent from a file read as the reference reads them (e.g. population.max_pop
// absent -> 0, GPUbounce.cu:2011).
#pragma once

#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace jsonlite {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    const Value& operator[](const std::string& key) const {
        static const Value null_value;
        if (kind != Object) return null_value;
        for (const auto& kv : obj)
            if (kv.first == key) return kv.second;
        return null_value;
    }
    bool has(const std::string& key) const {
        if (kind != Object) return false;
        for (const auto& kv : obj)
            if (kv.first == key) return true;
        return false;
    }
    bool is_null() const { return kind == Null; }
    // jsoncpp Value::asInt / asFloat / asBool / asString for the value kinds in use
    int as_int() const {
        if (kind == Number) return (int)num;  // truncation toward zero, as asInt of a real
        if (kind == Bool) return b ? 1 : 0;
        return 0;
    }
    float as_float() const {
        if (kind == Number) return (float)num;
        if (kind == Bool) return b ? 1.0f : 0.0f;
        return 0.0f;
    }
    bool as_bool() const {
        if (kind == Bool) return b;
        if (kind == Number) return num != 0.0;
        return false;
    }
    std::string as_string() const {
        if (kind == String) return str;
        return std::string();
    }
};

class Parser {
public:
    explicit Parser(const std::string& text) : s_(text) {}

    Value parse() {
        Value v = value();
        skip();
        if (i_ != s_.size()) error("trailing characters");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void error(const std::string& what) const {
        size_t line = 1;
        for (size_t k = 0; k < i_ && k < s_.size(); ++k)
            if (s_[k] == '\n') ++line;
        throw std::runtime_error("JSON parse error at line " + std::to_string(line) + ": " + what);
    }

    void skip() {
        for (;;) {
            while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '/') {
                while (i_ < s_.size() && s_[i_] != '\n') ++i_;
                continue;
            }
            if (i_ + 1 < s_.size() && s_[i_] == '/' && s_[i_ + 1] == '*') {
                const size_t e = s_.find("*/", i_ + 2);
                if (e == std::string::npos) error("unterminated comment");
                i_ = e + 2;
                continue;
            }
            return;
        }
    }

    bool take(char c) {
        skip();
        if (i_ < s_.size() && s_[i_] == c) {
            ++i_;
            return true;
        }
        return false;
    }

    void expect(char c) {
        if (!take(c)) error(std::string("expected '") + c + "'");
    }

    bool word(const char* w) {
        size_t n = 0;
        while (w[n]) ++n;
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }

    std::string string_body() {
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) error("bad escape");
                const char e = s_[i_++];
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
                        if (i_ + 4 > s_.size()) error("bad \\u escape");
                        const unsigned cp = (unsigned)std::strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) {
                            out += (char)cp;
                        } else if (cp < 0x800) {
                            out += (char)(0xC0 | (cp >> 6));
                            out += (char)(0x80 | (cp & 0x3F));
                        } else {
                            out += (char)(0xE0 | (cp >> 12));
                            out += (char)(0x80 | ((cp >> 6) & 0x3F));
                            out += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: error("bad escape");
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) error("unterminated string");
        ++i_;
        return out;
    }

    Value value() {
        skip();
        if (i_ >= s_.size()) error("unexpected end of input");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            ++i_;
            v.kind = Value::Object;
            if (take('}')) return v;
            do {
                skip();
                if (i_ >= s_.size() || s_[i_] != '"') error("expected a key");
                ++i_;
                std::string key = string_body();
                expect(':');
                v.obj.emplace_back(std::move(key), value());
            } while (take(','));
            expect('}');
        } else if (c == '[') {
            ++i_;
            v.kind = Value::Array;
            if (take(']')) return v;
            do v.arr.push_back(value());
            while (take(','));
            expect(']');
        } else if (c == '"') {
            ++i_;
            v.kind = Value::String;
            v.str = string_body();
        } else if (word("true")) {
            v.kind = Value::Bool;
            v.b = true;
        } else if (word("false")) {
            v.kind = Value::Bool;
            v.b = false;
        } else if (word("null")) {
            v.kind = Value::Null;
        } else {
            const char* start = s_.c_str() + i_;
            char* end = nullptr;
            v.num = std::strtod(start, &end);
            if (end == start) error("unexpected character");
            i_ += (size_t)(end - start);
            v.kind = Value::Number;
        }
        return v;
    }
};

inline Value parse(const std::string& text) { return Parser(text).parse(); }

}  // namespace jsonlite
