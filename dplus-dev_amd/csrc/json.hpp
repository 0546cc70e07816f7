// json.hpp -- minimal JSON value, parser and writer for the backend's state
// documents and replies.  Objects keep their member order.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "device.hpp"

namespace dplus {
namespace json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::vector<std::pair<std::string, Value>> obj;

    bool is_object() const { return kind == Object; }
    bool is_array() const { return kind == Array; }
    bool is_number() const { return kind == Number; }
    bool is_string() const { return kind == String; }
    bool is_bool() const { return kind == Bool; }

    const Value *find(const std::string &k) const {
        if (kind != Object) return nullptr;
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
    const Value &at(const std::string &k) const {
        const Value *v = find(k);
        if (!v) throw Error(kErrInvalidParamTree, "missing key \"" + k + "\" in the state");
        return *v;
    }
    double number(double def = 0.0) const {
        if (kind == Number) return num;
        if (kind == Bool) return b ? 1.0 : 0.0;
        if (kind == String) {
            if (str == "inf") return INFINITY;
            if (str == "-inf") return -INFINITY;
            char *e = nullptr;
            double d = strtod(str.c_str(), &e);
            if (e && e != str.c_str()) return d;
        }
        return def;
    }
    bool truth(bool def = false) const {
        if (kind == Bool) return b;
        if (kind == Number) return num != 0.0;
        return def;
    }
    std::string text(const std::string &def = "") const { return kind == String ? str : def; }
};

class Parser {
   public:
    explicit Parser(const std::string &s) : s_(s) {}
    Value parse() {
        Value v = value();
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

   private:
    const std::string &s_;
    size_t i_ = 0;
    [[noreturn]] void fail(const char *what) {
        throw Error(kErrIllegalJson, std::string("illegal JSON (") + what + ") at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
    }
    bool lit(const char *w) {
        size_t n = 0;
        while (w[n]) n++;
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    Value value() {
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        const char c = s_[i_];
        Value v;
        if (c == '{') {
            v.kind = Value::Object;
            i_++;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') {
                i_++;
                return v;
            }
            while (true) {
                ws();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected key");
                std::string k = string();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
                i_++;
                v.obj.emplace_back(std::move(k), value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    i_++;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    i_++;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            i_++;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') {
                i_++;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    i_++;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    i_++;
                    return v;
                }
                fail("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = Value::String;
            v.str = string();
            return v;
        }
        if (lit("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (lit("null")) return v;
        if (lit("NaN")) {
            v.kind = Value::Number;
            v.num = NAN;
            return v;
        }
        if (lit("Infinity")) {
            v.kind = Value::Number;
            v.num = INFINITY;
            return v;
        }
        if (lit("-Infinity")) {
            v.kind = Value::Number;
            v.num = -INFINITY;
            return v;
        }
        const char *b = s_.c_str() + i_;
        char *e = nullptr;
        double d = strtod(b, &e);
        if (e == b) fail("unexpected character");
        i_ += (size_t)(e - b);
        v.kind = Value::Number;
        v.num = d;
        return v;
    }
    std::string string() {
        std::string o;
        i_++;  // opening quote
        while (i_ < s_.size()) {
            char c = s_[i_++];
            if (c == '"') return o;
            if (c == '\\') {
                if (i_ >= s_.size()) fail("bad escape");
                char e = s_[i_++];
                switch (e) {
                    case 'n': o += '\n'; break;
                    case 't': o += '\t'; break;
                    case 'r': o += '\r'; break;
                    case 'b': o += '\b'; break;
                    case 'f': o += '\f'; break;
                    case 'u': {
                        if (i_ + 4 > s_.size()) fail("bad \\u escape");
                        unsigned cp = (unsigned)strtoul(s_.substr(i_, 4).c_str(), nullptr, 16);
                        i_ += 4;
                        if (cp < 0x80) {
                            o += (char)cp;
                        } else if (cp < 0x800) {
                            o += (char)(0xC0 | (cp >> 6));
                            o += (char)(0x80 | (cp & 0x3F));
                        } else {
                            o += (char)(0xE0 | (cp >> 12));
                            o += (char)(0x80 | ((cp >> 6) & 0x3F));
                            o += (char)(0x80 | (cp & 0x3F));
                        }
                        break;
                    }
                    default: o += e;
                }
            } else {
                o += c;
            }
        }
        fail("unterminated string");
    }
};

inline Value parse(const std::string &s) { return Parser(s).parse(); }

inline std::string num(double d) {
    if (std::isnan(d)) return "NaN";
    if (std::isinf(d)) return d > 0 ? "Infinity" : "-Infinity";
    char b[40];
    snprintf(b, sizeof b, "%.17g", d);
    return b;
}

inline std::string quote(const std::string &s) { return "\"" + json_escape(s) + "\""; }

}  // namespace json
}  // namespace dplus
