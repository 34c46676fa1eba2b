// json_min.hpp — a small JSON reader (RFC 8259 values: objects, arrays,
// strings, numbers, true/false/null) for map files.  Numbers are kept as
// double plus an "is an integer" flag; strings support the standard escapes
// (\uXXXX outside ASCII becomes '?': map files only use ASCII keys).
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace msim_host {

struct JsonValue {
    enum Kind { kNull, kBool, kNumber, kString, kArray, kObject } kind = kNull;
    bool b = false;
    double num = 0.0;
    bool integral = false;  // the literal had no fraction or exponent
    std::string str;
    std::vector<JsonValue> arr;
    std::map<std::string, JsonValue> obj;

    bool has(const std::string& k) const { return kind == kObject && obj.count(k) != 0; }
    const JsonValue& at(const std::string& k) const { return obj.at(k); }
};

class JsonReader {
 public:
    explicit JsonReader(const std::string& text) : s_(text) {}

    JsonValue parse() {
        JsonValue v = value(0);
        ws();
        if (i_ != s_.size()) fail("trailing characters");
        return v;
    }

 private:
    const std::string& s_;
    size_t i_ = 0;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string("JSON parse error at offset ") + std::to_string(i_) + ": " + what);
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) i_++;
    }
    bool eat(char c) {
        ws();
        if (i_ < s_.size() && s_[i_] == c) {
            i_++;
            return true;
        }
        return false;
    }
    void expect(char c) {
        if (!eat(c)) fail("unexpected character");
    }
    bool lit(const char* w) {
        size_t k = 0;
        while (w[k]) k++;
        if (s_.compare(i_, k, w) == 0) {
            i_ += k;
            return true;
        }
        return false;
    }
    JsonValue value(int depth) {
        if (depth > 256) fail("nesting too deep");
        ws();
        if (i_ >= s_.size()) fail("unexpected end");
        JsonValue v;
        const char c = s_[i_];
        if (c == '{') {
            i_++;
            v.kind = JsonValue::kObject;
            if (eat('}')) return v;
            do {
                ws();
                std::string k = string();
                expect(':');
                v.obj[k] = value(depth + 1);
            } while (eat(','));
            expect('}');
        } else if (c == '[') {
            i_++;
            v.kind = JsonValue::kArray;
            if (eat(']')) return v;
            do {
                v.arr.push_back(value(depth + 1));
            } while (eat(','));
            expect(']');
        } else if (c == '"') {
            v.kind = JsonValue::kString;
            v.str = string();
        } else if (lit("true")) {
            v.kind = JsonValue::kBool;
            v.b = true;
        } else if (lit("false")) {
            v.kind = JsonValue::kBool;
        } else if (lit("null")) {
            v.kind = JsonValue::kNull;
        } else {
            number(v);
        }
        return v;
    }
    std::string string() {
        if (i_ >= s_.size() || s_[i_] != '"') fail("expected a string");
        i_++;
        std::string out;
        while (true) {
            if (i_ >= s_.size()) fail("unterminated string");
            char c = s_[i_++];
            if (c == '"') break;
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (i_ >= s_.size()) fail("bad escape");
            char e = s_[i_++];
            switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    if (i_ + 4 > s_.size()) fail("bad \\u escape");
                    const unsigned long cp = std::stoul(s_.substr(i_, 4), nullptr, 16);
                    i_ += 4;
                    out.push_back(cp < 0x80 ? static_cast<char>(cp) : '?');
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    void number(JsonValue& v) {
        const size_t b = i_;
        if (i_ < s_.size() && s_[i_] == '-') i_++;
        bool digits = false, frac = false;
        while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') {
            i_++;
            digits = true;
        }
        if (i_ < s_.size() && s_[i_] == '.') {
            frac = true;
            i_++;
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') i_++;
        }
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            frac = true;
            i_++;
            if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) i_++;
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') i_++;
        }
        if (!digits) fail("expected a value");
        v.kind = JsonValue::kNumber;
        v.num = std::stod(s_.substr(b, i_ - b));
        v.integral = !frac;
    }
};

}  // namespace msim_host
