// Minimal JSON values for the serving loop's request/response function: a
// recursive-descent reader of RFC 8259 text (objects, arrays, strings with
// escapes, numbers, true/false/null) and a writer.  Numbers are doubles, as in
// the request bodies the reference's HTTP front end reads.
#pragma once
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

namespace tblaeq_json {

struct Value {
    enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool isNumber() const { return kind == Number; }
    bool isString() const { return kind == String; }
    bool isBool() const { return kind == Bool; }
    bool isArray() const { return kind == Array; }
    bool isObject() const { return kind == Object; }
    const Value* find(const std::string& key) const {
        if (kind != Object) return nullptr;
        auto it = obj.find(key);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class Reader {
public:
    explicit Reader(const std::string& text) : s_(text) {}

    // Parses the whole text as one value; throws std::runtime_error with the
    // byte offset of the first problem.
    Value parse() {
        Value v = value(0);
        skip();
        if (i_ != s_.size()) fail("unexpected text after the value");
        return v;
    }

private:
    const std::string& s_;
    size_t i_ = 0;
    static constexpr int kMaxDepth = 64;

    [[noreturn]] void fail(const char* what) const {
        throw std::runtime_error(std::string(what) + " at byte " + std::to_string(i_));
    }
    void skip() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r')) ++i_;
    }
    bool eat(const char* word) {
        size_t n = 0;
        while (word[n]) ++n;
        if (s_.compare(i_, n, word) != 0) return false;
        i_ += n;
        return true;
    }
    Value value(int depth) {
        if (depth > kMaxDepth) fail("nesting too deep");
        skip();
        if (i_ >= s_.size()) fail("unexpected end of text");
        Value v;
        const char c = s_[i_];
        if (c == '{') {
            v.kind = Value::Object;
            ++i_;
            skip();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return v;
            }
            for (;;) {
                skip();
                if (i_ >= s_.size() || s_[i_] != '"') fail("expected a string key");
                std::string key = string();
                skip();
                if (i_ >= s_.size() || s_[i_] != ':') fail("expected ':'");
                ++i_;
                v.obj[key] = value(depth + 1);
                skip();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    return v;
                }
                fail("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = Value::Array;
            ++i_;
            skip();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value(depth + 1));
                skip();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
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
        if (eat("true")) {
            v.kind = Value::Bool;
            v.b = true;
            return v;
        }
        if (eat("false")) {
            v.kind = Value::Bool;
            return v;
        }
        if (eat("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            v.kind = Value::Number;
            v.num = number();
            return v;
        }
        fail("unexpected character");
    }
    static unsigned hexDigit(char c) {
        if (c >= '0' && c <= '9') return static_cast<unsigned>(c - '0');
        if (c >= 'a' && c <= 'f') return static_cast<unsigned>(c - 'a' + 10);
        if (c >= 'A' && c <= 'F') return static_cast<unsigned>(c - 'A' + 10);
        return 16;
    }
    unsigned hex4() {
        if (i_ + 4 > s_.size()) fail("short \\u escape");
        unsigned u = 0;
        for (int k = 0; k < 4; ++k) {
            const unsigned d = hexDigit(s_[i_++]);
            if (d > 15) fail("bad \\u escape");
            u = u * 16 + d;
        }
        return u;
    }
    static void utf8(std::string& out, unsigned cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    std::string string() {
        ++i_;  // opening quote
        std::string out;
        for (;;) {
            if (i_ >= s_.size()) fail("unterminated string");
            const char c = s_[i_++];
            if (c == '"') return out;
            if (static_cast<unsigned char>(c) < 0x20) fail("control character in string");
            if (c != '\\') {
                out += c;
                continue;
            }
            if (i_ >= s_.size()) fail("unterminated escape");
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
                    unsigned cp = hex4();
                    if (cp >= 0xD800 && cp < 0xDC00) {
                        if (!(i_ + 1 < s_.size() && s_[i_] == '\\' && s_[i_ + 1] == 'u')) fail("lone surrogate");
                        i_ += 2;
                        const unsigned lo = hex4();
                        if (lo < 0xDC00 || lo >= 0xE000) fail("bad surrogate pair");
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    } else if (cp >= 0xDC00 && cp < 0xE000) {
                        fail("lone surrogate");
                    }
                    utf8(out, cp);
                    break;
                }
                default: fail("bad escape");
            }
        }
    }
    double number() {
        const size_t b = i_;
        if (s_[i_] == '-') ++i_;
        auto digits = [&] {
            const size_t d = i_;
            while (i_ < s_.size() && s_[i_] >= '0' && s_[i_] <= '9') ++i_;
            return i_ - d;
        };
        if (i_ < s_.size() && s_[i_] == '0') {
            ++i_;
        } else if (digits() == 0) {
            fail("bad number");
        }
        if (i_ < s_.size() && s_[i_] == '.') {
            ++i_;
            if (digits() == 0) fail("bad fraction");
        }
        if (i_ < s_.size() && (s_[i_] == 'e' || s_[i_] == 'E')) {
            ++i_;
            if (i_ < s_.size() && (s_[i_] == '+' || s_[i_] == '-')) ++i_;
            if (digits() == 0) fail("bad exponent");
        }
        return std::strtod(s_.substr(b, i_ - b).c_str(), nullptr);
    }
};

// Writer: appends JSON text to a string.  Doubles print with 17 significant
// digits (integers up to 2^53 print without a fraction).
class Writer {
public:
    std::string out;
    void beginObject() { sep(), out += '{', first_ = true; }
    void endObject() { out += '}', first_ = false; }
    void beginArray() { sep(), out += '[', first_ = true; }
    void endArray() { out += ']', first_ = false; }
    void key(const std::string& k) {
        sep();
        quote(k);
        out += ':';
        first_ = true;
    }
    void str(const std::string& v) { sep(), quote(v); }
    void boolean(bool v) { sep(), out += v ? "true" : "false"; }
    void number(double v) {
        sep();
        if (!std::isfinite(v)) {
            out += "null";
            return;
        }
        char buf[40];
        if (v == std::floor(v) && std::fabs(v) < 9007199254740992.0)
            std::snprintf(buf, sizeof buf, "%.0f", v);
        else
            std::snprintf(buf, sizeof buf, "%.17g", v);
        out += buf;
    }
    void integer(uint64_t v) { sep(), out += std::to_string(v); }
    template <class T>
    void integers(const std::vector<T>& v) {
        beginArray();
        for (const T& x : v) integer(static_cast<uint64_t>(x));
        endArray();
    }

private:
    bool first_ = true;
    void sep() {
        if (!first_) out += ',';
        first_ = false;
    }
    void quote(const std::string& s) {
        out += '"';
        for (const char c : s) {
            const unsigned char u = static_cast<unsigned char>(c);
            if (c == '"') out += "\\\"";
            else if (c == '\\') out += "\\\\";
            else if (c == '\n') out += "\\n";
            else if (c == '\r') out += "\\r";
            else if (c == '\t') out += "\\t";
            else if (u < 0x20) {
                char buf[8];
                std::snprintf(buf, sizeof buf, "\\u%04x", u);
                out += buf;
            } else {
                out += c;
            }
        }
        out += '"';
    }
};

}  // namespace tblaeq_json
