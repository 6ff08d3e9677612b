// Minimal JSON reader for scene descriptions (Models.json, Parameters.json,
// colouring category files). Numbers are held as double; no writer.
#pragma once
#include <cctype>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace gvbd {
namespace json {

struct Value {
    enum Type { Null, Bool, Number, String, Array, Object } type = Null;
    bool b = false;
    double num = 0.0;
    std::string str;
    std::vector<Value> arr;
    std::map<std::string, Value> obj;

    bool has(const std::string& k) const { return type == Object && obj.count(k) != 0; }
    const Value* get(const std::string& k) const {
        if (type != Object) return nullptr;
        auto it = obj.find(k);
        return it == obj.end() ? nullptr : &it->second;
    }
};

class Parser {
  public:
    explicit Parser(const std::string& s) : s_(s) {}
    bool parse(Value& out, std::string& err) {
        pos_ = 0;
        if (!value(out)) {
            err = "JSON parse error near offset " + std::to_string(pos_);
            return false;
        }
        ws();
        if (pos_ != s_.size()) {
            err = "trailing characters after JSON value at offset " + std::to_string(pos_);
            return false;
        }
        return true;
    }

  private:
    void ws() {
        while (pos_ < s_.size() && std::isspace((unsigned char)s_[pos_])) ++pos_;
    }
    bool lit(const char* w) {
        size_t n = std::char_traits<char>::length(w);
        if (s_.compare(pos_, n, w) != 0) return false;
        pos_ += n;
        return true;
    }
    bool value(Value& v) {
        ws();
        if (pos_ >= s_.size()) return false;
        char c = s_[pos_];
        if (c == '{') return object(v);
        if (c == '[') return array(v);
        if (c == '"') {
            v.type = Value::String;
            return string(v.str);
        }
        if (lit("true")) {
            v.type = Value::Bool;
            v.b = true;
            return true;
        }
        if (lit("false")) {
            v.type = Value::Bool;
            v.b = false;
            return true;
        }
        if (lit("null")) {
            v.type = Value::Null;
            return true;
        }
        return number(v);
    }
    bool number(Value& v) {
        const char* start = s_.c_str() + pos_;
        char* end = nullptr;
        double d = std::strtod(start, &end);
        if (end == start) return false;
        pos_ += (size_t)(end - start);
        v.type = Value::Number;
        v.num = d;
        return true;
    }
    bool string(std::string& out) {
        if (s_[pos_] != '"') return false;
        ++pos_;
        out.clear();
        while (pos_ < s_.size()) {
            char c = s_[pos_++];
            if (c == '"') return true;
            if (c == '\\') {
                if (pos_ >= s_.size()) return false;
                char e = s_[pos_++];
                switch (e) {
                case 'n': out += '\n'; break;
                case 't': out += '\t'; break;
                case 'r': out += '\r'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'u': {
                    if (pos_ + 4 > s_.size()) return false;
                    unsigned cp = (unsigned)std::strtoul(s_.substr(pos_, 4).c_str(), nullptr, 16);
                    pos_ += 4;
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
                default: out += e; break;
                }
            } else {
                out += c;
            }
        }
        return false;
    }
    bool array(Value& v) {
        v.type = Value::Array;
        ++pos_;
        ws();
        if (pos_ < s_.size() && s_[pos_] == ']') {
            ++pos_;
            return true;
        }
        for (;;) {
            Value e;
            if (!value(e)) return false;
            v.arr.push_back(std::move(e));
            ws();
            if (pos_ >= s_.size()) return false;
            if (s_[pos_] == ',') {
                ++pos_;
                continue;
            }
            if (s_[pos_] == ']') {
                ++pos_;
                return true;
            }
            return false;
        }
    }
    bool object(Value& v) {
        v.type = Value::Object;
        ++pos_;
        ws();
        if (pos_ < s_.size() && s_[pos_] == '}') {
            ++pos_;
            return true;
        }
        for (;;) {
            ws();
            std::string key;
            if (pos_ >= s_.size() || !string(key)) return false;
            ws();
            if (pos_ >= s_.size() || s_[pos_] != ':') return false;
            ++pos_;
            Value e;
            if (!value(e)) return false;
            v.obj[key] = std::move(e);
            ws();
            if (pos_ >= s_.size()) return false;
            if (s_[pos_] == ',') {
                ++pos_;
                continue;
            }
            if (s_[pos_] == '}') {
                ++pos_;
                return true;
            }
            return false;
        }
    }

    const std::string& s_;
    size_t pos_ = 0;
};

}  // namespace json
}  // namespace gvbd
