// Minimal JSON values for processor specs: objects keep their keys sorted
// (as QJsonObject does) and print indented by four spaces per level.
#pragma once

#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace msamd_host {

class Json {
public:
    enum Kind { Null, Bool, Number, String, Array, Object };

    Json() = default;
    Json(bool b) : kind_(Bool), b_(b) {}
    Json(int v) : kind_(Number), d_(v) {}
    Json(long long v) : kind_(Number), d_(static_cast<double>(v)) {}
    Json(double v) : kind_(Number), d_(v) {}
    Json(const char* s) : kind_(String), s_(s) {}
    Json(const std::string& s) : kind_(String), s_(s) {}

    static Json array() {
        Json j;
        j.kind_ = Array;
        return j;
    }
    static Json object() {
        Json j;
        j.kind_ = Object;
        return j;
    }

    Json& operator[](const std::string& key) {
        kind_ = Object;
        return obj_[key];
    }
    void push_back(const Json& v) {
        kind_ = Array;
        arr_.push_back(v);
    }
    bool has(const std::string& key) const { return kind_ == Object && obj_.count(key); }
    const Json& at(const std::string& key) const { return obj_.at(key); }
    const std::string& str() const { return s_; }

    std::string dump(int indent = 0) const {
        std::string out;
        write(out, indent);
        return out;
    }

private:
    static std::string pad(int n) { return std::string(4 * n, ' '); }

    static std::string number(double d) {
        if (std::isfinite(d) && d == std::floor(d) && std::fabs(d) < 1e15) {
            char buf[32];
            snprintf(buf, sizeof buf, "%lld", static_cast<long long>(d));
            return buf;
        }
        char buf[64];
        for (int p = 1; p <= 17; ++p) {
            snprintf(buf, sizeof buf, "%.*g", p, d);
            if (strtod(buf, nullptr) == d) break;
        }
        return buf;
    }

    static std::string quote(const std::string& s) {
        std::string o = "\"";
        for (char c : s) {
            switch (c) {
                case '"': o += "\\\""; break;
                case '\\': o += "\\\\"; break;
                case '\n': o += "\\n"; break;
                case '\t': o += "\\t"; break;
                case '\r': o += "\\r"; break;
                default:
                    if (static_cast<unsigned char>(c) < 0x20) {
                        char buf[8];
                        snprintf(buf, sizeof buf, "\\u%04x", c);
                        o += buf;
                    } else {
                        o += c;
                    }
            }
        }
        return o + "\"";
    }

    void write(std::string& out, int level) const {
        switch (kind_) {
            case Null: out += "null"; break;
            case Bool: out += b_ ? "true" : "false"; break;
            case Number: out += number(d_); break;
            case String: out += quote(s_); break;
            case Array: {
                out += "[\n";
                for (size_t i = 0; i < arr_.size(); ++i) {
                    out += pad(level + 1);
                    arr_[i].write(out, level + 1);
                    out += i + 1 < arr_.size() ? ",\n" : "\n";
                }
                out += pad(level) + "]";
                break;
            }
            case Object: {
                out += "{\n";
                size_t i = 0;
                for (const auto& kv : obj_) {
                    out += pad(level + 1) + quote(kv.first) + ": ";
                    kv.second.write(out, level + 1);
                    out += ++i < obj_.size() ? ",\n" : "\n";
                }
                out += pad(level) + "}";
                break;
            }
        }
    }

    Kind kind_ = Null;
    bool b_ = false;
    double d_ = 0.0;
    std::string s_;
    std::vector<Json> arr_;
    std::map<std::string, Json> obj_;
};

}  // namespace msamd_host
