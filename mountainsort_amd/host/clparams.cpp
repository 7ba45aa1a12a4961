// Command-line parameter parsing (see clparams.h).
#include "clparams.h"

#include <cerrno>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>

namespace msamd_host {

namespace {

std::string trimmed(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

bool parse_long(const std::string& str, long long& out) {
    const std::string t = trimmed(str);
    if (t.empty()) return false;
    errno = 0;
    char* end = nullptr;
    const long long v = strtoll(t.c_str(), &end, 10);
    if (errno == ERANGE || end != t.c_str() + t.size()) return false;
    out = v;
    return true;
}

// QString::toDouble in the C locale: decimal or exponent form, inf/nan; no hex.
bool parse_double(const std::string& str, double& out) {
    const std::string t = trimmed(str);
    if (t.empty() || t.find_first_of("xXpP") != std::string::npos) return false;
    errno = 0;
    char* end = nullptr;
    const double v = strtod(t.c_str(), &end);
    if (end != t.c_str() + t.size()) return false;
    if (errno == ERANGE && std::isfinite(v) && v != 0.0) return false;
    out = v;
    return true;
}

// QString::toFloat: toDouble, then out of range for float is a failure.
bool parse_float(const std::string& str, float& out) {
    double d;
    if (!parse_double(str, d)) return false;
    if (std::isfinite(d) && std::fabs(d) > FLT_MAX) return false;
    out = static_cast<float>(d);
    return true;
}

std::string shortest_float(float f) {
    char buf[64];
    for (int p = 1; p <= 9; ++p) {
        snprintf(buf, sizeof buf, "%.*g", p, static_cast<double>(f));
        if (strtof(buf, nullptr) == f) return buf;
    }
    return buf;
}

}  // namespace

Value Value::of_string(const std::string& str) {
    Value v;
    v.kind = Str;
    v.s = str;
    return v;
}

Value Value::of_long(long long x) {
    Value v;
    v.kind = Long;
    v.l = x;
    return v;
}

Value Value::of_double(double x) {
    Value v;
    v.kind = Float;
    v.f = static_cast<float>(x);
    return v;
}

Value Value::from_string(const std::string& str) {
    long long l;
    if (parse_long(str, l)) return of_long(l);  // covers the 32-bit integer case too
    float f;
    if (parse_float(str, f)) {
        Value v;
        v.kind = Float;
        v.f = f;
        return v;
    }
    if (str.size() >= 2 && str.front() == '[' && str.back() == ']') {
        const std::string inner = str.substr(1, str.size() - 2);
        std::vector<std::string> parts;
        size_t pos = 0;
        while (true) {
            const size_t k = inner.find("][", pos);
            if (k == std::string::npos) {
                parts.push_back(inner.substr(pos));
                break;
            }
            parts.push_back(inner.substr(pos, k - pos));
            pos = k + 2;
        }
        if (parts.size() == 1) return of_string(parts[0]);
        Value v;
        v.kind = List;
        for (const auto& p : parts) v.list.push_back(of_string(p));
        return v;
    }
    return of_string(str);
}

double Value::to_double() const {
    switch (kind) {
        case Long: return static_cast<double>(l);
        case Float: return static_cast<double>(f);
        case Str: {
            double d;
            return parse_double(s, d) ? d : 0.0;
        }
        default: return 0.0;
    }
}

long long Value::to_long() const {
    switch (kind) {
        case Long: return l;
        case Float: {
            // qRound64
            const double d = static_cast<double>(f);
            return d >= 0.0 ? static_cast<long long>(d + 0.5) : static_cast<long long>(d - 0.5);
        }
        case Str: {
            long long x;
            return parse_long(s, x) ? x : 0;
        }
        default: return 0;
    }
}

std::string Value::to_string() const {
    switch (kind) {
        case Str: return s;
        case Long: return std::to_string(l);
        case Float: return shortest_float(f);
        default: return std::string();
    }
}

std::vector<std::string> Value::to_string_list() const {
    std::vector<std::string> out;
    if (kind == List) {
        for (const auto& v : list) out.push_back(v.to_string());
    } else {
        const std::string t = to_string();
        if (!t.empty()) out.push_back(t);
    }
    return out;
}

CLParams::CLParams(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        const std::string str = argv[i];
        if (str.rfind("--", 0) == 0) {
            std::string name, val;
            const size_t eq = str.find('=');
            if (eq != std::string::npos) {
                name = str.substr(2, eq - 2);
                val = str.substr(eq + 1);
            } else {
                name = str.substr(2);
                if (i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0) {
                    val = argv[i + 1];
                    ++i;
                }
            }
            if (name.empty()) {
                success = false;
                error = "Problem with parameter: " + str;
                return;
            }
            Value v = Value::from_string(val);
            auto it = named.find(name);
            if (it == named.end()) {
                named[name] = v;
            } else {
                Value lst;
                lst.kind = Value::List;
                if (it->second.kind == Value::List) lst.list = it->second.list;
                else lst.list.push_back(it->second);
                if (v.kind == Value::List) lst.list.insert(lst.list.end(), v.list.begin(), v.list.end());
                else lst.list.push_back(v);
                it->second = lst;
            }
        } else {
            unnamed.push_back(str);
        }
    }
}

std::vector<int> int_list(const std::string& s) {
    std::vector<int> out;
    size_t pos = 0;
    while (pos <= s.size()) {
        size_t k = s.find(',', pos);
        if (k == std::string::npos) k = s.size();
        const std::string part = s.substr(pos, k - pos);
        if (!part.empty()) {
            long long x;
            out.push_back(parse_long(part, x) ? static_cast<int>(x) : 0);
        }
        pos = k + 1;
    }
    return out;
}

}  // namespace msamd_host
