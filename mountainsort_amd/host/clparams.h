// Command-line parameters of a MountainLab processor executable.
//
// Same grammar as the reference's CLParams (packages/ms3/common/clparams.cpp:
// 20-72): "--name=value" or "--name value" (the next token, when it does not
// start with "--"); anything else is an unnamed parameter.  A value is coerced
// like clp_string_to_variant (:100-120): a 64-bit integer, else a 32-bit
// integer, else a single-precision float, else "[a][b]" becomes a list, else a
// string.  A repeated name collects its values into a list (:54-68).
#pragma once

#include <map>
#include <string>
#include <vector>

namespace msamd_host {

struct Value {
    enum Kind { None, Str, Long, Float, List } kind = None;
    std::string s;
    long long l = 0;
    float f = 0.0f;
    std::vector<Value> list;

    static Value from_string(const std::string& str);  // with coercion
    static Value of_string(const std::string& str);    // no coercion
    static Value of_long(long long v);
    static Value of_double(double v);  // kept as a float, like the parsed form

    bool valid() const { return kind != None; }
    double to_double() const;        // QVariant::toDouble
    long long to_long() const;       // QVariant::toLongLong (rounds floats)
    int to_int() const { return static_cast<int>(to_long()); }
    std::string to_string() const;   // QVariant::toString
    std::vector<std::string> to_string_list() const;  // MLUtil::toStringList
};

struct CLParams {
    std::map<std::string, Value> named;
    std::vector<std::string> unnamed;
    bool success = true;
    std::string error;

    CLParams(int argc, char** argv);
    bool has(const std::string& k) const { return named.count(k) != 0; }
    Value get(const std::string& k) const {
        auto it = named.find(k);
        return it == named.end() ? Value() : it->second;
    }
};

// "3,1,2" -> {3, 1, 2} (QString::split(",", SkipEmptyParts) then toInt).
std::vector<int> int_list(const std::string& s);

}  // namespace msamd_host
