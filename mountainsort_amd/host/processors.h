// Processor registry of the msamd.mp package executable.
#pragma once

#include <functional>
#include <string>
#include <vector>

#include "clparams.h"
#include "json.h"

namespace msamd_host {

struct ParamSpec {
    std::string name;
    std::string description;
    bool optional = false;
    bool has_default = false;
    Json default_json;   // as printed in the spec
    Value default_value; // as seen by the processor when the parameter is absent
};

struct FileSpec {
    std::string name;
    std::string description;
    bool optional = false;
};

struct Ctx;

struct ProcessorSpec {
    std::string name;
    std::string version;
    std::string description;
    std::vector<FileSpec> inputs, outputs;
    std::vector<ParamSpec> params;
    bool can_return_requirements = false;
    std::function<bool(Ctx&)> run;

    ProcessorSpec& input(const std::string& n, const std::string& d = "", bool opt = false) {
        inputs.push_back({n, d, opt});
        return *this;
    }
    ProcessorSpec& output(const std::string& n, const std::string& d = "", bool opt = false) {
        outputs.push_back({n, d, opt});
        return *this;
    }
    ProcessorSpec& required(const std::string& n, const std::string& d = "") {
        ParamSpec p;
        p.name = n;
        p.description = d;
        params.push_back(p);
        return *this;
    }
    ProcessorSpec& optional(const std::string& n, const std::string& d = "") {
        ParamSpec p;
        p.name = n;
        p.description = d;
        p.optional = true;
        params.push_back(p);
        return *this;
    }
    ProcessorSpec& optional(const std::string& n, const std::string& d, long long def) {
        ParamSpec p;
        p.name = n;
        p.description = d;
        p.optional = true;
        p.has_default = true;
        p.default_json = Json(def);
        p.default_value = Value::of_long(def);
        params.push_back(p);
        return *this;
    }
    ProcessorSpec& optional(const std::string& n, const std::string& d, double def) {
        ParamSpec p;
        p.name = n;
        p.description = d;
        p.optional = true;
        p.has_default = true;
        p.default_json = Json(def);
        p.default_value = Value::of_double(def);
        params.push_back(p);
        return *this;
    }
    ProcessorSpec& optional(const std::string& n, const std::string& d, const char* def) {
        ParamSpec p;
        p.name = n;
        p.description = d;
        p.optional = true;
        p.has_default = true;
        p.default_json = Json(def);
        p.default_value = Value::from_string(def);
        params.push_back(p);
        return *this;
    }
};

// Per-invocation context: parameters with the spec defaults applied.
struct Ctx {
    const CLParams& clp;
    const ProcessorSpec& spec;
    bool requirements_only = false;
    double peak_ram_mb = 0.0;
    std::string err;

    Ctx(const CLParams& c, const ProcessorSpec& s) : clp(c), spec(s) {}
    Value param(const std::string& name) const;
    std::string str(const std::string& name) const { return param(name).to_string(); }
    double num(const std::string& name) const { return param(name).to_double(); }
    long long integer(const std::string& name) const { return param(name).to_long(); }
    bool fail(const std::string& msg) {
        err = msg;
        return false;
    }
};

const std::vector<ProcessorSpec>& processors();
const ProcessorSpec* find_processor(const std::string& name);

// {"processors": [...]} with exe_command / requirements_command built from exe_path.
Json spec_json(const std::string& exe_path);
Json processor_json(const ProcessorSpec& p, const std::string& exe_path);

}  // namespace msamd_host
