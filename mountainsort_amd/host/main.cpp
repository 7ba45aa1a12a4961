// msamd.mp -- MountainLab processor package executable for the MI355X front end.
//
// Command forms of the reference packages (packages/ms3/mountainsort3_main.cpp:
// 385-436,812-820):
//   msamd.mp spec                      all processor specs as JSON
//   msamd.mp spec <name>               one spec, or {} and exit code -1
//   msamd.mp requirements <name> ...   {"peak_ram_mb": x}
//   msamd.mp <name> --param=value ...  run; exit code 0 on success, -1 on failure
#include <unistd.h>

#include <climits>
#include <cstdio>
#include <string>

#include "clparams.h"
#include "processors.h"

using namespace msamd_host;

static std::string exe_path(const char* argv0) {
    char buf[PATH_MAX];
    const ssize_t n = readlink("/proc/self/exe", buf, sizeof buf - 1);
    if (n > 0) {
        buf[n] = '\0';
        return buf;
    }
    return argv0;
}

int main(int argc, char** argv) {
    CLParams clp(argc, argv);
    if (!clp.success) {
        fprintf(stderr, "%s\n", clp.error.c_str());
        return -1;
    }
    const std::string exe = exe_path(argv[0]);
    const std::string arg1 = clp.unnamed.size() > 0 ? clp.unnamed[0] : "";
    const std::string arg2 = clp.unnamed.size() > 1 ? clp.unnamed[1] : "";
    if (arg1 == "spec") {
        if (arg2.empty()) {
            printf("%s\n", spec_json(exe).dump().c_str());
            return 0;
        }
        const ProcessorSpec* p = find_processor(arg2);
        if (!p) {
            printf("{}\n");
            return -1;
        }
        printf("%s\n", processor_json(*p, exe).dump().c_str());
        return 0;
    }
    const bool requirements_only = arg1 == "requirements";
    const std::string name = requirements_only ? arg2 : arg1;
    const ProcessorSpec* p = find_processor(name);
    if (!p) {
        fprintf(stderr, "Unexpected processor name: %s\n", name.c_str());
        return -1;
    }
    if (requirements_only && !p->can_return_requirements) {
        fprintf(stderr, "Processor %s does not return requirements\n", name.c_str());
        return -1;
    }
    for (const auto& q : p->params)
        if (!q.optional && !clp.has(q.name)) {
            fprintf(stderr, "%s: missing required parameter %s\n", name.c_str(), q.name.c_str());
            return -1;
        }
    Ctx ctx(clp, *p);
    ctx.requirements_only = requirements_only;
    const bool ok = p->run(ctx);
    if (requirements_only) {
        if (ok) {
            Json r = Json::object();
            r["peak_ram_mb"] = ctx.peak_ram_mb;
            printf("%s\n", r.dump().c_str());
        } else {
            printf("Error in processor while trying to retrieve requirements.\n");
        }
    }
    if (!ok) {
        fprintf(stderr, "%s: %s\n", name.c_str(), ctx.err.c_str());
        return -1;
    }
    return 0;
}
