// Shared helpers of the gfx950 HIP library behind include/msamd.h.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>

#include "../../include/msamd.h"

namespace msamd {

// Native 4-wide float vector (usable with the nontemporal builtins).
typedef float vf4 __attribute__((ext_vector_type(4)));

// Thread-local last-error string (see msamd_last_error).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
void clear_error();

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }
inline int64_t align_up(int64_t a, int64_t b) { return ceil_div(a, b) * b; }

// Carves a caller-owned workspace into 256-byte aligned sub-buffers.  The
// same sequence of take() calls on a null base computes the required size.
struct Carver {
    char* base;
    int64_t off = 0;
    explicit Carver(void* b) : base(static_cast<char*>(b)) {}
    template <typename T>
    T* take(int64_t count) {
        off = align_up(off, 256);
        T* p = base ? reinterpret_cast<T*>(base + off) : nullptr;
        off += count * static_cast<int64_t>(sizeof(T));
        return p;
    }
    int64_t bytes() const { return align_up(off, 256); }
};

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

}  // namespace msamd

#define MSAMD_HIP(expr)                                                                       \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess) {                                                               \
            msamd::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,               \
                             hipGetErrorString(e_));                                          \
            return MSAMD_ERR_HIP;                                                             \
        }                                                                                     \
    } while (0)

#define MSAMD_LAUNCH_CHECK()                                                                  \
    do {                                                                                      \
        hipError_t e_ = hipGetLastError();                                                    \
        if (e_ != hipSuccess) {                                                               \
            msamd::set_error("%s:%d: kernel launch failed: %s", __FILE__, __LINE__,           \
                             hipGetErrorString(e_));                                          \
            return MSAMD_ERR_HIP;                                                             \
        }                                                                                     \
    } while (0)

#define MSAMD_REQUIRE(cond, ...)                                                              \
    do {                                                                                      \
        if (!(cond)) {                                                                        \
            msamd::set_error(__VA_ARGS__);                                                    \
            return MSAMD_ERR_ARG;                                                             \
        }                                                                                     \
    } while (0)
