// Artifact masking (a7).
//
// Reference: p_mask_out_artifacts (packages/ms3/p_mask_out_artifacts.cpp:24-104):
//   norms[m, i] = sqrt(sum_{s in interval i} x[m, s]^2)            (:39-56)
//   per channel: mean0, sigma0 of norms[m, :] (MLCompute, :62-67)
//   norms[m, i] > mean0 + sigma0*threshold  clears use_it[i-1..i+1] (:68-75)
//   kept intervals are copied into a zeroed output                (:80-99)
// Every double sum runs in the reference's order (one thread per sum), so
// the interval decisions are exact.  use_it[-1] (written by the reference
// for i == 0, out of bounds) is not represented.
#include "common.h"

#pragma clang fp contract(off)

namespace msamd {
namespace {

__device__ __forceinline__ int64_t ceil_div_dev(int64_t a, int64_t b) { return (a + b - 1) / b; }

struct MaskBuffers {
    double* norms;     // [nint * M], interval-major
    double* thr;       // [M]
    int32_t* bad;      // [nint]
    int32_t* use_it;   // [nint + 1]
};

inline MaskBuffers mask_carve(void* ws, int64_t M, int64_t nint, int64_t* bytes) {
    Carver c(ws);
    MaskBuffers b{};
    b.norms = c.take<double>(nint * M);
    b.thr = c.take<double>(M);
    b.bad = c.take<int32_t>(nint + 2);
    b.use_it = c.take<int32_t>(nint + 1);
    if (bytes) *bytes = c.bytes();
    return b;
}

__global__ void __launch_bounds__(256) mask_norms_kernel(const float* __restrict__ X, int64_t M,
                                                         int64_t nint, int64_t isz,
                                                         double* __restrict__ norms) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= nint * M) return;
    const int64_t m = g % M;
    const int64_t i = g / M;
    const float* xp = X + i * isz * M + m;
    double s = 0.0;
    int64_t a = 0;
    constexpr int U = 16;
    for (; a + U <= isz; a += U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(xp + (a + u) * M);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double d = static_cast<double>(v[u]);
            s = s + d * d;
        }
    }
    for (; a < isz; ++a) {
        const double d = static_cast<double>(xp[a * M]);
        s = s + d * d;
    }
    norms[g] = sqrt(s);
}

// One thread per channel: MLCompute::stdev / mean of the channel's norms
// (mlcompute.cpp:62-85) and the decision threshold mean0 + sigma0*threshold.
__global__ void mask_stats_kernel(const double* __restrict__ norms, int64_t M, int64_t nint,
                                  double threshold, double* __restrict__ thr) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (m >= M) return;
    double s = 0.0, ss = 0.0;
    int64_t i = 0;
    constexpr int U = 16;
    for (; i + U <= nint; i += U) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = norms[(i + u) * M + m];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            s = s + v[u];
            ss = ss + v[u] * v[u];
        }
    }
    for (; i < nint; ++i) {
        const double v = norms[i * M + m];
        s = s + v;
        ss = ss + v * v;
    }
    const double mean0 = nint ? s / static_cast<double>(nint) : 0.0;
    const int ct = static_cast<int>(nint);
    const double sigma0 = (ct >= 2) ? sqrt((ss - s * s / ct) / (ct - 1)) : 0.0;
    thr[m] = mean0 + sigma0 * threshold;
}

// One wavefront per interval: bad[i] = any channel over its threshold.
__global__ void __launch_bounds__(256) mask_flag_kernel(const double* __restrict__ norms,
                                                        const double* __restrict__ thr,
                                                        int64_t M, int64_t nint,
                                                        int32_t* __restrict__ bad) {
    const int lane = threadIdx.x & 63;
    const int64_t i = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (i >= nint) return;
    bool over = false;
    for (int64_t m = lane; m < M; m += 64) over |= (norms[i * M + m] > thr[m]);
    const bool any = __any(over);
    if (lane == 0) bad[i] = any ? 1 : 0;
}

__global__ void mask_useit_kernel(const int32_t* __restrict__ bad, int64_t nint,
                                  int32_t* __restrict__ use_it) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i > nint) return;
    const bool b0 = (i < nint) && bad[i];
    const bool bl = (i >= 1) && (i - 1 < nint) && bad[i - 1];
    const bool br = (i + 1 < nint) && bad[i + 1];
    use_it[i] = (b0 || bl || br) ? 0 : 1;
}

__global__ void __launch_bounds__(256) mask_copy_kernel(const float* __restrict__ X, int64_t M,
                                                        int64_t N, int64_t nint, int64_t isz,
                                                        const int32_t* __restrict__ use_it,
                                                        float* __restrict__ Y) {
    const int64_t total = M * N;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    if (M % 4 == 0) {
        const int64_t n4 = total / 4;
        const vf4* x4 = reinterpret_cast<const vf4*>(X);
        vf4* y4 = reinterpret_cast<vf4*>(Y);
        for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < n4;
             q += stride) {
            const int64_t t = (q * 4) / M;
            const int64_t i = t / isz;
            const bool keep = (i < nint) && use_it[i];
            vf4 v = {0.f, 0.f, 0.f, 0.f};
            if (keep) v = __builtin_nontemporal_load(x4 + q);
            __builtin_nontemporal_store(v, y4 + q);
        }
    } else {
        for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < total;
             q += stride) {
            const int64_t t = q / M;
            const int64_t i = t / isz;
            const bool keep = (i < nint) && use_it[i];
            Y[q] = keep ? X[q] : 0.0f;
        }
    }
}

// Row bitmap of the decisions: bit t%32 of word t/32 is set when row t lies in
// a kept interval (the tail N mod interval_size is never kept).
__global__ void __launch_bounds__(256) mask_rowbits_kernel(const int32_t* __restrict__ use_it,
                                                           int64_t N, int64_t nint, int64_t isz,
                                                           uint32_t* __restrict__ rows) {
    const int64_t w = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (w >= ceil_div_dev(N, 32)) return;
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
        const int64_t t = 32 * w + b;
        const int64_t i = t / isz;
        if (t < N && i < nint && use_it[i]) bits |= 1u << b;
    }
    rows[w] = bits;
}

}  // namespace
}  // namespace msamd

using namespace msamd;

extern "C" int64_t msamd_mask_workspace_size(int64_t M, int64_t N, int64_t interval_size) {
    if (M < 1 || N < 0 || interval_size < 1) return -1;
    int64_t bytes = 0;
    mask_carve(nullptr, M, N / interval_size, &bytes);
    return bytes;
}

namespace {

int mask_impl(const float* X, int64_t M, int64_t N, double threshold, int64_t interval_size,
              float* Y, uint32_t* row_keep, int32_t* use_it, void* workspace,
              int64_t workspace_bytes, void* stream) {
    MSAMD_REQUIRE(threshold != 0 && interval_size != 0,
                  "mask_out_artifacts: threshold or interval_size is zero");
    MSAMD_REQUIRE(interval_size > 0 && M >= 1 && N >= 0, "mask_out_artifacts: bad shape");
    const int64_t nint = N / interval_size;
    int64_t need = 0;
    MaskBuffers b = mask_carve(workspace, M, nint, &need);
    if (!workspace || workspace_bytes < need) {
        set_error("mask_out_artifacts: workspace too small (%lld < %lld)",
                  (long long)workspace_bytes, (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    if (nint > 0) {
        mask_norms_kernel<<<static_cast<unsigned>(ceil_div(nint * M, 256)), 256, 0, st>>>(
            X, M, nint, interval_size, b.norms);
        MSAMD_LAUNCH_CHECK();
        mask_stats_kernel<<<static_cast<unsigned>(ceil_div(M, 64)), 64, 0, st>>>(
            b.norms, M, nint, threshold, b.thr);
        MSAMD_LAUNCH_CHECK();
        mask_flag_kernel<<<static_cast<unsigned>(ceil_div(nint * 64, 256)), 256, 0, st>>>(
            b.norms, b.thr, M, nint, b.bad);
        MSAMD_LAUNCH_CHECK();
    }
    mask_useit_kernel<<<static_cast<unsigned>(ceil_div(nint + 1, 256)), 256, 0, st>>>(
        b.bad, nint, b.use_it);
    MSAMD_LAUNCH_CHECK();
    if (row_keep && N > 0) {
        mask_rowbits_kernel<<<static_cast<unsigned>(ceil_div(ceil_div(N, 32), 256)), 256, 0, st>>>(
            b.use_it, N, nint, interval_size, row_keep);
        MSAMD_LAUNCH_CHECK();
    }
    if (Y && M * N > 0) {
        int64_t blocks = ceil_div(M * N / 4 + 1, 256);
        if (blocks > 16384) blocks = 16384;
        mask_copy_kernel<<<static_cast<unsigned>(blocks), 256, 0, st>>>(X, M, N, nint,
                                                                        interval_size, b.use_it, Y);
        MSAMD_LAUNCH_CHECK();
    }
    if (use_it) {
        MSAMD_HIP(hipMemcpyAsync(use_it, b.use_it, sizeof(int32_t) * (nint + 1),
                                 hipMemcpyDeviceToDevice, st));
    }
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_mask_out_artifacts(const float* X, int64_t M, int64_t N, double threshold,
                                        int64_t interval_size, float* Y, int32_t* use_it,
                                        void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(Y, "mask_out_artifacts: null output");
    return mask_impl(X, M, N, threshold, interval_size, Y, nullptr, use_it, workspace,
                     workspace_bytes, stream);
}

extern "C" int msamd_mask_rows(const float* X, int64_t M, int64_t N, double threshold,
                               int64_t interval_size, uint32_t* row_keep, int32_t* use_it,
                               void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(row_keep, "mask_rows: null row_keep");
    return mask_impl(X, M, N, threshold, interval_size, nullptr, row_keep, use_it, workspace,
                     workspace_bytes, stream);
}
