// Exact whitening matrix: W = C^{-1/2} in FP64 on the GPU.
//
// The reference's W (whitening_matrix_from_XXt, packages/mountainsortalg/
// pca.cpp:634-650) is C diag(sigma^-1/2) C' built from 10 unconverged power
// steps per component, which is chaotic in rounding for larger M.  This mode
// computes the matrix that W approximates, the symmetric inverse square root
// of the covariance, by the coupled Newton-Schulz iteration
//     A = C / ||C||_F,  Y_0 = A,  Z_0 = I,
//     T_k = (3 I - Z_k Y_k) / 2,  Y_{k+1} = Y_k T_k,  Z_{k+1} = T_k Z_k,
// for which Y_k -> A^{1/2} and Z_k -> A^{-1/2} (quadratically once
// ||I - Z_k Y_k|| < 1; the eigenvalues of A lie in (0, 1]), so
// W = Z / sqrt(||C||_F).  Channels whose covariance diagonal is zero (all-zero
// data; their rows and columns of C vanish) get identity rows in the
// iteration and zero rows and columns in W, as the reference's sigma == 0
// rule gives them.  Every product is an FP64 GEMM on the VALU.  The
// iteration stops when ||T_k - I||_F is below 1e-12 sqrt(M) or stops
// decreasing once below 1e-8 (rounding level).
#include <cmath>

#include "common.h"

namespace msamd {
namespace {

struct NsState {
    int done;
    int iters;
    int bad;        // non-finite residual
    int pad;
    double resid;
    double resid_prev;
    double scale;   // ||C'||_F
};

constexpr int kNsTile = 32;
constexpr int kNsThreads = 256;

// One workgroup: ||C'||_F of C with identity rows/columns for channels whose
// diagonal is zero; active[i] = (C_ii > 0).
__global__ void __launch_bounds__(kNsThreads) ns_scale_kernel(const double* __restrict__ C,
                                                             int M, int* __restrict__ active,
                                                             NsState* st) {
    __shared__ double red[kNsThreads / 64];
    for (int i = threadIdx.x; i < M; i += blockDim.x) active[i] = C[i + static_cast<int64_t>(M) * i] > 0.0;
    __syncthreads();
    double s = 0.0;
    for (int64_t e = threadIdx.x; e < static_cast<int64_t>(M) * M; e += blockDim.x) {
        const int i = static_cast<int>(e % M), j = static_cast<int>(e / M);
        const double v = (active[i] && active[j]) ? C[e] : (i == j ? 1.0 : 0.0);
        s += v * v;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kNsThreads / 64; ++w) t += red[w];
        st->scale = sqrt(t);
        st->done = 0;
        st->iters = 0;
        st->bad = !(t > 0.0) || !isfinite(t);
        st->resid = 0.0;
        st->resid_prev = 0.0;
    }
}

// Y = C'/s, Z = I.
__global__ void ns_init_kernel(const double* __restrict__ C, int M, const int* __restrict__ active,
                               const NsState* st, double* __restrict__ Y, double* __restrict__ Z) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(M) * M) return;
    const int i = static_cast<int>(e % M), j = static_cast<int>(e / M);
    const double v = (active[i] && active[j]) ? C[e] : (i == j ? 1.0 : 0.0);
    Y[e] = v / st->scale;
    Z[e] = i == j ? 1.0 : 0.0;
}

// Out = alpha * A * B + diag * I (column-major M x M); 32 x 32 output tile per
// workgroup, 2 x 2 outputs per thread, K staged through LDS in 32-wide slabs.
// Skipped once the iteration has converged.
__global__ void __launch_bounds__(kNsThreads) ns_gemm_kernel(const double* __restrict__ A,
                                                            const double* __restrict__ B,
                                                            double* __restrict__ Out, int M,
                                                            double alpha, double diag,
                                                            const NsState* st) {
    if (st->done || st->bad) return;
    __shared__ double As[kNsTile][kNsTile + 1];  // [row][k]
    __shared__ double Bs[kNsTile][kNsTile + 1];  // [k][col]
    const int r0 = blockIdx.x * kNsTile, c0 = blockIdx.y * kNsTile;
    const int tr = threadIdx.x % 16, tc = threadIdx.x / 16;  // outputs (r0 + tr + 16a, c0 + tc + 16b)
    double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    for (int k0 = 0; k0 < M; k0 += kNsTile) {
        for (int q = threadIdx.x; q < kNsTile * kNsTile; q += blockDim.x) {
            const int a = q % kNsTile, b = q / kNsTile;
            const int ra = r0 + a, ka = k0 + b;   // A(ra, ka): column-major, rows fastest
            As[a][b] = (ra < M && ka < M) ? A[ra + static_cast<int64_t>(M) * ka] : 0.0;
            const int kb = k0 + a, cb = c0 + b;   // B(kb, cb)
            Bs[a][b] = (kb < M && cb < M) ? B[kb + static_cast<int64_t>(M) * cb] : 0.0;
        }
        __syncthreads();
#pragma unroll 8
        for (int k = 0; k < kNsTile; ++k) {
            const double a0 = As[tr][k], a1 = As[tr + 16][k];
            const double b0 = Bs[k][tc], b1 = Bs[k][tc + 16];
            acc[0][0] = fma(a0, b0, acc[0][0]);
            acc[0][1] = fma(a0, b1, acc[0][1]);
            acc[1][0] = fma(a1, b0, acc[1][0]);
            acc[1][1] = fma(a1, b1, acc[1][1]);
        }
        __syncthreads();
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = r0 + tr + 16 * a, c = c0 + tc + 16 * b;
            if (r < M && c < M)
                Out[r + static_cast<int64_t>(M) * c] = alpha * acc[a][b] + (r == c ? diag : 0.0);
        }
}

// ||T - I||_F and the stopping rule (one workgroup).
__global__ void __launch_bounds__(kNsThreads) ns_check_kernel(const double* __restrict__ T, int M,
                                                             NsState* st, double tol) {
    if (st->done || st->bad) return;
    __shared__ double red[kNsThreads / 64];
    double s = 0.0;
    for (int64_t e = threadIdx.x; e < static_cast<int64_t>(M) * M; e += blockDim.x) {
        const int i = static_cast<int>(e % M), j = static_cast<int>(e / M);
        const double d = T[e] - (i == j ? 1.0 : 0.0);
        s += d * d;
    }
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kNsThreads / 64; ++w) t += red[w];
        const double r = sqrt(t);
        st->iters += 1;
        if (!isfinite(r)) {
            st->bad = 1;
        } else {
            const double prev = st->iters > 1 ? st->resid : INFINITY;
            st->resid_prev = prev;
            st->resid = r;
            // converged, or quadratic convergence has reached rounding level
            if (r < tol || (r < 1e-8 && r > 0.5 * prev)) st->done = 1;
        }
    }
}

// Y <- U, Z <- V (skipped once converged, like the products that fill U, V).
__global__ void ns_copy_kernel(const double* __restrict__ U, const double* __restrict__ V,
                               double* __restrict__ Y, double* __restrict__ Z, int64_t n,
                               const NsState* st) {
    if (st->done || st->bad) return;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    Y[e] = U[e];
    Z[e] = V[e];
}

// W = sym(Z) / sqrt(s) on active channels, zero elsewhere.
__global__ void ns_finish_kernel(const double* __restrict__ Z, int M,
                                 const int* __restrict__ active, const NsState* st,
                                 double* __restrict__ W) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= static_cast<int64_t>(M) * M) return;
    const int i = static_cast<int>(e % M), j = static_cast<int>(e / M);
    const double z = 0.5 * (Z[e] + Z[j + static_cast<int64_t>(M) * i]);
    W[e] = (active[i] && active[j]) ? z / sqrt(st->scale) : 0.0;
}

constexpr int kNsMaxIters = 100;
constexpr int kNsBatch = 6;   // iterations enqueued between host checks

}  // namespace
}  // namespace msamd

using namespace msamd;

extern "C" int64_t msamd_whitening_matrix_exact_workspace_size(int64_t M) {
    if (M < 1 || M > 4096) return -1;
    Carver c(nullptr);
    c.take<NsState>(1);
    c.take<int>(M);
    for (int k = 0; k < 5; ++k) c.take<double>(M * M);
    return c.bytes();
}

extern "C" int msamd_whitening_matrix_exact(const double* XXt, int64_t M, double* W,
                                            void* workspace, int64_t workspace_bytes,
                                            void* stream, int32_t* iterations_out,
                                            double* residual_out) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= 4096, "whitening_matrix_exact: M=%lld outside 1..4096",
                  (long long)M);
    MSAMD_REQUIRE(XXt && W, "whitening_matrix_exact: null matrix");
    Carver c(workspace);
    NsState* st = c.take<NsState>(1);
    int* active = c.take<int>(M);
    double* Y = c.take<double>(M * M);
    double* Z = c.take<double>(M * M);
    double* T = c.take<double>(M * M);
    double* U = c.take<double>(M * M);
    double* V = c.take<double>(M * M);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("whitening_matrix_exact: workspace too small (%lld < %lld)",
                  (long long)workspace_bytes, (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t s = as_stream(stream);
    const int m = static_cast<int>(M);
    const int64_t n = M * M;
    const unsigned ew = static_cast<unsigned>(ceil_div(n, 256));
    ns_scale_kernel<<<1, kNsThreads, 0, s>>>(XXt, m, active, st);
    MSAMD_LAUNCH_CHECK();
    ns_init_kernel<<<ew, 256, 0, s>>>(XXt, m, active, st, Y, Z);
    MSAMD_LAUNCH_CHECK();
    const dim3 gg(static_cast<unsigned>(ceil_div(M, kNsTile)),
                  static_cast<unsigned>(ceil_div(M, kNsTile)));
    const double tol = 1e-12 * std::sqrt(static_cast<double>(M));
    NsState h{};
    for (int it = 0; it < kNsMaxIters; it += kNsBatch) {
        for (int b = 0; b < kNsBatch; ++b) {
            // T = 1.5 I - 0.5 Z Y; stop test; U = Y T; V = T Z; (Y, Z) = (U, V).
            // Every kernel after the stop test returns at once once it is met,
            // so (Y, Z) hold the last iterate.
            ns_gemm_kernel<<<gg, kNsThreads, 0, s>>>(Z, Y, T, m, -0.5, 1.5, st);
            MSAMD_LAUNCH_CHECK();
            ns_check_kernel<<<1, kNsThreads, 0, s>>>(T, m, st, tol);
            MSAMD_LAUNCH_CHECK();
            ns_gemm_kernel<<<gg, kNsThreads, 0, s>>>(Y, T, U, m, 1.0, 0.0, st);
            MSAMD_LAUNCH_CHECK();
            ns_gemm_kernel<<<gg, kNsThreads, 0, s>>>(T, Z, V, m, 1.0, 0.0, st);
            MSAMD_LAUNCH_CHECK();
            ns_copy_kernel<<<ew, 256, 0, s>>>(U, V, Y, Z, n, st);
            MSAMD_LAUNCH_CHECK();
        }
        MSAMD_HIP(hipMemcpyAsync(&h, st, sizeof(NsState), hipMemcpyDeviceToHost, s));
        MSAMD_HIP(hipStreamSynchronize(s));
        if (h.done || h.bad) break;
    }
    if (iterations_out) *iterations_out = h.iters;
    if (residual_out) *residual_out = h.resid;
    if (h.bad) {
        set_error("whitening_matrix_exact: non-finite iterate (covariance not positive "
                  "definite on its nonzero channels?)");
        return MSAMD_ERR_UNSUPPORTED;
    }
    if (!h.done) {
        set_error("whitening_matrix_exact: no convergence in %d iterations (residual %g)",
                  h.iters, h.resid);
        return MSAMD_ERR_UNSUPPORTED;
    }
    ns_finish_kernel<<<ew, 256, 0, s>>>(Z, m, active, st, W);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}
