// Bandpass filter windows of any length n (the reference's window length
// n = chunk + 2*overlap follows its thread count and channel count,
// p_bandpass_filter.cpp:147-155, and is rarely a product of small primes).
//
// The n-point transform of Kernel_runner::apply (p_bandpass_filter.cpp:77-101)
// is computed by Bluestein's chirp-z identity with c_j = exp(i pi j^2 / n):
//   DFT_n(z)_k  = conj(c_k) (a (*) b)_k,   a_j = z_j conj(c_j),  b_m = c_m,
//   IDFT_n(V)_j = c_j (v (*) conj(b))_j,   v_k = V_k c_k,
// with (*) a linear convolution done as a length-P cyclic one, P = 2^k >=
// 2n - 1.  With V = G Z (G_k = H_k / n, the reference's gain) the chirps
// between the two convolutions cancel: v_k = G_k (a (*) b)_k.  So one window
// (two channels packed as z = x_a + i x_b, as in the four-step filter) is
//   u = G . IFFT_P(FFT_P(a) . B1) / P            (k < n, zero above)
//   y = c . IFFT_P(FFT_P(u) . B2) / P            (the core samples kept)
// with B1 = FFT_P(b), B2 = FFT_P(conj(b)) computed once per call.  Each
// FFT_P is a four-step P = P1 * P2: column passes of P1-point FFTs (stride
// P2) with the W_P twiddles, and row passes of P2-point FFTs that stay in
// the permuted [k1][k2] order -- the pointwise products with B1, B2 are
// taken in that order, so no pass ever restores natural order in between.
// Every short FFT runs in LDS as an iterative radix-2 transform (decimation
// in time on bit-reversed input forward, decimation in frequency inverse).
// Chirps, gains and twiddles are computed on the host in FP64 (the chirp
// phase from j^2 mod 2n in integers) and rounded once to FP32.
#include <cmath>
#include <vector>

#include "common.h"
#include "fft_radix.h"

namespace msamd {
namespace {

using fft::cf;
using fft::cadd;
using fft::csub;
using fft::cmul;
using fft::cconjmul;

constexpr int kBsThreads = 256;
constexpr int kBsTwLo = 8192;        // four-step twiddle W_P^e = hi[e >> 13] * lo[e & 8191]
constexpr int kBsMaxLog = 26;        // P <= 2^26 (P1, P2 <= 8192)
constexpr int64_t kBsBudget = 512LL << 20;  // bytes of window buffers per batch

struct BsPlan {
    int64_t M, N, cs, ov, n, N_out;
    int32_t k_sub, out_i16, in_i16;
    double qscale;
    int32_t lg1, lg2, P1, P2;
    int64_t P;
    int32_t pairs;
    int64_t nWin, nItems;
    int32_t U;    // window-pairs per batch
    int32_t CW;   // columns per workgroup in the column passes
    int64_t w0;   // first window of this call (nWin windows from there)
    int64_t xt0;  // recording time of input row 0
    int64_t yt0;  // output row of Y row 0
};

__device__ __forceinline__ int brev(int x, int lg) {
    return lg ? static_cast<int>(__brev(static_cast<unsigned>(x)) >> (32 - lg)) : 0;
}

// In-place radix-2 DIT over ncols columns of length L = 2^lg held at d[col*L],
// input in bit-reversed order, output natural; tw[m] = exp(-2 pi i m / L).
__device__ void lds_fft_dit(cf* d, int lg, int ncols, const cf* __restrict__ tw) {
    const int L = 1 << lg;
    const int half = L >> 1;
    for (int s = 1; s <= lg; ++s) {
        const int m = 1 << (s - 1);
        const int step = L >> s;
        for (int q = threadIdx.x; q < half * ncols; q += blockDim.x) {
            const int col = q >> (lg - 1);
            const int bb = q & (half - 1);
            const int p = bb & (m - 1);
            const int i1 = ((bb >> (s - 1)) << s) + p;
            cf* c = d + col * L;
            const cf t = cmul(c[i1 + m], tw[p * step]);
            const cf a = c[i1];
            c[i1] = cadd(a, t);
            c[i1 + m] = csub(a, t);
        }
        __syncthreads();
    }
}

// In-place radix-2 DIF, inverse direction (conjugate twiddles): natural input,
// bit-reversed output.
__device__ void lds_ifft_dif(cf* d, int lg, int ncols, const cf* __restrict__ tw) {
    const int L = 1 << lg;
    const int half = L >> 1;
    for (int s = lg; s >= 1; --s) {
        const int m = 1 << (s - 1);
        const int step = L >> s;
        for (int q = threadIdx.x; q < half * ncols; q += blockDim.x) {
            const int col = q >> (lg - 1);
            const int bb = q & (half - 1);
            const int p = bb & (m - 1);
            const int i1 = ((bb >> (s - 1)) << s) + p;
            cf* c = d + col * L;
            const cf a = c[i1], b = c[i1 + m];
            c[i1] = cadd(a, b);
            c[i1 + m] = cconjmul(csub(a, b), tw[p * step]);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ cf tw4(const cf* __restrict__ hi, const cf* __restrict__ lo, int64_t e) {
    return cmul(hi[e >> 13], lo[e & (kBsTwLo - 1)]);
}

// Column pass, forward: for columns n2 of one window-pair, the P1-point FFT
// over n1 (element j = n1*P2 + n2), times W_P^{n2*k1}, stored at [k1][n2].
// kFromInput: the window samples times conj(chirp) (zero past n and outside
// the recording), else the buffer's natural-order contents.
template <bool kFromInput, bool kI16>
__global__ void __launch_bounds__(kBsThreads) bs_col_fwd_kernel(
    const void* __restrict__ Xv, BsPlan p, int64_t item0, const cf* __restrict__ chirp,
    const cf* __restrict__ tw1, const cf* __restrict__ twHi, const cf* __restrict__ twLo,
    cf* __restrict__ buf) {
    extern __shared__ cf sd[];
    const int ngrp = p.P2 / p.CW;
    const int64_t item = item0 + blockIdx.x / ngrp;
    const int c0 = (blockIdx.x % ngrp) * p.CW;
    cf* B = buf + (item - item0) * p.P;
    const int pair = static_cast<int>(item % p.pairs);
    const int64_t w = p.w0 + item / p.pairs;
    const int64_t tbase = w * p.cs - p.ov;
    const int64_t ch0 = 2 * static_cast<int64_t>(pair);
    for (int q = threadIdx.x; q < p.CW * p.P1; q += blockDim.x) {
        const int col = q / p.P1, n1 = q % p.P1;
        const int64_t j = static_cast<int64_t>(n1) * p.P2 + c0 + col;
        cf v = {0.f, 0.f};
        if (kFromInput) {
            if (j < p.n) {
                const int64_t t = tbase + j;
                float a = 0.f, c = 0.f;
                if (t >= 0 && t < p.N) {
                    if (kI16) {
                        const int16_t* X = static_cast<const int16_t*>(Xv) + (t - p.xt0) * p.M + ch0;
                        a = static_cast<float>(X[0]);
                        if (ch0 + 1 < p.M) c = static_cast<float>(X[1]);
                    } else {
                        const float* X = static_cast<const float*>(Xv) + (t - p.xt0) * p.M + ch0;
                        a = X[0];
                        if (ch0 + 1 < p.M) c = X[1];
                    }
                }
                v = cconjmul(cf{a, c}, chirp[j]);
            }
        } else {
            v = B[j];
        }
        sd[col * p.P1 + brev(n1, p.lg1)] = v;
    }
    __syncthreads();
    lds_fft_dit(sd, p.lg1, p.CW, tw1);
    for (int q = threadIdx.x; q < p.CW * p.P1; q += blockDim.x) {
        const int col = q / p.P1, k1 = q % p.P1;
        const int n2 = c0 + col;
        B[static_cast<int64_t>(k1) * p.P2 + n2] =
            cmul(sd[col * p.P1 + k1], tw4(twHi, twLo, static_cast<int64_t>(n2) * k1));
    }
}

// Row pass: forward P2-point FFT of row k1, times the kernel spectrum K
// ([k1][k2] order), inverse P2-point FFT, stored back.  With K == nullptr
// the forward spectrum is stored to `spec` instead (kernel setup).
__global__ void __launch_bounds__(kBsThreads) bs_row_kernel(BsPlan p, int64_t item0,
                                                           const cf* __restrict__ tw2,
                                                           const cf* __restrict__ K,
                                                           cf* __restrict__ buf,
                                                           cf* __restrict__ spec) {
    extern __shared__ cf sd[];
    const int64_t item = item0 + blockIdx.x / p.P1;
    const int k1 = blockIdx.x % p.P1;
    cf* R = buf + (item - item0) * p.P + static_cast<int64_t>(k1) * p.P2;
    for (int q = threadIdx.x; q < p.P2; q += blockDim.x) sd[brev(q, p.lg2)] = R[q];
    __syncthreads();
    lds_fft_dit(sd, p.lg2, 1, tw2);
    if (!K) {
        for (int q = threadIdx.x; q < p.P2; q += blockDim.x)
            spec[static_cast<int64_t>(k1) * p.P2 + q] = sd[q];
        return;
    }
    const cf* Kr = K + static_cast<int64_t>(k1) * p.P2;
    for (int q = threadIdx.x; q < p.P2; q += blockDim.x) sd[q] = cmul(sd[q], Kr[q]);
    __syncthreads();
    lds_ifft_dif(sd, p.lg2, 1, tw2);
    for (int q = threadIdx.x; q < p.P2; q += blockDim.x) R[q] = sd[brev(q, p.lg2)];
}

// Column pass, inverse: times W_P^{-n2*k1}, P1-point inverse FFT over k1.
// kOutput false: u_j = G_j * value / P for j < n (zero above), stored in
// natural order.  kOutput true: y_j = chirp_j * value / P for the core
// samples j in [ov, ov + cs), written to the output (the two channels of the
// pair; subsampling and int16 quantization as the four-step filter).
template <bool kOutput, bool kOutI16>
__global__ void __launch_bounds__(kBsThreads) bs_col_inv_kernel(
    BsPlan p, int64_t item0, const cf* __restrict__ chirp, const float* __restrict__ G,
    const cf* __restrict__ tw1, const cf* __restrict__ twHi, const cf* __restrict__ twLo,
    cf* __restrict__ buf, void* __restrict__ Yv) {
    extern __shared__ cf sd[];
    const int ngrp = p.P2 / p.CW;
    const int64_t item = item0 + blockIdx.x / ngrp;
    const int c0 = (blockIdx.x % ngrp) * p.CW;
    cf* B = buf + (item - item0) * p.P;
    for (int q = threadIdx.x; q < p.CW * p.P1; q += blockDim.x) {
        const int col = q / p.P1, k1 = q % p.P1;
        const int n2 = c0 + col;
        sd[col * p.P1 + k1] = cconjmul(B[static_cast<int64_t>(k1) * p.P2 + n2],
                                       tw4(twHi, twLo, static_cast<int64_t>(n2) * k1));
    }
    __syncthreads();
    lds_ifft_dif(sd, p.lg1, p.CW, tw1);
    const float invP = 1.0f / static_cast<float>(p.P);   // a power of two: exact
    const int pair = static_cast<int>(item % p.pairs);
    const int64_t w = p.w0 + item / p.pairs;
    const int64_t ch0 = 2 * static_cast<int64_t>(pair);
    for (int q = threadIdx.x; q < p.CW * p.P1; q += blockDim.x) {
        const int col = q / p.P1, n1 = q % p.P1;
        const int64_t j = static_cast<int64_t>(n1) * p.P2 + c0 + col;
        const cf v = sd[col * p.P1 + brev(n1, p.lg1)];
        if (!kOutput) {
            cf u = {0.f, 0.f};
            if (j < p.n) {
                const float g = G[j] * invP;
                u = {v.x * g, v.y * g};
            }
            B[j] = u;
            continue;
        }
        const int64_t tl = j - p.ov;   // core-relative
        if (tl < 0 || tl >= p.cs) continue;
        int64_t to;
        if (p.k_sub == 1) {
            to = w * p.cs + tl;
        } else {
            if (tl % p.k_sub) continue;
            if (tl / p.k_sub >= p.cs / p.k_sub) continue;
            to = (w * p.cs) / p.k_sub + tl / p.k_sub;
        }
        if (to >= p.N_out) continue;
        const cf y = cmul(v, chirp[j]);
        const float vals[2] = {y.x * invP, y.y * invP};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int64_t ch = ch0 + c;
            if (ch >= p.M) break;
            if (kOutI16) {
                // multiply_by_factor (:207-209, float *= double), then the
                // float -> int16 conversion of the writer (truncation)
                const float s = static_cast<float>(static_cast<double>(vals[c]) * p.qscale);
                static_cast<int16_t*>(Yv)[(to - p.yt0) * p.M + ch] = static_cast<int16_t>(static_cast<int>(s));
            } else {
                static_cast<float*>(Yv)[(to - p.yt0) * p.M + ch] = vals[c];
            }
        }
    }
}

bool bs_make_plan(int64_t M, int64_t N, int64_t cs, int64_t ov, int k_sub, double qunit,
                  bool in_i16, BsPlan* out, int64_t w0 = 0, int64_t nwin = -1) {
    BsPlan p{};
    p.M = M;
    p.N = N;
    p.cs = cs;
    p.ov = ov;
    p.n = cs + 2 * ov;
    p.k_sub = k_sub;
    p.N_out = N / k_sub;
    p.qscale = qunit != 0 ? 1.0 / qunit : 0.0;
    p.out_i16 = qunit != 0;
    p.in_i16 = in_i16;
    int lg = 1;
    while ((int64_t{1} << lg) < 2 * p.n - 1) ++lg;
    if (lg > kBsMaxLog) return false;
    p.P = int64_t{1} << lg;
    p.lg1 = (lg + 1) / 2;
    p.lg2 = lg / 2;
    p.P1 = 1 << p.lg1;
    p.P2 = 1 << p.lg2;
    p.CW = 16384 / p.P1;
    if (p.CW > 4) p.CW = 4;
    if (p.CW < 1) p.CW = 1;
    if (p.CW > p.P2) p.CW = p.P2;
    p.pairs = static_cast<int32_t>((M + 1) / 2);
    p.w0 = w0;
    p.nWin = nwin >= 0 ? nwin : ceil_div(N, cs);
    p.nItems = p.nWin * p.pairs;
    int64_t U = kBsBudget / (p.P * static_cast<int64_t>(sizeof(cf)));
    if (U < 1) U = 1;
    if (U > p.nItems) U = p.nItems;
    p.U = static_cast<int32_t>(U);
    *out = p;
    return true;
}

struct BsBuffers {
    cf *buf, *K1, *K2, *chirp, *tw1, *tw2, *twHi, *twLo;
    float* G;
};

BsBuffers bs_carve(void* ws, const BsPlan& p, int64_t* bytes) {
    Carver c(ws);
    BsBuffers b{};
    b.buf = c.take<cf>(static_cast<int64_t>(p.U) * p.P);
    b.K1 = c.take<cf>(p.P);
    b.K2 = c.take<cf>(p.P);
    b.chirp = c.take<cf>(p.n);
    b.G = c.take<float>(p.n);
    b.tw1 = c.take<cf>(p.P1 / 2 > 0 ? p.P1 / 2 : 1);
    b.tw2 = c.take<cf>(p.P2 / 2 > 0 ? p.P2 / 2 : 1);
    b.twHi = c.take<cf>(p.P / kBsTwLo > 0 ? p.P / kBsTwLo : 1);
    b.twLo = c.take<cf>(kBsTwLo);
    if (bytes) *bytes = c.bytes();
    return b;
}

}  // namespace

// Entry points used by bandpass.hip for window lengths without a compiled
// four-step factorization.
int64_t bluestein_workspace_size(int64_t M, int64_t N, int64_t cs, int64_t ov) {
    BsPlan p;
    if (!bs_make_plan(M, N > 0 ? N : 1, cs, ov, 1, 0.0, false, &p)) return -1;
    int64_t bytes = 0;
    bs_carve(nullptr, p, &bytes);
    return bytes;
}

bool bluestein_supported(int64_t n) { return n >= 1 && 2 * n - 1 <= (int64_t{1} << kBsMaxLog); }

int bluestein_filter(const void* X, bool in_i16, int64_t M, int64_t N, int64_t xt0, int64_t w0,
                     int64_t w1, double samplerate, double freq_min, double freq_max,
                     double freq_wid, int64_t cs, int64_t ov, int k_sub, double qunit, void* Y,
                     int64_t yt0, void* workspace, int64_t workspace_bytes, hipStream_t st) {
    BsPlan p;
    if (!bs_make_plan(M, N, cs, ov, k_sub, qunit, in_i16, &p, w0, w1 - w0)) {
        set_error("bandpass: window length %lld is too long", (long long)(cs + 2 * ov));
        return MSAMD_ERR_UNSUPPORTED;
    }
    int64_t need = 0;
    BsBuffers b = bs_carve(workspace, p, &need);
    if (!workspace || workspace_bytes < need) {
        set_error("bandpass: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    p.xt0 = xt0;
    p.yt0 = yt0;
    if (N == 0 || p.nItems == 0) return MSAMD_OK;
    {
        auto lds_attr = [](const void* f, size_t bytes) {
            return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       static_cast<int>(bytes));
        };
        const size_t lc = sizeof(cf) * p.CW * p.P1, lr = sizeof(cf) * p.P2;
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_col_fwd_kernel<true, true>), lc));
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_col_fwd_kernel<true, false>), lc));
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_col_fwd_kernel<false, false>), lc));
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_col_inv_kernel<false, false>), lc));
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_col_inv_kernel<true, true>), lc));
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_col_inv_kernel<true, false>), lc));
        MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&bs_row_kernel), lr));
    }
    const long double kPi = 3.14159265358979323846264338327950288L;
    {
        // host tables: chirp, gain H/n, kernel sequence b, twiddles
        std::vector<cf> chirp(p.n);
        for (int64_t j = 0; j < p.n; ++j) {
            const int64_t r = static_cast<int64_t>((static_cast<unsigned __int128>(j) * j) %
                                                   static_cast<unsigned __int128>(2 * p.n));
            const long double ang = kPi * static_cast<long double>(r) / static_cast<long double>(p.n);
            chirp[j] = {static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang))};
        }
        std::vector<double> H(p.n);
        msamd_bandpass_kernel(p.n, samplerate, freq_min, freq_max, freq_wid, H.data());
        std::vector<float> G(p.n);
        for (int64_t k = 0; k < p.n; ++k) G[k] = static_cast<float>(H[k] * (1.0 / p.n));
        auto table = [&](int64_t len, int64_t period, int64_t step) {
            std::vector<cf> t(len > 0 ? len : 1);
            for (int64_t m = 0; m < len; ++m) {
                const long double ang = -2.0L * kPi *
                                        static_cast<long double>((m * step) % period) / period;
                t[m] = {static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang))};
            }
            return t;
        };
        const std::vector<cf> t1 = table(p.P1 / 2, p.P1, 1);
        const std::vector<cf> t2 = table(p.P2 / 2, p.P2, 1);
        const std::vector<cf> th = table(p.P / kBsTwLo, p.P, kBsTwLo);
        const std::vector<cf> tl = table(kBsTwLo, p.P, 1);
        // b_m = c_m for |m| < n, wrapped modulo P; its conjugate for the inverse
        std::vector<cf> kb(p.P, cf{0.f, 0.f});
        for (int64_t m = 0; m < p.n; ++m) {
            kb[m] = chirp[m];
            if (m > 0) kb[p.P - m] = chirp[m];
        }
        MSAMD_HIP(hipMemcpyAsync(b.chirp, chirp.data(), sizeof(cf) * p.n, hipMemcpyHostToDevice, st));
        MSAMD_HIP(hipMemcpyAsync(b.G, G.data(), sizeof(float) * p.n, hipMemcpyHostToDevice, st));
        MSAMD_HIP(hipMemcpyAsync(b.tw1, t1.data(), sizeof(cf) * t1.size(), hipMemcpyHostToDevice, st));
        MSAMD_HIP(hipMemcpyAsync(b.tw2, t2.data(), sizeof(cf) * t2.size(), hipMemcpyHostToDevice, st));
        MSAMD_HIP(hipMemcpyAsync(b.twHi, th.data(), sizeof(cf) * th.size(), hipMemcpyHostToDevice, st));
        MSAMD_HIP(hipMemcpyAsync(b.twLo, tl.data(), sizeof(cf) * tl.size(), hipMemcpyHostToDevice, st));
        // kernel spectra K1 = FFT(b), K2 = FFT(conj b), in the permuted order
        const size_t lds_col = sizeof(cf) * p.CW * p.P1;
        const size_t lds_row = sizeof(cf) * p.P2;
        const unsigned gcol = static_cast<unsigned>(p.P2 / p.CW);
        for (int which = 0; which < 2; ++which) {
            if (which == 1)
                for (auto& v : kb) v.y = -v.y;
            MSAMD_HIP(hipMemcpyAsync(b.buf, kb.data(), sizeof(cf) * p.P, hipMemcpyHostToDevice, st));
            bs_col_fwd_kernel<false, false><<<gcol, kBsThreads, lds_col, st>>>(
                nullptr, p, 0, b.chirp, b.tw1, b.twHi, b.twLo, b.buf);
            MSAMD_LAUNCH_CHECK();
            bs_row_kernel<<<static_cast<unsigned>(p.P1), kBsThreads, lds_row, st>>>(
                p, 0, b.tw2, nullptr, b.buf, which == 0 ? b.K1 : b.K2);
            MSAMD_LAUNCH_CHECK();
            MSAMD_HIP(hipStreamSynchronize(st));  // kb is rewritten / released
        }
    }
    const size_t lds_col = sizeof(cf) * p.CW * p.P1;
    const size_t lds_row = sizeof(cf) * p.P2;
    for (int64_t i0 = 0; i0 < p.nItems; i0 += p.U) {
        const int64_t nu = (p.nItems - i0) < p.U ? (p.nItems - i0) : p.U;
        const unsigned gcol = static_cast<unsigned>(nu * (p.P2 / p.CW));
        const unsigned grow = static_cast<unsigned>(nu * p.P1);
        if (in_i16)
            bs_col_fwd_kernel<true, true><<<gcol, kBsThreads, lds_col, st>>>(
                X, p, i0, b.chirp, b.tw1, b.twHi, b.twLo, b.buf);
        else
            bs_col_fwd_kernel<true, false><<<gcol, kBsThreads, lds_col, st>>>(
                X, p, i0, b.chirp, b.tw1, b.twHi, b.twLo, b.buf);
        MSAMD_LAUNCH_CHECK();
        bs_row_kernel<<<grow, kBsThreads, lds_row, st>>>(p, i0, b.tw2, b.K1, b.buf, nullptr);
        MSAMD_LAUNCH_CHECK();
        bs_col_inv_kernel<false, false><<<gcol, kBsThreads, lds_col, st>>>(
            p, i0, b.chirp, b.G, b.tw1, b.twHi, b.twLo, b.buf, nullptr);
        MSAMD_LAUNCH_CHECK();
        bs_col_fwd_kernel<false, false><<<gcol, kBsThreads, lds_col, st>>>(
            nullptr, p, i0, b.chirp, b.tw1, b.twHi, b.twLo, b.buf);
        MSAMD_LAUNCH_CHECK();
        bs_row_kernel<<<grow, kBsThreads, lds_row, st>>>(p, i0, b.tw2, b.K2, b.buf, nullptr);
        MSAMD_LAUNCH_CHECK();
        if (p.out_i16)
            bs_col_inv_kernel<true, true><<<gcol, kBsThreads, lds_col, st>>>(
                p, i0, b.chirp, b.G, b.tw1, b.twHi, b.twLo, b.buf, Y);
        else
            bs_col_inv_kernel<true, false><<<gcol, kBsThreads, lds_col, st>>>(
                p, i0, b.chirp, b.G, b.tw1, b.twHi, b.twLo, b.buf, Y);
        MSAMD_LAUNCH_CHECK();
    }
    return MSAMD_OK;
}

}  // namespace msamd
