// Whitening (a4-a6): covariance X X^T / (N-1), the whitening matrix W by the
// reference's deflated power method, and Y = quantize(W X, 1e-4).
//
// Reference: p_whiten (packages/ms3/p_whiten.cpp:45-206) and
// whitening_matrix_from_XXt / pca_from_XXt / iterate_XXt_to_get_top_component
// / subtract_out_rank_1_from_XXt (packages/mountainsortalg/pca.cpp:171-193,
// 408-430, 541-557, 634-650).
//
// Contractions run on MFMA as split-BF16 products: x = hi + lo with
// hi = bf16(x), lo = bf16(x - hi); x*y ~ hi*hi' + hi*lo' + lo*hi' (the lo*lo'
// term is dropped), accumulated in FP32 by v_mfma_f32_16x16x32_bf16.  The
// covariance sums FP32 block partials in FP64; the whitening matrix is FP64.
#include <cstdint>
#include <cstring>
#include <mutex>

#include "common.h"

#pragma clang fp contract(off)

namespace msamd {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split8(const float* v, bf16x8& h, bf16x8& l) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 hb = static_cast<__bf16>(v[e]);
        h[e] = hb;
        l[e] = static_cast<__bf16>(v[e] - static_cast<float>(hb));
    }
}

__device__ __forceinline__ f32x4 mfma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh,
                                       const bf16x8& bl, f32x4 acc) {
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    return acc;
}

// ---------------------------------------------------------------- covariance
// Operand fragment of v_mfma_f32_16x16x32_bf16 for a block of 16 channels
// over 32 samples: lane (i + 16*b) holds samples 8b..8b+7 of channel i.  The
// A and B operands use the same sample order, so any consistent assignment
// of samples to (b, e) yields the full sum; the D tile holds element
// (4*(lane/16) + r, lane % 16) in register r.

constexpr int kCovTPW = 16;  // output tiles per wavefront
constexpr int kCovWaves = 4;

// Streaming form for nB <= 24 (M <= 384): G workgroups of 12 waves share a
// time span and hold the output tiles of the upper triangle between them
// (up to 13 per wave, 168 registers: three waves per SIMD, which hide the
// sample loads and the per-step barrier better than two waves of 20 tiles);
// each streams the span once, staging the fragments of all channel blocks
// per 32 samples in a double-buffered LDS area while the next 32 samples
// load.  The G workgroups of one span run on the same XCD, so the span is
// read from HBM about once.
constexpr int kCovAllWaves = 12;
constexpr int kCovAllTPW = 13;
constexpr int kCovAllKSteps = 1200;  // 32-sample steps per workgroup (FP32 partial length)
constexpr int kCovAllItems = 2;      // (channel, 8-sample group) items per thread (4*Mp <= 1536)

struct CovPlan {
    int64_t M, N, span, nsplit;
    int32_t Mp, nB, nTiles, G;
    int32_t all;  // single-pass kernel
    const uint32_t* rowkeep;  // optional: bit t%32 of word t/32 = row t is kept
};

// Rows whose keep bit is clear read as zero (the artifact mask applied on the
// fly); no bitmap keeps every row.
__device__ __forceinline__ bool row_kept(const uint32_t* rk, int64_t t) {
    return !rk || ((rk[t >> 5] >> (t & 31)) & 1u);
}

inline CovPlan cov_plan(int64_t M, int64_t N) {
    CovPlan p{};
    p.M = M;
    p.N = N;
    p.Mp = static_cast<int32_t>(align_up(M, 16));
    p.nB = p.Mp / 16;
    p.nTiles = p.nB * (p.nB + 1) / 2;
    const int64_t ksteps = ceil_div(N, 32);
    p.all = p.nB <= 24 && 4 * p.Mp <= kCovAllItems * 64 * kCovAllWaves;
    if (p.all) {
        p.G = static_cast<int32_t>(ceil_div(p.nTiles, kCovAllWaves * kCovAllTPW));
        int64_t nsplit = ceil_div(ksteps, kCovAllKSteps);
        if (nsplit < 256) nsplit = 256;
        if (nsplit > ksteps) nsplit = ksteps;
        if (nsplit < 1) nsplit = 1;
        p.span = align_up(ceil_div(N, nsplit), 32);
        p.nsplit = ceil_div(N, p.span);
        if (p.nsplit < 1) p.nsplit = 1;
        return p;
    }
    p.G = static_cast<int32_t>(ceil_div(p.nTiles, kCovTPW * kCovWaves));
    int64_t nsplit = ceil_div(2048, p.G);
    if (nsplit > ksteps) nsplit = ksteps;
    if (nsplit < 1) nsplit = 1;
    p.span = align_up(ceil_div(N, nsplit), 32);
    p.nsplit = ceil_div(N, p.span);
    if (p.nsplit < 1) p.nsplit = 1;
    return p;
}

__device__ __forceinline__ void tile_ij(int t, int nB, int& I, int& J) {
    // row-major enumeration of the upper triangle (I <= J)
    int i = 0;
    int rowlen = nB;
    while (t >= rowlen) {
        t -= rowlen;
        ++i;
        --rowlen;
    }
    I = i;
    J = i + t;
}

__global__ void __launch_bounds__(256) cov_kernel(const float* __restrict__ X, CovPlan p,
                                                  float* __restrict__ part) {
    extern __shared__ bf16x8 lds[];
    bf16x8* fh = lds;
    bf16x8* fl = lds + p.nB * 64;
    const int g = static_cast<int>(blockIdx.x % p.G);
    const int64_t split = blockIdx.x / p.G;
    const int64_t tb = split * p.span;
    int64_t te = tb + p.span;
    if (te > p.N) te = p.N;
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int tile0 = (g * kCovWaves + wid) * kCovTPW;
    int tI[kCovTPW], tJ[kCovTPW];
    {
        int I = 0, J = 0;
        if (tile0 < p.nTiles) tile_ij(tile0, p.nB, I, J);
#pragma unroll
        for (int k = 0; k < kCovTPW; ++k) {
            tI[k] = I;
            tJ[k] = J;
            if (++J == p.nB) {
                ++I;
                J = I;
            }
        }
    }
    f32x4 acc[kCovTPW];
#pragma unroll
    for (int k = 0; k < kCovTPW; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nitems = 4 * p.Mp;
    for (int64_t t0 = tb; t0 < te; t0 += 32) {
        // t0 is a multiple of 32: one keep word covers the step's rows
        const uint32_t kw = p.rowkeep ? p.rowkeep[t0 >> 5] : ~0u;
        for (int q = threadIdx.x; q < nitems; q += blockDim.x) {
            const int m = q % p.Mp;
            const int b = q / p.Mp;
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t t = t0 + 8 * b + e;
                v[e] = (m < p.M && t < te && ((kw >> (8 * b + e)) & 1u))
                           ? __builtin_nontemporal_load(X + t * p.M + m)
                           : 0.0f;
            }
            bf16x8 h, l;
            split8(v, h, l);
            const int idx = (m >> 4) * 64 + (m & 15) + 16 * b;
            fh[idx] = h;
            fl[idx] = l;
        }
        __syncthreads();
        int lastI = -1;
        bf16x8 ah = {}, al = {};
#pragma unroll
        for (int k = 0; k < kCovTPW; ++k) {
            if (tile0 + k < p.nTiles) {
                if (tI[k] != lastI) {
                    ah = fh[tI[k] * 64 + lane];
                    al = fl[tI[k] * 64 + lane];
                    lastI = tI[k];
                }
                const bf16x8 bh = fh[tJ[k] * 64 + lane];
                const bf16x8 bl = fl[tJ[k] * 64 + lane];
                acc[k] = mfma3(ah, al, bh, bl, acc[k]);
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kCovTPW; ++k) {
        if (tile0 + k < p.nTiles) {
            float* dst = part + (split * p.nTiles + tile0 + k) * 256;
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc[k][r];
        }
    }
}

template <int TPW>
__global__ void __launch_bounds__(kCovAllWaves * 64) cov_all_kernel(const float* __restrict__ X,
                                                                   CovPlan p,
                                                                   float* __restrict__ part) {
    extern __shared__ bf16x8 lds[];  // [buffer][hi, lo][nB][64]
    // block b runs on XCD b % 8; the G tile groups of a span share that XCD
    const int64_t r = blockIdx.x / 8;
    const int g = static_cast<int>(r % p.G);
    const int64_t split = (r / p.G) * 8 + blockIdx.x % 8;
    if (split >= p.nsplit) return;
    const int64_t tb = split * p.span;
    int64_t te = tb + p.span;
    if (te > p.N) te = p.N;
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int tile0 = (g * kCovAllWaves + wid) * TPW;
    int I0 = 0, J0 = 0;
    if (tile0 < p.nTiles) tile_ij(tile0, p.nB, I0, J0);
    f32x4 acc[TPW];
#pragma unroll
    for (int k = 0; k < TPW; ++k) acc[k] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int nitems = 4 * p.Mp;
    const int plane = p.nB * 64;
    float v[kCovAllItems][8];
    // Each item's byte offset within a step (rows 8b..8b+7 of channel m);
    // items past the channels get an offset beyond any step's range.
    const uint32_t um4 = 4u * static_cast<uint32_t>(p.M);
    uint32_t vo[kCovAllItems];
    int vb[kCovAllItems];
#pragma unroll
    for (int it = 0; it < kCovAllItems; ++it) {
        const int q = threadIdx.x + it * kCovAllWaves * 64;
        const int m = q % p.Mp;
        const int b = q / p.Mp;
        const bool ok = q < nitems && m < p.M;
        vo[it] = ok ? static_cast<uint32_t>(8 * b) * um4 + 4u * static_cast<uint32_t>(m)
                    : 0x80000000u;
        vb[it] = ok ? 8 * b : 0;
    }
    // The keep words (bit t%32 of word t/32 set when row t is kept) come in
    // through a buffer load too, issued with the step's samples and applied
    // when the step is split, after its loads have landed.  Without a bitmap
    // the range is empty, the load returns 0 and nokeep sets every bit.
    const __amdgpu_buffer_rsrc_t rk = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<uint32_t*>(p.rowkeep), 0,
        p.rowkeep ? static_cast<int>(4 * ((p.N + 31) >> 5)) : 0, 0x00020000);
    const uint32_t nokeep = p.rowkeep ? 0u : ~0u;
    uint32_t kwv = 0;
    auto load = [&](int64_t t0) {
        // One buffer range per step: the step's rows below te.  A raw buffer
        // load past the range returns 0.0f, which covers the rows past te and
        // the items past the channels.
        const int64_t left = te - t0;
        const uint32_t rows = left < 32 ? static_cast<uint32_t>(left) : 32u;
        const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(X + t0 * p.M), 0, static_cast<int>(rows * um4), 0x00020000);
        // t0 is a multiple of 32: one keep word covers the step's rows
        kwv = __builtin_amdgcn_raw_buffer_load_b32(rk, static_cast<uint32_t>(t0 >> 5) * 4u, 0, 0);
#pragma unroll
        for (int it = 0; it < kCovAllItems; ++it) {
#pragma unroll
            for (int e = 0; e < 8; ++e)
                v[it][e] = __builtin_bit_cast(
                    float, __builtin_amdgcn_raw_buffer_load_b32(rs, vo[it] + e * um4, 0, 2));
        }
    };
    // rows outside the kept intervals read as zero
    auto keep = [&]() {
        const uint32_t kw = kwv | nokeep;
#pragma unroll
        for (int it = 0; it < kCovAllItems; ++it) {
            uint32_t kb = kw >> vb[it];
            asm volatile("" : "+v"(kb));  // one shift per step, not 8 masks held live
#pragma unroll
            for (int e = 0; e < 8; ++e) v[it][e] = ((kb >> e) & 1u) ? v[it][e] : 0.0f;
        }
    };
    load(tb);
    int buf = 0;
    for (int64_t t0 = tb; t0 < te; t0 += 32, buf ^= 1) {
        bf16x8* fh = lds + buf * 2 * plane;
        bf16x8* fl = fh + plane;
        keep();
#pragma unroll
        for (int it = 0; it < kCovAllItems; ++it) {
            const int q = threadIdx.x + it * kCovAllWaves * 64;
            if (q < nitems) {
                const int m = q % p.Mp;
                const int b = q / p.Mp;
                bf16x8 h, l;
                split8(v[it], h, l);
                const int idx = (m >> 4) * 64 + (m & 15) + 16 * b;
                fh[idx] = h;
                fl[idx] = l;
            }
        }
        __syncthreads();
        if (t0 + 32 < te) load(t0 + 32);  // next step's samples load during the MFMAs
        int I = I0, J = J0;
        bool needA = true;
        bf16x8 ah = {}, al = {};
#pragma unroll
        for (int k = 0; k < TPW; ++k) {
            if (tile0 + k < p.nTiles) {
                if (needA) {
                    ah = fh[I * 64 + lane];
                    al = fl[I * 64 + lane];
                    needA = false;
                }
                const bf16x8 bh = fh[J * 64 + lane];
                const bf16x8 bl = fl[J * 64 + lane];
                acc[k] = mfma3(ah, al, bh, bl, acc[k]);
            }
            if (++J == p.nB) {
                ++I;
                J = I;
                needA = true;
            }
        }
        // the other buffer is written next; the barrier after that write keeps
        // every wave's reads of this buffer ahead of its next overwrite
    }
#pragma unroll
    for (int k = 0; k < TPW; ++k) {
        if (tile0 + k < p.nTiles) {
            float* dst = part + (split * p.nTiles + tile0 + k) * 256;
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(4 * (lane >> 4) + r) * 16 + (lane & 15)] = acc[k][r];
        }
    }
}

// FP64 sum of the block partials in split order, / (N-1) when normalize,
// added to C when accumulate, mirrored.
__global__ void __launch_bounds__(256) cov_reduce_kernel(const float* __restrict__ part,
                                                         CovPlan p, double* __restrict__ C,
                                                         int accumulate, int normalize) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= static_cast<int64_t>(p.nTiles) * 256) return;
    const int tile = static_cast<int>(q >> 8);
    const int e = static_cast<int>(q & 255);
    int I, J;
    tile_ij(tile, p.nB, I, J);
    const int64_t ii = I * 16 + (e >> 4);
    const int64_t jj = J * 16 + (e & 15);
    if (ii >= p.M || jj >= p.M) return;
    if (I == J && ii > jj) return;  // diagonal tiles: upper half only, mirrored below
    // partials summed in split order; 8 loads in flight ahead of their adds
    double s = 0.0;
    for (int64_t sp0 = 0; sp0 < p.nsplit; sp0 += 8) {
        float t8[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (sp0 + k < p.nsplit) t8[k] = part[((sp0 + k) * p.nTiles + tile) * 256 + e];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (sp0 + k < p.nsplit) s = s + static_cast<double>(t8[k]);
    }
    if (normalize && p.N > 1) s = s / static_cast<double>(p.N - 1);
    if (accumulate) s = C[ii * p.M + jj] + s;
    C[ii * p.M + jj] = s;
    C[jj * p.M + ii] = s;
}

// C /= (N-1) for N > 1 (the accumulated sums of msamd_whiten_covariance_accumulate).
__global__ void __launch_bounds__(256) cov_finish_kernel(double* __restrict__ C, int64_t MM,
                                                         int64_t N) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q < MM && N > 1) C[q] = C[q] / static_cast<double>(N - 1);
}

// ------------------------------------------------------- whitening matrix
// The power method is a chain of dependent matrix-vector products (10 per
// component), so its speed is set by how fast one product and one exchange
// complete.  G is split by rows over NP workgroups of one wavefront each and
// held in their LDS (4 rows of 384 doubles in each of 96 workgroups at
// M = 384).
//
// Each row is owned by a group of L lanes (L = 16 up to M = 384): lane q of
// the group holds columns q, q+L, ... of the row in LDS and, in registers,
// the same columns of the current vector v.  A power step is
//   * the row's dot product with v: per-lane FMAs, then a DPP reduction over
//     the L lanes (every lane ends with the same sum);
//   * publication of the row's value of w = G v as one tagged 64-bit word:
//     the value rounded to 48 fraction bits with the step number (mod 16)
//     in the 4 low bits, written by a relaxed agent-scope atomic store;
//   * every lane polls its columns of the published w with relaxed
//     agent-scope atomic loads until each word carries the current tag.  A
//     64-bit atomic word is never torn, so a word with the right tag holds
//     this step's value: no flags, fences or workgroup barriers are needed;
//   * every row group forms |w|^2 by the same per-lane sums and the same DPP
//     reduction, so all row groups of all workgroups hold bit-identical
//     v = w / |w|.
// Published words alternate between two slots by step parity.  A lane
// waiting for step e can find step e-2 in its slot (a different tag) but
// never step e+2: the owner of a row publishes step e+1 only after its whole
// wavefront has read step e, and step e+2 needs every row of step e+1.
// Row groups without a row (padding of the last workgroup) neither publish
// nor poll.  The 2^-48 rounding of the exchanged values is far below the
// power method's own sensitivity to rounding.
//
// Deflation G <- (I-vv')G(I-vv') is applied to each workgroup's rows in the
// rank-1 form G - v u' - u v' + (v'u) v v' (u = G v), equal to the
// reference's two M^3 products in exact arithmetic.  Its exchange (u) also
// carries y = G v0, from which the first power step of the next component
// follows without another exchange:
//   G' v0 = y - v (u'v0) - u (v'v0) + (v'u) v (v'v0).
// So a component costs 10 exchanges instead of 11.
//
// The launch is an ordinary one (a cooperative launch is refused while
// another process on the GPU holds the cooperative queue); the NP groups
// are few enough to be co-resident on an otherwise idle device, and every
// wait is bounded: on timeout the error flag is raised, later waits end at
// once, every wavefront runs to the end and the call returns an error.
constexpr int kWMaxThreads = 512;
constexpr unsigned kWSpinLimit = 1u << 20;

// Polls a wait may spin before it gives up: kWSpinLimit, or
// MSAMD_W_SPIN_LIMIT from the environment (read at every call; a small
// value makes waits give up early, which tests use to drive the timeout
// path while every wavefront still runs to its end).
inline unsigned w_spin_limit() {
    if (const char* e = std::getenv("MSAMD_W_SPIN_LIMIT")) {
        const long long v = std::atoll(e);
        if (v >= 0 && v < static_cast<long long>(kWSpinLimit)) return static_cast<unsigned>(v);
    }
    return kWSpinLimit;
}

struct WPlan {
    int M, Mp, S, L, J, RW, NP, T;
};

inline WPlan w_plan(int64_t M) {
    WPlan p{};
    p.M = static_cast<int>(M);
    // lanes per row and columns per lane: one kernel instance per (L, J)
    if (M <= 64) { p.L = 16; p.J = 4; }
    else if (M <= 128) { p.L = 16; p.J = 8; }
    else if (M <= 256) { p.L = 16; p.J = 16; }
    else if (M <= 384) { p.L = 16; p.J = 24; }
    else if (M <= 768) { p.L = 32; p.J = 24; }
    else { p.L = 64; p.J = 16; }
    p.Mp = p.L * p.J;                 // zero-padded vector length
    p.S = p.Mp + 16;                  // LDS row stride: adjacent rows 128 B apart
    // one wavefront per workgroup (measured fastest at M = 384: 4 rows per
    // workgroup, 96 workgroups), at most 128 workgroups
    int rw = 64 / p.L;
    if (rw < ceil_div(M, 128)) rw = static_cast<int>(ceil_div(M, 128));
    const int cap_lds = static_cast<int>((150 * 1024 / 8 - 3 * p.Mp) / p.S);
    const int cap_thr = kWMaxThreads / p.L;
    if (rw > cap_lds) rw = cap_lds;
    if (rw > cap_thr) rw = cap_thr;
    if (rw < 1) rw = 1;
    p.RW = rw;
    p.NP = static_cast<int>(ceil_div(M, rw));
    p.T = static_cast<int>(align_up(static_cast<int64_t>(rw) * p.L, 64));
    return p;
}

inline size_t w_lds_bytes(const WPlan& p) {
    return sizeof(double) * (static_cast<size_t>(p.RW) * p.S + 3 * static_cast<size_t>(p.Mp));
}

// x read from the lane selected by the DPP control CTRL (all rows, all banks)
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double x) {
    const long long b = __builtin_bit_cast(long long, x);
    const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xF, 0xF, false);
    const int hi =
        __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) |
                                          static_cast<unsigned>(lo));
}

// Sum over the L lanes of a row group; every lane of the group gets the same
// value (each step adds a lane's value and its partner's, a + b == b + a).
template <int L>
__device__ __forceinline__ double group_sum(double x) {
    x = x + dpp_f64<0xB1>(x);   // quad_perm [1,0,3,2]
    x = x + dpp_f64<0x4E>(x);   // quad_perm [2,3,0,1]
    x = x + dpp_f64<0x141>(x);  // row_half_mirror: quads 0<->1, 2<->3
    x = x + dpp_f64<0x140>(x);  // row_mirror: half rows 0<->1
    if (L >= 32) x = x + __shfl_xor(x, 16, 64);
    if (L >= 64) x = x + __shfl_xor(x, 32, 64);
    return x;
}

typedef unsigned long long u64;

// x rounded to 48 fraction bits (nearest, ties away from zero) with tag in
// the 4 low bits
__device__ __forceinline__ u64 w_tagged(double x, unsigned tag) {
    const u64 b = __builtin_bit_cast(u64, x);
    return ((b + 8ull) & ~15ull) | tag;
}

// kLocal: every workgroup of the exchange runs on one XCD, so a word need
// only reach that XCD's L2 (a workgroup-scope store writes through the CU's
// L1 into it); the polls are device-scope loads, which miss the L1 and find
// the line in that L2.  Otherwise the store is device scope too.
template <bool kLocal>
__device__ __forceinline__ void w_publish(u64* slot, double x, unsigned tag) {
    __hip_atomic_store(slot, w_tagged(x, tag), __ATOMIC_RELAXED,
                       kLocal ? __HIP_MEMORY_SCOPE_WORKGROUP : __HIP_MEMORY_SCOPE_AGENT);
}

// Lane q's columns q + L*j of a published vector (pub points at column q of
// the slot); words beyond M are never written and read as 0.  Returns false
// after a timeout or when another wavefront has raised the error flag.
template <int L, int J, bool kLocal>
__device__ __forceinline__ bool w_fetch(const u64* pub, int q, int M, unsigned tag, double* out,
                                        int* err, unsigned limit) {
    unsigned spins = 0;
    for (;;) {
        bool ready = true;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const u64 b = __hip_atomic_load(const_cast<u64*>(pub) + j * L, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
            ready = ready && (j * L + q >= M || (b & 15ull) == tag);
            out[j] = __builtin_bit_cast(double, b & ~15ull);
        }
        if (ready) return true;
        ++spins;
        if (((spins & 63u) == 0 || spins > limit) &&
            (spins > limit ||
             __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

struct WBufs {
    u64* wbuf;  // [2][Mp] published rows of G v (tagged words)
    u64* ybuf;  // [2][Mp] published rows of G v0 (deflation exchange)
    int* err;     // 1: a wait timed out; 2: the XCD-local form found its workgroups on several XCDs
    u64* place;   // XCD-local form: working workgroups counted per XCD, 8 bits each
    unsigned spin_limit;
    int stride;   // workgroup b works as row block b / stride when b % stride == 0
    int spread;   // tests: the XCD-local form reports its workgroups as spread (error 2)
};

template <int L, int J, bool kLocal>
__global__ void __launch_bounds__(kWMaxThreads) whitening_power_kernel(
    const double* __restrict__ XXt, WPlan p, const double* __restrict__ v0g,
    double* __restrict__ comp, double* __restrict__ sig, WBufs bufs) {
    extern __shared__ double sh[];
    // XCD-local form: workgroups are dispatched round-robin over the XCDs, so
    // the workgroups b = 0, stride, 2 stride, ... all run on the first one;
    // the others end here
    if (kLocal && blockIdx.x % bufs.stride != 0) return;
    const int blk = kLocal ? static_cast<int>(blockIdx.x) / bufs.stride : static_cast<int>(blockIdx.x);
    const int M = p.M, S = p.S, Mp = p.Mp;
    const int tid = threadIdx.x;
    const int rg = tid / L;   // row group = local row
    const int q = tid % L;
    const int r0 = blk * p.RW;
    const int nrows = (M - r0) < p.RW ? (M - r0) : p.RW;
    const int row = r0 + rg;
    const bool has_row = rg < nrows;
    double* Gl = sh;                                   // [RW][S]
    double* v0l = Gl + static_cast<int64_t>(p.RW) * S; // [Mp]
    double* vl = v0l + Mp;                             // [Mp] staging of v (deflation)
    double* ul = vl + Mp;                              // [Mp] staging of u (deflation)
    for (int64_t e = tid; e < static_cast<int64_t>(p.RW) * S; e += blockDim.x) {
        const int r = static_cast<int>(e / S);
        const int c = static_cast<int>(e - static_cast<int64_t>(r) * S);
        Gl[e] = (r < nrows && c < M) ? XXt[static_cast<int64_t>(r0 + r) * M + c] : 0.0;
    }
    for (int i = tid; i < Mp; i += blockDim.x) v0l[i] = i < M ? v0g[i] : 0.0;
    if (kLocal) {
        // The exchange below relies on every working workgroup running on one
        // XCD.  Each counts itself in its XCD's 8-bit field of one word and
        // waits (bounded) until all NP have counted; if they are not all on
        // one XCD every workgroup sees the same word and stops with error 2
        // (the host then runs the device-scope form).
        __shared__ int placed;
        if (tid == 0) {
            const unsigned xcc = static_cast<unsigned>(__builtin_amdgcn_s_getreg((15 << 11) | 20)) & 7u;
            __hip_atomic_fetch_add(bufs.place, 1ull << (8 * xcc), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            int ok = 0;
            unsigned spins = 0;
            for (;;) {
                const u64 w = __hip_atomic_load(bufs.place, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                unsigned n = 0;
#pragma unroll
                for (int x = 0; x < 8; ++x) n += static_cast<unsigned>((w >> (8 * x)) & 0xffull);
                if (n >= static_cast<unsigned>(p.NP)) {
                    ok = (w == (static_cast<u64>(p.NP) << (8 * xcc)) && !bufs.spread) ? 1 : 2;
                    break;
                }
                ++spins;
                if (((spins & 63u) == 0 || spins > bufs.spin_limit) &&
                    (spins > bufs.spin_limit ||
                     __hip_atomic_load(bufs.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                    ok = 3;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            if (ok == 2) __hip_atomic_store(bufs.err, 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if (ok == 3) __hip_atomic_store(bufs.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            placed = ok;
        }
        __syncthreads();
        if (placed != 1) return;
    }
    __syncthreads();
    double* Grow = Gl + static_cast<int64_t>(has_row ? rg : 0) * S + q;
    u64* wpub = bufs.wbuf + q;
    u64* ypub = bufs.ybuf + q;

    double v[J];
#pragma unroll
    for (int j = 0; j < J; ++j) v[j] = v0l[j * L + q];
    unsigned epoch = 0;
    double sigma = 0.0;
    int s_first = 1;
    for (int k = 0; k < M; ++k) {
        for (int s = s_first; s <= 10; ++s) {
            ++epoch;
            const int par = epoch & 1;
            const unsigned tag = epoch & 15u;
            if (has_row) {
                // the row's J values first (independent LDS reads in flight together)
                double g[J];
#pragma unroll
                for (int j = 0; j < J; ++j) g[j] = Grow[j * L];
                double acc = 0.0;
#pragma unroll
                for (int j = 0; j < J; ++j) acc = fma(g[j], v[j], acc);
                acc = group_sum<L>(acc);
                if (q == 0) w_publish<kLocal>(bufs.wbuf + par * Mp + row, acc, tag);
                w_fetch<L, J, kLocal>(wpub + par * Mp, q, M, tag, v, bufs.err, bufs.spin_limit);
            }
            double ss = 0.0;
#pragma unroll
            for (int j = 0; j < J; ++j) ss = fma(v[j], v[j], ss);
            ss = group_sum<L>(ss);
            sigma = sqrt(ss);
            if (sigma != 0.0) {
                const double inv = 1.0 / sigma;
#pragma unroll
                for (int j = 0; j < J; ++j) v[j] = v[j] * inv;
            }
        }
        // component k (pca.cpp:549-557): v and its last norm
        if (rg == 0) {
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const int n = j * L + q;
                if (n >= r0 && n < r0 + nrows) comp[static_cast<int64_t>(n) * M + k] = v[j];
            }
        }
        if (blk == 0 && tid == 0) sig[k] = sigma;
        if (k == M - 1) break;

        // deflation exchange: u = G v and y = G v0
        ++epoch;
        const int par = epoch & 1;
        const unsigned tag = epoch & 15u;
        double u[J], y[J];
#pragma unroll
        for (int j = 0; j < J; ++j) u[j] = y[j] = 0.0;
        if (has_row) {
            double au = 0.0, ay = 0.0;
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const double g = Grow[j * L];
                au = fma(g, v[j], au);
                ay = fma(g, v0l[j * L + q], ay);
            }
            au = group_sum<L>(au);
            ay = group_sum<L>(ay);
            if (q == 0) {
                w_publish<kLocal>(bufs.wbuf + par * Mp + row, au, tag);
                w_publish<kLocal>(bufs.ybuf + par * Mp + row, ay, tag);
            }
            w_fetch<L, J, kLocal>(wpub + par * Mp, q, M, tag, u, bufs.err, bufs.spin_limit);
            w_fetch<L, J, kLocal>(ypub + par * Mp, q, M, tag, y, bufs.err, bufs.spin_limit);
        }
        double d_uv0 = 0.0, d_vv0 = 0.0, alpha = 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const double v0j = v0l[j * L + q];
            d_uv0 = fma(u[j], v0j, d_uv0);
            d_vv0 = fma(v[j], v0j, d_vv0);
            alpha = fma(v[j], u[j], alpha);
        }
        d_uv0 = group_sum<L>(d_uv0);
        d_vv0 = group_sum<L>(d_vv0);
        alpha = group_sum<L>(alpha);
        if (rg == 0) {
#pragma unroll
            for (int j = 0; j < J; ++j) {
                vl[j * L + q] = v[j];
                ul[j * L + q] = u[j];
            }
        }
        __syncthreads();
        if (has_row) {
            const double vi = vl[row], ui = ul[row];
#pragma unroll
            for (int j = 0; j < J; ++j) {
                const double g = Grow[j * L];
                const double cross = vi * u[j] + ui * v[j];
                Grow[j * L] = (g - cross) + alpha * (vi * v[j]);
            }
        }
        // staging reads complete before the next component's staging write
        __syncthreads();
        // first power step of component k+1: w = G' v0 from the exchanged vectors
        const double c_v = alpha * d_vv0 - d_uv0;
        double ss = 0.0;
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const double z = (y[j] + c_v * v[j]) - d_vv0 * u[j];
            y[j] = z;
            ss = fma(z, z, ss);
        }
        ss = group_sum<L>(ss);
        sigma = sqrt(ss);
        const double inv = sigma != 0.0 ? 1.0 / sigma : 1.0;
#pragma unroll
        for (int j = 0; j < J; ++j) v[j] = y[j] * inv;
        s_first = 2;
    }
}

template <int L, int J>
int launch_power(dim3 grid, dim3 block, size_t lds, hipStream_t st, const double* XXt,
                 const WPlan& wp, const double* v0, double* comp, double* sig, WBufs wb,
                 bool want_local) {
    const auto kloc = &whitening_power_kernel<L, J, true>;
    const auto kdev = &whitening_power_kernel<L, J, false>;
    MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kloc),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(kdev),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    // the XCD-local form when every workgroup fits on one XCD at once
    wb.stride = 1;
    if (want_local) {
        int dev = 0, nx = 1, cus = 0, per_cu = 0;  // (8 bits per XCD in the placement word)
        MSAMD_HIP(hipGetDevice(&dev));
        MSAMD_HIP(hipDeviceGetAttribute(&nx, hipDeviceAttributeNumberOfXccs, dev));
        MSAMD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        MSAMD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, reinterpret_cast<const void*>(kloc), static_cast<int>(block.x), lds));
        if (nx > 1 && nx <= 8 && grid.x < 256 &&
            static_cast<int64_t>(per_cu) * (cus / nx) >= static_cast<int64_t>(grid.x))
            wb.stride = nx;
    }
    if (wb.stride > 1)
        kloc<<<dim3(grid.x * static_cast<unsigned>(wb.stride)), block, lds, st>>>(XXt, wp, v0, comp,
                                                                                 sig, wb);
    else
        kdev<<<grid, block, lds, st>>>(XXt, wp, v0, comp, sig, wb);
    return MSAMD_OK;
}

// W[m, n] = sum_l (C[m,l] d_l) C[n,l], l ascending (mult_AB then
// mult_ABtrans, pca.cpp:648-649), d_l = 1/sqrt(sigma_l) or 0.
// comp is row-major C[i*M + k]; W is written column-major like the Mda.
// A 16x16 tile of W per workgroup: the rows of comp for its 16 m and 16 n
// are staged through LDS 32 components at a time, with d_l computed once per
// component; every element keeps the sum over l in ascending order and the
// same operations, ((C[m,l] d_l) C[n,l]) added to the running sum.
__global__ void __launch_bounds__(256) whitening_assemble_kernel(
    const double* __restrict__ comp, const double* __restrict__ sig, int M,
    double* __restrict__ W) {
    __shared__ double cmT[32][17], cnT[32][17], ds[32];
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int m0 = blockIdx.x * 16, n0 = blockIdx.y * 16;
    const int m = m0 + tx, n = n0 + ty;
    double acc = 0.0;
    for (int l0 = 0; l0 < M; l0 += 32) {
        // stage comp[m0 + r][l0 + c] and comp[n0 + r][l0 + c], r < 16, c < 32
        for (int e = threadIdx.x; e < 16 * 32; e += 256) {
            const int r = e >> 5, c = e & 31;
            const int l = l0 + c;
            cmT[c][r] = (m0 + r < M && l < M) ? comp[static_cast<int64_t>(m0 + r) * M + l] : 0.0;
            cnT[c][r] = (n0 + r < M && l < M) ? comp[static_cast<int64_t>(n0 + r) * M + l] : 0.0;
        }
        if (threadIdx.x < 32) {
            const int l = l0 + threadIdx.x;
            const double sv = l < M ? sig[l] : 0.0;
            ds[threadIdx.x] = sv != 0.0 ? 1.0 / sqrt(sv) : 0.0;
        }
        __syncthreads();
        const int lc = M - l0 < 32 ? M - l0 : 32;
        for (int c = 0; c < lc; ++c) acc = acc + (cmT[c][tx] * ds[c]) * cnT[c][ty];
        __syncthreads();
    }
    if (m < M && n < M) W[m + static_cast<int64_t>(M) * n] = acc;
}

// ------------------------------------------------------------------- apply
// Y[i, t] = sum_j X[j, t] W(j, i) with W(j, i) = Wcm[j + M*i] (the reference
// walks W column-major, p_whiten.cpp:174-182).  MFMA view: D(16 samples x
// 16 outputs) += A(16 samples x 32 channels) B(32 channels x 16 outputs).
//
// Precision: the 1e-4 output quantum of the reference needs products close
// to FP32 accuracy (bf16 hi+lo keeps 16 bits, which with cancellation inside
// W x reaches two quanta).  Both operands are split into two fp16 pieces,
// v = h + l with h = fp16(v), l = fp16(v - h), which keep 22 significant
// bits (|v - h - l| <= 2^-22 |v| while l is a normal number), and the three
// products of order >= 2^-11 are accumulated in FP32: hh + hl + lh (the
// dropped ll is O(2^-22)); fp16 x fp16 products are exact in FP32.  fp16 has
// a narrow exponent range, so both operands are scaled by powers of two
// first: W by 2^(14 - e_W) with 2^e_W <= max|W| < 2^(e_W+1), and X by
// 2^(14 - e_X) with e_X from a bound on |x| -- the covariance diagonal when
// the caller passes it (x_tj^2 <= sum_t x_tj^2 = (N-1) C_jj) or else the
// maximum found by a pass over X.  Scaled values stay below 2^15; values
// far below the bound lose relative precision in their low piece, but their
// absolute error stays below 2^-24 of the scale, far under the quantum.  The
// epilogue divides by the two scales exactly.
//
// Each workgroup keeps the two-piece fragments of CT column tiles (16*CT
// outputs) of W in LDS and streams 8 waves x 64 samples per pass over its
// 2048-sample time block.  Workgroups are numbered so that all column
// blocks of one time block run back to back on the same XCD (blockIdx % 8
// selects the XCD), and the time block is read once from HBM into that
// XCD's L2.
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

constexpr int kApplyWaves = 8;
constexpr int kApplyRowTiles = 3;    // 16-sample row tiles per wave per pass
constexpr int kApplyPasses = 6;      // passes per workgroup
constexpr int kApplyTB = kApplyWaves * kApplyRowTiles * 16 * kApplyPasses;  // 2304 samples
constexpr int kXcd = 8;
constexpr int kApplyPieces = 2;

struct ApplyPlan {
    int64_t M, N, nTB, nWG;
    int32_t Mp, nK, nColTiles, nCB, CT;
    double qunit;
    const uint32_t* rowkeep;  // optional row mask, as in CovPlan
    int32_t raw;              // 1: no 1e-4 quantize (the clip form)
};

inline ApplyPlan apply_plan(int64_t M, int64_t N, double qunit) {
    ApplyPlan p{};
    p.M = M;
    p.N = N;
    p.Mp = static_cast<int32_t>(align_up(M, 64));
    p.nK = p.Mp / 32;
    p.nColTiles = p.Mp / 16;
    // column tiles per workgroup: as many as the LDS holds (fewer column blocks
    // means fewer re-splits of every sample)
    const auto fits = [&](int ct) { return kApplyPieces * ct * p.nK * 1024 <= 150 * 1024; };
    if (p.nColTiles % 6 == 0 && fits(6))
        p.CT = 6;
    else {
        p.CT = 4;
        while (p.CT > 1 && !fits(p.CT)) p.CT >>= 1;
    }
    p.nCB = p.nColTiles / p.CT;
    p.qunit = qunit;
    p.nTB = ceil_div(N, kApplyTB);
    p.nWG = align_up(p.nTB, kXcd) * p.nCB;
    return p;
}

__device__ __forceinline__ void split2_8(const float* v, f16x8& h, f16x8& l) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const _Float16 hh = static_cast<_Float16>(v[e]);
        h[e] = hh;
        l[e] = static_cast<_Float16>(v[e] - static_cast<float>(hh));  // the difference is exact
    }
}

// 2^(14 - e) for 2^e <= b < 2^(e+1); 1 for b == 0 or a non-finite bound
__device__ __forceinline__ float pow2_scale(double b) {
    if (!(b > 0.0) || !isfinite(b)) return 1.0f;
    return ldexpf(1.0f, 14 - ilogb(b));
}

// Scale words (workspace): [0] W scale (float), [1] bound on |x| as float
// bits, [2] X scale (float).
// One workgroup: max|W| -> W scale; with the covariance, the |x| bound
// sqrt((N-1) max_j C_jj); without it the bound word is cleared for the
// X pass to raise.
__global__ void __launch_bounds__(256) apply_scale_kernel(const double* __restrict__ Wcm,
                                                          int64_t M, const double* __restrict__ C,
                                                          int64_t N, uint32_t* __restrict__ sc) {
    __shared__ double red[2][4];
    double wm = 0.0, cm = 0.0;
    {
        // 16 loads per thread in flight (the max is exact in any order)
        constexpr int kB = 16;
        double wk[kB];
#pragma unroll
        for (int k = 0; k < kB; ++k) wk[k] = 0.0;
        const int64_t bd = blockDim.x;
        for (int64_t e0 = threadIdx.x; e0 < M * M; e0 += kB * bd) {
            double v[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) v[k] = Wcm[e0 + k * bd < M * M ? e0 + k * bd : 0];
#pragma unroll
            for (int k = 0; k < kB; ++k) wk[k] = fmax(wk[k], fabs(v[k]));  // element 0 again past the end
        }
#pragma unroll
        for (int k = 0; k < kB; ++k) wm = fmax(wm, wk[k]);
    }
    if (C)
        for (int64_t j = threadIdx.x; j < M; j += blockDim.x) cm = fmax(cm, C[j + M * j]);
    for (int o = 32; o > 0; o >>= 1) {
        wm = fmax(wm, __shfl_xor(wm, o, 64));
        cm = fmax(cm, __shfl_xor(cm, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        red[0][threadIdx.x >> 6] = wm;
        red[1][threadIdx.x >> 6] = cm;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) {
            wm = fmax(wm, red[0][w]);
            cm = fmax(cm, red[1][w]);
        }
        sc[0] = __float_as_uint(pow2_scale(wm));
        // the bound rounded up to float so that it stays a bound
        const double xb = C ? sqrt(cm * static_cast<double>(N > 1 ? N - 1 : 1)) * (1.0 + 1e-6) : 0.0;
        sc[1] = __float_as_uint(static_cast<float>(xb));
    }
}

// max|x| over the kept rows, raised into sc[1] (non-negative floats order as
// their bit patterns).
__global__ void __launch_bounds__(256) apply_xmax_kernel(const float* __restrict__ X, int64_t M,
                                                         int64_t N, const uint32_t* __restrict__ rk,
                                                         uint32_t* __restrict__ sc) {
    // a workgroup takes groups of 32 rows (one keep word): 32*M contiguous
    // floats, read as float4 when M % 4 == 0; a row's samples count only when
    // its keep bit is set (all set: no per-element test)
    float m = 0.0f;
    const int64_t ngroups = (N + 31) / 32;
    const bool v4 = (M % 4) == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0;
    for (int64_t gq = blockIdx.x; gq < ngroups; gq += gridDim.x) {
        const int64_t t0 = 32 * gq;
        const int rows = N - t0 < 32 ? static_cast<int>(N - t0) : 32;
        const uint32_t kw = rk ? rk[gq] : ~0u;
        if (rows == 32 && kw == 0u) continue;
        const bool all = rows == 32 && kw == ~0u;
        const float* base = X + t0 * M;
        const int n = static_cast<int>(rows * M);  // <= 32 * 1024
        const int Mi = static_cast<int>(M);
        if (v4) {
            const float4* b4 = reinterpret_cast<const float4*>(base);
            const int M4 = Mi / 4;
            for (int f = threadIdx.x; f < n / 4; f += blockDim.x) {
                if (!all && !((kw >> (f / M4)) & 1u)) continue;
                const float4 v = b4[f];
                m = fmaxf(m, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
            }
        } else {
            for (int f = threadIdx.x; f < n; f += blockDim.x) {
                if (!all && !((kw >> (f / Mi)) & 1u)) continue;
                m = fmaxf(m, fabsf(base[f]));
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0f) atomicMax(sc + 1, __float_as_uint(m));
}

// X scale from the bound word.
__global__ void apply_xscale_kernel(uint32_t* __restrict__ sc) {
    if (threadIdx.x == 0) sc[2] = __float_as_uint(pow2_scale(__uint_as_float(sc[1])));
}

// Fragments of W: index ((ct * nK + kk) * 64 + lane) within each piece array;
// lane (i + 16b) of column tile ct holds W(32kk + 8b + e, 16ct + i), e = 0..7,
// scaled by the W scale.
__global__ void __launch_bounds__(256) apply_prep_kernel(const double* __restrict__ Wcm,
                                                         ApplyPlan p, const uint32_t* __restrict__ sc,
                                                         f16x8* __restrict__ fr) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = static_cast<int64_t>(p.nColTiles) * p.nK * 64;
    if (q >= total) return;
    const double sw = static_cast<double>(__uint_as_float(sc[0]));
    const int lane = static_cast<int>(q & 63);
    const int64_t tk = q >> 6;
    const int kk = static_cast<int>(tk % p.nK);
    const int ct = static_cast<int>(tk / p.nK);
    const int64_t i = ct * 16 + (lane & 15);
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int64_t j = kk * 32 + 8 * (lane >> 4) + e;
        v[e] = (i < p.M && j < p.M) ? static_cast<float>(Wcm[j + p.M * i] * sw) : 0.0f;
    }
    f16x8 h, l;
    split2_8(v, h, l);
    fr[q] = h;
    fr[total + q] = l;
}

template <bool kVec, int CT>
__global__ void __launch_bounds__(kApplyWaves * 64) apply_kernel(
    const float* __restrict__ X, ApplyPlan p, const f16x8* __restrict__ gfr,
    const uint32_t* __restrict__ sc, float* __restrict__ Y, int16_t* __restrict__ Yq) {
    extern __shared__ f16x8 alds[];
    const int nfr = CT * p.nK * 64;  // fragments of this column block, per piece
    // XCD-aware numbering: all column blocks of a time block share blockIdx % 8
    const int64_t w = blockIdx.x;
    const int64_t xcd = w % kXcd;
    const int64_t jx = w / kXcd;
    const int cb = static_cast<int>(jx % p.nCB);
    const int64_t tb = (jx / p.nCB) * kXcd + xcd;
    if (tb >= p.nTB) return;
    f16x8* fh = alds;
    f16x8* fl = alds + nfr;
    {
        const int64_t total = static_cast<int64_t>(p.nColTiles) * p.nK * 64;
        const int64_t base = static_cast<int64_t>(cb) * nfr;
        // four fragments of each piece per thread in flight before their
        // stores (the 9 rounds at 6 column tiles would otherwise each wait
        // out a load)
        constexpr int kB = 4;
        for (int q0 = threadIdx.x; q0 < nfr; q0 += kB * blockDim.x) {
            f16x8 h[kB], l[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const int q = q0 + k * static_cast<int>(blockDim.x);
                if (q < nfr) {
                    h[k] = gfr[base + q];
                    l[k] = gfr[total + base + q];
                }
            }
#pragma unroll
            for (int k = 0; k < kB; ++k) {
                const int q = q0 + k * static_cast<int>(blockDim.x);
                if (q < nfr) {
                    fh[q] = h[k];
                    fl[q] = l[k];
                }
            }
        }
    }
    const float sx = __uint_as_float(sc[2]);
    // 1 / (sx * sw): a power of two, so the unscaling is exact
    const double unscale = 1.0 / (static_cast<double>(sx) * static_cast<double>(__uint_as_float(sc[0])));
    // For a float y, y / 1e-4 in double equals y * 10000 exactly: y * 10000 has
    // at most 24 + 10 significant bits, and 1e-4 in double is 1e-4 * (1 + 4.8e-17),
    // a relative change below half an ulp, so the correctly rounded quotient is
    // y * 10000 itself.  The quantizer's product acc * (unscale * 10000) is
    // therefore the quotient of the FP64 formula, with no rounding.
    const double qscale = unscale * 10000.0;
    __syncthreads();
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int li = lane & 15;
    const int lb = lane >> 4;
    const int64_t tblock = tb * kApplyTB;
    for (int pass = 0; pass < kApplyPasses; ++pass) {
        const int64_t row0 = tblock + (static_cast<int64_t>(pass) * kApplyWaves + wid) * (kApplyRowTiles * 16);
        if (row0 >= p.N) break;
        // Loads are unconditional: rows past the end read the last row and
        // columns past the end read the row's last columns, which meet zero
        // rows of W; masked rows are computed and replaced by zero on output.
        const float* rowp[kApplyRowTiles];
#pragma unroll
        for (int rt = 0; rt < kApplyRowTiles; ++rt) {
            const int64_t t = row0 + rt * 16 + li;
            rowp[rt] = X + (t < p.N ? t : p.N - 1) * p.M;
        }
        typedef vf4 Samples[kApplyRowTiles][2];  // one step: 8 samples per row tile
        const auto load_step = [&](int kk, Samples& nx) {
            const int64_t j0 = static_cast<int64_t>(kk) * 32 + 8 * lb;
#pragma unroll
            for (int rt = 0; rt < kApplyRowTiles; ++rt) {
                if (kVec) {
                    // M % 8 == 0: j0 < M implies j0 + 8 <= M
                    const vf4* src = reinterpret_cast<const vf4*>(rowp[rt] + (j0 < p.M ? j0 : p.M - 8));
                    nx[rt][0] = src[0];
                    nx[rt][1] = src[1];
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        nx[rt][e >> 2][e & 3] = rowp[rt][j0 + e < p.M ? j0 + e : p.M - 1];
                }
            }
        };
        // two FP16 pieces of x * sx (exact: a power of two); each piece is one
        // fused multiply-add rounded once to FP16
        const auto split_step = [&](const Samples& nx, f16x8* h, f16x8* l) {
#pragma unroll
            for (int rt = 0; rt < kApplyRowTiles; ++rt) {
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float x = nx[rt][e >> 2][e & 3];
                    const _Float16 hh = static_cast<_Float16>(__builtin_fmaf(x, sx, 0.0f));
                    h[rt][e] = hh;
                    l[rt][e] = static_cast<_Float16>(__builtin_fmaf(x, sx, -static_cast<float>(hh)));
                }
            }
        };
        f32x4 acc[kApplyRowTiles][CT];
#pragma unroll
        for (int rt = 0; rt < kApplyRowTiles; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
        const auto mfma_step = [&](int kk, const f16x8* ah, const f16x8* al) {
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int fi = (ct * p.nK + kk) * 64 + lane;
                const f16x8 bh = fh[fi];
                const f16x8 bl = fl[fi];
#pragma unroll
                for (int rt = 0; rt < kApplyRowTiles; ++rt) {
                    // smallest terms first
                    f32x4 a = acc[rt][ct];
                    a = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[rt], bh, a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[rt], bl, a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[rt], bh, a, 0, 0, 0);
                    acc[rt][ct] = a;
                }
            }
        };
        // Two sample buffers in turn: the samples of step kk+2 are loaded
        // before the MFMAs of step kk, and step kk+1 is split after them
        // (steps past the end load the last step again and are not used).
        // The scheduling barriers keep the loads ahead of the MFMAs and the
        // splits after them; the scheduler would otherwise bring loads and
        // uses together to shorten register lifetimes.
        const int last = p.nK - 1;
        f16x8 h[kApplyRowTiles], l[kApplyRowTiles];
        Samples s0, s1;
        load_step(0, s0);
        load_step(last < 1 ? last : 1, s1);
        split_step(s0, h, l);
        // nK = Mp / 32 is even (Mp is a multiple of 64), so the steps come in
        // pairs and the loop body has no exit between a load and its use
        const auto step_pair = [&](int kk) {
            load_step(kk + 2 < last ? kk + 2 : last, s0);
            __builtin_amdgcn_sched_barrier(0);
            mfma_step(kk, h, l);
            __builtin_amdgcn_sched_barrier(0);
            split_step(s1, h, l);
            __builtin_amdgcn_sched_barrier(0);
            load_step(kk + 3 < last ? kk + 3 : last, s1);
            __builtin_amdgcn_sched_barrier(0);
            mfma_step(kk + 1, h, l);
            __builtin_amdgcn_sched_barrier(0);
            split_step(s0, h, l);
        };
        for (int kk = 0; kk < p.nK; kk += 2) step_pair(kk);
        // epilogue: undo the scales and quantize(y, 1e-4) =
        // floor(y/1e-4 + 0.5)*1e-4 in double, stored as float (masked rows:
        // zero); optional int16 (int)(y/q + 0.5).
#pragma unroll
        for (int rt = 0; rt < kApplyRowTiles; ++rt) {
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int64_t t = row0 + rt * 16 + 4 * lb + r;
                if (t >= p.N) continue;
                const bool kept = row_kept(p.rowkeep, t);
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    const int64_t i = static_cast<int64_t>(cb) * (CT * 16) + ct * 16 + li;
                    if (i >= p.M) continue;
                    const double a = static_cast<double>(acc[rt][ct][r]);
                    float yf = p.raw ? static_cast<float>(a * unscale)
                                     : static_cast<float>(floor(a * qscale + 0.5) * 0.0001);
                    if (!kept) yf = 0.0f;
                    if (Y) Y[t * p.M + i] = yf;
                    if (Yq) {
                        const double s = static_cast<double>(yf) / p.qunit + 0.5;
                        Yq[t * p.M + i] = static_cast<int16_t>(static_cast<int>(s));
                    }
                }
            }
        }
    }
}

}  // namespace
}  // namespace msamd

using namespace msamd;

extern "C" int64_t msamd_whiten_covariance_workspace_size(int64_t M, int64_t N) {
    if (M < 1 || N < 0) return -1;
    const CovPlan p = cov_plan(M, N > 0 ? N : 1);
    Carver c(nullptr);
    c.take<float>(p.nsplit * p.nTiles * 256);
    return c.bytes();
}

namespace {

int covariance_impl(const float* X, int64_t M, int64_t N, const uint32_t* rowkeep, double* XXt,
                    void* workspace, int64_t workspace_bytes, void* stream,
                    bool accumulate = false, bool normalize = true) {
    MSAMD_REQUIRE(M >= 1 && M <= 1024, "whiten_covariance: M=%lld outside 1..1024", (long long)M);
    MSAMD_REQUIRE(N >= 0, "whiten_covariance: negative N");
    hipStream_t st = as_stream(stream);
    if (N == 0) {
        if (!accumulate) MSAMD_HIP(hipMemsetAsync(XXt, 0, sizeof(double) * M * M, st));
        return MSAMD_OK;
    }
    CovPlan p = cov_plan(M, N);
    p.rowkeep = rowkeep;
    Carver c(workspace);
    float* part = c.take<float>(p.nsplit * p.nTiles * 256);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("whiten_covariance: workspace too small (%lld < %lld)",
                  (long long)workspace_bytes, (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    if (p.all) {
        const size_t lds = sizeof(bf16x8) * 4 * p.nB * 64;
        const unsigned grid = static_cast<unsigned>(ceil_div(p.nsplit, 8) * 8 * p.G);
        const int tpw = static_cast<int>(ceil_div(p.nTiles, kCovAllWaves * p.G));
#define MSAMD_COV_ALL(T)                                                                     \
    do {                                                                                     \
        MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&cov_all_kernel<T>),      \
                                      hipFuncAttributeMaxDynamicSharedMemorySize,            \
                                      static_cast<int>(lds)));                               \
        cov_all_kernel<T><<<grid, kCovAllWaves * 64, lds, st>>>(X, p, part);                  \
    } while (0)
        if (tpw <= 5) MSAMD_COV_ALL(5);
        else if (tpw <= 10) MSAMD_COV_ALL(10);
        else MSAMD_COV_ALL(kCovAllTPW);
#undef MSAMD_COV_ALL
        MSAMD_LAUNCH_CHECK();
    } else {
        const size_t lds = sizeof(bf16x8) * 2 * p.nB * 64;
        const unsigned grid = static_cast<unsigned>(p.nsplit * p.G);
        MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&cov_kernel),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(lds)));
        cov_kernel<<<grid, 256, lds, st>>>(X, p, part);
        MSAMD_LAUNCH_CHECK();
    }
    const int64_t nred = static_cast<int64_t>(p.nTiles) * 256;
    cov_reduce_kernel<<<static_cast<unsigned>(ceil_div(nred, 256)), 256, 0, st>>>(
        part, p, XXt, accumulate ? 1 : 0, normalize ? 1 : 0);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_whiten_covariance(const float* X, int64_t M, int64_t N, double* XXt,
                                       void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    return covariance_impl(X, M, N, nullptr, XXt, workspace, workspace_bytes, stream);
}

extern "C" int msamd_whiten_covariance_accumulate(const float* X, int64_t M, int64_t N,
                                                  const uint32_t* row_keep, double* XXt_sum,
                                                  int accumulate, void* workspace,
                                                  int64_t workspace_bytes, void* stream) {
    clear_error();
    return covariance_impl(X, M, N, row_keep, XXt_sum, workspace, workspace_bytes, stream,
                           accumulate != 0, false);
}

extern "C" int msamd_whiten_covariance_finish(double* XXt, int64_t M, int64_t N_total,
                                              void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= 1024 && N_total >= 0, "whiten_covariance_finish: bad shape");
    const int64_t MM = M * M;
    cov_finish_kernel<<<static_cast<unsigned>(ceil_div(MM, 256)), 256, 0, as_stream(stream)>>>(
        XXt, MM, N_total);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int msamd_whiten_covariance_rows(const float* X, int64_t M, int64_t N,
                                            const uint32_t* row_keep, double* XXt,
                                            void* workspace, int64_t workspace_bytes,
                                            void* stream) {
    clear_error();
    MSAMD_REQUIRE(row_keep, "whiten_covariance_rows: null row_keep");
    return covariance_impl(X, M, N, row_keep, XXt, workspace, workspace_bytes, stream);
}

extern "C" int64_t msamd_whitening_matrix_workspace_size(int64_t M) {
    if (M < 1) return -1;
    const WPlan wp = w_plan(M);
    Carver c(nullptr);
    c.take<double>(M * M);       // components
    c.take<double>(M);           // sigma
    c.take<double>(M);           // v0
    c.take<u64>(4 * wp.Mp);      // published rows of G v and of G v0, two slots each
    c.take<u64>(1);              // placement word
    c.take<int>(1);              // error flag
    return c.bytes();
}

extern "C" int msamd_whitening_matrix(const double* XXt, int64_t M, double* W, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= 1024, "whitening_matrix: M=%lld outside 1..1024", (long long)M);
    const WPlan wp = w_plan(M);
    Carver c(workspace);
    double* comp = c.take<double>(M * M);
    double* sig = c.take<double>(M);
    double* v0 = c.take<double>(M);
    u64* wbuf = c.take<u64>(4 * wp.Mp);
    u64* ybuf = wbuf + 2 * wp.Mp;
    u64* place = c.take<u64>(1);
    int* err = c.take<int>(1);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("whitening_matrix: workspace too small (%lld < %lld)",
                  (long long)workspace_bytes, (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    // Start vector normalize(sin(1..M)) (pca.cpp:545-548), computed on the host
    // with the C library's sin and a sequential double norm.
    double* h = static_cast<double*>(malloc(sizeof(double) * M));
    if (!h) {
        set_error("whitening_matrix: host allocation failed");
        return MSAMD_ERR_ARG;
    }
    double s2 = 0.0;
    for (int64_t i = 0; i < M; ++i) {
        h[i] = sin(static_cast<double>(i + 1));
        s2 = s2 + h[i] * h[i];
    }
    const double nrm = sqrt(s2);
    if (nrm != 0.0)
        for (int64_t i = 0; i < M; ++i) h[i] = h[i] / nrm;
    hipError_t e1 = hipMemcpyAsync(v0, h, sizeof(double) * M, hipMemcpyHostToDevice, st);
    hipError_t e2 = hipStreamSynchronize(st);
    free(h);
    MSAMD_HIP(e1);
    MSAMD_HIP(e2);
    // XXt is symmetric: row-major and column-major views coincide.
    MSAMD_REQUIRE(wp.T <= kWMaxThreads && wp.Mp >= M,
                  "whitening_matrix: plan out of range (M=%lld)", (long long)M);
    WBufs wb{};
    wb.wbuf = wbuf;
    wb.ybuf = ybuf;
    wb.err = err;
    wb.spin_limit = w_spin_limit();
    wb.place = place;
    // The exchange runs XCD-local (all workgroups on one XCD, words meeting in
    // its L2) when they fit there; MSAMD_W_XCD_LOCAL=0 keeps the device-scope
    // form.  If the local form finds its workgroups spread over several XCDs
    // it stops before any step and the device-scope form runs instead.
    const char* xl = std::getenv("MSAMD_W_XCD_LOCAL");
    const bool want_local = !(xl && atoi(xl) == 0);
    // MSAMD_W_XCD_LOCAL=spread (tests): the local form finds its workgroups
    // spread over XCDs, so the fallback to the device-scope form runs
    wb.spread = (xl && std::strcmp(xl, "spread") == 0) ? 1 : 0;
    const size_t lds = w_lds_bytes(wp);
    const dim3 grid(static_cast<unsigned>(wp.NP)), block(static_cast<unsigned>(wp.T));
    int herr = 0;
    for (int attempt = 0; attempt < 2; ++attempt) {
        const bool local = want_local && attempt == 0;
        // published words start at 0.  A word is first written at step 1 or 2
        // (G v slots) or 11 (G v0 slot), whose tags are not 0; later a waiting
        // lane can only find the tag of an earlier use (step e-2, or e-10 for
        // the G v0 slot), never equal to e mod 16.
        MSAMD_HIP(hipMemsetAsync(wbuf, 0, 4 * sizeof(u64) * wp.Mp, st));
        MSAMD_HIP(hipMemsetAsync(place, 0, sizeof(u64), st));
        MSAMD_HIP(hipMemsetAsync(err, 0, sizeof(int), st));
        int rc = MSAMD_OK;
        switch (wp.L * 100 + wp.J) {
        case 1604: rc = launch_power<16, 4>(grid, block, lds, st, XXt, wp, v0, comp, sig, wb, local); break;
        case 1608: rc = launch_power<16, 8>(grid, block, lds, st, XXt, wp, v0, comp, sig, wb, local); break;
        case 1616: rc = launch_power<16, 16>(grid, block, lds, st, XXt, wp, v0, comp, sig, wb, local); break;
        case 1624: rc = launch_power<16, 24>(grid, block, lds, st, XXt, wp, v0, comp, sig, wb, local); break;
        case 3224: rc = launch_power<32, 24>(grid, block, lds, st, XXt, wp, v0, comp, sig, wb, local); break;
        default: rc = launch_power<64, 16>(grid, block, lds, st, XXt, wp, v0, comp, sig, wb, local); break;
        }
        if (rc != MSAMD_OK) return rc;
        MSAMD_LAUNCH_CHECK();
        herr = 0;
        MSAMD_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, st));
        MSAMD_HIP(hipStreamSynchronize(st));
        if (herr != 2 || !local) break;
    }
    if (herr) {
        set_error("whitening_matrix: the exchange between the %d workgroups of the power method "
                  "timed out (other work on the device kept some of them from running); "
                  "msamd_whitening_matrix_exact needs no co-resident workgroups",
                  wp.NP);
        return MSAMD_ERR_TIMEOUT;
    }
    const dim3 ag(static_cast<unsigned>(ceil_div(M, 16)), static_cast<unsigned>(ceil_div(M, 16)));
    whitening_assemble_kernel<<<ag, 256, 0, st>>>(comp, sig, static_cast<int>(M), W);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int64_t msamd_whiten_apply_workspace_size(int64_t M) {
    if (M < 1) return -1;
    const ApplyPlan p = apply_plan(M, 1, 0.0);
    Carver c(nullptr);
    c.take<f16x8>(kApplyPieces * static_cast<int64_t>(p.nColTiles) * p.nK * 64);
    c.take<uint32_t>(4);
    return c.bytes();
}

namespace {

int apply_impl(const float* X, int64_t M, int64_t N, const double* W, const double* C,
               const uint32_t* rowkeep, float* Y, double quantization_unit, int16_t* Yq,
               void* workspace, int64_t workspace_bytes, void* stream, bool raw = false) {
    MSAMD_REQUIRE(M >= 1 && M <= 1024, "whiten_apply: M=%lld outside 1..1024", (long long)M);
    MSAMD_REQUIRE(N >= 0, "whiten_apply: negative N");
    MSAMD_REQUIRE(Y || Yq, "whiten_apply: no output");
    MSAMD_REQUIRE(!Yq || quantization_unit > 0, "whiten_apply: int16 output needs quantization_unit > 0");
    ApplyPlan p = apply_plan(M, N, quantization_unit);
    p.rowkeep = rowkeep;
    p.raw = raw ? 1 : 0;
    Carver c(workspace);
    const int64_t nfr = static_cast<int64_t>(p.nColTiles) * p.nK * 64;
    f16x8* fr = c.take<f16x8>(kApplyPieces * nfr);
    uint32_t* sc = c.take<uint32_t>(4);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("whiten_apply: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    apply_scale_kernel<<<1, 256, 0, st>>>(W, M, C, N, sc);
    MSAMD_LAUNCH_CHECK();
    if (!C && N > 0) {
        int64_t blocks = ceil_div(N, 32);
        if (blocks > 4096) blocks = 4096;
        apply_xmax_kernel<<<static_cast<unsigned>(blocks), 256, 0, st>>>(X, M, N, rowkeep, sc);
        MSAMD_LAUNCH_CHECK();
    }
    apply_xscale_kernel<<<1, 64, 0, st>>>(sc);
    MSAMD_LAUNCH_CHECK();
    apply_prep_kernel<<<static_cast<unsigned>(ceil_div(nfr, 256)), 256, 0, st>>>(W, p, sc, fr);
    MSAMD_LAUNCH_CHECK();
    if (N == 0) return MSAMD_OK;
    const size_t lds = sizeof(f16x8) * kApplyPieces * p.CT * p.nK * 64;
    const dim3 grid(static_cast<unsigned>(p.nWG));
    const bool vec = (M % 8 == 0);
#define MSAMD_APPLY_LAUNCH(V, CTv)                                                            \
    do {                                                                                      \
        MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&apply_kernel<V, CTv>),    \
                                      hipFuncAttributeMaxDynamicSharedMemorySize,             \
                                      static_cast<int>(lds)));                                \
        apply_kernel<V, CTv><<<grid, kApplyWaves * 64, lds, st>>>(X, p, fr, sc, Y, Yq);       \
    } while (0)
    if (p.CT == 6) {
        if (vec) MSAMD_APPLY_LAUNCH(true, 6); else MSAMD_APPLY_LAUNCH(false, 6);
    } else if (p.CT == 4) {
        if (vec) MSAMD_APPLY_LAUNCH(true, 4); else MSAMD_APPLY_LAUNCH(false, 4);
    } else if (p.CT == 2) {
        if (vec) MSAMD_APPLY_LAUNCH(true, 2); else MSAMD_APPLY_LAUNCH(false, 2);
    } else {
        if (vec) MSAMD_APPLY_LAUNCH(true, 1); else MSAMD_APPLY_LAUNCH(false, 1);
    }
#undef MSAMD_APPLY_LAUNCH
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_whiten_apply(const float* X, int64_t M, int64_t N, const double* W, float* Y,
                                  double quantization_unit, int16_t* Yq, void* workspace,
                                  int64_t workspace_bytes, void* stream) {
    clear_error();
    return apply_impl(X, M, N, W, nullptr, nullptr, Y, quantization_unit, Yq, workspace,
                      workspace_bytes, stream);
}

extern "C" int msamd_whiten_apply_rows(const float* X, int64_t M, int64_t N, const double* W,
                                       const uint32_t* row_keep, float* Y,
                                       double quantization_unit, int16_t* Yq, void* workspace,
                                       int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(row_keep, "whiten_apply_rows: null row_keep");
    return apply_impl(X, M, N, W, nullptr, row_keep, Y, quantization_unit, Yq, workspace,
                      workspace_bytes, stream);
}

extern "C" int msamd_whiten_clips(const float* clips, int64_t M, int64_t T, int64_t L,
                                  const double* W, float* clips_out, double quantization_unit,
                                  int16_t* clips_out_q, void* workspace, int64_t workspace_bytes,
                                  void* stream) {
    clear_error();
    MSAMD_REQUIRE(T >= 0 && L >= 0 && (T == 0 || L <= INT64_MAX / T), "whiten_clips: bad shape");
    if (T * L == 0) return MSAMD_OK;  // no clips: nothing to write
    // M x T x L clips are an M x (T*L) array: the apply without the 1e-4 quantize
    return apply_impl(clips, M, T * L, W, nullptr, nullptr, clips_out, quantization_unit,
                      clips_out_q, workspace, workspace_bytes, stream, true);
}

extern "C" int msamd_whiten_apply_cov(const float* X, int64_t M, int64_t N, const double* W,
                                      const double* XXt, const uint32_t* row_keep, float* Y,
                                      double quantization_unit, int16_t* Yq, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(XXt, "whiten_apply_cov: null covariance");
    return apply_impl(X, M, N, W, XXt, row_keep, Y, quantization_unit, Yq, workspace,
                      workspace_bytes, stream);
}

// ------------------------------------------------- apply without allocation
// msamd_whiten_apply_cov_noalloc keeps its scratch (the two-piece W
// fragments and the scale words) in this library-owned device array instead
// of a caller workspace.  Sized for M <= kNoAllocMaxM.  In host code the
// symbol is only a stand-in: its device address comes from
// hipGetSymbolAddress, once per device, and only that address is ever passed
// to a kernel.  Calls on one device share the array, so calls that overlap
// in time on one device (different streams, or threads) race; calls that are
// ordered on one stream do not.  Nothing in it has to start at zero: every
// call's scale kernel writes each word the later kernels read.
namespace {

constexpr int64_t kNoAllocMaxM = 512;
constexpr int64_t kNoAllocBytes = (int64_t{1} << 20) + 4096;  // fragments of M = 512 + scale words

}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// Device address of `scratch` on the current device (looked up on the first
// call per device).
int scratch_address(void** out) {
    int dev = -1;
    MSAMD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kNoAllocMaxDevices) {
        set_error("whiten_apply_cov_noalloc: device %d outside 0..%d", dev,
                  kNoAllocMaxDevices - 1);
        return MSAMD_ERR_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        MSAMD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_whiten_apply_cov_noalloc(const float* X, int64_t M, int64_t N,
                                              const double* W, const double* XXt,
                                              const uint32_t* row_keep, float* Y,
                                              double quantization_unit, int16_t* Yq,
                                              void* stream) {
    clear_error();
    MSAMD_REQUIRE(XXt, "whiten_apply_cov_noalloc: null covariance");
    if (M < 1 || M > kNoAllocMaxM) {
        set_error("whiten_apply_cov_noalloc: M=%lld outside 1..%lld (the library's scratch "
                  "holds the W fragments of at most %lld channels; use msamd_whiten_apply_cov "
                  "with a workspace)",
                  (long long)M, (long long)kNoAllocMaxM, (long long)kNoAllocMaxM);
        return MSAMD_ERR_CAPACITY;
    }
    const int64_t need = msamd_whiten_apply_workspace_size(M);
    if (need < 0 || need > kNoAllocBytes) {
        set_error("whiten_apply_cov_noalloc: scratch of %lld bytes needed, %lld held",
                  (long long)need, (long long)kNoAllocBytes);
        return MSAMD_ERR_CAPACITY;
    }
    void* ws = nullptr;
    const int rc = scratch_address(&ws);
    if (rc != MSAMD_OK) return rc;
    return apply_impl(X, M, N, W, XXt, row_keep, Y, quantization_unit, Yq, ws, kNoAllocBytes,
                      stream);
}
