// FFT bandpass filter (a1-a3) with the reference's exact windows.
//
// Reference: p_bandpass_filter (packages/ms3/p_bandpass_filter.cpp:115-229),
// Kernel_runner::apply (:77-101), define_kernel (:421-447).  For every core
// chunk [t, t+cs) the window [t-ov, t+cs+ov) (zero outside the recording) is
// transformed with an n = cs + 2*ov point DFT, scaled by H[i]/n, inverse
// transformed, and the core kept.  The gain is even (H[i] == H[n-i] for even
// n), so the filter maps real to real: two channels are packed into one
// complex sequence z = x_a + i*x_b and come back as Re / Im.
//
// The n-point transform is a four-step FFT, n = N1*N2 with n = n1*N2 + n2 and
// k = k1 + N1*k2:
//   stage A  for each (n2, pair): N1-point DFT over n1, times W_n^{n2*k1}
//   stage B  for each (k1, pair): N2-point DFT over n2, times H[k1+N1*k2]/n,
//            inverse N2-point DFT
//   stage C  for each (n2, pair): times W_n^{-n2*k1}, inverse N1-point DFT,
//            core samples written out.
// Each sub-transform N = R1*R2 runs as two Stockham radix passes held in
// registers (fft_radix.h) with one LDS exchange between them.  Windows are
// processed in units of (channel group, window) whose scratch spectra stay
// in the MI355X last-level cache between the three stages.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

#include "common.h"
#include "fft_radix.h"

namespace msamd {
namespace {

using fft::cf;
using fft::cmul;
using fft::cconjmul;

// Tuning (compile-time; the defaults are the measured best on MI355X).
#ifndef MSAMD_BP_PBA
#define MSAMD_BP_PBA 8
#endif
#ifndef MSAMD_BP_PBB
#define MSAMD_BP_PBB 8
#endif
#ifndef MSAMD_BP_PBC
#define MSAMD_BP_PBC 8
#endif
#ifndef MSAMD_BP_CGMAX
#define MSAMD_BP_CGMAX 32
#endif
#ifndef MSAMD_BP_SCRATCH_MB
#define MSAMD_BP_SCRATCH_MB 128
#endif
#ifndef MSAMD_BP_ORDER
#define MSAMD_BP_ORDER 2
#endif
#ifndef MSAMD_BP_LDS_PAD
#define MSAMD_BP_LDS_PAD 1
#endif
#ifndef MSAMD_BP_TW_LDS
#define MSAMD_BP_TW_LDS 1
#endif
constexpr int kPBA = MSAMD_BP_PBA;  // channel pairs per workgroup, stage A
constexpr int kPBB = MSAMD_BP_PBB;  // channel pairs per workgroup, stage B
constexpr int kPBC = MSAMD_BP_PBC;  // channel pairs per workgroup, stage C
constexpr int kCGMax = MSAMD_BP_CGMAX;  // channel pairs per unit (a multiple of every kPB*)
// Block order of stages A and C: n2 fastest (0), or the pair blocks of one
// (unit, n2) adjacent (1) so that they share the same input rows and output
// rows in time.
constexpr int kOrder = MSAMD_BP_ORDER;
// LDS rows of the register-pass exchange padded to an odd length: the rows a
// half-wavefront writes then start on different banks.
__host__ __device__ constexpr int lds_row(int R) { return MSAMD_BP_LDS_PAD ? (R | 1) : R; }
// Four-step twiddles W_n^{n2*k1} of a column n2 at k1 = j + r*R1 are formed as
// W_n^{n2*j} * W_n^{n2*R1*r}: one table product per thread plus a per-workgroup
// LDS table over r (instead of two table gathers per element).
constexpr bool kTwLds = MSAMD_BP_TW_LDS != 0;
constexpr int kCGAlign = kPBA > kPBB ? (kPBA > kPBC ? kPBA : kPBC) : (kPBB > kPBC ? kPBB : kPBC);
static_assert(kCGMax % kPBA == 0 && kCGMax % kPBB == 0 && kCGMax % kPBC == 0 &&
                  kCGAlign % kPBA == 0 && kCGAlign % kPBB == 0 && kCGAlign % kPBC == 0,
              "pairs per unit must be a multiple of the pairs per workgroup of every stage");
constexpr int kTwLo = 512;       // four-step twiddle: W = hi[e >> 9] * lo[e & 511]

struct BpPlan {
    int64_t M, N, cs, ov, n, N_out;
    int32_t k_sub;
    int32_t out_i16;
    int32_t in_i16;
    double qscale;   // 1/quantization_unit (0 -> float output)
    int32_t N1, N2;
    int32_t P;       // channel pairs
    int32_t CG;      // pairs per unit (multiple of kPB)
    int32_t nGroups;
    int64_t nWin, nUnits;
    int32_t U;       // units per batch
    int32_t nTwHi;
    int32_t streams; // batches go round-robin over this many streams, one spectra buffer each
    int64_t w0;      // first window of this call (nWin windows from there)
    int64_t xt0;     // recording time of input row 0
    int64_t yt0;     // output row of Y row 0
};

// Factorizations with kernels compiled in: n = (R1*R2) * (Q1*Q2).
struct Shape {
    int R1, R2, Q1, Q2;
};
constexpr Shape kShapes[] = {
    {20, 20, 24, 25},  // 400 x 600 = 240000 (chunk 120000, overlap 60000)
};

// Workgroup index -> (unit within the batch, pair block, column n2).
// Order 2 is XCD-aware: workgroups are dispatched round-robin over the eight
// XCDs, so the pair blocks of one column are given ids that differ by
// multiples of 8 inside one run of 8*npb ids -- the same XCD (and L2) at
// nearly the same time.  The grid is rounded up to whole runs; ids past the
// last column return *u = -1.
__device__ __forceinline__ void block_unit_pb_col(int64_t b, int npb, int N2, int64_t* u, int* pb,
                                                  int* n2) {
    if (kOrder == 2) {
        const int64_t run = b / (8 * npb);
        const int q = static_cast<int>(b % (8 * npb));
        *pb = q / 8;
        const int64_t col = run * 8 + (q % 8);
        *n2 = static_cast<int>(col % N2);
        *u = col / N2;
    } else if (kOrder == 1) {
        *pb = static_cast<int>(b % npb);
        b /= npb;
        *n2 = static_cast<int>(b % N2);
        *u = b / N2;
    } else {
        *n2 = static_cast<int>(b % N2);
        b /= N2;
        *pb = static_cast<int>(b % npb);
        *u = b / npb;
    }
}

// ------------------------------------------------------------------ stage A
// Workgroup: one n2, kPB pairs of one unit; thread (j, pl) with pl fastest.
// pass 1: radix R1 over x[j + r*R2] (R2 tasks), pass 2: radix R2 (R1 tasks).
template <int R1, int R2, bool kI16, int kPB>
__global__ void __launch_bounds__(kPB * (R1 > R2 ? R1 : R2)) bp_stage_a(
    const void* __restrict__ Xv, BpPlan p, int64_t unit0, int64_t nu, const cf* __restrict__ twN1,
    const cf* __restrict__ twHi, const cf* __restrict__ twLo, cf* __restrict__ S) {
    constexpr int N1 = R1 * R2;
    constexpr int RA = lds_row(R1);
    __shared__ cf buf[R2 * RA * kPB];
    __shared__ cf tw1[N1];
    const int tid = threadIdx.x;
    for (int q = tid; q < N1; q += blockDim.x) tw1[q] = twN1[q];
    const int pl = tid % kPB;
    const int j = tid / kPB;
    // block -> (unit, pair block, n2)
    const int npb = p.CG / kPB;
    int n2, pb;
    int64_t u;
    block_unit_pb_col(blockIdx.x, npb, p.N2, &u, &pb, &n2);
    if (u >= nu) return;  // whole workgroup: past the last column of the batch
    u += unit0;
    __shared__ cf stp[R2];  // W_n^{n2*R1*r}, r < R2 (n2*R1*r < n)
    if (kTwLds && tid < R2) {
        const int e = n2 * R1 * tid;
        stp[tid] = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
    }
    const int g = static_cast<int>(u / p.nWin);
    const int64_t w = p.w0 + u % p.nWin;
    const int pair = g * p.CG + pb * kPB + pl;   // global pair index
    const int ch0 = 2 * pair;
    const int64_t tbase = w * p.cs - p.ov;       // time of window sample 0
    const bool even_m = (p.M % 2) == 0;
    cf v[R1];
    if (j < R2) {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const int64_t tl = static_cast<int64_t>(j + r * R2) * p.N2 + n2;
            const int64_t t = tbase + tl;
            float a = 0.f, c = 0.f;
            if (pair < p.P && t >= 0 && t < p.N) {
                if (kI16) {
                    const int16_t* X = static_cast<const int16_t*>(Xv) + (t - p.xt0) * p.M + ch0;
                    if (even_m) {  // both channels in one aligned 32-bit word
                        const uint32_t w2 = *reinterpret_cast<const uint32_t*>(X);
                        a = static_cast<float>(static_cast<int16_t>(w2 & 0xffffu));
                        c = static_cast<float>(static_cast<int16_t>(w2 >> 16));
                    } else {
                        a = static_cast<float>(X[0]);
                        if (ch0 + 1 < p.M) c = static_cast<float>(X[1]);
                    }
                } else {
                    const float* X = static_cast<const float*>(Xv) + (t - p.xt0) * p.M + ch0;
                    if (even_m) {
                        const float2 v2 = *reinterpret_cast<const float2*>(X);
                        a = v2.x;
                        c = v2.y;
                    } else {
                        a = X[0];
                        if (ch0 + 1 < p.M) c = X[1];
                    }
                }
            }
            v[r] = {a, c};
        }
        fft::Dft<R1, -1>::run(v);
#pragma unroll
        for (int r = 0; r < R1; ++r) buf[(j * RA + r) * kPB + pl] = v[r];
    }
    __syncthreads();
    if (j < R1) {
        cf y[R2];
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const cf x = buf[(r * RA + j) * kPB + pl];
            y[r] = (r == 0 || j == 0) ? x : cmul(x, tw1[r * j]);
        }
        fft::Dft<R2, -1>::run(y);
        cf* Su = S + (u - unit0) * static_cast<int64_t>(p.n) * p.CG;
        cf wj = {1.f, 0.f};
        if (kTwLds) {
            const int e = n2 * j;
            wj = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
        }
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const int k1 = j + r * R1;
            cf w4;
            if (kTwLds) {
                w4 = cmul(wj, stp[r]);
            } else {
                const int e = n2 * k1;  // < n, no reduction needed
                w4 = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
            }
            Su[(static_cast<int64_t>(k1) * p.N2 + n2) * p.CG + pb * kPB + pl] = cmul(y[r], w4);
        }
    }
}

// Stage A for int16 input with M % 8 == 0: one workgroup per (unit, pair
// block of kWideP pairs, column n2).  The N1 input rows of the column are
// staged through LDS in 16-byte pieces (4 pairs), kWideP/4 lanes per row
// (64 or 128 contiguous bytes), where the kPB-pair form reads a 32-byte piece
// of every row per workgroup and its loads wait on the L1's pending misses.
// The same LDS then holds the radix exchange.
#ifndef MSAMD_BP_WIDEP
#define MSAMD_BP_WIDEP 32
#endif
constexpr int kWideP = MSAMD_BP_WIDEP;
static_assert(kWideP % 4 == 0 && kCGMax % kWideP == 0, "wide stage A: pairs per workgroup");

template <int R1, int R2>
constexpr size_t wide_a_lds_bytes() {
    constexpr size_t stage = sizeof(uint32_t) * R1 * R2 * kWideP;
    constexpr size_t exch = sizeof(cf) * R2 * lds_row(R1) * kWideP;
    return stage > exch ? stage : exch;
}

template <int R1, int R2>
__global__ void __launch_bounds__(kWideP * (R1 > R2 ? R1 : R2)) bp_stage_a_wide(
    const int16_t* __restrict__ X, BpPlan p, int64_t unit0, int64_t nu, const cf* __restrict__ twN1,
    const cf* __restrict__ twHi, const cf* __restrict__ twLo, cf* __restrict__ S) {
    constexpr int N1 = R1 * R2;
    constexpr int RA = lds_row(R1);
    extern __shared__ cf wbuf[];  // staging [N1][kWideP] uint32, then exchange [R2][RA][kWideP]
    __shared__ cf tw1[N1];
    __shared__ cf stp[R2];  // W_n^{n2*R1*r}, r < R2
    const int tid = threadIdx.x;
    for (int q = tid; q < N1; q += blockDim.x) tw1[q] = twN1[q];
    const int npb = p.CG / kWideP;
    int n2, pb;
    int64_t u;
    block_unit_pb_col(blockIdx.x, npb, p.N2, &u, &pb, &n2);
    if (u >= nu) return;  // whole workgroup: past the last column of the batch
    u += unit0;
    if (kTwLds && tid < R2) {
        const int e = n2 * R1 * tid;
        stp[tid] = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
    }
    const int g = static_cast<int>(u / p.nWin);
    const int64_t w = p.w0 + u % p.nWin;
    const int64_t tbase = w * p.cs - p.ov;  // time of window sample 0
    const int64_t ch0 = 2 * (static_cast<int64_t>(g) * p.CG + pb * kWideP);
    uint32_t* stg = reinterpret_cast<uint32_t*>(wbuf);
    // 16-byte chunk c of row n1 holds pairs 4c..4c+3; with M % 8 == 0 a chunk
    // is wholly inside or wholly past the last channel
    constexpr int kCh = kWideP / 4;  // chunks per row
    {
        // thread q takes rows n1 = q / kCh + k * (threads / kCh) of chunk q % kCh:
        // the row time and the source address advance by fixed steps.  All of
        // a thread's loads are issued before its first LDS store, so their
        // latencies overlap instead of adding up.
        constexpr int kThreads = kWideP * (R1 > R2 ? R1 : R2);  // the launch's block size
        constexpr int kDn1 = kThreads / kCh;
        constexpr int kIt = (N1 + kDn1 - 1) / kDn1;
        const int c = tid % kCh;
        const int n1 = tid / kCh;
        int64_t t = tbase + static_cast<int64_t>(n1) * p.N2 + n2;
        const int64_t dt = static_cast<int64_t>(kDn1) * p.N2;
        const bool ch_ok = ch0 + 8 * c < p.M;
        const int16_t* src = X + (t - p.xt0) * p.M + ch0 + 8 * c;
        const int64_t dsrc = dt * p.M;
        uint4 rv[kIt];
#pragma unroll
        for (int k = 0; k < kIt; ++k, t += dt, src += dsrc) {
            rv[k] = uint4{0u, 0u, 0u, 0u};
            if (n1 + k * kDn1 < N1 && ch_ok && t >= 0 && t < p.N)
                rv[k] = *reinterpret_cast<const uint4*>(src);
        }
#pragma unroll
        for (int k = 0; k < kIt; ++k)
            if (n1 + k * kDn1 < N1)
                *reinterpret_cast<uint4*>(stg + (n1 + k * kDn1) * kWideP + 4 * c) = rv[k];
    }
    __syncthreads();
    const int pl = tid % kWideP;
    const int j = tid / kWideP;
    cf v[R1];
    if (j < R2) {
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const uint32_t w2 = stg[(j + r * R2) * kWideP + pl];
            v[r] = {static_cast<float>(static_cast<int16_t>(w2 & 0xffffu)),
                    static_cast<float>(static_cast<int16_t>(w2 >> 16))};
        }
    }
    __syncthreads();  // the staging is read before the exchange overwrites it
    if (j < R2) {
        fft::Dft<R1, -1>::run(v);
#pragma unroll
        for (int r = 0; r < R1; ++r) wbuf[(j * RA + r) * kWideP + pl] = v[r];
    }
    __syncthreads();
    if (j < R1) {
        cf y[R2];
#pragma unroll
        for (int r = 0; r < R2; ++r) {
            const cf x = wbuf[(r * RA + j) * kWideP + pl];
            y[r] = (r == 0 || j == 0) ? x : cmul(x, tw1[r * j]);
        }
        fft::Dft<R2, -1>::run(y);
        // spectra element (k1 = j + r R1, n2) of this pair; k1 advances by R1
        cf* dst = S + (u - unit0) * static_cast<int64_t>(p.n) * p.CG +
                  (static_cast<int64_t>(j) * p.N2 + n2) * p.CG + pb * kWideP + pl;
        const int64_t dstep = static_cast<int64_t>(R1) * p.N2 * p.CG;
        cf wj = {1.f, 0.f};
        if (kTwLds) {
            const int e = n2 * j;
            wj = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
        }
#pragma unroll
        for (int r = 0; r < R2; ++r, dst += dstep) {
            const int k1 = j + r * R1;
            cf w4;
            if (kTwLds) {
                w4 = cmul(wj, stp[r]);
            } else {
                const int e = n2 * k1;
                w4 = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
            }
            *dst = cmul(y[r], w4);
        }
    }
}

// ------------------------------------------------------------------ stage B
// Workgroup: one k1, kPB pairs of one unit.  Forward N2 = Q1*Q2 transform
// (pass radix Q1 then Q2), gain, inverse (radix Q2 on the same registers,
// then radix Q1), written back in place.
template <int Q1, int Q2, int kPB>
__global__ void __launch_bounds__(kPB * (Q1 > Q2 ? Q1 : Q2)) bp_stage_b(
    BpPlan p, int64_t unit0, int64_t nu, const cf* __restrict__ twN2, const float* __restrict__ Ht,
    cf* __restrict__ S) {
    constexpr int N2 = Q1 * Q2;
    constexpr int QA = lds_row(Q1), QB = lds_row(Q2);
    constexpr int kBuf = (Q2 * QA > Q1 * QB ? Q2 * QA : Q1 * QB) * kPB;
    __shared__ cf buf[kBuf];
    __shared__ cf tw2[N2];
    const int tid = threadIdx.x;
    for (int q = tid; q < N2; q += blockDim.x) tw2[q] = twN2[q];
    const int pl = tid % kPB;
    const int j = tid / kPB;
    const int npb = p.CG / kPB;
    int k1, pb;
    int64_t u;
    block_unit_pb_col(blockIdx.x, npb, p.N1, &u, &pb, &k1);
    if (u >= nu) return;  // whole workgroup: past the last row of the batch
    u += unit0;
    cf* Srow = S + (u - unit0) * static_cast<int64_t>(p.n) * p.CG +
               static_cast<int64_t>(k1) * N2 * p.CG + pb * kPB + pl;
    // element n2 = j + r*Q2 of the row: Srow + j*CG, then steps of Q2*CG
    cf* const Sj = Srow + static_cast<int64_t>(j) * p.CG;
    const int64_t rstep = static_cast<int64_t>(Q2) * p.CG;
    // forward pass 1: radix Q1 over n2 = j + r*Q2 (Q2 tasks)
    if (j < Q2) {
        cf v[Q1];
        const cf* src = Sj;
#pragma unroll
        for (int r = 0; r < Q1; ++r, src += rstep) v[r] = *src;
        fft::Dft<Q1, -1>::run(v);
#pragma unroll
        for (int r = 0; r < Q1; ++r) buf[(j * QA + r) * kPB + pl] = v[r];
    }
    __syncthreads();
    // forward pass 2: radix Q2 (Q1 tasks) -> k2 = j + r*Q1; gain; inverse pass 1
    // (radix Q2 over the same elements, Ns = 1) -> positions j*Q2 + r
    cf y[Q2];
    if (j < Q1) {
#pragma unroll
        for (int r = 0; r < Q2; ++r) {
            const cf x = buf[(r * QA + j) * kPB + pl];
            y[r] = (r == 0 || j == 0) ? x : cmul(x, tw2[r * j]);
        }
        fft::Dft<Q2, -1>::run(y);
        const float* Hrow = Ht + static_cast<int64_t>(k1) * N2;
#pragma unroll
        for (int r = 0; r < Q2; ++r) {
            const float h = Hrow[j + r * Q1];
            y[r] = {y[r].x * h, y[r].y * h};
        }
        fft::Dft<Q2, 1>::run(y);
    }
    __syncthreads();
    if (j < Q1) {
#pragma unroll
        for (int r = 0; r < Q2; ++r) buf[(j * QB + r) * kPB + pl] = y[r];
    }
    __syncthreads();
    // inverse pass 2: radix Q1, Ns = Q2 (Q2 tasks): inputs j + r*Q2, twiddle
    // exp(+2 pi i r j / N2), outputs n2 = j + r*Q2
    if (j < Q2) {
        cf v[Q1];
#pragma unroll
        for (int r = 0; r < Q1; ++r) {
            const cf x = buf[(r * QB + j) * kPB + pl];
            v[r] = (r == 0 || j == 0) ? x : cconjmul(x, tw2[r * j]);
        }
        fft::Dft<Q1, 1>::run(v);
        cf* dst = Sj;
#pragma unroll
        for (int r = 0; r < Q1; ++r, dst += rstep) *dst = v[r];
    }
}

// ------------------------------------------------------------------ stage C
// Workgroup: one n2, kPB pairs.  Inverse over k1: pass 1 radix R2 over
// k1 = j + r*R1 (R1 tasks, after the conjugate four-step twiddle), pass 2
// radix R1 (Ns = R2, R2 tasks) -> n1 = j + r*R2.
template <int R1, int R2, bool kOutI16, bool kSub1, int kPB>
__global__ void __launch_bounds__(kPB * (R1 > R2 ? R1 : R2)) bp_stage_c(
    BpPlan p, int64_t unit0, int64_t nu, const cf* __restrict__ twN1, const cf* __restrict__ twHi,
    const cf* __restrict__ twLo, const cf* __restrict__ S, void* __restrict__ Yv) {
    constexpr int N1 = R1 * R2;
    constexpr int RC = lds_row(R2);
    __shared__ cf buf[R1 * RC * kPB];
    __shared__ cf tw1[N1];
    const int tid = threadIdx.x;
    for (int q = tid; q < N1; q += blockDim.x) tw1[q] = twN1[q];
    const int pl = tid % kPB;
    const int j = tid / kPB;
    const int npb = p.CG / kPB;
    int n2, pb;
    int64_t u;
    block_unit_pb_col(blockIdx.x, npb, p.N2, &u, &pb, &n2);
    if (u >= nu) return;  // whole workgroup: past the last column of the batch
    u += unit0;
    const int g = static_cast<int>(u / p.nWin);
    const int64_t w = p.w0 + u % p.nWin;
    const int pair = g * p.CG + pb * kPB + pl;
    const int ch0 = 2 * pair;
    const cf* Su = S + (u - unit0) * static_cast<int64_t>(p.n) * p.CG;
    __shared__ cf stp[R2];  // W_n^{n2*R1*r}, r < R2 (n2*R1*r < n)
    if (kTwLds) {
        if (tid < R2) {
            const int e = n2 * R1 * tid;
            stp[tid] = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
        }
        __syncthreads();
    }
    if (j < R1) {
        cf v[R2];
        cf wj = {1.f, 0.f};
        if (kTwLds) {
            const int e = n2 * j;
            wj = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
        }
        const cf* src = Su + (static_cast<int64_t>(j) * p.N2 + n2) * p.CG + pb * kPB + pl;
        const int64_t sstep = static_cast<int64_t>(R1) * p.N2 * p.CG;  // k1 advances by R1
#pragma unroll
        for (int r = 0; r < R2; ++r, src += sstep) {
            const int k1 = j + r * R1;
            cf w4;
            if (kTwLds) {
                w4 = cmul(wj, stp[r]);
            } else {
                const int e = n2 * k1;
                w4 = cmul(twHi[e / kTwLo], twLo[e % kTwLo]);
            }
            v[r] = cconjmul(*src, w4);
        }
        fft::Dft<R2, 1>::run(v);
#pragma unroll
        for (int r = 0; r < R2; ++r) buf[(j * RC + r) * kPB + pl] = v[r];
    }
    __syncthreads();
    if (j < R2) {
        cf y[R1];
#pragma unroll
        for (int r = 0; r < R1; ++r) {
            const cf x = buf[(r * RC + j) * kPB + pl];
            y[r] = (r == 0 || j == 0) ? x : cconjmul(x, tw1[r * j]);
        }
        fft::Dft<R1, 1>::run(y);
        if (kSub1 && pair < p.P && (p.M % 2) == 0) {
            // output row to = w*cs + tl, tl = (j + r R2) N2 + n2 - ov (|tl| < n <
            // 2^31): the row is kept when 0 <= tl < min(cs, N_out - w*cs), one
            // unsigned 32-bit compare per r, and the store address advances by a
            // fixed step
            const int step = R2 * p.N2;
            const int tl0 = j * p.N2 + n2 - static_cast<int>(p.ov);
            const int64_t rest = p.N_out - w * p.cs;
            const uint32_t lim = static_cast<uint32_t>(
                rest <= 0 ? 0 : (rest < p.cs ? rest : p.cs));
            const int64_t off0 = (w * p.cs + tl0 - p.yt0) * p.M + ch0;
            const int64_t ostep = static_cast<int64_t>(step) * p.M;
            if (kOutI16) {
                uint32_t* dst = reinterpret_cast<uint32_t*>(static_cast<int16_t*>(Yv) + off0);
#pragma unroll
                for (int r = 0; r < R1; ++r, dst += ostep / 2) {
                    if (static_cast<uint32_t>(tl0 + r * step) >= lim) continue;
                    const float s0 = static_cast<float>(static_cast<double>(y[r].x) * p.qscale);
                    const float s1 = static_cast<float>(static_cast<double>(y[r].y) * p.qscale);
                    const uint32_t lo = static_cast<uint16_t>(static_cast<int16_t>(static_cast<int>(s0)));
                    const uint32_t hi = static_cast<uint16_t>(static_cast<int16_t>(static_cast<int>(s1)));
                    *dst = lo | (hi << 16);
                }
            } else {
                float2* dst = reinterpret_cast<float2*>(static_cast<float*>(Yv) + off0);
#pragma unroll
                for (int r = 0; r < R1; ++r, dst += ostep / 2) {
                    if (static_cast<uint32_t>(tl0 + r * step) >= lim) continue;
                    *dst = make_float2(y[r].x, y[r].y);
                }
            }
        } else if (pair < p.P) {
#pragma unroll
            for (int r = 0; r < R1; ++r) {
                const int64_t n1 = j + r * R2;
                const int64_t tl = n1 * p.N2 + n2 - p.ov;   // core-relative
                if (tl < 0 || tl >= p.cs) continue;
                int64_t to;
                if (kSub1) {
                    to = w * p.cs + tl;
                } else {
                    if (tl % p.k_sub) continue;
                    if (tl / p.k_sub >= p.cs / p.k_sub) continue;
                    to = (w * p.cs) / p.k_sub + tl / p.k_sub;
                }
                if (to >= p.N_out) continue;
                const float vals[2] = {y[r].x, y[r].y};
                if ((p.M % 2) == 0) {  // both channels of the pair in one store
                    if (kOutI16) {
                        const float s0 = static_cast<float>(static_cast<double>(vals[0]) * p.qscale);
                        const float s1 = static_cast<float>(static_cast<double>(vals[1]) * p.qscale);
                        const uint32_t lo = static_cast<uint16_t>(static_cast<int16_t>(static_cast<int>(s0)));
                        const uint32_t hi = static_cast<uint16_t>(static_cast<int16_t>(static_cast<int>(s1)));
                        *reinterpret_cast<uint32_t*>(static_cast<int16_t*>(Yv) + (to - p.yt0) * p.M + ch0) = lo | (hi << 16);
                    } else {
                        *reinterpret_cast<float2*>(static_cast<float*>(Yv) + (to - p.yt0) * p.M + ch0) =
                            make_float2(vals[0], vals[1]);
                    }
                    continue;
                }
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
    }
}

bool pick_shape(int64_t n, Shape* s) {
    for (const Shape& sh : kShapes) {
        if (static_cast<int64_t>(sh.R1) * sh.R2 * sh.Q1 * sh.Q2 == n) {
            *s = sh;
            return true;
        }
    }
    return false;
}

BpPlan make_plan(int64_t M, int64_t N, int64_t cs, int64_t ov, int k_sub, double qunit,
                 bool in_i16, const Shape& sh, int64_t scratch_budget, int streams,
                 int64_t w0 = 0, int64_t nwin = -1) {
    BpPlan p{};
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
    p.N1 = sh.R1 * sh.R2;
    p.N2 = sh.Q1 * sh.Q2;
    p.P = static_cast<int32_t>((M + 1) / 2);
    p.CG = static_cast<int32_t>(align_up(p.P < kCGMax ? p.P : kCGMax, kCGAlign));
    p.nGroups = static_cast<int32_t>(ceil_div(p.P, p.CG));
    p.w0 = w0;
    p.nWin = nwin >= 0 ? nwin : ceil_div(N, cs);
    p.nUnits = p.nWin * p.nGroups;
    const int64_t per_unit = p.n * p.CG * static_cast<int64_t>(sizeof(cf));
    int64_t U = scratch_budget / streams / per_unit;
    if (U < 1) U = 1;
    if (U > p.nUnits) U = p.nUnits;
    p.U = static_cast<int32_t>(U);
    const int64_t nbatch = ceil_div(p.nUnits, U);
    p.streams = static_cast<int32_t>(nbatch < streams ? nbatch : streams);
    p.nTwHi = static_cast<int32_t>(ceil_div(p.n, kTwLo));
    return p;
}

constexpr int kMaxStreams = 4;

struct BpBuffers {
    cf* S[kMaxStreams];
    float* Ht;
    cf* twN1;
    cf* twN2;
    cf* twHi;
    cf* twLo;
};

BpBuffers bp_carve(void* ws, const BpPlan& p, int64_t* bytes) {
    Carver c(ws);
    BpBuffers b{};
    for (int i = 0; i < p.streams; ++i) b.S[i] = c.take<cf>(static_cast<int64_t>(p.U) * p.n * p.CG);
    b.Ht = c.take<float>(p.n);
    b.twN1 = c.take<cf>(p.N1);
    b.twN2 = c.take<cf>(p.N2);
    b.twHi = c.take<cf>(p.nTwHi);
    b.twLo = c.take<cf>(kTwLo);
    if (bytes) *bytes = c.bytes();
    return b;
}

constexpr int64_t kScratchBudget = static_cast<int64_t>(MSAMD_BP_SCRATCH_MB) << 20;  // spectra kept in the last-level cache

// Batches go round-robin over the caller's stream and library-owned side
// streams (one spectra buffer each), so that one batch's memory-bound stages
// A and C run beside another's arithmetic-bound stage B.  MSAMD_BP_STREAMS
// (1..4) in the environment sets the number of streams.
#ifndef MSAMD_BP_NSTREAMS
#define MSAMD_BP_NSTREAMS 2
#endif
int bp_streams() {
    static const int n = [] {
        const char* e = std::getenv("MSAMD_BP_STREAMS");
        const int v = e ? std::atoi(e) : MSAMD_BP_NSTREAMS;
        return v < 1 ? 1 : (v > kMaxStreams ? kMaxStreams : v);
    }();
    return n;
}

// Per-device side streams and fork/join events, created on first use.  The
// issue of one filter call's launches holds the device's lock, so the event
// pairs of calls from several host threads do not interleave.
// Host tables of one filter shape (the gain H/n transposed to [k1][k2],
// the Stockham and four-step twiddles), computed once per parameter set in
// pinned host memory, so that every later call uploads them with
// asynchronous copies on its stream instead of recomputing 240000 erf
// values and waiting for pageable copies.  Entries live until the process
// ends (at most kMaxTables parameter sets; beyond that the tables are built
// per call).
struct BpTables {
    int64_t n;
    int32_t N1, N2, nTwHi;
    double fs, fmin, fmax, fwid;
    char* host;       // pinned: Ht | twN1 | twN2 | twHi | twLo
    size_t off[5], len[5];
};
constexpr int kMaxTables = 32;
std::mutex g_tables_mu;
std::vector<BpTables*> g_tables;

void fill_tables(const BpPlan& p, double fs, double fmin, double fmax, double fwid, BpTables* t) {
    std::vector<double> H(p.n);
    msamd_bandpass_kernel(p.n, fs, fmin, fmax, fwid, H.data());
    const double factor = 1.0 / p.n;
    float* Ht = reinterpret_cast<float*>(t->host + t->off[0]);
    for (int64_t k1 = 0; k1 < p.N1; ++k1)
        for (int64_t k2 = 0; k2 < p.N2; ++k2)
            Ht[k1 * p.N2 + k2] = static_cast<float>(H[k1 + p.N1 * k2] * factor);
    auto table = [](cf* dst, int64_t len, int64_t period, int64_t step) {
        for (int64_t m = 0; m < len; ++m) {
            const long double ang = -2.0L * 3.14159265358979323846264338327950288L *
                                    static_cast<long double>((m * step) % period) / period;
            dst[m] = {static_cast<float>(std::cos(ang)), static_cast<float>(std::sin(ang))};
        }
    };
    table(reinterpret_cast<cf*>(t->host + t->off[1]), p.N1, p.N1, 1);
    table(reinterpret_cast<cf*>(t->host + t->off[2]), p.N2, p.N2, 1);
    table(reinterpret_cast<cf*>(t->host + t->off[3]), p.nTwHi, p.n, kTwLo);
    table(reinterpret_cast<cf*>(t->host + t->off[4]), kTwLo, p.n, 1);
}

void layout_tables(const BpPlan& p, BpTables* t) {
    const size_t len[5] = {sizeof(float) * p.n, sizeof(cf) * p.N1, sizeof(cf) * p.N2,
                           sizeof(cf) * p.nTwHi, sizeof(cf) * kTwLo};
    size_t o = 0;
    for (int i = 0; i < 5; ++i) {
        t->off[i] = o;
        t->len[i] = len[i];
        o += align_up(static_cast<int64_t>(len[i]), 256);
    }
}

size_t tables_bytes(const BpTables& t) { return t.off[4] + t.len[4]; }

// Enqueues the upload of the tables into b (asynchronous when they are cached).
int upload_tables(const BpPlan& p, double fs, double fmin, double fmax, double fwid,
                  const BpBuffers& b, hipStream_t st) {
    BpTables* hit = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_tables_mu);
        for (BpTables* t : g_tables)
            if (t->n == p.n && t->N1 == p.N1 && t->N2 == p.N2 && t->nTwHi == p.nTwHi &&
                t->fs == fs && t->fmin == fmin && t->fmax == fmax && t->fwid == fwid) {
                hit = t;
                break;
            }
        if (!hit && static_cast<int>(g_tables.size()) < kMaxTables) {
            BpTables* t = new BpTables{};
            t->n = p.n;
            t->N1 = p.N1;
            t->N2 = p.N2;
            t->nTwHi = p.nTwHi;
            t->fs = fs;
            t->fmin = fmin;
            t->fmax = fmax;
            t->fwid = fwid;
            layout_tables(p, t);
            void* host = nullptr;
            if (hipHostMalloc(&host, tables_bytes(*t), hipHostMallocDefault) == hipSuccess) {
                t->host = static_cast<char*>(host);
                fill_tables(p, fs, fmin, fmax, fwid, t);
                g_tables.push_back(t);
                hit = t;
            } else {
                delete t;
                (void)hipGetLastError();  // no pinned memory: build per call below
            }
        }
    }
    void* const dst[5] = {b.Ht, b.twN1, b.twN2, b.twHi, b.twLo};
    if (hit) {
        for (int i = 0; i < 5; ++i)
            MSAMD_HIP(hipMemcpyAsync(dst[i], hit->host + hit->off[i], hit->len[i],
                                     hipMemcpyHostToDevice, st));
        return MSAMD_OK;
    }
    BpTables t{};
    layout_tables(p, &t);
    std::vector<char> host(tables_bytes(t));
    t.host = host.data();
    fill_tables(p, fs, fmin, fmax, fwid, &t);
    for (int i = 0; i < 5; ++i)
        MSAMD_HIP(hipMemcpyAsync(dst[i], t.host + t.off[i], t.len[i], hipMemcpyHostToDevice, st));
    MSAMD_HIP(hipStreamSynchronize(st));  // pageable source goes out of scope
    return MSAMD_OK;
}

struct SideStream {
    hipStream_t stream[kMaxStreams - 1] = {};
    hipEvent_t fork = nullptr, join[kMaxStreams - 1] = {};
    std::mutex mu;
};
constexpr int kMaxDevices = 64;
SideStream g_side[kMaxDevices];
std::mutex g_side_init;

int side_stream(SideStream** out) {
    int dev = 0;
    MSAMD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("bandpass: device %d beyond the supported %d", dev, kMaxDevices);
        return MSAMD_ERR_UNSUPPORTED;
    }
    SideStream& s = g_side[dev];
    std::lock_guard<std::mutex> lock(g_side_init);
    if (!s.fork) {
        for (int i = 0; i < kMaxStreams - 1; ++i) {
            MSAMD_HIP(hipEventCreateWithFlags(&s.join[i], hipEventDisableTiming));
            MSAMD_HIP(hipStreamCreateWithFlags(&s.stream[i], hipStreamNonBlocking));
        }
        MSAMD_HIP(hipEventCreateWithFlags(&s.fork, hipEventDisableTiming));
    }
    *out = &s;
    return MSAMD_OK;
}

}  // namespace
}  // namespace msamd

namespace msamd {
// Windows of any other length (bandpass_bluestein.hip).
int64_t bluestein_workspace_size(int64_t M, int64_t N, int64_t cs, int64_t ov);
bool bluestein_supported(int64_t n);
int bluestein_filter(const void* X, bool in_i16, int64_t M, int64_t N, int64_t xt0, int64_t w0,
                     int64_t w1, double samplerate, double freq_min, double freq_max,
                     double freq_wid, int64_t cs, int64_t ov, int k_sub, double qunit, void* Y,
                     int64_t yt0, void* workspace, int64_t workspace_bytes, hipStream_t st);
}  // namespace msamd

using namespace msamd;

extern "C" int64_t msamd_bandpass_chunk_size(int64_t M, int64_t num_threads) {
    // p_bandpass_filter.cpp:147-155
    if (M < 1) M = 1;
    if (num_threads < 1) num_threads = 1;
    const int overhead = 6;
    const int64_t overlap_size = 60000;
    const int64_t min_chunk_size = overlap_size * 2;
    const double target_ram_bytes = 1.0 * (1024 * 1024 * 1024);
    const double target = target_ram_bytes / (overhead * sizeof(float) * M * num_threads) - 2 * overlap_size;
    int64_t chunk = static_cast<int64_t>(target);
    if (chunk < min_chunk_size) chunk = min_chunk_size;
    return chunk;
}

extern "C" int msamd_bandpass_kernel(int64_t n_fft, double samplerate, double freq_min,
                                     double freq_max, double freq_wid, double* kernel_host) {
    // define_kernel, p_bandpass_filter.cpp:421-447 (host, FP64, C library erf)
    clear_error();
    MSAMD_REQUIRE(n_fft >= 1 && kernel_host, "bandpass_kernel: bad arguments");
    const double T = n_fft / samplerate;
    const double df = 1 / T;
    const double relwid = 3.0;
    for (int64_t i = 0; i < n_fft; ++i) {
        const double fgrid = (i <= (n_fft + 1) / 2) ? df * i : df * (i - n_fft);
        const double absf = std::fabs(fgrid);
        double val = 1.0;
        if (freq_min != 0) {
            if (i == 0)
                val = 0.0;
            else
                val *= (1 + std::erf(relwid * (absf - freq_min) / freq_min)) / 2;
        }
        if (freq_max != 0) val *= (1 - std::erf((absf - freq_max) / freq_wid)) / 2;
        kernel_host[i] = std::sqrt(val);
    }
    return MSAMD_OK;
}

extern "C" int msamd_bandpass_supported(int64_t chunk_size, int64_t overlap) {
    Shape s;
    if (chunk_size < 1 || overlap < 0) return 0;
    return (pick_shape(chunk_size + 2 * overlap, &s) || bluestein_supported(chunk_size + 2 * overlap))
               ? 1
               : 0;
}

extern "C" int64_t msamd_bandpass_workspace_size(int64_t M, int64_t N, int64_t chunk_size,
                                                 int64_t overlap) {
    Shape sh;
    if (M < 1 || N < 0 || chunk_size < 1 || overlap < 0) return -1;
    if (!pick_shape(chunk_size + 2 * overlap, &sh))
        return bluestein_supported(chunk_size + 2 * overlap)
                   ? bluestein_workspace_size(M, N, chunk_size, overlap)
                   : -1;
    const BpPlan p = make_plan(M, N > 0 ? N : 1, chunk_size, overlap, 1, 0.0, false, sh,
                               kScratchBudget, bp_streams());
    int64_t bytes = 0;
    bp_carve(nullptr, p, &bytes);
    return bytes;
}

namespace {

// Windows [w0, w1) of the filter.  X holds recording rows [xt0, xt0 + x_rows)
// and must cover those windows' input (clipped to [0, N)); Y row 0 is output
// row yt0.  The whole filter is w0 = 0, w1 = nWin, xt0 = yt0 = 0.
int bandpass_run(const void* X, int dtype, int64_t M, int64_t N, int64_t xt0, int64_t x_rows,
                 int64_t w0, int64_t w1, double samplerate, double freq_min, double freq_max,
                 double freq_wid, int64_t chunk_size, int64_t overlap, int subsample_factor,
                 double quantization_unit, void* Y, int64_t yt0, int64_t y_rows,
                 void* workspace, int64_t workspace_bytes, void* stream) {
    MSAMD_REQUIRE(M >= 1 && N >= 0, "bandpass: bad shape");
    MSAMD_REQUIRE(dtype == MSAMD_DTYPE_FLOAT32 || dtype == MSAMD_DTYPE_INT16,
                  "bandpass: input dtype %d not supported (float32 or int16)", dtype);
    MSAMD_REQUIRE(freq_max != 0, "bandpass: freq_max == 0 is the copy path of the processor");
    MSAMD_REQUIRE(chunk_size >= 1 && overlap >= 0, "bandpass: bad chunk plan");
    MSAMD_REQUIRE(samplerate > 0, "bandpass: samplerate must be positive");
    const int k_sub = subsample_factor > 0 ? subsample_factor : 1;
    const int64_t nwin_all = ceil_div(N, chunk_size);
    MSAMD_REQUIRE(0 <= w0 && w0 <= w1 && w1 <= nwin_all,
                  "bandpass: windows [%lld, %lld) outside the %lld windows of the recording",
                  (long long)w0, (long long)w1, (long long)nwin_all);
    {
        const int64_t lo = std::max<int64_t>(0, w0 * chunk_size - overlap);
        const int64_t hi = std::min<int64_t>(N, w1 * chunk_size + overlap);
        MSAMD_REQUIRE(w0 == w1 || (xt0 <= lo && xt0 + x_rows >= hi),
                      "bandpass: input rows [%lld, %lld) do not cover the windows' [%lld, %lld)",
                      (long long)xt0, (long long)(xt0 + x_rows), (long long)lo, (long long)hi);
        const int64_t ylo = (w0 * chunk_size) / k_sub;
        const int64_t yhi = std::min<int64_t>(N / k_sub, (w1 * chunk_size) / k_sub);
        MSAMD_REQUIRE(w0 == w1 || (yt0 <= ylo && yt0 + y_rows >= yhi),
                      "bandpass: output rows [%lld, %lld) do not cover the windows' [%lld, %lld)",
                      (long long)yt0, (long long)(yt0 + y_rows), (long long)ylo, (long long)yhi);
    }
    if (k_sub > 1 && chunk_size % k_sub != 0 && y_rows > 0) {
        // Window w writes its cs/k_sub outputs from (w*cs)/k_sub on, which
        // leaves a row between windows when k_sub does not divide cs; the
        // reference's output file is zero there (DiskWriteMda pre-zeroes).
        const int64_t bytes = y_rows * M * (quantization_unit != 0 ? 2 : 4);
        MSAMD_HIP(hipMemsetAsync(Y, 0, static_cast<size_t>(bytes), as_stream(stream)));
    }
    Shape sh;
    if (!pick_shape(chunk_size + 2 * overlap, &sh)) {
        // no four-step factorization: Bluestein's chirp-z form of the same window filter
        if (!bluestein_supported(chunk_size + 2 * overlap)) {
            set_error("bandpass: FFT length %lld is beyond the supported maximum (2^25)",
                      (long long)(chunk_size + 2 * overlap));
            return MSAMD_ERR_UNSUPPORTED;
        }
        return bluestein_filter(X, dtype == MSAMD_DTYPE_INT16, M, N, xt0, w0, w1, samplerate,
                                freq_min, freq_max, freq_wid, chunk_size, overlap, k_sub,
                                quantization_unit, Y, yt0, workspace, workspace_bytes,
                                as_stream(stream));
    }
    BpPlan p = make_plan(M, N, chunk_size, overlap, k_sub, quantization_unit,
                         dtype == MSAMD_DTYPE_INT16, sh, kScratchBudget, bp_streams(), w0,
                         w1 - w0);
    p.xt0 = xt0;
    p.yt0 = yt0;
    int64_t need = 0;
    BpBuffers b = bp_carve(workspace, p, &need);
    if (!workspace || workspace_bytes < need) {
        set_error("bandpass: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    if (N == 0 || w0 == w1) return MSAMD_OK;
    hipStream_t st = as_stream(stream);
    // host tables: gain H/n transposed to [k1][k2], Stockham twiddles,
    // four-step twiddle factors (all computed in double, rounded once)
    {
        const int rc = upload_tables(p, samplerate, freq_min, freq_max, freq_wid, b, st);
        if (rc != MSAMD_OK) return rc;
    }
    // whole-line staging of the int16 input (stage A) when a unit holds 32
    // pairs and rows are 16-byte aligned
    const bool wide_a = p.in_i16 && p.M % 8 == 0 && p.CG % kWideP == 0 && sh.R1 == 20 && sh.R2 == 20;
    if (wide_a)
        MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&bp_stage_a_wide<20, 20>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize,
                                      static_cast<int>(wide_a_lds_bytes<20, 20>())));
    SideStream* side = nullptr;
    std::unique_lock<std::mutex> side_lock;
    if (p.streams > 1) {
        const int rc = side_stream(&side);
        if (rc != MSAMD_OK) return rc;
        side_lock = std::unique_lock<std::mutex>(side->mu);
        MSAMD_HIP(hipEventRecord(side->fork, st));
        for (int i = 0; i + 1 < p.streams; ++i)
            MSAMD_HIP(hipStreamWaitEvent(side->stream[i], side->fork, 0));
    }
    int64_t batch = 0;
    for (int64_t u0 = 0; u0 < p.nUnits; u0 += p.U, ++batch) {
        const int64_t nu = (p.nUnits - u0) < p.U ? (p.nUnits - u0) : p.U;
        const int si = static_cast<int>(batch % p.streams);
        hipStream_t bs = si == 0 ? st : side->stream[si - 1];
        cf* S = b.S[si];
        auto grid = [&](int pb_per_wg, int64_t cols_per_unit) {
            const int64_t cols = nu * cols_per_unit;
            const int64_t npb = p.CG / pb_per_wg;
            return static_cast<unsigned>(kOrder == 2 ? ceil_div(cols, 8) * 8 * npb : cols * npb);
        };
        const unsigned gridA = grid(kPBA, p.N2);
        const unsigned gridC = grid(kPBC, p.N2);
        const unsigned gridB = grid(kPBB, p.N1);
        if (sh.R1 == 20 && sh.R2 == 20 && sh.Q1 == 24 && sh.Q2 == 25) {
            if (wide_a) {
                bp_stage_a_wide<20, 20><<<grid(kWideP, p.N2), kWideP * 20,
                                          wide_a_lds_bytes<20, 20>(), bs>>>(
                    static_cast<const int16_t*>(X), p, u0, nu, b.twN1, b.twHi, b.twLo, S);
            } else if (p.in_i16)
                bp_stage_a<20, 20, true, kPBA><<<gridA, kPBA * 20, 0, bs>>>(X, p, u0, nu, b.twN1,
                                                                             b.twHi, b.twLo, S);
            else
                bp_stage_a<20, 20, false, kPBA><<<gridA, kPBA * 20, 0, bs>>>(X, p, u0, nu, b.twN1,
                                                                              b.twHi, b.twLo, S);
            MSAMD_LAUNCH_CHECK();
            bp_stage_b<24, 25, kPBB><<<gridB, kPBB * 25, 0, bs>>>(p, u0, nu, b.twN2, b.Ht, S);
            MSAMD_LAUNCH_CHECK();
            const bool s1 = p.k_sub == 1;
            if (p.out_i16 && s1)
                bp_stage_c<20, 20, true, true, kPBC><<<gridC, kPBC * 20, 0, bs>>>(
                    p, u0, nu, b.twN1, b.twHi, b.twLo, S, Y);
            else if (p.out_i16)
                bp_stage_c<20, 20, true, false, kPBC><<<gridC, kPBC * 20, 0, bs>>>(
                    p, u0, nu, b.twN1, b.twHi, b.twLo, S, Y);
            else if (s1)
                bp_stage_c<20, 20, false, true, kPBC><<<gridC, kPBC * 20, 0, bs>>>(
                    p, u0, nu, b.twN1, b.twHi, b.twLo, S, Y);
            else
                bp_stage_c<20, 20, false, false, kPBC><<<gridC, kPBC * 20, 0, bs>>>(
                    p, u0, nu, b.twN1, b.twHi, b.twLo, S, Y);
            MSAMD_LAUNCH_CHECK();
        } else {
            set_error("bandpass: shape not instantiated");
            return MSAMD_ERR_UNSUPPORTED;
        }
    }
    for (int i = 0; i + 1 < p.streams; ++i) {  // the caller's stream waits for the side streams
        MSAMD_HIP(hipEventRecord(side->join[i], side->stream[i]));
        MSAMD_HIP(hipStreamWaitEvent(st, side->join[i], 0));
    }
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_bandpass_filter(const void* X, int dtype, int64_t M, int64_t N,
                                     double samplerate, double freq_min, double freq_max,
                                     double freq_wid, int64_t chunk_size, int64_t overlap,
                                     int subsample_factor, double quantization_unit, void* Y,
                                     void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(chunk_size >= 1, "bandpass: bad chunk plan");
    const int k_sub = subsample_factor > 0 ? subsample_factor : 1;
    return bandpass_run(X, dtype, M, N, 0, N, 0, N > 0 ? ceil_div(N, chunk_size) : 0, samplerate,
                        freq_min, freq_max, freq_wid, chunk_size, overlap, subsample_factor,
                        quantization_unit, Y, 0, N / k_sub, workspace, workspace_bytes, stream);
}

extern "C" int64_t msamd_bandpass_windows(int64_t N, int64_t chunk_size) {
    if (N < 0 || chunk_size < 1) return -1;
    return ceil_div(N, chunk_size);
}

extern "C" int msamd_bandpass_filter_windows(const void* X, int dtype, int64_t M, int64_t N,
                                             int64_t x_t0, int64_t x_rows, int64_t w0, int64_t w1,
                                             double samplerate, double freq_min, double freq_max,
                                             double freq_wid, int64_t chunk_size, int64_t overlap,
                                             int subsample_factor, double quantization_unit,
                                             void* Y, int64_t y_t0, int64_t y_rows,
                                             void* workspace, int64_t workspace_bytes,
                                             void* stream) {
    clear_error();
    return bandpass_run(X, dtype, M, N, x_t0, x_rows, w0, w1, samplerate, freq_min, freq_max,
                        freq_wid, chunk_size, overlap, subsample_factor, quantization_unit, Y,
                        y_t0, y_rows, workspace, workspace_bytes, stream);
}
