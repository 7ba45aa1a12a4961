// Learned spike-presence detector: a 3-layer temporal convolution network
// scoring 32-sample windows of every channel.
//
// This component has no counterpart in the reference.  It proposes the
// windows in which detection (detect_events.cpp:69-135) has to look; the
// plain rule decides every sample of a proposed window (detect.hip, gated
// scan), so skipped windows are the only source of missed events.
//
// Network (shared by all channels; exact index conventions):
//   x[ci][t] = X[t, c-3+ci]  ci = 0..6, zero outside the recording / [0, M)
//   h1[o][j] = relu(b1[o] + sum_{ci,k} W1[o][ci][k] x[ci][8j+4+k])     stride 8
//   h2[o][p] = relu(b2[o] + sum_{i,k}  W2[o][i][k]  h1[i][4p-3+k])     stride 4
//   z[w]     = b3 + sum_{i,k} W3[i][k] h2[i][w-4+k]                   window w
// with 32 hidden channels and 9 taps; z[w] scores samples [32w, 32w+32).
// Weights are used as bf16 (rounded from fp32 master copies), sums are fp32.
//
// Kernels:
//   tcn_prep_kernel      bf16 weight fragments in MFMA lane order
//   tcn_classify_kernel  inference streamed along time: layers 1-2 on
//                        v_mfma_f32_16x16x32_bf16 from LDS-staged bf16 input
//                        rows, layer 3 folded into per-position sums
//   tcn_grad_kernel      training forward + backward of one segment per
//                        workgroup (fp32 VALU), per-segment gradients
//   tcn_reduce_kernel    ordered sum of the segment gradients
//   tcn_adam_kernel      Adam update of the fp32 master parameters
#include "tcn.h"

namespace msamd {
namespace tcn {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bfr(float w) {
    return static_cast<float>(static_cast<__bf16>(w));
}

// splitmix64: counter-based random bits for initialisation and sampling.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit_open(uint64_t r) {
    return (static_cast<double>(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// ------------------------------------------------------------ fragments
// v_mfma_f32_16x16x32_bf16 operands: A (16 x 32): lane l holds row l%16,
// k = 8*(l/16)..+7; B (32 x 16): lane l holds column l%16, the same k range;
// D (16 x 16): lane l, register r holds (4*(l/16) + r, l%16).
//
// Layer 1 as a GEMM over K = 64 = 8 groups of 8: group g < 7 pairs taps 0..7
// of input channel g with x[g][8j+4 .. 8j+11]; group 7 pairs tap 8 of input
// channels 0..6 (and a zero) with x[0..6][8j+12].  K-step s holds groups
// 4s .. 4s+3.  Layer 2: K-step s = tap s, group b = hidden channels 8b..8b+7.
// Layer 3 (per-position sums u_k = sum_o W3[o][k] h2[o]): one K-step of 32
// hidden channels ordered as the layer-2 accumulators hold them -- lane
// group g's k = 8g + e is channel 4g + e (e < 4) or 16 + 4g + e - 4 -- and
// rows k = tap (0..8; rows 9..15 zero), so the accumulators are the B
// operand as they stand.
__host__ __device__ inline int l3_channel(int g, int e) { return e < 4 ? 4 * g + e : 12 + 4 * g + e; }

struct Frag {
    bf16x8 a1[2][2][64];
    bf16x8 a2[2][9][64];
    bf16x8 a3[64];
    float b1[kH];
    float b2[kH];
    float w3[kH * kK];  // bf16-rounded, [o][k] (training layout)
    float b3;
};

__global__ void tcn_prep_kernel(const float* __restrict__ P, Frag* __restrict__ F) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < 2 * 2 * 64) {
        const int lane = q & 63, s = (q >> 6) & 1, mt = q >> 7;
        const int o = 16 * mt + (lane & 15);
        const int g = 4 * s + (lane >> 4);
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            float w = 0.0f;
            if (g < kCI) w = P[kOffW1 + (o * kCI + g) * kK + e];
            else if (e < kCI) w = P[kOffW1 + (o * kCI + e) * kK + 8];
            v[e] = static_cast<__bf16>(w);
        }
        F->a1[mt][s][lane] = v;
    }
    if (q < 2 * 9 * 64) {
        const int lane = q & 63, s = (q >> 6) % 9, mt = (q >> 6) / 9;
        const int o = 16 * mt + (lane & 15);
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int i = 8 * (lane >> 4) + e;
            v[e] = static_cast<__bf16>(P[kOffW2 + (o * kH + i) * kK + s]);
        }
        F->a2[mt][s][lane] = v;
    }
    if (q < kH) {
        F->b1[q] = P[kOffB1 + q];
        F->b2[q] = P[kOffB2 + q];
    }
    if (q < 64) {
        const int k = q & 15, g = q >> 4;
        bf16x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e)
            v[e] = static_cast<__bf16>(k < kK ? P[kOffW3 + l3_channel(g, e) * kK + k] : 0.0f);
        F->a3[q] = v;
    }
    if (q < kH * kK) F->w3[q] = bfr(P[kOffW3 + q]);
    if (q == 0) F->b3 = P[kOffB3];
}

// ------------------------------------------------------------ inference
// A workgroup (8 waves) scores kCB central channels over a range of kRB
// blocks of 16 windows, streaming along time: block b (first window w0)
// computes layer 1 at the 64 new positions j in [4w0+18, 4w0+82) from the
// input samples [32w0+148, 32w0+661), layer 2 at the 16 positions
// p in [w0+4, w0+20), and the logits of windows [w0, w0+16), whose
// receptive fields end inside that block.  The layer-1 outputs (bf16) of the
// last 5 positions and the layer-3 partial sums of the last 8 layer-2
// positions are carried to the next block, so no position is computed
// twice; one extra block before the range fills the carries (its logits
// are dropped).  The input rows of a block (kCB + 8 channels, bf16) are
// staged in LDS once for all waves and double-buffered: the next block's
// samples are loaded into registers before this block's arithmetic.  Each
// wave owns two central channels and their carried state, so the only
// barrier is one per block.
//
// Layer 3 is folded through the per-position sums u_k[p] = sum_o W3[o][k]
// h2[o][p], one MFMA on the layer-2 outputs as they sit in the accumulators
// (split into two bf16 pieces, so h2 keeps ~16 bits), and
// z[w] = b3 + sum_k u_k[w - 4 + k].
constexpr int kSB = 16;                       // windows per block
constexpr int kCB = 16;                       // central channels per workgroup
constexpr int kRB = 64;                       // blocks per workgroup range
// Channel group cb holds centrals 16 cb - 12 .. 16 cb + 3, so its 24 input
// rows 16 cb - 16 .. 16 cb + 7 fill 96 bytes of one 128-byte line of a
// sample row for every even cb (one line fetched instead of two): 25% fewer
// lines through the L1 than groups starting at 16 cb, for one extra group.
constexpr int kCShift = 12;
// input rows: channels c0-4 .. c0+kCB+3 (central cc uses rows cc+1 .. cc+7),
// starting on a 16-byte boundary so that rows load as float4 when M % 4 == 0
constexpr int kRows = kCB + 8;                // 24 input rows
constexpr int kXN = 8 * 64 + 1;               // 513 input samples per block
constexpr int kXStride = 520;                 // bf16 per input row (1040 bytes)
constexpr int kH1Pos = 72;                    // 5 carried + 64 new (+ slack)
constexpr int kH1S = 40;                      // bf16 per layer-1 row (80 bytes)
constexpr int kUPos = 24;                     // 8 carried + 16 new layer-2 positions
constexpr int kCThreads = 512;
constexpr int kCWaves = kCThreads / 64;
constexpr int kXStep = kCThreads / kRows;    // 21 staging threads per input row (scalar form)
constexpr int kXPer = (kXN + kXStep - 1) / kXStep;  // staged samples per thread (scalar form)
constexpr int kQ4 = kRows / 4;                // float4 pieces per sample (vector form)
constexpr int kXPer4 = (kQ4 * kXN + kCThreads - 1) / kCThreads;  // pieces per thread
constexpr size_t kXBytes = sizeof(__bf16) * kRows * kXStride;
constexpr size_t kH1Bytes = sizeof(__bf16) * kH1Pos * kH1S;
constexpr size_t kUBytes = sizeof(float) * kK * kUPos;
constexpr size_t kZBytes = sizeof(float) * kSB * kCB;
constexpr size_t kClassifyLds =
    2 * kXBytes + kCB * (kH1Bytes + kUBytes) + 2 * kZBytes;
static_assert(kCB == 2 * kCWaves, "two central channels per wave");
static_assert(kClassifyLds <= 160 * 1024, "LDS");
static_assert((kXBytes % 16) == 0 && (kH1Bytes % 16) == 0 && (kXStride * 2) % 16 == 0,
              "alignment");

// Orders a wavefront's LDS accesses across lanes: the LDS performs one
// wave's operations in issue order, so only the compiler must not move
// loads and stores across this point.  (A wavefront-scope fence would also
// wait for the block's outstanding global prefetch.)
__device__ __forceinline__ void wave_sync() { asm volatile("" ::: "memory"); }

template <bool kVec>
__global__ void __launch_bounds__(kCThreads) tcn_classify_kernel(
    const float* __restrict__ X, int64_t M, int64_t N, int64_t nW, int nCB,
    const Frag* __restrict__ F, float lo, float hi, float* __restrict__ logits,
    uint8_t* __restrict__ classes) {
    extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
    __bf16* xbuf = reinterpret_cast<__bf16*>(lds_raw);                      // [2][kRows][kXStride]
    __bf16* h1all = reinterpret_cast<__bf16*>(lds_raw + 2 * kXBytes);       // [kCB][kH1Pos][kH1S]
    float* uall = reinterpret_cast<float*>(lds_raw + 2 * kXBytes + kCB * kH1Bytes);  // [kCB][kK][kUPos]
    float* zt = uall + kCB * kK * kUPos;                                      // [2][kSB][kCB]

    const int cb = static_cast<int>(blockIdx.x % nCB);
    const int64_t range = blockIdx.x / nCB;
    const int64_t c0 = static_cast<int64_t>(cb) * kCB - kCShift;
    const int64_t wbase = range * (kRB * kSB);
    int nb = kRB;  // real blocks of this range
    if (wbase + static_cast<int64_t>(nb) * kSB > nW) nb = static_cast<int>((nW - wbase + kSB - 1) / kSB);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wv = tid >> 6;
    const int n = lane & 15;
    const int grp = lane >> 4;

    // weights into registers
    bf16x8 a1[2][2], a2[2][9];
    const bf16x8 a3 = F->a3[lane];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
        for (int s = 0; s < 2; ++s) a1[mt][s] = F->a1[mt][s][lane];
#pragma unroll
        for (int s = 0; s < 9; ++s) a2[mt][s] = F->a2[mt][s][lane];
    }
    float b1r[2][4], b2r[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            b1r[mt][r] = F->b1[16 * mt + 4 * grp + r];
            b2r[mt][r] = F->b2[16 * mt + 4 * grp + r];
        }
    const float b3 = F->b3;

    // input samples of block ib (first window wbase + 16 (ib - 1)): registers,
    // then LDS.  Vector form (M % 4 == 0): thread (piece q, sample s) loads
    // channels c0-4+4q .. +3 of one sample as a float4, consecutive threads
    // taking consecutive pieces, so a load instruction reads whole 96-byte
    // runs of channels.  Scalar form: thread (row fr, first sample fs) stages
    // samples fs + 21 i of row fr.
    constexpr int kPer = kVec ? 4 * kXPer4 : kXPer;
    float pre[kPer];
    const int fr = tid % kRows;
    const int fs = tid / kRows;
    const int64_t fm = c0 - 4 + fr;
    const bool fok = fs < kXStep && fm >= 0 && fm < M;
    const float* frow = X + (fok ? fm : 0);
    const uint64_t M64 = static_cast<uint64_t>(M);
    const auto fetch = [&](int ib) {
        const int64_t tb = 32 * (wbase + kSB * static_cast<int64_t>(ib - 1)) + 148;
        if (kVec) {
#pragma unroll
            for (int i = 0; i < kXPer4; ++i) {
                const int e = tid + kCThreads * i;
                const int q = e % kQ4;
                const int s = e / kQ4;
                const int64_t t = tb + s;
                const int64_t m = c0 - 4 + 4 * q;  // M % 4 == 0: the piece is wholly in or out
                const bool in = s < kXN && t >= 0 && t < N && m >= 0 && m < M;
                const uint32_t tt = in ? static_cast<uint32_t>(t) : 0u;
                const float4 v = *reinterpret_cast<const float4*>(X + tt * M64 + (in ? m : 0));
                pre[4 * i + 0] = in ? v.x : 0.0f;
                pre[4 * i + 1] = in ? v.y : 0.0f;
                pre[4 * i + 2] = in ? v.z : 0.0f;
                pre[4 * i + 3] = in ? v.w : 0.0f;
            }
        } else {
#pragma unroll
            for (int i = 0; i < kXPer; ++i) {
                const int64_t t = tb + fs + kXStep * i;
                const bool in = fok && fs + kXStep * i < kXN && t >= 0 && t < N;
                const uint32_t tt = in ? static_cast<uint32_t>(t) : 0u;
                const float v = frow[tt * M64];
                pre[i] = in ? v : 0.0f;
            }
        }
    };
    const auto stash = [&](int buf) {
        __bf16* xb = xbuf + buf * (kRows * kXStride);
        if (kVec) {
#pragma unroll
            for (int i = 0; i < kXPer4; ++i) {
                const int e = tid + kCThreads * i;
                const int q = e % kQ4;
                const int s = e / kQ4;
                if (s < kXN) {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        xb[(4 * q + k) * kXStride + s] = static_cast<__bf16>(pre[4 * i + k]);
                }
            }
        } else {
            xb += fr * kXStride + fs;
#pragma unroll
            for (int i = 0; i < kXPer; ++i)
                if (fs < kXStep && fs + kXStep * i < kXN) xb[kXStep * i] = static_cast<__bf16>(pre[i]);
        }
    };
    fetch(0);
    stash(0);
    __syncthreads();

    for (int ib = 0; ib <= nb + 1; ++ib) {
        // logits of block ib - 1 (block 0 fills the carries and is not written)
        if (ib >= 2 && tid < kSB * kCB) {
            const int wl = tid / kCB, cc = tid % kCB;
            const int64_t w = wbase + kSB * static_cast<int64_t>(ib - 2) + wl;
            const int64_t c = c0 + cc;
            if (w < nW && c >= 0 && c < M) {
                const float z = zt[((ib - 1) & 1) * kSB * kCB + wl * kCB + cc];
                if (logits) logits[w * M + c] = z;
                if (classes) classes[w * M + c] = z < lo ? 0 : (z > hi ? 2 : 1);
            }
        }
        if (ib == nb + 1) break;
        const bool more = ib < nb;
        if (more) fetch(ib + 1);
        const __bf16* xb = xbuf + (ib & 1) * (kRows * kXStride);
        // the wave's two central channels, phase by phase (independent MFMA
        // chains of both in flight together)
        __bf16* h1c[2];
        float* uc[2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            h1c[h] = h1all + (2 * wv + h) * (kH1Pos * kH1S);
            uc[h] = uall + (2 * wv + h) * (kK * kUPos);
        }
        // carries: layer-1 positions 64..68 -> 0..4, layer-3 sums 16..23 -> 0..7
        if (lane < 40) {
            const int h = lane / 20, q = (lane % 20) >> 2, part = lane & 3;
            uint2* dst = reinterpret_cast<uint2*>(h1c[h] + q * kH1S) + part * 2;
            const uint2* src = reinterpret_cast<const uint2*>(h1c[h] + (64 + q) * kH1S) + part * 2;
            const uint2 zero = make_uint2(0u, 0u);
            const uint2 v0 = ib > 0 ? src[0] : zero, v1 = ib > 0 ? src[1] : zero;
            dst[0] = v0;
            dst[1] = v1;
        }
        for (int q = lane; q < 2 * kK * 8; q += 64) {
            const int h = q / (kK * 8), k = (q % (kK * 8)) >> 3, pp = q & 7;
            float* u = uc[h];
            u[k * kUPos + pp] = ib > 0 ? u[k * kUPos + 16 + pp] : 0.0f;
        }
        wave_sync();
        // ---- layer 1: 4 tiles of 16 new positions for each central
#pragma unroll
        for (int t = 0; t < 4; ++t) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int cc = 2 * wv + h;
                const int jj = 16 * t + n;
                const bf16x8 bA = *reinterpret_cast<const bf16x8*>(xb + (cc + 1 + grp) * kXStride + 8 * jj);
                bf16x8 bB;
                if (grp < 3) {
                    bB = *reinterpret_cast<const bf16x8*>(xb + (cc + 5 + grp) * kXStride + 8 * jj);
                } else {
#pragma unroll
                    for (int ci = 0; ci < kCI; ++ci) bB[ci] = xb[(cc + 1 + ci) * kXStride + 8 * jj + 8];
                    bB[7] = static_cast<__bf16>(0.0f);
                }
                f32x4 acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[0][0], bA, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[1][0], bA, acc1, 0, 0, 0);
                acc0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[0][1], bB, acc0, 0, 0, 0);
                acc1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1[1][1], bB, acc1, 0, 0, 0);
                bf16x4 o0, o1;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    o0[r] = static_cast<__bf16>(fmaxf(acc0[r] + b1r[0][r], 0.0f));
                    o1[r] = static_cast<__bf16>(fmaxf(acc1[r] + b1r[1][r], 0.0f));
                }
                __bf16* dst = h1c[h] + (5 + jj) * kH1S + 4 * grp;
                *reinterpret_cast<bf16x4*>(dst) = o0;
                *reinterpret_cast<bf16x4*>(dst + 16) = o1;
            }
        }
        wave_sync();
        // ---- layer 2: positions w0 + 4 + n, layer-1 rows 4n + s
        f32x4 acc[2][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) acc[h][0] = acc[h][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < kK; ++s) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const bf16x8 b = *reinterpret_cast<const bf16x8*>(h1c[h] + (4 * n + s) * kH1S + 8 * grp);
                acc[h][0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[0][s], b, acc[h][0], 0, 0, 0);
                acc[h][1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2[1][s], b, acc[h][1], 0, 0, 0);
            }
        }
        // ---- layer 3: u_k at positions w0 + 4 + n (rows k = 4 grp + r < 9)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            bf16x8 hh, hl;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const float v0 = fmaxf(acc[h][0][r] + b2r[0][r], 0.0f);
                const float v1 = fmaxf(acc[h][1][r] + b2r[1][r], 0.0f);
                hh[r] = static_cast<__bf16>(v0);
                hh[4 + r] = static_cast<__bf16>(v1);
                hl[r] = static_cast<__bf16>(v0 - static_cast<float>(hh[r]));
                hl[4 + r] = static_cast<__bf16>(v1 - static_cast<float>(hh[4 + r]));
            }
            f32x4 uacc = {0.f, 0.f, 0.f, 0.f};
            uacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, hl, uacc, 0, 0, 0);
            uacc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, hh, uacc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r)
                if (4 * grp + r < kK) uc[h][(4 * grp + r) * kUPos + 8 + n] = uacc[r];
        }
        wave_sync();
        // logits: lanes 0..15 the first central, 16..31 the second
        if (lane < 2 * kSB) {
            const int h = lane >> 4, wl = lane & 15;
            const float* u = uc[h];
            float z = b3;
#pragma unroll
            for (int k = 0; k < kK; ++k) z += u[k * kUPos + wl + k];
            zt[(ib & 1) * kSB * kCB + wl * kCB + 2 * wv + h] = z;
        }
        if (more) stash((ib + 1) & 1);
        __syncthreads();
    }
}

// ------------------------------------------------------------- training
struct SegDims {
    int S, nj, np, nt;
};

__host__ __device__ inline SegDims seg_dims(int S) {
    SegDims d;
    d.S = S;
    d.np = S + 8;
    d.nj = 4 * (S + 7) + 9;
    d.nt = 8 * (d.nj - 1) + 9;
    return d;
}

inline size_t grad_lds_bytes(int S) {
    const SegDims d = seg_dims(S);
    return sizeof(float) * (static_cast<size_t>(kCI) * d.nt + 2 * kH * d.nj + 2 * kH * d.np + 2 * S);
}

__device__ __forceinline__ float softplus(float x) {
    return x > 0.0f ? x + log1pf(expf(-x)) : log1pf(expf(x));
}

// One workgroup per segment (central channel c, windows w0 .. w0+S-1):
// forward, weighted BCE (mean over batch*S windows via inv_count), backward.
// Every sum runs in a fixed sequential order.
__global__ void __launch_bounds__(256) tcn_grad_kernel(
    const float* __restrict__ X, int64_t M, int64_t N, const uint8_t* __restrict__ labels,
    int64_t nW, const float* __restrict__ P, const int32_t* __restrict__ seg_c,
    const int64_t* __restrict__ seg_w, int S, float pos_weight, float inv_count,
    float* __restrict__ gpart, float* __restrict__ lpart) {
    extern __shared__ float sm[];
    const SegDims d = seg_dims(S);
    float* x = sm;                          // [7][nt]
    float* pre1 = x + kCI * d.nt;           // [32][nj]
    float* dp1 = pre1 + kH * d.nj;          // [32][nj]
    float* pre2 = dp1 + kH * d.nj;          // [32][np]
    float* dp2 = pre2 + kH * d.np;          // [32][np]
    float* zs = dp2 + kH * d.np;            // [S]
    float* dzs = zs + S;                    // [S]
    const int tid = threadIdx.x;
    const int nth = blockDim.x;
    const int64_t seg = blockIdx.x;
    const int64_t c = seg_c[seg];
    const int64_t w0 = seg_w[seg];
    const int64_t jlo = 4 * (w0 - 4) - 3;
    const int64_t t0 = 8 * jlo + 4;
    float* g = gpart + seg * kNumParams;

    for (int q = tid; q < kCI * d.nt; q += nth) {
        const int ci = q / d.nt;
        const int tau = q - ci * d.nt;
        const int64_t t = t0 + tau;
        const int64_t m = c - 3 + ci;
        x[q] = (t >= 0 && t < N && m >= 0 && m < M) ? X[t * M + m] : 0.0f;
    }
    __syncthreads();
    for (int q = tid; q < kH * d.nj; q += nth) {
        const int o = q / d.nj;
        const int jl = q - o * d.nj;
        float acc = P[kOffB1 + o];
        for (int ci = 0; ci < kCI; ++ci)
            for (int k = 0; k < kK; ++k)
                acc += bfr(P[kOffW1 + (o * kCI + ci) * kK + k]) * x[ci * d.nt + 8 * jl + k];
        pre1[q] = acc;
    }
    __syncthreads();
    for (int q = tid; q < kH * d.np; q += nth) {
        const int o = q / d.np;
        const int pl = q - o * d.np;
        float acc = P[kOffB2 + o];
        for (int i = 0; i < kH; ++i)
            for (int k = 0; k < kK; ++k)
                acc += bfr(P[kOffW2 + (o * kH + i) * kK + k]) *
                       fmaxf(pre1[i * d.nj + 4 * pl + k], 0.0f);
        pre2[q] = acc;
    }
    __syncthreads();
    for (int wl = tid; wl < S; wl += nth) {
        float z = P[kOffB3];
        for (int i = 0; i < kH; ++i)
            for (int k = 0; k < kK; ++k)
                z += bfr(P[kOffW3 + i * kK + k]) * fmaxf(pre2[i * d.np + wl + k], 0.0f);
        const int64_t w = w0 + wl;
        const float y = (w >= 0 && w < nW) ? static_cast<float>(labels[w * M + c]) : 0.0f;
        const float s = 1.0f / (1.0f + expf(-z));
        zs[wl] = pos_weight * y * softplus(-z) + (1.0f - y) * softplus(z);
        dzs[wl] = (pos_weight * y * (s - 1.0f) + (1.0f - y) * s) * inv_count;
    }
    __syncthreads();
    if (tid == 0) {
        float l = 0.0f;
        for (int wl = 0; wl < S; ++wl) l += zs[wl];
        lpart[seg] = l;
    }
    // layer 3 gradients and dh2
    for (int q = tid; q < kH * kK + 1; q += nth) {
        float acc = 0.0f;
        if (q < kH * kK) {
            const int i = q / kK, k = q - (q / kK) * kK;
            for (int wl = 0; wl < S; ++wl) acc += dzs[wl] * fmaxf(pre2[i * d.np + wl + k], 0.0f);
            g[kOffW3 + q] = acc;
        } else {
            for (int wl = 0; wl < S; ++wl) acc += dzs[wl];
            g[kOffB3] = acc;
        }
    }
    for (int q = tid; q < kH * d.np; q += nth) {
        const int i = q / d.np;
        const int pl = q - i * d.np;
        float acc = 0.0f;
        for (int k = 0; k < kK; ++k) {
            const int wl = pl - k;
            if (wl >= 0 && wl < S) acc += bfr(P[kOffW3 + i * kK + k]) * dzs[wl];
        }
        dp2[q] = pre2[q] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // layer 2 gradients and dh1
    for (int q = tid; q < kH * kH * kK; q += nth) {
        const int o = q / (kH * kK);
        const int i = (q / kK) % kH;
        const int k = q % kK;
        float acc = 0.0f;
        for (int pl = 0; pl < d.np; ++pl)
            acc += dp2[o * d.np + pl] * fmaxf(pre1[i * d.nj + 4 * pl + k], 0.0f);
        g[kOffW2 + q] = acc;
    }
    for (int o = tid; o < kH; o += nth) {
        float acc = 0.0f;
        for (int pl = 0; pl < d.np; ++pl) acc += dp2[o * d.np + pl];
        g[kOffB2 + o] = acc;
    }
    for (int q = tid; q < kH * d.nj; q += nth) {
        const int i = q / d.nj;
        const int jl = q - i * d.nj;
        float acc = 0.0f;
        for (int o = 0; o < kH; ++o)
            for (int k = jl & 3; k < kK; k += 4) {
                const int pl = (jl - k) >> 2;
                if (jl - k >= 0 && pl < d.np)
                    acc += bfr(P[kOffW2 + (o * kH + i) * kK + k]) * dp2[o * d.np + pl];
            }
        dp1[q] = pre1[q] > 0.0f ? acc : 0.0f;
    }
    __syncthreads();
    // layer 1 gradients
    for (int q = tid; q < kH * kCI * kK; q += nth) {
        const int o = q / (kCI * kK);
        const int ci = (q / kK) % kCI;
        const int k = q % kK;
        float acc = 0.0f;
        for (int jl = 0; jl < d.nj; ++jl) acc += dp1[o * d.nj + jl] * x[ci * d.nt + 8 * jl + k];
        g[kOffW1 + q] = acc;
    }
    for (int o = tid; o < kH; o += nth) {
        float acc = 0.0f;
        for (int jl = 0; jl < d.nj; ++jl) acc += dp1[o * d.nj + jl];
        g[kOffB1 + o] = acc;
    }
}

// grad[q] = sum over segments in order; loss = sum of segment losses * inv_count.
__global__ void __launch_bounds__(256) tcn_reduce_kernel(const float* __restrict__ gpart,
                                                         const float* __restrict__ lpart, int B,
                                                         float inv_count, float* __restrict__ grad,
                                                         float* __restrict__ loss) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < kNumParams) {
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s += gpart[static_cast<int64_t>(b) * kNumParams + q];
        grad[q] = s;
    }
    if (q == 0 && loss) {
        float s = 0.0f;
        for (int b = 0; b < B; ++b) s += lpart[b];
        *loss = s * inv_count;
    }
}

__global__ void __launch_bounds__(256) tcn_adam_kernel(float* __restrict__ P,
                                                       float* __restrict__ m1,
                                                       float* __restrict__ m2,
                                                       const float* __restrict__ grad, float lr,
                                                       float c1, float c2) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kNumParams) return;
    const float g = grad[q];
    const float a = 0.9f * m1[q] + 0.1f * g;
    const float b = 0.999f * m2[q] + 0.001f * g * g;
    m1[q] = a;
    m2[q] = b;
    P[q] -= lr * (a * c1) / (sqrtf(b * c2) + 1e-8f);
}

// Segments: half at random windows, half around a random positive window.
__global__ void tcn_sample_kernel(int64_t M, int64_t nW, int S, const int64_t* __restrict__ pos,
                                  int64_t npos, uint64_t seed, int64_t step, int B,
                                  int32_t* __restrict__ seg_c, int64_t* __restrict__ seg_w) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= B) return;
    const uint64_t r1 = mix64(seed ^ mix64(static_cast<uint64_t>(step) * 0x100000001B3ull + b));
    const uint64_t r2 = mix64(r1);
    const uint64_t r3 = mix64(r2);
    const int64_t span = nW - S + 1 > 1 ? nW - S + 1 : 1;
    if ((b & 1) && npos > 0) {
        const int64_t e = pos[r1 % static_cast<uint64_t>(npos)];
        const int64_t w = e / M;
        int64_t w0 = w - static_cast<int64_t>(r2 % static_cast<uint64_t>(S));
        if (w0 < 0) w0 = 0;
        if (w0 > span - 1) w0 = span - 1;
        seg_c[b] = static_cast<int32_t>(e % M);
        seg_w[b] = w0;
    } else {
        seg_c[b] = static_cast<int32_t>(r1 % static_cast<uint64_t>(M));
        seg_w[b] = static_cast<int64_t>(r3 % static_cast<uint64_t>(span));
    }
}

__global__ void tcn_init_kernel(float* __restrict__ P, uint64_t seed) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kNumParams) return;
    float std_ = 0.0f;
    if (q < kOffB1) std_ = sqrtf(2.0f / (kCI * kK));
    else if (q >= kOffW2 && q < kOffB2) std_ = sqrtf(2.0f / (kH * kK));
    else if (q >= kOffW3 && q < kOffB3) std_ = sqrtf(1.0f / (kH * kK));
    float v = 0.0f;
    if (std_ > 0.0f) {
        const uint64_t r = mix64(seed ^ mix64(static_cast<uint64_t>(q)));
        const double u1 = unit_open(r), u2 = unit_open(mix64(r));
        v = static_cast<float>(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2)) * std_;
    }
    if (q == kOffB3) v = -3.0f;
    P[q] = v;
}

// labels[w][c] = 1 for windows holding an event; positives[i] = w * M + c.
__global__ void tcn_label_kernel(const int64_t* __restrict__ times,
                                 const int64_t* __restrict__ offsets, int32_t M,
                                 uint8_t* __restrict__ labels, int64_t* __restrict__ positives) {
    const int c = blockIdx.x;
    for (int64_t i = offsets[c] + threadIdx.x; i < offsets[c + 1]; i += blockDim.x) {
        const int64_t w = times[i] / kWin;
        labels[w * M + c] = 1;
        if (positives) positives[i] = w * M + c;
    }
}

struct TrainBufs {
    float* gpart;
    float* lpart;
    float* grad;
    int32_t* seg_c;
    int64_t* seg_w;
};

inline TrainBufs train_carve(void* ws, int B, int64_t* bytes) {
    Carver c(ws);
    TrainBufs t{};
    t.gpart = c.take<float>(static_cast<int64_t>(B) * kNumParams);
    t.lpart = c.take<float>(B);
    t.grad = c.take<float>(kNumParams);
    t.seg_c = c.take<int32_t>(B);
    t.seg_w = c.take<int64_t>(B);
    if (bytes) *bytes = c.bytes();
    return t;
}

int launch_grad(const float* X, int64_t M, int64_t N, const uint8_t* labels, const float* P,
                const int32_t* seg_c, const int64_t* seg_w, int B, int S, float pos_weight,
                const TrainBufs& t, float* grad_out, float* loss_out, hipStream_t st) {
    const size_t lds = grad_lds_bytes(S);
    const float inv = 1.0f / (static_cast<float>(B) * S);
    MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&tcn_grad_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(lds)));
    tcn_grad_kernel<<<B, 256, lds, st>>>(X, M, N, labels, num_windows(N), P, seg_c, seg_w, S,
                                         pos_weight, inv, t.gpart, t.lpart);
    MSAMD_LAUNCH_CHECK();
    tcn_reduce_kernel<<<static_cast<unsigned>(ceil_div(kNumParams, 256)), 256, 0, st>>>(
        t.gpart, t.lpart, B, inv, grad_out, loss_out);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

}  // namespace

int64_t frag_bytes() { return align_up(sizeof(Frag), 256); }

int prepare(const float* params, void* frag, hipStream_t s) {
    tcn_prep_kernel<<<static_cast<unsigned>(ceil_div(2 * 9 * 64, 256)), 256, 0, s>>>(
        params, static_cast<Frag*>(frag));
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

int classify(const float* X, int64_t M, int64_t N, const void* frag, float lo_logit,
             float hi_logit, float* logits, uint8_t* classes, hipStream_t s) {
    const int64_t nW = num_windows(N);
    if (nW == 0 || M == 0) return MSAMD_OK;
    const int nCB = static_cast<int>(ceil_div(M + kCShift, kCB));
    const int64_t nR = ceil_div(nW, static_cast<int64_t>(kRB) * kSB);
    MSAMD_REQUIRE(nR * nCB < (1LL << 31), "tcn classify: recording too long");
    const bool vec = M % 4 == 0 && reinterpret_cast<uintptr_t>(X) % 16 == 0;
    const void* fn = vec ? reinterpret_cast<const void*>(&tcn_classify_kernel<true>)
                         : reinterpret_cast<const void*>(&tcn_classify_kernel<false>);
    MSAMD_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(kClassifyLds)));
    const unsigned grid = static_cast<unsigned>(nCB * nR);
    if (vec)
        tcn_classify_kernel<true><<<grid, kCThreads, kClassifyLds, s>>>(
            X, M, N, nW, nCB, static_cast<const Frag*>(frag), lo_logit, hi_logit, logits, classes);
    else
        tcn_classify_kernel<false><<<grid, kCThreads, kClassifyLds, s>>>(
            X, M, N, nW, nCB, static_cast<const Frag*>(frag), lo_logit, hi_logit, logits, classes);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

}  // namespace tcn
}  // namespace msamd

using namespace msamd;
using namespace msamd::tcn;

extern "C" int32_t msamd_tcn_num_params(void) { return kNumParams; }

extern "C" int msamd_tcn_init(float* params, uint64_t seed, void* stream) {
    clear_error();
    MSAMD_REQUIRE(params != nullptr, "tcn_init: null parameters");
    tcn_init_kernel<<<static_cast<unsigned>(ceil_div(kNumParams, 256)), 256, 0,
                      as_stream(stream)>>>(params, seed);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int msamd_tcn_labels(const int64_t* times, const int64_t* offsets, int32_t M,
                                int64_t N, uint8_t* labels, int64_t* positives, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 0 && labels, "tcn_labels: bad arguments");
    hipStream_t st = as_stream(stream);
    MSAMD_HIP(hipMemsetAsync(labels, 0, static_cast<size_t>(num_windows(N) * M), st));
    tcn_label_kernel<<<M, 256, 0, st>>>(times, offsets, M, labels, positives);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int64_t msamd_tcn_train_workspace_size(int32_t batch, int32_t seg_windows) {
    if (batch < 1 || seg_windows < 1 || seg_windows > 64) return -1;
    int64_t bytes = 0;
    train_carve(nullptr, batch, &bytes);
    return bytes;
}

extern "C" int msamd_tcn_grad(const float* X, int64_t M, int64_t N, const uint8_t* labels,
                              const float* params, const int32_t* seg_channel,
                              const int64_t* seg_window, int32_t batch, int32_t seg_windows,
                              double pos_weight, float* grad, float* loss, void* workspace,
                              int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 1 && batch >= 1 && seg_windows >= 1 && seg_windows <= 64,
                  "tcn_grad: bad arguments");
    int64_t need = 0;
    TrainBufs t = train_carve(workspace, batch, &need);
    if (!workspace || workspace_bytes < need) {
        set_error("tcn_grad: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    return launch_grad(X, M, N, labels, params, seg_channel, seg_window, batch, seg_windows,
                       static_cast<float>(pos_weight), t, grad, loss, as_stream(stream));
}

extern "C" int msamd_tcn_train(const float* X, int64_t M, int64_t N, const uint8_t* labels,
                               const int64_t* positives, int64_t num_positives, float* params,
                               float* adam_state, int32_t steps, int32_t step0, int32_t batch,
                               int32_t seg_windows, double lr, double pos_weight, uint64_t seed,
                               float* loss_trace, void* workspace, int64_t workspace_bytes,
                               void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 1 && batch >= 1 && seg_windows >= 1 && seg_windows <= 64 &&
                      steps >= 0 && step0 >= 0 && params && adam_state,
                  "tcn_train: bad arguments");
    int64_t need = 0;
    TrainBufs t = train_carve(workspace, batch, &need);
    if (!workspace || workspace_bytes < need) {
        set_error("tcn_train: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    const int64_t nW = num_windows(N);
    float* m1 = adam_state;
    float* m2 = adam_state + kNumParams;
    for (int32_t i = 0; i < steps; ++i) {
        const int64_t step = static_cast<int64_t>(step0) + i + 1;
        tcn_sample_kernel<<<static_cast<unsigned>(ceil_div(batch, 256)), 256, 0, st>>>(
            M, nW, seg_windows, positives, num_positives, seed, step, batch, t.seg_c, t.seg_w);
        MSAMD_LAUNCH_CHECK();
        const int rc = launch_grad(X, M, N, labels, params, t.seg_c, t.seg_w, batch, seg_windows,
                                   static_cast<float>(pos_weight), t, t.grad,
                                   loss_trace ? loss_trace + i : nullptr, st);
        if (rc != MSAMD_OK) return rc;
        const float c1 = static_cast<float>(1.0 / (1.0 - pow(0.9, static_cast<double>(step))));
        const float c2 = static_cast<float>(1.0 / (1.0 - pow(0.999, static_cast<double>(step))));
        tcn_adam_kernel<<<static_cast<unsigned>(ceil_div(kNumParams, 256)), 256, 0, st>>>(
            params, m1, m2, t.grad, static_cast<float>(lr), c1, c2);
        MSAMD_LAUNCH_CHECK();
    }
    return MSAMD_OK;
}

extern "C" int64_t msamd_tcn_score_workspace_size(void) { return frag_bytes(); }

extern "C" int msamd_tcn_score(const float* X, int64_t M, int64_t N, const float* params,
                               float* logits, void* workspace, int64_t workspace_bytes,
                               void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 0 && params && logits, "tcn_score: bad arguments");
    if (!workspace || workspace_bytes < frag_bytes()) {
        set_error("tcn_score: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)frag_bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    int rc = prepare(params, workspace, st);
    if (rc != MSAMD_OK) return rc;
    return classify(X, M, N, workspace, 0.0f, 0.0f, logits, nullptr, st);
}
