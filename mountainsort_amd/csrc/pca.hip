// Two-stage per-neighborhood PCA features (a12-a15), batched over all
// neighborhoods.
//
// Reference: NeighborhoodSorterPrivate::dimension_reduce_clips
// (packages/mountainsortalg/neighborhoodsorter.cpp:236-276) -> pca_subsampled
// (pca.cpp:99-127) -> pca (pca.cpp:68-97), then the stage-2 pca_subsampled of
// sort_clips_subset (sort_clips.cpp:58-70) whose features go to isosplit5.
//
// Stage 1, per neighborhood g and channel w: A = clips[w, :, :] (T x L);
// stage 2, per g: R = reduced clips viewed as (nf1*W) x L.  Each PCA:
//   columns: all L, or the pca_subsampled indices (int64)i for
//            i = 0, inc, 2inc, ... < L, inc = L/max (double accumulation)
//   Gram:    G = A_s A_s^T  -- two-piece FP16 MFMA after a power-of-two scale
//            per 32-column step (3 product terms), FP32 within a step, FP64
//            across steps
//   power:   per component k < K: v = normalize(sin(1..n)); 10 times
//            v <- G v / |G v| (sigma = |G v|); deflate G <- (I-vv')G(I-vv')
//            (FP64; the reference iterates X X^T v on the data, equal in
//            exact arithmetic)
//   project: F = C^T A on all L columns (FP32 FMA).
// Phases run as one launch each over every neighborhood (items sized for
// load balance: Gram and power per neighborhood / channel, projections per
// chunk of events).  The stage-1 Gram and projection read either a clip
// array or, fused with the clip gather, the timeseries itself (GatherSrc).
#include <algorithm>
#include <cstring>
#include <mutex>
#include <cmath>
#include <vector>

#include "common.h"

#pragma clang fp contract(off)

namespace msamd {
namespace {

typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kKC = 32;        // subsampled columns per staging step
constexpr int kMaxW = 128;     // channels per neighborhood supported
constexpr int kMaxT = 64;      // clip length supported
constexpr int kGroupW = 12;    // stage-1 channels per Gram work item (at most)
constexpr int kProjW = 16;     // stage-1 channels per projection work item (at most)
constexpr int kMaxTiles = 10;  // Gram tiles per wavefront (12 waves per workgroup)
constexpr int kGramThreads = 768;
constexpr int kMaxR = 2048;    // stage-2 dimension nf1 * W supported
constexpr int kSinTab = kMaxR; // sin(i + 1), i < kSinTab
// every staged element has a thread (stage 1 kGroupW * T <= 768 for
// T <= 64; stage 2 at most 256 rows), and every Gram tile of a work item
// has a wave slot: stage 1 up to 10 tiles (T <= 64) for each channel of a
// group; stage 2 up to 120 tiles (one band of up to 240 rows) or 64 (a pair
// of 128-row bands)
static_assert((kGramThreads / 64) * kMaxTiles >= 10 * kGroupW &&
                  (kGramThreads / 64) * kMaxTiles >= 120,
              "Gram tiles exceed the wave slots");
static_assert(kGroupW * 64 <= kGramThreads && 256 <= kGramThreads,
              "staged elements exceed the threads");

// Stage-2 Gram rows are split into bands: one band when nf1 * W <= 240,
// else bands of 128 rows; a work item covers one band pair (A <= B).
__host__ __device__ inline int band_rows(int R) { return R <= 240 ? 240 : 128; }
__host__ __device__ inline int num_bands(int R) { return R <= 240 ? 1 : (R + 127) / 128; }

struct PcaPlan {
    int32_t G, T, nf1, nf2;
    int64_t max_s;
    int32_t maxW;
    int32_t gw;          // stage-1 channels per Gram work item (<= kGroupW, sized to the LDS)
    int32_t pw;          // stage-1 channels per projection work item (<= kProjW)
    int64_t idx_stride;  // idx_stride_for(max_s)
    // tiles (chosen per workload, see msamd_pca_autotune)
    int32_t gram_chunk;  // subsampled columns per Gram work item (multiple of kKC)
    int32_t pe1;         // events per stage-1 projection item
    int32_t pe2;         // events per stage-2 projection item
};

struct PcaTables {
    const int64_t* clip_off;   // [G] floats
    const int64_t* red_off;    // [G] floats (W*nf1*L)
    const int64_t* feat_off;   // [G] floats (nf2*L)
    const int32_t* L;          // [G]
    const int32_t* W;          // [G]
    const int64_t* g1_off;     // [G] floats: W*T*T
    const int64_t* g2_off;     // [G] floats: (nf1*W)^2
    const int64_t* c1_off;     // [G] floats: W*T*nf1
    const int64_t* c2_off;     // [G] floats: nf1*W*nf2
    const int32_t* p1_items;   // [n1][2] (g, w)
    const int32_t* j1_items;   // [nj1][3] (g, first event, first channel of the group)
    const int32_t* j2_items;   // [nj2][2] (g, first event)
    int32_t* ns;               // [G] subsampled column count
    int32_t* idx;              // [G][idx_stride]
    double* g1;                // Gram stage 1
    double* g2;                // Gram stage 2
    float* c1;                 // components stage 1, [w][t][a]
    float* c2;                 // components stage 2, [r][a]
    const double* sin_tab;     // sin(i+1), i < kSinTab
    double* work2;             // FP64 stage-2 Gram copies when n > kPowMaxN
    const int64_t* w2_off;     // [G] doubles
    const int32_t* gram1_items;  // [n][3] (g, chunk, channel group)
    const int32_t* gram2_items;  // [n][3] (g, chunk, band pair A << 16 | B)
    const int32_t* gram_first;   // [G+1] first chunk of each neighborhood
    const int64_t* gp1_off;      // [nchunks] floats: stage-1 partial of each chunk
    const int64_t* gp2_off;      // [nchunks] floats: stage-2 partial of each chunk
    const int64_t* ev_off;       // [G+1] first event of each neighborhood (CSR of L)
};

// Stage-1 input read straight from the timeseries instead of a clip array:
// clip element (w, t, i) of neighborhood g is X[ch_w + M (time_i - tmid + t)],
// zero outside [0, N) -- the same values extract_clips_batched writes, so
// the features are bit-identical to those of the clip form.
struct GatherSrc {
    const float* X;          // M x N, x[m + M t]
    int64_t M, N;
    const int64_t* times;    // event centres, neighborhood after neighborhood (ev_off)
    const int32_t* nc;       // neighborhood channels, 0-based, central first
    const int64_t* no;       // [G+1] channel offsets
    int32_t tmid;            // (T + 1) / 2 - 1
};

// ----------------------------------------------------- subsample indices
// pca_subsampled, pca.cpp:106-115: the indices (int64)i for i = 0, inc, 2inc,
// ... < L, accumulated sequentially in double.  One workgroup per
// neighborhood: thread 0 runs the sequential accumulation and keeps its
// value at every kSubSeg-th index (checkpoints in LDS); then every thread
// replays kSubSeg additions from one checkpoint and writes those indices.
// The replay repeats the same additions in the same order, so the indices
// are those of the one sequential loop.  Neighborhoods with more checkpoints
// than the LDS holds run the whole loop on thread 0.
constexpr int kSubSeg = 64;
constexpr int kSubCk = 4096;

__global__ void __launch_bounds__(256) pca_subsample_kernel(PcaPlan p, PcaTables tb) {
    __shared__ double ck[kSubCk];
    __shared__ int64_t nseg_s;
    const int g = blockIdx.x;
    const int64_t L = tb.L[g];
    if (L <= p.max_s) {
        if (threadIdx.x == 0) tb.ns[g] = static_cast<int32_t>(L);
        return;
    }
    const double increment = L * 1.0 / p.max_s;
    int32_t* out = tb.idx + g * p.idx_stride;
    const int64_t cap = p.idx_stride;
    const int64_t maxseg = (cap + kSubSeg - 1) / kSubSeg;
    if (maxseg > kSubCk) {  // one thread, the whole loop
        if (threadIdx.x == 0) {
            int64_t c = 0;
            for (double i = 0; i < L && c < cap; i += increment) out[c++] = static_cast<int32_t>(static_cast<int64_t>(i));
            tb.ns[g] = static_cast<int32_t>(c);
        }
        return;
    }
    // checkpoints: the accumulated value at indices 0, kSubSeg, 2 kSubSeg, ...
    // (the values increase, so a segment whose end value is below L lies
    // wholly below it; the segment that reaches L holds the last index)
    if (threadIdx.x == 0) {
        double i = 0;
        int64_t seg = 0;
        for (;;) {
            ck[seg++] = i;
            double j = i;
#pragma unroll
            for (int k = 0; k < kSubSeg; ++k) j += increment;
            if (!(j < L) || seg >= maxseg) break;
            i = j;
        }
        nseg_s = seg;
    }
    __syncthreads();
    const int64_t nseg = nseg_s;
    for (int64_t seg = threadIdx.x; seg < nseg; seg += blockDim.x) {
        double i = ck[seg];
        int64_t c = seg * kSubSeg;
        const int64_t c1 = c + kSubSeg < cap ? c + kSubSeg : cap;
        for (; c < c1 && i < L; ++c) {
            out[c] = static_cast<int32_t>(static_cast<int64_t>(i));
            i += increment;
        }
        if (seg == nseg - 1) tb.ns[g] = static_cast<int32_t>(c);
    }
}

// ------------------------------------------------------------------ Gram
// Work item = (neighborhood, chunk of up to gram_chunk subsampled columns,
// sub-block): a group of up to gw channels for stage 1, a band pair for
// stage 2.  The items of one chunk write disjoint parts of the chunk's FP64
// partial Gram, and pca_gram_reduce_kernel sums a neighborhood's chunk
// partials in chunk order.
//
// Precision.  Each 32-column step is scaled by a power of two s that puts
// the step's largest |x| in [2^14, 2^15) (exact), and every scaled value is
// split into two fp16 pieces, x s = h + l (h = fp16(x s), l = fp16(x s - h),
// 22 significant bits), of which the three products of weight >= 2^-11 are
// accumulated: hh + hl + lh (the dropped ll is below 2^-22 of |x y|).  fp16
// products are exact in FP32; the FP32 accumulator is cleared at every step
// and its sum, times 1/s^2 (exact), added to FP64 accumulators, so rounding
// grows with one step's sum, not with the whole chunk's.  (FP32
// accumulation over a whole chunk leaves 1e-6 .. 1e-5, which the power
// method's later components amplify past 1e-3 on some neighborhoods.)
//
// Each K step stages 32 columns -- the item's rows of each: T rows per
// channel of the group for stage 1, the rows of bands A and B of the nf1*W
// reduced vector for stage 2 -- in LDS as the two fp16 pieces laid out
// [sub-matrix][row][k]: thread o loads element o of the 32 columns
// (coalesced across threads; the column bases are uniform) and writes each
// piece of 8 columns as one 16-byte store, and an MFMA operand fragment
// (8 consecutive k of one row) is one 16-byte load.
// Rows have a 40-element stride and each sub-matrix an odd row count, so
// stores spread over the banks.  Operands of the last row block read up to
// 15 rows past a sub-matrix's end (the next sub-matrix, or 16 rows of slack
// after the last plane): such rows feed only output rows and columns that
// are not written, so the loads need no lane mask.  Waves own Gram tiles
// (q, I, J).
constexpr int kRowStride = 40;
constexpr int kPieces = 2;

// rows allocated per sub-matrix: odd
__host__ __device__ inline int gram_rows_alloc(int nrows) { return nrows | 1; }

template <int kStage, bool kGather = false>
__global__ void __launch_bounds__(kGramThreads) pca_gram_kernel(const float* __restrict__ src, PcaPlan p,
                                                       PcaTables tb, double* __restrict__ part,
                                                       GatherSrc gs) {
    static_assert(kStage == 1 || !kGather, "only stage 1 reads the timeseries");
    extern __shared__ _Float16 gl[];
    const int32_t* item = (kStage == 1 ? tb.gram1_items : tb.gram2_items) + 3 * blockIdx.x;
    const int g = item[0];
    if (g < 0) return;  // padding item (xcd_order)
    const int chunk = item[1];
    const int sub = item[2];
    const int L = tb.L[g];
    const int W = tb.W[g];
    const int ns = tb.ns[g];
    // rows of this item: stage 1 channels q0 .. q0 + nsub - 1 (T rows each);
    // stage 2 band A (nA rows) then, when B != A, band B (nB rows)
    int nsub, nrows, S, q0 = 0, A = 0, B = 0, bs = 0, nA = 0, nB = 0;
    if (kStage == 1) {
        q0 = sub * p.gw;
        nsub = W - q0 < p.gw ? W - q0 : p.gw;
        nrows = p.T;
        S = nsub * p.T;
    } else {
        const int R = p.nf1 * W;
        A = sub >> 16;
        B = sub & 0xffff;
        bs = band_rows(R);
        nA = R - A * bs < bs ? R - A * bs : bs;
        nB = A == B ? 0 : (R - B * bs < bs ? R - B * bs : bs);
        nsub = 1;
        nrows = nA + nB;
        S = nrows;
    }
    const int Sfull = kStage == 1 ? W * p.T : p.nf1 * W;  // floats per source column
    const int ra = gram_rows_alloc(nrows);
    const float* base = kGather ? nullptr : src + (kStage == 1 ? tb.clip_off[g] : tb.red_off[g]);
    const int32_t* idx = tb.idx + g * p.idx_stride;
    const bool sub_cols = L > p.max_s;
    // tiles: stage 1 and stage 2 with A == B: J >= I over nb row blocks;
    // stage 2 with A < B: every (I, J), I over band A, J over band B
    const bool pair = kStage == 2 && A != B;
    const int nbI = pair ? nA / 16 : (nrows + 15) / 16;
    const int nbJ = pair ? (nB + 15) / 16 : nbI;
    const int tpq = pair ? nbI * nbJ : nbI * (nbI + 1) / 2;
    const int ntiles = tpq * nsub;
    const int plane = nsub * ra * kRowStride;  // bf16 elements per piece
    // the wave index is uniform: scalar registers for the tile table
    const int wid = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int nw = blockDim.x >> 6;
    // packed (q << 16) | (I << 8) | J, J local to its band
    int tile[kMaxTiles];
#pragma unroll
    for (int k = 0; k < kMaxTiles; ++k) {
        const int t = wid + nw * k;
        int q = 0, I = 0, J = 0;
        if (t < ntiles) {
            q = t / tpq;
            int tt = t - q * tpq;
            if (pair) {
                I = tt / nbJ;
                J = tt - I * nbJ;
            } else {
                int rl = nbI;
                while (tt >= rl) {
                    tt -= rl;
                    ++I;
                    --rl;
                }
                J = I + tt;
            }
        }
        tile[k] = (q << 16) | (I << 8) | J;
    }
    f64x4 acc[kMaxTiles];
#pragma unroll
    for (int k = 0; k < kMaxTiles; ++k) acc[k] = f64x4{0.0, 0.0, 0.0, 0.0};
    const int bOff = pair ? nA : 0;  // local row of band B's first row
    const f16x8* pv = reinterpret_cast<const f16x8*>(gl);
    const int lfrag = (lane & 15) * (kRowStride / 8) + (lane >> 4);
    const int cbeg = chunk * p.gram_chunk;
    int cend = cbeg + p.gram_chunk;
    if (cend > ns) cend = ns;
    // staging slots: element o of a run of ng groups of 8 columns, ng = 1, 2
    // or 4 so that the 4 * S / ng slots fit the threads (S <= 768: stage 1
    // gw * T, stage 2 at most 256 rows); a slot's source offset and LDS row
    // are fixed for the whole item
    const int ng = 4 * S <= static_cast<int>(blockDim.x) ? 1 : (2 * S <= static_cast<int>(blockDim.x) ? 2 : 4);
    const int o = threadIdx.x % S;
    const int kg0 = (threadIdx.x / S) * ng;
    const bool stager = kg0 < 4;
    int soff = 0, srow = 0;
    // gathering: the slot's channel (as an offset into a sample row) and clip row
    int64_t gch = 0;
    const int64_t* gtimes = nullptr;
    if (kStage == 1) {
        const int q = o % nsub;
        const int r = o / nsub;
        soff = r * W + q0 + q;
        srow = q * ra + r;
        if (kGather) {
            gch = gs.nc[gs.no[g] + q0 + q];
            soff = r;
            gtimes = gs.times + tb.ev_off[g];
        }
    } else {
        soff = o < nA ? A * bs + o : B * bs + (o - nA);
        srow = o;
    }
    f16x8* stg = reinterpret_cast<f16x8*>(gl) + srow * (kRowStride / 8);
    // the next step's samples are loaded into registers before the current
    // step's MFMAs, so the load latency overlaps them
    float nxt[kKC];
    // Column indices of the step being fetched sit in colbuf: threads 0..31
    // load the indices two steps ahead into a register (issued with the
    // prefetch, after the split, so no wait before a barrier covers it) and
    // publish them a step later.  Indices are clamped into
    // the chunk (the sample is then replaced by zero), so loads need no branch.
    // (When gathering, colbuf holds each column's first sample time instead.)
    __shared__ int64_t colbuf[kKC];
    // per-wave max |x| of a step's samples, by step parity
    __shared__ float wmax[2][kGramThreads / 64];
    const auto colidx = [&](int c) -> int64_t {
        c = c < cend ? c : cend - 1;
        const int col = sub_cols ? idx[c] : c;
        if (kGather) return gtimes[col] - gs.tmid;
        return col;
    };
    const auto fetch = [&](int c0) {
        const int cb = c0 + 8 * kg0;
#pragma unroll
        for (int j = 0; j < kKC; ++j) {
            if (j >= 8 * ng) break;
            float x;
            if (kGather) {
                const int64_t tt = colbuf[8 * kg0 + j] + soff;
                const bool in = tt >= 0 && tt < gs.N;
                x = gs.X[(in ? tt : 0) * gs.M + gch];
                x = in ? x : 0.0f;
            } else {
                x = base[colbuf[8 * kg0 + j] * Sfull + soff];
            }
            nxt[j] = cb + j < cend ? x : 0.0f;
        }
    };
    // this wave's max |x| of the fetched step into wmax[par]
    const auto publish_max = [&](int par) {
        float m = 0.0f;
        if (stager) {
#pragma unroll
            for (int j = 0; j < kKC; ++j) {
                if (j >= 8 * ng) break;
                m = fmaxf(m, fabsf(nxt[j]));
            }
        }
        for (int o2 = 32; o2 > 0; o2 >>= 1) m = fmaxf(m, __shfl_xor(m, o2, 64));
        if (lane == 0) wmax[par][wid] = m;
    };
    int64_t cnext = 0;
    if (cbeg < cend) {
        if (threadIdx.x < kKC) {
            colbuf[threadIdx.x] = colidx(cbeg + threadIdx.x);
            cnext = colidx(cbeg + kKC + threadIdx.x);
        }
        __syncthreads();
        if (stager) fetch(cbeg);
        publish_max(0);
    }
    int par = 0;
    for (int c0 = cbeg; c0 < cend; c0 += kKC, par ^= 1) {
        __syncthreads();  // previous step's fragments and index reads consumed; wmax[par] written
        if (threadIdx.x < kKC) colbuf[threadIdx.x] = cnext;  // indices of step c0 + kKC
        // the step's scale 2^(14 - e), 2^e <= max |x| < 2^(e+1) (1 for an all-zero step)
        float smax = 0.0f;
        for (int w2 = 0; w2 < nw; ++w2) smax = fmaxf(smax, wmax[par][w2]);
        const float sc = smax > 0.0f ? ldexpf(1.0f, 14 - ilogbf(smax)) : 1.0f;
        const double unsc = 1.0 / (static_cast<double>(sc) * static_cast<double>(sc));
        if (stager) {
#pragma unroll
            for (int g8 = 0; g8 < 4; ++g8) {
                if (g8 >= ng) break;
                f16x8 h8, l8;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const float x = nxt[8 * g8 + e];
                    const _Float16 h = static_cast<_Float16>(__builtin_fmaf(x, sc, 0.0f));
                    h8[e] = h;
                    l8[e] = static_cast<_Float16>(__builtin_fmaf(x, sc, -static_cast<float>(h)));
                }
                stg[kg0 + g8] = h8;
                stg[plane / 8 + kg0 + g8] = l8;
            }
        }
        __syncthreads();
        if (threadIdx.x < kKC) cnext = colidx(c0 + 2 * kKC + threadIdx.x);
        const bool more = c0 + kKC < cend;
        if (stager && more) fetch(c0 + kKC);
        // keep the per-tile fragment addresses inside the loop: hoisted out of
        // it they would exceed the registers and spill
        int lf = lfrag;
        asm volatile("" : "+v"(lf));
#pragma unroll
        for (int k = 0; k < kMaxTiles; ++k) {
            if (wid + nw * k < ntiles) {
                // fragment of row block X: lane (i + 16b) reads row 16X + i, k = 8b..8b+7;
                // a per-lane part plus a scalar per-tile part (row stride 40 = 5 fragments)
                const int q = tile[k] >> 16, I = (tile[k] >> 8) & 255, J = tile[k] & 255;
                const int fa = lf + (q * ra + I * 16) * (kRowStride / 8);
                const int fb = lf + (q * ra + bOff + J * 16) * (kRowStride / 8);
                const f16x8 ah = pv[fa], al = pv[plane / 8 + fa];
                const f16x8 bh = pv[fb], bl = pv[plane / 8 + fb];
                f32x4 s = {0.f, 0.f, 0.f, 0.f};
                s = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, s, 0, 0, 0);
                s = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, s, 0, 0, 0);
                s = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, s, 0, 0, 0);
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[k][r] += static_cast<double>(s[r]) * unsc;
            }
        }
        // the next step's max, once its samples have arrived (after the MFMAs)
        if (more) publish_max(par ^ 1);
    }
    // this item's part of the chunk's partial Gram ([channel][row][row] for
    // stage 1, [row][row] for stage 2), each tile's transpose mirrored.
    // Accumulator element r of lane l holds row 4 (l / 16) + r, column l % 16.
    double* out = part + (kStage == 1 ? tb.gp1_off : tb.gp2_off)[tb.gram_first[g] + chunk];
    const int ldo = kStage == 1 ? p.T : p.nf1 * W;
#pragma unroll
    for (int k = 0; k < kMaxTiles; ++k) {
        if (wid + nw * k < ntiles) {
            const int q = tile[k] >> 16, I = (tile[k] >> 8) & 255, J = tile[k] & 255;
            double* oq = out + static_cast<int64_t>(kStage == 1 ? q0 + q : 0) * ldo * ldo;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int li = I * 16 + 4 * (lane >> 4) + r;  // local rows
                const int lj = J * 16 + (lane & 15);
                const bool ok = pair ? (lj < nB) : (li < nrows && lj < nrows && !(I == J && li > lj));
                if (ok) {
                    const int ii = kStage == 1 ? li : A * bs + li;
                    const int jj = kStage == 1 ? lj : B * bs + lj;
                    oq[static_cast<int64_t>(ii) * ldo + jj] = acc[k][r];
                    oq[static_cast<int64_t>(jj) * ldo + ii] = acc[k][r];
                }
            }
        }
    }
}

// Sum of a neighborhood's chunk partials (FP64, chunk order) into its Gram.
template <int kStage>
__global__ void __launch_bounds__(256) pca_gram_reduce_kernel(PcaPlan p, PcaTables tb,
                                                              const double* __restrict__ part) {
    const int g = blockIdx.y;
    const int W = tb.W[g];
    const int nrows = kStage == 1 ? p.T : p.nf1 * W;
    const int nsub = kStage == 1 ? W : 1;
    const int64_t n = static_cast<int64_t>(nsub) * nrows * nrows;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int i0 = tb.gram_first[g];
    const int nch = tb.gram_first[g + 1] - i0;
    const int64_t* offs = kStage == 1 ? tb.gp1_off : tb.gp2_off;
    // chunk order; 8 chunks' loads in flight ahead of their adds
    double s = 0.0;
    for (int c0 = 0; c0 < nch; c0 += 8) {
        double v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (c0 + k < nch) v[k] = part[offs[i0 + c0 + k] + e];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (c0 + k < nch) s = s + v[k];
    }
    double* out = kStage == 1 ? tb.g1 + tb.g1_off[g] : tb.g2 + tb.g2_off[g];
    out[e] = s;
}

// ----------------------------------------------------------- power method
// One workgroup per Gram matrix (n x n, FP64): held in registers for
// n <= 128 (power_method_reg), else in LDS or the workspace (power_method).
// Components are written as float [row][component].
constexpr int kPowMaxN = 128;
// Row stride of a Gram matrix held in LDS: n rounded up to odd.
__host__ __device__ inline int pow_ld(int n) { return n | 1; }

// Sum over i < n of f(i), computed by every wavefront of the workgroup on
// its own (lane l adds i = l, l + 64, ... in order, then a butterfly over the
// lanes): each wave gets the same value in the same order, with no barrier.
template <typename F>
__device__ __forceinline__ double wave_sum(int n, F f) {
    const int lane = threadIdx.x & 63;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += f(i);
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    return s;
}

// Gs is n x n with row stride ld (odd in LDS, so that the rows read
// together start on different banks).  Matrix-vector products give each row
// tpr adjacent threads (4 for n <= 64, 2 for n <= 128), each summing every
// tpr-th column into two accumulators, combined across the threads by
// shuffles; the norms and the deflation's v'Gv are summed by every wave on
// its own (wave_sum), so an iteration has two barriers.
__device__ __forceinline__ void power_method(double* Gs, int n, int ld, int K, const double* sin_tab, float* comp,
                             double* v, double* w) {
    const int nt = blockDim.x;
    const int tpr = n <= 64 ? 4 : (n <= 128 ? 2 : 1);
    const int sub = threadIdx.x % tpr;
    const int rpp = nt / tpr;  // rows per pass
    for (int k = 0; k < K; ++k) {
        // v = normalize(sin(1..n))
        const double nrm = sqrt(wave_sum(n, [&](int i) { return sin_tab[i] * sin_tab[i]; }));
        for (int i = threadIdx.x; i < n; i += nt) v[i] = nrm != 0.0 ? sin_tab[i] / nrm : sin_tab[i];
        __syncthreads();
        for (int it = 0; it < 11; ++it) {
            for (int r0 = 0; r0 < n; r0 += rpp) {
                const int r = r0 + static_cast<int>(threadIdx.x) / tpr;
                double a0 = 0.0, a1 = 0.0;
                if (r < n) {
                    const double* row = Gs + static_cast<int64_t>(r) * ld;
                    int c = sub;
                    for (; c + tpr < n; c += 2 * tpr) {
                        a0 += row[c] * v[c];
                        a1 += row[c + tpr] * v[c + tpr];
                    }
                    if (c < n) a0 += row[c] * v[c];
                }
                double a = a0 + a1;
                for (int o = 1; o < tpr; o <<= 1) a += __shfl_xor(a, o, 64);
                if (r < n && sub == 0) w[r] = a;
            }
            __syncthreads();
            if (it < 10) {
                const double sigma = sqrt(wave_sum(n, [&](int i) { return w[i] * w[i]; }));
                for (int i = threadIdx.x; i < n; i += nt) v[i] = sigma != 0.0 ? w[i] / sigma : w[i];
                __syncthreads();
            } else {
                // deflation with u = G v
                const double alpha = wave_sum(n, [&](int i) { return v[i] * w[i]; });
                for (int q = threadIdx.x; q < n * n; q += nt) {
                    const int i = q / n;
                    const int j = q - i * n;
                    double& gq = Gs[static_cast<int64_t>(i) * ld + j];
                    gq = (gq - (v[i] * w[j] + w[i] * v[j])) + alpha * (v[i] * v[j]);
                }
                __syncthreads();
            }
        }
        for (int i = threadIdx.x; i < n; i += nt) comp[i * K + k] = static_cast<float>(v[i]);
        __syncthreads();
    }
}

// The same power method for n <= 64 * kHalves with G held in registers: 4
// adjacent threads per row, each with a contiguous block of ceil(n/4) <=
// kCols columns of up to kHalves rows (rows r, r + 64).  A matrix-vector
// product is then register FMAs against v from LDS and two shuffles; the
// deflation updates the registers in place.  Only v and w (2n doubles) live
// in LDS, so many Gram matrices run on one CU at once.
template <int kHalves, int kCols>
__device__ __forceinline__ void power_method_reg(const double* __restrict__ Gsrc, int n, int K,
                                 const double* sin_tab, float* comp, double* v, double* w) {
    constexpr int kRegCols = kCols;
    const int nt = blockDim.x;
    const int sub = threadIdx.x & 3;
    const int cpt = (n + 3) / 4;
    const int c0 = sub * cpt;
    const int nc = c0 + cpt <= n ? cpt : (n > c0 ? n - c0 : 0);
    double g[kHalves][kRegCols];
    int rr[kHalves];
#pragma unroll
    for (int h = 0; h < kHalves; ++h) {
        rr[h] = h * 64 + static_cast<int>(threadIdx.x >> 2);
#pragma unroll
        for (int k = 0; k < kRegCols; ++k)
            g[h][k] = (rr[h] < n && k < nc) ? Gsrc[static_cast<int64_t>(rr[h]) * n + c0 + k] : 0.0;
    }
    for (int k = 0; k < K; ++k) {
        const double nrm = sqrt(wave_sum(n, [&](int i) { return sin_tab[i] * sin_tab[i]; }));
        for (int i = threadIdx.x; i < n; i += nt) v[i] = nrm != 0.0 ? sin_tab[i] / nrm : sin_tab[i];
        __syncthreads();
        for (int it = 0; it < 11; ++it) {
#pragma unroll
            for (int h = 0; h < kHalves; ++h) {
                double a0 = 0.0, a1 = 0.0;
#pragma unroll
                for (int q = 0; q < kRegCols; q += 2) {
                    if (q < nc) a0 += g[h][q] * v[c0 + q];
                    if (q + 1 < nc) a1 += g[h][q + 1] * v[c0 + q + 1];
                }
                double a = a0 + a1;
                a += __shfl_xor(a, 1, 64);
                a += __shfl_xor(a, 2, 64);
                if (rr[h] < n && sub == 0) w[rr[h]] = a;
            }
            __syncthreads();
            if (it < 10) {
                const double sigma = sqrt(wave_sum(n, [&](int i) { return w[i] * w[i]; }));
                for (int i = threadIdx.x; i < n; i += nt) v[i] = sigma != 0.0 ? w[i] / sigma : w[i];
                __syncthreads();
            } else {
                // deflation with u = G v, on the registers
                const double alpha = wave_sum(n, [&](int i) { return v[i] * w[i]; });
#pragma unroll
                for (int h = 0; h < kHalves; ++h) {
                    const bool row_ok = rr[h] < n;
                    const double vi = row_ok ? v[rr[h]] : 0.0, wi = row_ok ? w[rr[h]] : 0.0;
#pragma unroll
                    for (int q = 0; q < kRegCols; ++q) {
                        if (row_ok && q < nc) {
                            const double vj = v[c0 + q], wj = w[c0 + q];
                            g[h][q] = (g[h][q] - (vi * wj + wi * vj)) + alpha * (vi * vj);
                        }
                    }
                }
                __syncthreads();  // v and w are rewritten by the next component
            }
        }
        for (int i = threadIdx.x; i < n; i += nt) comp[i * K + k] = static_cast<float>(v[i]);
        __syncthreads();
    }
}

// kReg: 1 -> G in registers for n <= 64, 2 -> for n <= 128, 0 -> G in LDS
// or the workspace (any n).
template <int kStage, int kReg>
__global__ void __launch_bounds__(256) pca_power_kernel(PcaPlan p, PcaTables tb) {
    extern __shared__ double psh[];
    int g, q;
    if (kStage == 1) {
        g = tb.p1_items[2 * blockIdx.x];
        q = tb.p1_items[2 * blockIdx.x + 1];
    } else {
        g = blockIdx.x;
        q = 0;
    }
    const int W = tb.W[g];
    const int n = kStage == 1 ? p.T : p.nf1 * W;
    const int K = kStage == 1 ? p.nf1 : p.nf2;
    const double* Gsrc = kStage == 1 ? tb.g1 + tb.g1_off[g] + static_cast<int64_t>(q) * n * n
                                     : tb.g2 + tb.g2_off[g];
    float* comp = kStage == 1 ? tb.c1 + tb.c1_off[g] + static_cast<int64_t>(q) * n * K
                              : tb.c2 + tb.c2_off[g];
    if (kReg == 1) {  // G in registers, v and w in LDS
        power_method_reg<1, 16>(Gsrc, n, K, tb.sin_tab, comp, psh, psh + n);
        return;
    }
    if (kReg == 2) {
        power_method_reg<2, 32>(Gsrc, n, K, tb.sin_tab, comp, psh, psh + n);
        return;
    }
    // larger Gram matrices: the FP64 copy of G lives in LDS when it fits,
    // otherwise in the workspace
    const bool in_lds = kStage == 1 || p.nf1 * p.maxW <= kPowMaxN;  // decided per launch
    double* Gs = in_lds ? psh : tb.work2 + tb.w2_off[g];
    const int ld = in_lds ? pow_ld(n) : n;
    double* v = in_lds ? Gs + n * ld : psh;
    double* w = v + n;
    for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
        const int i = e / n;
        Gs[i * ld + (e - i * n)] = Gsrc[e];
    }
    __syncthreads();
    power_method(Gs, n, ld, K, tb.sin_tab, comp, v, w);
}

// ------------------------------------------------------------ projections
// Stage 1 item: up to pe1 events of one neighborhood and a group of up to pw
// of its channels.  The group's components (gw x T x nf1) sit in LDS with
// rows padded to kCPad floats; thread (event pair, channel) reads its two
// clips' samples straight from global memory (consecutive threads ->
// consecutive channels -> contiguous addresses) and forms nf1 dot products
// over T for each, so every component read from LDS serves two events.
// Component rows are padded to kCPad floats (a multiple of 4): 12 when
// nf1 <= 12 (the default 10 features), else 16.
inline int cpad_for(int nf1) { return nf1 <= 12 ? 12 : 16; }

template <int kCPad, bool kGather = false>
__global__ void __launch_bounds__(256) pca_project1_kernel(const float* __restrict__ clips,
                                                           PcaPlan p, PcaTables tb,
                                                           float* __restrict__ reduced,
                                                           GatherSrc gs) {
    extern __shared__ float sp[];
    const int g = tb.j1_items[3 * blockIdx.x];
    if (g < 0) return;  // padding item (xcd_order)
    const int e0 = tb.j1_items[3 * blockIdx.x + 1];
    const int q0 = tb.j1_items[3 * blockIdx.x + 2];
    const int L = tb.L[g];
    const int W = tb.W[g];
    const int Wg = W - q0 < p.pw ? W - q0 : p.pw;
    const int T = p.T;
    const int nf = p.nf1;
    const int ne = (L - e0) < p.pe1 ? (L - e0) : p.pe1;
    float* Cs = sp;  // [Wg][T][kCPad]
    const float* cg = tb.c1 + tb.c1_off[g] + static_cast<int64_t>(q0) * T * nf;
    for (int e = threadIdx.x; e < Wg * T * kCPad; e += blockDim.x) {
        const int a = e % kCPad;
        const int wt = e / kCPad;
        Cs[e] = a < nf ? cg[wt * nf + a] : 0.0f;
    }
    __syncthreads();
    const float* xg = kGather ? nullptr : clips + tb.clip_off[g] + static_cast<int64_t>(e0) * W * T + q0;
    // thread (event pair, channel): the two events share every component read
    const int np = (ne + 1) / 2;
    for (int o = threadIdx.x; o < np * Wg; o += blockDim.x) {
        const int ev = 2 * (o / Wg);
        const int wc = o % Wg;
        const bool two = ev + 1 < ne;
        float acc0[kCPad], acc1[kCPad];
#pragma unroll
        for (int a = 0; a < kCPad; ++a) acc0[a] = acc1[a] = 0.0f;
        // sample t of the two clips: x0[t * xs], x1[t * xs]; when gathering
        // from the timeseries, samples outside [0, N) read as zero (clips near
        // the recording's ends take the bounds-checked loads)
        const float* x0;
        const float* x1;
        int64_t xs = W;
        int64_t ta = 0, tb2 = 0;
        bool edge = false;
        if (kGather) {
            const int64_t* te = gs.times + tb.ev_off[g] + e0;
            const int64_t ch = gs.nc[gs.no[g] + q0 + wc];
            ta = te[ev] - gs.tmid;
            tb2 = two ? te[ev + 1] - gs.tmid : ta;
            edge = ta < 0 || tb2 < 0 || ta + T > gs.N || tb2 + T > gs.N;
            xs = gs.M;
            x0 = gs.X + ch + (edge ? 0 : ta * gs.M);
            x1 = gs.X + ch + (edge ? 0 : tb2 * gs.M);
        } else {
            x0 = xg + static_cast<int64_t>(ev) * W * T + wc;
            x1 = two ? x0 + W * T : x0;
        }
        const auto ld = [&](const float* xp, int64_t t0c, int t) -> float {
            if (!kGather || !edge) return xp[t * xs];
            const int64_t tt = t0c + t;
            const bool in = tt >= 0 && tt < gs.N;
            const float v = xp[(in ? tt : 0) * xs];
            return in ? v : 0.0f;
        };
        const float* c = Cs + wc * T * kCPad;
        constexpr int U = 10;
        int t = 0;
        for (; t + U <= T; t += U) {
            float xa[U], xb[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                xa[u] = ld(x0, ta, t + u);
                xb[u] = ld(x1, tb2, t + u);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const float4* c4 = reinterpret_cast<const float4*>(c + (t + u) * kCPad);
#pragma unroll
                for (int q = 0; q < kCPad / 4; ++q) {
                    const float4 cc = c4[q];
                    acc0[4 * q + 0] = __builtin_fmaf(cc.x, xa[u], acc0[4 * q + 0]);
                    acc0[4 * q + 1] = __builtin_fmaf(cc.y, xa[u], acc0[4 * q + 1]);
                    acc0[4 * q + 2] = __builtin_fmaf(cc.z, xa[u], acc0[4 * q + 2]);
                    acc0[4 * q + 3] = __builtin_fmaf(cc.w, xa[u], acc0[4 * q + 3]);
                    acc1[4 * q + 0] = __builtin_fmaf(cc.x, xb[u], acc1[4 * q + 0]);
                    acc1[4 * q + 1] = __builtin_fmaf(cc.y, xb[u], acc1[4 * q + 1]);
                    acc1[4 * q + 2] = __builtin_fmaf(cc.z, xb[u], acc1[4 * q + 2]);
                    acc1[4 * q + 3] = __builtin_fmaf(cc.w, xb[u], acc1[4 * q + 3]);
                }
            }
        }
        for (; t < T; ++t) {
            const float xa = ld(x0, ta, t);
            const float xb = ld(x1, tb2, t);
#pragma unroll
            for (int a = 0; a < kCPad; ++a) {
                acc0[a] = __builtin_fmaf(c[t * kCPad + a], xa, acc0[a]);
                acc1[a] = __builtin_fmaf(c[t * kCPad + a], xb, acc1[a]);
            }
        }
        float* out = reduced + tb.red_off[g] + static_cast<int64_t>(e0 + ev) * W * nf + q0 + wc;
#pragma unroll
        for (int a = 0; a < kCPad; ++a)
            if (a < nf) out[W * a] = acc0[a];
        if (two) {
#pragma unroll
            for (int a = 0; a < kCPad; ++a)
                if (a < nf) out[W * nf + W * a] = acc1[a];
        }
    }
}

// Stage 2 item: pe2 events; thread (event, feature) dots nf1*W inputs.
__global__ void __launch_bounds__(256) pca_project2_kernel(const float* __restrict__ reduced,
                                                           PcaPlan p, PcaTables tb,
                                                           float* __restrict__ features) {
    extern __shared__ float sq[];
    const int g = tb.j2_items[2 * blockIdx.x];
    const int e0 = tb.j2_items[2 * blockIdx.x + 1];
    const int L = tb.L[g];
    const int W = tb.W[g];
    const int R = p.nf1 * W;
    const int nf = p.nf2;
    const int ne = (L - e0) < p.pe2 ? (L - e0) : p.pe2;
    float* Cs = sq;            // [R][nf]
    float* Xs = sq + R * nf;   // [ne][R]
    // the copies into LDS keep 8 loads per thread in flight before their
    // stores (a plain load-store loop waits out every load)
    const auto stage = [&](float* dst, const float* src, int n) {
        constexpr int kB = 8;
        const int bd = static_cast<int>(blockDim.x);
        for (int e0i = threadIdx.x; e0i < n; e0i += kB * bd) {
            float t8[kB];
#pragma unroll
            for (int k = 0; k < kB; ++k)
                if (e0i + k * bd < n) t8[k] = src[e0i + k * bd];
#pragma unroll
            for (int k = 0; k < kB; ++k)
                if (e0i + k * bd < n) dst[e0i + k * bd] = t8[k];
        }
    };
    stage(Cs, tb.c2 + tb.c2_off[g], R * nf);
    stage(Xs, reduced + tb.red_off[g] + static_cast<int64_t>(e0) * R, ne * R);
    __syncthreads();
    for (int o = threadIdx.x; o < ne * nf; o += blockDim.x) {
        const int ev = o / nf;
        const int a = o - ev * nf;
        float acc = 0.0f;
        const float* x = Xs + ev * R;
        for (int r = 0; r < R; ++r) acc = __builtin_fmaf(Cs[r * nf + a], x[r], acc);
        features[tb.feat_off[g] + static_cast<int64_t>(e0 + ev) * nf + a] = acc;
    }
}

// ------------------------------------------------------------- host plan
struct HostPlan {
    std::vector<int64_t> clip_off, red_off, feat_off, g1_off, g2_off, c1_off, c2_off;
    std::vector<int32_t> L, W, p1, j1, j2;
    std::vector<int64_t> w2_off, gp1_off, gp2_off, ev_off;
    std::vector<int32_t> gram1_items, gram2_items, gram_first;
    int64_t g1_total = 0, g2_total = 0, c1_total = 0, c2_total = 0, w2_total = 0;
    int64_t gp1_total = 0, gp2_total = 0;
    int32_t maxW = 0;
};

struct Tiles {
    int32_t gram_chunk = 2560, pe1 = 64, pe2 = 64;
};

constexpr int64_t kMaxLds = 160 * 1024;

// Stage-1 channels per work item: at most kGroupW, fewer when that many
// channels' staged T-row sub-matrices would not fit the LDS.
inline int32_t group_width(int32_t T, int32_t maxW) {
    const int64_t per = static_cast<int64_t>(sizeof(_Float16)) * kPieces * gram_rows_alloc(T) * kRowStride;
    int64_t gw = (kMaxLds - 2048) / per;
    if (gw > kGroupW) gw = kGroupW;
    if (gw > maxW) gw = maxW;
    return static_cast<int32_t>(gw < 1 ? 1 : gw);
}

// Reorders stage-1 work items (3 ints each, neighborhood first) for the
// last-level caches of the eight XCDs, to which workgroups are dispatched
// round-robin: the neighborhoods are split into eight contiguous ranges of
// about equal item counts, range k's items go to the block indices that
// land on XCD k, and each range's items run in the order of pos (the
// relative position of the item's events in its neighborhood's time-sorted
// list, in 1024 buckets, stable).  The workgroups that run together on an
// XCD then read clips of nearby channels at nearby times.  Ranges are padded
// to one length with items whose neighborhood is -1 (the kernels return at
// once).  Items write disjoint outputs (the Gram chunk partials are summed
// in chunk order), so the order changes no result.  Linear time: the plan
// is built on the host for every batch.
void xcd_order(std::vector<int32_t>& items, const std::vector<float>& pos, int32_t G) {
    constexpr int kXcd = 8;
    constexpr int kBuckets = 1024;
    const size_t n = items.size() / 3;
    if (n < 4 * kXcd) return;
    std::vector<int64_t> cnt(G, 0);
    for (size_t i = 0; i < n; ++i) ++cnt[items[3 * i]];
    std::vector<int32_t> xcd_of(G);
    int64_t acc = 0;
    for (int32_t g = 0; g < G; ++g) {
        const int64_t k = (2 * acc + cnt[g]) * kXcd / (2 * static_cast<int64_t>(n));
        xcd_of[g] = static_cast<int32_t>(k < kXcd - 1 ? k : kXcd - 1);
        acc += cnt[g];
    }
    // counting sort by (XCD, bucket of pos), stable
    std::vector<int64_t> start(kXcd * kBuckets + 1, 0);
    std::vector<int32_t> key(n);
    for (size_t i = 0; i < n; ++i) {
        int b = static_cast<int>(pos[i] * kBuckets);
        b = b < 0 ? 0 : (b >= kBuckets ? kBuckets - 1 : b);
        key[i] = xcd_of[items[3 * i]] * kBuckets + b;
        ++start[key[i] + 1];
    }
    for (int k = 0; k < kXcd * kBuckets; ++k) start[k + 1] += start[k];
    size_t len = 0;
    for (int k = 0; k < kXcd; ++k) {
        const size_t lk = static_cast<size_t>(start[(k + 1) * kBuckets] - start[k * kBuckets]);
        len = lk > len ? lk : len;
    }
    std::vector<int32_t> out(3 * kXcd * len);
    for (size_t j = 0; j < kXcd * len; ++j) {
        out[3 * j] = -1;
        out[3 * j + 1] = out[3 * j + 2] = 0;
    }
    std::vector<int64_t> fill(start.begin(), start.end() - 1);
    for (size_t i = 0; i < n; ++i) {
        const int k = key[i] / kBuckets;
        const int64_t r = fill[key[i]]++ - start[k * kBuckets];  // rank within XCD k
        int32_t* d = &out[3 * (kXcd * static_cast<size_t>(r) + k)];
        d[0] = items[3 * i];
        d[1] = items[3 * i + 1];
        d[2] = items[3 * i + 2];
    }
    items.swap(out);
}

HostPlan host_plan(int32_t G, const int64_t* num_events, const int32_t* widths, int32_t T,
                   int32_t nf1, int32_t nf2, int64_t max_s, const Tiles& tl, int32_t gw,
                   int32_t pw) {
    HostPlan h;
    int32_t maxW = 0;
    for (int32_t g = 0; g < G; ++g) maxW = widths[g] > maxW ? widths[g] : maxW;
    const bool w2_global = static_cast<int64_t>(nf1) * maxW > kPowMaxN;
    int64_t co = 0, ro = 0, fo = 0, eo = 0;
    int32_t nchunks = 0;
    {
        int64_t n1 = 0, n2 = 0;
        for (int32_t g = 0; g < G; ++g) {
            n1 += ceil_div(num_events[g], static_cast<int64_t>(tl.pe1)) * ceil_div(widths[g], pw);
            n2 += ceil_div(num_events[g], static_cast<int64_t>(tl.pe2));
        }
        h.j1.reserve(3 * n1);
        h.j2.reserve(2 * n2);
    }
    for (int32_t g = 0; g < G; ++g) {
        const int64_t L = num_events[g];
        const int32_t W = widths[g];
        h.L.push_back(static_cast<int32_t>(L));
        h.W.push_back(W);
        h.maxW = W > h.maxW ? W : h.maxW;
        h.ev_off.push_back(eo);
        eo += L;
        h.clip_off.push_back(co);
        h.red_off.push_back(ro);
        h.feat_off.push_back(fo);
        co += static_cast<int64_t>(W) * T * L;
        ro += static_cast<int64_t>(W) * nf1 * L;
        fo += static_cast<int64_t>(nf2) * L;
        h.g1_off.push_back(h.g1_total);
        h.g1_total += static_cast<int64_t>(W) * T * T;
        const int64_t R = static_cast<int64_t>(nf1) * W;
        h.g2_off.push_back(h.g2_total);
        h.g2_total += R * R;
        h.w2_off.push_back(h.w2_total);
        if (w2_global) h.w2_total += R * R;
        // Gram chunks: the subsampled column count is min(L, max_s + 1)
        const int64_t ns_bound = L < max_s ? L : max_s + 1;
        const int64_t nch = ns_bound > 0 ? ceil_div(ns_bound, tl.gram_chunk) : 1;
        const int32_t ngroups = static_cast<int32_t>(ceil_div(W, gw));
        const int32_t nb2 = num_bands(static_cast<int>(R));
        h.gram_first.push_back(nchunks);
        for (int64_t c = 0; c < nch; ++c) {
            for (int32_t q = 0; q < ngroups; ++q) {
                h.gram1_items.push_back(g);
                h.gram1_items.push_back(static_cast<int32_t>(c));
                h.gram1_items.push_back(q);
            }
            for (int32_t a = 0; a < nb2; ++a)
                for (int32_t b = a; b < nb2; ++b) {
                    h.gram2_items.push_back(g);
                    h.gram2_items.push_back(static_cast<int32_t>(c));
                    h.gram2_items.push_back((a << 16) | b);
                }
            h.gp1_off.push_back(h.gp1_total);
            h.gp1_total += static_cast<int64_t>(W) * T * T;
            h.gp2_off.push_back(h.gp2_total);
            h.gp2_total += R * R;
            ++nchunks;
        }
        h.c1_off.push_back(h.c1_total);
        h.c1_total += static_cast<int64_t>(W) * T * nf1;
        h.c2_off.push_back(h.c2_total);
        h.c2_total += R * nf2;
        for (int32_t w = 0; w < W; ++w) {
            h.p1.push_back(g);
            h.p1.push_back(w);
        }
        for (int64_t e = 0; e < L; e += tl.pe1)
            for (int32_t q0 = 0; q0 < W; q0 += pw) {
                h.j1.push_back(g);
                h.j1.push_back(static_cast<int32_t>(e));
                h.j1.push_back(q0);
            }
        for (int64_t e = 0; e < L; e += tl.pe2) {
            h.j2.push_back(g);
            h.j2.push_back(static_cast<int32_t>(e));
        }
    }
    h.gram_first.push_back(nchunks);
    h.ev_off.push_back(eo);
    {
        std::vector<float> pos(h.gram1_items.size() / 3);
        for (size_t i = 0; i < pos.size(); ++i) {
            const int32_t g = h.gram1_items[3 * i];
            const int32_t nch = h.gram_first[g + 1] - h.gram_first[g];
            pos[i] = static_cast<float>(h.gram1_items[3 * i + 1]) / static_cast<float>(nch);
        }
        xcd_order(h.gram1_items, pos, G);
        pos.assign(h.j1.size() / 3, 0.0f);
        for (size_t i = 0; i < pos.size(); ++i) {
            const int32_t g = h.j1[3 * i];
            pos[i] = static_cast<float>(h.j1[3 * i + 1]) / static_cast<float>(h.L[g]);
        }
        xcd_order(h.j1, pos, G);
    }
    return h;
}

template <typename T>
T* take_copy(Carver& c, const std::vector<T>& v) {
    return c.take<T>(static_cast<int64_t>(v.size() > 0 ? v.size() : 1));
}

}  // namespace
}  // namespace msamd

using namespace msamd;

namespace {

struct PcaLayout {
    PcaTables tb{};
    int64_t bytes = 0;
    // device destinations of the host tables
    int64_t* d_clip_off = nullptr;
    int64_t* d_red_off = nullptr;
    int64_t* d_feat_off = nullptr;
    int32_t* d_L = nullptr;
    int32_t* d_W = nullptr;
    int64_t* d_g1_off = nullptr;
    int64_t* d_g2_off = nullptr;
    int64_t* d_c1_off = nullptr;
    int64_t* d_c2_off = nullptr;
    int32_t* d_p1 = nullptr;
    int32_t* d_j1 = nullptr;
    int32_t* d_j2 = nullptr;
    double* d_sin = nullptr;
    int64_t* d_w2_off = nullptr;
    int32_t* d_gram1_items = nullptr;
    int32_t* d_gram2_items = nullptr;
    int32_t* d_gram_first = nullptr;
    int64_t* d_gp1_off = nullptr;
    int64_t* d_gp2_off = nullptr;
    int64_t* d_ev_off = nullptr;
    double* part = nullptr;
    int64_t table_bytes = 0;
};

// Room for the subsample list of one neighborhood (at most max_s + 1
// indices), rounded up to a multiple of 4 for 16-byte stores.
inline int64_t idx_stride_for(int64_t max_s) { return align_up(max_s + 2, 4); }

PcaLayout pca_layout(void* ws, const HostPlan& h, int32_t G, int64_t idx_stride) {
    PcaLayout l;
    Carver c(ws);
    l.d_clip_off = take_copy(c, h.clip_off);
    l.d_red_off = take_copy(c, h.red_off);
    l.d_feat_off = take_copy(c, h.feat_off);
    l.d_L = take_copy(c, h.L);
    l.d_W = take_copy(c, h.W);
    l.d_g1_off = take_copy(c, h.g1_off);
    l.d_g2_off = take_copy(c, h.g2_off);
    l.d_c1_off = take_copy(c, h.c1_off);
    l.d_c2_off = take_copy(c, h.c2_off);
    l.d_p1 = take_copy(c, h.p1);
    l.d_j1 = take_copy(c, h.j1);
    l.d_j2 = take_copy(c, h.j2);
    l.d_sin = c.take<double>(kSinTab);
    l.d_w2_off = take_copy(c, h.w2_off);
    l.d_gram1_items = take_copy(c, h.gram1_items);
    l.d_gram2_items = take_copy(c, h.gram2_items);
    l.d_gram_first = take_copy(c, h.gram_first);
    l.d_gp1_off = take_copy(c, h.gp1_off);
    l.d_gp2_off = take_copy(c, h.gp2_off);
    l.d_ev_off = take_copy(c, h.ev_off);
    l.table_bytes = c.bytes();  // the host tables above are uploaded as one image
    double* work2 = c.take<double>(h.w2_total > 0 ? h.w2_total : 1);
    l.part = c.take<double>(std::max<int64_t>(std::max(h.gp1_total, h.gp2_total), 1));
    int32_t* ns = c.take<int32_t>(G);
    int32_t* idx = c.take<int32_t>(static_cast<int64_t>(G) * idx_stride);
    double* g1 = c.take<double>(h.g1_total > 0 ? h.g1_total : 1);
    double* g2 = c.take<double>(h.g2_total > 0 ? h.g2_total : 1);
    float* c1 = c.take<float>(h.c1_total > 0 ? h.c1_total : 1);
    float* c2 = c.take<float>(h.c2_total > 0 ? h.c2_total : 1);
    l.bytes = c.bytes();
    l.tb = PcaTables{l.d_clip_off, l.d_red_off, l.d_feat_off, l.d_L, l.d_W, l.d_g1_off,
                     l.d_g2_off, l.d_c1_off, l.d_c2_off, l.d_p1, l.d_j1, l.d_j2, ns, idx,
                     g1, g2, c1, c2, l.d_sin, work2, l.d_w2_off, l.d_gram1_items,
                     l.d_gram2_items, l.d_gram_first, l.d_gp1_off, l.d_gp2_off, l.d_ev_off};
    return l;
}

// vectors: the stage-2 PCA alone on vectors of widths[g] entries (T = nf1 = 1).
int check_pca_args(int32_t G, const int64_t* num_events, const int32_t* widths, int32_t T,
                   int32_t nf1, int32_t nf2, int64_t max_s, bool vectors = false) {
    MSAMD_REQUIRE(G >= 1 && num_events && widths, "pca_features: bad neighborhood tables");
    MSAMD_REQUIRE(T >= 1 && T <= kMaxT, "pca_features: clip_size %d outside 1..%d", T, kMaxT);
    MSAMD_REQUIRE(nf1 >= 1 && nf1 <= 16 && nf2 >= 1 && nf2 <= 16,
                  "pca_features: feature counts must be 1..16");
    MSAMD_REQUIRE(nf1 <= T, "pca_features: num_features_per_channel > clip_size");
    MSAMD_REQUIRE(max_s >= 1 && max_s < (1LL << 30), "pca_features: bad max_pca_samples");
    for (int32_t g = 0; g < G; ++g) {
        MSAMD_REQUIRE(widths[g] >= 1 && widths[g] <= (vectors ? kMaxR : kMaxW),
                      "pca_features: neighborhood %d has %d channels (1..%d supported)", g,
                      widths[g], vectors ? kMaxR : kMaxW);
        MSAMD_REQUIRE(static_cast<int64_t>(nf1) * widths[g] <= kMaxR,
                      "pca_features: neighborhood %d: stage-2 dimension %d exceeds %d", g,
                      nf1 * widths[g], kMaxR);
        MSAMD_REQUIRE(num_events[g] >= 0 && num_events[g] < (1LL << 31),
                      "pca_features: bad event count");
        MSAMD_REQUIRE(nf2 <= nf1 * widths[g], "pca_features: num_features > stage-2 dimension");
    }
    MSAMD_REQUIRE(T <= kPowMaxN, "pca_features: clip_size too large");
    return MSAMD_OK;
}

// Dynamic LDS of the stage-2 projection for a tile choice.
int64_t project2_lds(int32_t nf1, int32_t maxW, int32_t nf2, int32_t pe2) {
    return static_cast<int64_t>(sizeof(float)) * (static_cast<int64_t>(nf1) * maxW * nf2 +
                                                  static_cast<int64_t>(pe2) * nf1 * maxW);
}

// The stage-2 events per item that fit the LDS next to the components
// (wide neighborhoods take fewer than the requested pe2).
int32_t fit_pe2(int32_t pe2, int32_t nf1, int32_t maxW, int32_t nf2) {
    while (pe2 > 1 && project2_lds(nf1, maxW, nf2, pe2) > kMaxLds) pe2 = (pe2 + 1) / 2;
    return pe2;
}

int check_tiles(const Tiles& t) {
    MSAMD_REQUIRE(t.gram_chunk >= kKC && t.gram_chunk <= (1 << 20) && t.gram_chunk % kKC == 0,
                  "pca tiles: gram_chunk %d must be a multiple of %d in [%d, 2^20]",
                  t.gram_chunk, kKC, kKC);
    MSAMD_REQUIRE(t.pe1 >= 1 && t.pe1 <= 4096, "pca tiles: stage-1 events per item %d", t.pe1);
    MSAMD_REQUIRE(t.pe2 >= 1 && t.pe2 <= 4096, "pca tiles: stage-2 events per item %d", t.pe2);
    return MSAMD_OK;
}

Tiles tiles_from(const int32_t* t) {
    Tiles r;
    if (t) {
        r.gram_chunk = t[0];
        r.pe1 = t[1];
        r.pe2 = t[2];
    }
    return r;
}

}  // namespace

// A batched PCA over fixed neighborhood shapes: host tables, their device
// copies in the caller's workspace, and the launch sequence.  Built once,
// run many times; the run itself only enqueues kernels.
struct msamd_pca_plan {
    HostPlan h;
    PcaPlan p{};
    PcaLayout l;
    Tiles tiles;
    int pow2_reg = 0;  // stage-2 power kernel form (pca_power_kernel's kReg)
    size_t lds_gram1 = 0, lds_power1 = 0, lds_proj1 = 0, lds_gram2 = 0, lds_power2 = 0,
           lds_proj2 = 0;
    // hipGraph of the launch sequence, keyed by the buffers it was captured with
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t capture = nullptr;
    const float* g_clips = nullptr;
    float* g_reduced = nullptr;
    float* g_features = nullptr;
    // the stage-2 PCA alone, on vectors held where the reduced clips would be
    bool vectors = false;
    // tables uploaded into a workspace (false after a deferred create)
    bool attached = false;
};

namespace {

int plan_build(msamd_pca_plan* P, int32_t G, const int64_t* L, const int32_t* W, int32_t T,
               int32_t nf1, int32_t nf2, int64_t max_s, const Tiles& tl, bool vectors = false) {
    int rc = check_pca_args(G, L, W, T, nf1, nf2, max_s, vectors);
    P->vectors = vectors;
    if (rc != MSAMD_OK) return rc;
    int32_t maxW = 0;
    for (int32_t g = 0; g < G; ++g) maxW = W[g] > maxW ? W[g] : maxW;
    rc = check_tiles(tl);
    if (rc != MSAMD_OK) return rc;
    Tiles t = tl;
    t.pe2 = fit_pe2(tl.pe2, nf1, maxW, nf2);
    const int32_t gw = group_width(T, maxW);
    const int32_t pw = maxW < kProjW ? maxW : kProjW;
    P->h = host_plan(G, L, W, T, nf1, nf2, max_s, t, gw, pw);
    P->tiles = t;
    P->p = PcaPlan{G, T, nf1, nf2, max_s, P->h.maxW, gw, pw, idx_stride_for(max_s),
                   t.gram_chunk, t.pe1, t.pe2};
    P->l = pca_layout(nullptr, P->h, G, idx_stride_for(max_s));
    const int n2 = nf1 * P->h.maxW;
    // stage-2 Gram items stage one band (up to 240 rows) or a pair of 128-row bands
    const int rows2 = n2 <= 240 ? n2 : 256;
    P->lds_gram1 = sizeof(_Float16) * (kPieces * gw * gram_rows_alloc(T) + 16) * kRowStride;
    // stage 1 (n = T <= 64): G in registers, v and w in LDS
    P->lds_power1 = sizeof(double) * 2 * T;
    P->lds_proj1 = sizeof(float) * pw * T * cpad_for(nf1);
    P->lds_gram2 = sizeof(_Float16) * (kPieces * gram_rows_alloc(rows2) + 16) * kRowStride;
    // stage 2: G in registers for n2 <= 128 (v and w in LDS), else G in LDS or
    // the workspace
    P->pow2_reg = n2 <= 64 ? 1 : (n2 <= kPowMaxN ? 2 : 0);
    P->lds_power2 = sizeof(double) * (P->pow2_reg ? 2 * n2 : 2 * n2 + 16);
    P->lds_proj2 = static_cast<size_t>(project2_lds(nf1, P->h.maxW, nf2, t.pe2));
    return MSAMD_OK;
}

// Pinned staging for the plan's table image, one grow-only buffer per
// device: the upload is then an asynchronous copy on the caller's stream, and
// the next upload waits for that copy (an event) before refilling the
// buffer.
struct PlanStaging {
    std::mutex mu;
    char* buf = nullptr;
    size_t cap = 0;
    hipEvent_t done = nullptr;
};
constexpr int kMaxDevices = 16;
PlanStaging g_staging[kMaxDevices];

// Copies the host tables into the carved workspace (stream-ordered; the
// tables live as long as the plan).
int plan_upload(msamd_pca_plan* P, void* workspace, int64_t workspace_bytes, hipStream_t st) {
    const HostPlan& h = P->h;
    P->l = pca_layout(workspace, h, P->p.G, P->p.idx_stride);
    PcaLayout& l = P->l;
    if (!workspace || workspace_bytes < l.bytes) {
        set_error("pca_features: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)l.bytes);
        return MSAMD_ERR_WORKSPACE;
    }
    static const std::vector<double> sins = [] {
        std::vector<double> v(kSinTab);
        for (int i = 0; i < kSinTab; ++i) v[i] = std::sin(static_cast<double>(i + 1));
        return v;
    }();
    // one host image of every table at its workspace offset, one copy: from
    // the device's pinned staging buffer (asynchronous), else from pageable
    // memory (waited for)
    const size_t need = static_cast<size_t>(l.table_bytes);
    int dev = 0;
    MSAMD_HIP(hipGetDevice(&dev));
    PlanStaging* stg = dev >= 0 && dev < kMaxDevices ? &g_staging[dev] : nullptr;
    std::unique_lock<std::mutex> stg_lock;
    std::vector<char> pageable;
    char* img = nullptr;
    if (stg) {
        stg_lock = std::unique_lock<std::mutex>(stg->mu);
        if (stg->done) MSAMD_HIP(hipEventSynchronize(stg->done));  // the last copy has read it
        if (stg->cap < need) {
            if (stg->buf) (void)hipHostFree(stg->buf);
            stg->buf = nullptr;
            stg->cap = 0;
            void* b = nullptr;
            if (hipHostMalloc(&b, need, hipHostMallocDefault) == hipSuccess) {
                stg->buf = static_cast<char*>(b);
                stg->cap = need;
            } else {
                (void)hipGetLastError();
            }
        }
        if (stg->buf && !stg->done &&
            hipEventCreateWithFlags(&stg->done, hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            stg->done = nullptr;
        }
        if (stg->buf && stg->done) img = stg->buf;
    }
    if (!img) {
        pageable.resize(need);
        img = pageable.data();
    }
    char* const wsb = static_cast<char*>(workspace);
    auto put = [&](const void* dst, const void* src, size_t n) {
        if (n) std::memcpy(img + (static_cast<const char*>(dst) - wsb), src, n);
    };
    put(l.d_clip_off, h.clip_off.data(), h.clip_off.size() * 8);
    put(l.d_red_off, h.red_off.data(), h.red_off.size() * 8);
    put(l.d_feat_off, h.feat_off.data(), h.feat_off.size() * 8);
    put(l.d_L, h.L.data(), h.L.size() * 4);
    put(l.d_W, h.W.data(), h.W.size() * 4);
    put(l.d_g1_off, h.g1_off.data(), h.g1_off.size() * 8);
    put(l.d_g2_off, h.g2_off.data(), h.g2_off.size() * 8);
    put(l.d_c1_off, h.c1_off.data(), h.c1_off.size() * 8);
    put(l.d_c2_off, h.c2_off.data(), h.c2_off.size() * 8);
    put(l.d_p1, h.p1.data(), h.p1.size() * 4);
    put(l.d_j1, h.j1.data(), h.j1.size() * 4);
    put(l.d_j2, h.j2.data(), h.j2.size() * 4);
    put(l.d_sin, sins.data(), sins.size() * sizeof(double));
    put(l.d_w2_off, h.w2_off.data(), h.w2_off.size() * 8);
    put(l.d_gram1_items, h.gram1_items.data(), h.gram1_items.size() * 4);
    put(l.d_gram2_items, h.gram2_items.data(), h.gram2_items.size() * 4);
    put(l.d_gram_first, h.gram_first.data(), h.gram_first.size() * 4);
    put(l.d_gp1_off, h.gp1_off.data(), h.gp1_off.size() * 8);
    put(l.d_gp2_off, h.gp2_off.data(), h.gp2_off.size() * 8);
    put(l.d_ev_off, h.ev_off.data(), h.ev_off.size() * 8);
    MSAMD_HIP(hipMemcpyAsync(workspace, img, need, hipMemcpyHostToDevice, st));
    if (img == pageable.data())
        MSAMD_HIP(hipStreamSynchronize(st));  // pageable source: the copy completes before return
    else
        MSAMD_HIP(hipEventRecord(stg->done, st));
    auto lds_attr = [](const void* f, size_t bytes) {
        return hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   static_cast<int>(bytes));
    };
    MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&pca_gram_kernel<1>), P->lds_gram1));
    MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&pca_gram_kernel<1, true>), P->lds_gram1));
    MSAMD_HIP(lds_attr(cpad_for(P->p.nf1) == 12
                           ? reinterpret_cast<const void*>(&pca_project1_kernel<12>)
                           : reinterpret_cast<const void*>(&pca_project1_kernel<16>),
                       P->lds_proj1));
    MSAMD_HIP(lds_attr(cpad_for(P->p.nf1) == 12
                           ? reinterpret_cast<const void*>(&pca_project1_kernel<12, true>)
                           : reinterpret_cast<const void*>(&pca_project1_kernel<16, true>),
                       P->lds_proj1));
    MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&pca_gram_kernel<2>), P->lds_gram2));
    MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&pca_power_kernel<2, 0>), P->lds_power2));
    MSAMD_HIP(lds_attr(reinterpret_cast<const void*>(&pca_project2_kernel), P->lds_proj2));
    return MSAMD_OK;
}

// Stage 2 of the launch sequence: Gram, FP64 sums, power method, projection.
int enqueue_stage2(const msamd_pca_plan* P, float* reduced, float* features, hipStream_t st) {
    const HostPlan& h = P->h;
    const PcaPlan& p = P->p;
    const PcaLayout& l = P->l;
    const int32_t G = p.G, nf1 = p.nf1;
    // stage 2 Gram
    const unsigned ng2 = static_cast<unsigned>(h.gram2_items.size() / 3);
    pca_gram_kernel<2><<<ng2, kGramThreads, P->lds_gram2, st>>>(reduced, p, l.tb, l.part,
                                                               GatherSrc{});
    MSAMD_LAUNCH_CHECK();
    {
        const int R = nf1 * h.maxW;
        const dim3 rg(static_cast<unsigned>(ceil_div(static_cast<int64_t>(R) * R, 256)),
                      static_cast<unsigned>(G));
        pca_gram_reduce_kernel<2><<<rg, 256, 0, st>>>(p, l.tb, l.part);
        MSAMD_LAUNCH_CHECK();
    }
    // stage 2 power method
    if (P->pow2_reg == 1)
        pca_power_kernel<2, 1><<<G, 256, P->lds_power2, st>>>(p, l.tb);
    else if (P->pow2_reg == 2)
        pca_power_kernel<2, 2><<<G, 256, P->lds_power2, st>>>(p, l.tb);
    else
        pca_power_kernel<2, 0><<<G, 256, P->lds_power2, st>>>(p, l.tb);
    MSAMD_LAUNCH_CHECK();
    if (!h.j2.empty()) {  // stage 2 projections
        pca_project2_kernel<<<static_cast<unsigned>(h.j2.size() / 2), 256, P->lds_proj2, st>>>(
            reduced, p, l.tb, features);
        MSAMD_LAUNCH_CHECK();
    }
    return MSAMD_OK;
}


// The launch sequence: nine kernels, each one launch over every neighborhood.
// With gs, the stage-1 kernels read the clips' samples from the timeseries
// (clips is then unused).
int plan_enqueue(const msamd_pca_plan* P, const float* clips, float* reduced, float* features,
                 hipStream_t st, const GatherSrc* gs = nullptr) {
    const HostPlan& h = P->h;
    const PcaPlan& p = P->p;
    const PcaLayout& l = P->l;
    const int32_t G = p.G, T = p.T, nf1 = p.nf1;
    pca_subsample_kernel<<<static_cast<unsigned>(G), 256, 0, st>>>(p, l.tb);
    MSAMD_LAUNCH_CHECK();
    if (P->vectors) return enqueue_stage2(P, reduced, features, st);
    // stage 1 Gram: chunk partials, then per-neighborhood FP64 sums
    const unsigned ng1 = static_cast<unsigned>(h.gram1_items.size() / 3);
    if (gs)
        pca_gram_kernel<1, true><<<ng1, kGramThreads, P->lds_gram1, st>>>(clips, p, l.tb, l.part,
                                                                          *gs);
    else
        pca_gram_kernel<1><<<ng1, kGramThreads, P->lds_gram1, st>>>(clips, p, l.tb, l.part,
                                                                    GatherSrc{});
    MSAMD_LAUNCH_CHECK();
    {
        const dim3 rg(static_cast<unsigned>(ceil_div(static_cast<int64_t>(h.maxW) * T * T, 256)),
                      static_cast<unsigned>(G));
        pca_gram_reduce_kernel<1><<<rg, 256, 0, st>>>(p, l.tb, l.part);
        MSAMD_LAUNCH_CHECK();
    }
    // stage 1 power method
    pca_power_kernel<1, 1><<<static_cast<unsigned>(h.p1.size() / 2), 256, P->lds_power1, st>>>(p, l.tb);
    MSAMD_LAUNCH_CHECK();
    if (!h.j1.empty()) {  // stage 1 projections
        const unsigned g1 = static_cast<unsigned>(h.j1.size() / 3);
        const bool c12 = cpad_for(nf1) == 12;
        if (gs && c12)
            pca_project1_kernel<12, true><<<g1, 256, P->lds_proj1, st>>>(clips, p, l.tb, reduced, *gs);
        else if (gs)
            pca_project1_kernel<16, true><<<g1, 256, P->lds_proj1, st>>>(clips, p, l.tb, reduced, *gs);
        else if (c12)
            pca_project1_kernel<12><<<g1, 256, P->lds_proj1, st>>>(clips, p, l.tb, reduced,
                                                                    GatherSrc{});
        else
            pca_project1_kernel<16><<<g1, 256, P->lds_proj1, st>>>(clips, p, l.tb, reduced,
                                                                    GatherSrc{});
        MSAMD_LAUNCH_CHECK();
    }
    return enqueue_stage2(P, reduced, features, st);
}

void plan_free_graph(msamd_pca_plan* P) {
    if (P->exec) (void)hipGraphExecDestroy(P->exec);  // teardown: nothing to report
    if (P->graph) (void)hipGraphDestroy(P->graph);
    P->exec = nullptr;
    P->graph = nullptr;
}

}  // namespace

extern "C" int64_t msamd_pca_workspace_size(int32_t G, const int64_t* num_events_host,
                                            const int32_t* widths_host, int32_t T,
                                            int32_t num_features_per_channel,
                                            int32_t num_features, int64_t max_pca_samples) {
    return msamd_pca_plan_workspace_size(G, num_events_host, widths_host, T,
                                         num_features_per_channel, num_features, max_pca_samples,
                                         nullptr);
}

extern "C" int msamd_pca_features(const float* clips, int32_t G, const int64_t* num_events_host,
                                  const int32_t* widths_host, int32_t T,
                                  int32_t num_features_per_channel, int32_t num_features,
                                  int64_t max_pca_samples, float* reduced, float* features,
                                  void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    msamd_pca_plan P;
    int rc = plan_build(&P, G, num_events_host, widths_host, T, num_features_per_channel,
                        num_features, max_pca_samples, Tiles{});
    if (rc != MSAMD_OK) return rc;
    hipStream_t st = as_stream(stream);
    rc = plan_upload(&P, workspace, workspace_bytes, st);
    if (rc != MSAMD_OK) return rc;
    return plan_enqueue(&P, clips, reduced, features, st);
}

extern "C" int64_t msamd_pca_plan_workspace_size(int32_t G, const int64_t* num_events_host,
                                                 const int32_t* widths_host, int32_t T,
                                                 int32_t num_features_per_channel,
                                                 int32_t num_features, int64_t max_pca_samples,
                                                 const int32_t* tiles) {
    msamd_pca_plan P;
    if (plan_build(&P, G, num_events_host, widths_host, T, num_features_per_channel,
                   num_features, max_pca_samples, tiles_from(tiles)) != MSAMD_OK)
        return -1;
    return P.l.bytes;
}

extern "C" int msamd_pca_plan_create(int32_t G, const int64_t* num_events_host,
                                     const int32_t* widths_host, int32_t T,
                                     int32_t num_features_per_channel, int32_t num_features,
                                     int64_t max_pca_samples, const int32_t* tiles,
                                     void* workspace, int64_t workspace_bytes, void* stream,
                                     msamd_pca_plan** plan) {
    clear_error();
    MSAMD_REQUIRE(plan, "pca_plan_create: null plan pointer");
    *plan = nullptr;
    msamd_pca_plan* P = new msamd_pca_plan();
    int rc = plan_build(P, G, num_events_host, widths_host, T, num_features_per_channel,
                        num_features, max_pca_samples, tiles_from(tiles));
    const bool deferred = !workspace && workspace_bytes == 0;
    if (rc == MSAMD_OK && !deferred) rc = plan_upload(P, workspace, workspace_bytes, as_stream(stream));
    if (rc != MSAMD_OK) {
        delete P;
        return rc;
    }
    P->attached = !deferred;
    *plan = P;
    return MSAMD_OK;
}

extern "C" int64_t msamd_pca_plan_workspace_bytes(const msamd_pca_plan* plan) {
    return plan ? plan->l.bytes : -1;
}

extern "C" int msamd_pca_plan_attach(msamd_pca_plan* plan, void* workspace,
                                     int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(plan, "pca_plan_attach: null plan");
    msamd_pca_plan* P = plan;
    plan_free_graph(P);  // a graph captured on the old workspace is stale
    P->attached = false;
    const int rc = plan_upload(P, workspace, workspace_bytes, as_stream(stream));
    if (rc != MSAMD_OK) return rc;
    P->attached = true;
    return MSAMD_OK;
}

extern "C" int msamd_pca_plan_run(msamd_pca_plan* plan, const float* clips, float* reduced,
                                  float* features, void* stream) {
    clear_error();
    MSAMD_REQUIRE(plan && clips && reduced && features, "pca_plan_run: null argument");
    MSAMD_REQUIRE(plan->attached, "pca_plan_run: plan has no workspace (msamd_pca_plan_attach)");
    return plan_enqueue(plan, clips, reduced, features, as_stream(stream));
}

namespace {

int gather_source(const msamd_pca_plan* P, const float* X, int64_t M, int64_t N,
                  const int64_t* times, const int32_t* nbhd_channels,
                  const int64_t* nbhd_offsets, GatherSrc* gs) {
    MSAMD_REQUIRE(X && times && nbhd_channels && nbhd_offsets,
                  "pca timeseries: null argument");
    MSAMD_REQUIRE(M >= 1 && N >= 0, "pca timeseries: bad shape");
    MSAMD_REQUIRE(!P || !P->vectors, "pca timeseries: plan holds the stage-2 PCA alone");
    *gs = GatherSrc{X, M, N, times, nbhd_channels, nbhd_offsets,
                    P ? (P->p.T + 1) / 2 - 1 : 0};
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_pca_plan_run_timeseries(msamd_pca_plan* plan, const float* X, int64_t M,
                                             int64_t N, const int64_t* times,
                                             const int32_t* nbhd_channels,
                                             const int64_t* nbhd_offsets, float* reduced,
                                             float* features, void* stream) {
    clear_error();
    MSAMD_REQUIRE(plan && reduced && features, "pca_plan_run_timeseries: null argument");
    MSAMD_REQUIRE(plan->attached,
                  "pca_plan_run_timeseries: plan has no workspace (msamd_pca_plan_attach)");
    GatherSrc gs;
    const int rc = gather_source(plan, X, M, N, times, nbhd_channels, nbhd_offsets, &gs);
    if (rc != MSAMD_OK) return rc;
    return plan_enqueue(plan, nullptr, reduced, features, as_stream(stream), &gs);
}

extern "C" int msamd_pca_plan_launch_graph(msamd_pca_plan* plan, const float* clips,
                                           float* reduced, float* features, void* stream) {
    clear_error();
    MSAMD_REQUIRE(plan && clips && reduced && features, "pca_plan_launch_graph: null argument");
    MSAMD_REQUIRE(plan->attached, "pca_plan_launch_graph: plan has no workspace (msamd_pca_plan_attach)");
    msamd_pca_plan* P = plan;
    if (!P->exec || P->g_clips != clips || P->g_reduced != reduced || P->g_features != features) {
        plan_free_graph(P);
        if (!P->capture) MSAMD_HIP(hipStreamCreateWithFlags(&P->capture, hipStreamNonBlocking));
        MSAMD_HIP(hipStreamBeginCapture(P->capture, hipStreamCaptureModeThreadLocal));
        const int rc = plan_enqueue(P, clips, reduced, features, P->capture);
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(P->capture, &g);
        if (rc != MSAMD_OK) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        MSAMD_HIP(ec);
        P->graph = g;
        MSAMD_HIP(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
        P->g_clips = clips;
        P->g_reduced = reduced;
        P->g_features = features;
    }
    MSAMD_HIP(hipGraphLaunch(P->exec, as_stream(stream)));
    return MSAMD_OK;
}

extern "C" void msamd_pca_plan_tiles(const msamd_pca_plan* plan, int32_t* tiles) {
    if (!plan || !tiles) return;
    tiles[0] = plan->tiles.gram_chunk;
    tiles[1] = plan->tiles.pe1;
    tiles[2] = plan->tiles.pe2;
}

extern "C" void msamd_pca_plan_destroy(msamd_pca_plan* plan) {
    if (!plan) return;
    plan_free_graph(plan);
    if (plan->capture) (void)hipStreamDestroy(plan->capture);
    delete plan;
}

// ------------------------------------------------------------- autotuner
namespace {

constexpr int32_t kTuneGram[] = {1280, 2560, 5120, 10240};
constexpr int32_t kTunePe1[] = {32, 64, 128, 256};
constexpr int32_t kTunePe2[] = {16, 32, 64, 128};

template <typename F>
void for_each_candidate(const Tiles& base, F&& f) {
    for (int32_t v : kTuneGram) { Tiles t = base; t.gram_chunk = v; f(0, t); }
    for (int32_t v : kTunePe1) { Tiles t = base; t.pe1 = v; f(1, t); }
    for (int32_t v : kTunePe2) { Tiles t = base; t.pe2 = v; f(2, t); }
}

}  // namespace

extern "C" int64_t msamd_pca_autotune_workspace_size(int32_t G, const int64_t* num_events_host,
                                                     const int32_t* widths_host, int32_t T,
                                                     int32_t num_features_per_channel,
                                                     int32_t num_features,
                                                     int64_t max_pca_samples) {
    int64_t best = -1;
    // every candidate of every coordinate, on top of any base the search can reach
    for (int32_t gc : kTuneGram)
        for (int32_t e1 : kTunePe1)
            for (int32_t e2 : kTunePe2) {
                const int32_t t[3] = {gc, e1, e2};
                const int64_t b = msamd_pca_plan_workspace_size(
                    G, num_events_host, widths_host, T, num_features_per_channel, num_features,
                    max_pca_samples, t);
                if (b < 0) continue;  // does not fit this shape: the search skips it too
                best = b > best ? b : best;
            }
    return best;
}

namespace {

int autotune(const float* clips, const GatherSrc* gs, int32_t G, const int64_t* num_events_host,
             const int32_t* widths_host, int32_t T, int32_t num_features_per_channel,
             int32_t num_features, int64_t max_pca_samples, float* reduced, float* features,
             void* workspace, int64_t workspace_bytes, void* stream, int32_t* tiles_out,
             float* best_ms_out) {
    MSAMD_REQUIRE(tiles_out, "pca_autotune: null tiles_out");
    hipStream_t st = as_stream(stream);
    hipEvent_t e0, e1;
    MSAMD_HIP(hipEventCreate(&e0));
    MSAMD_HIP(hipEventCreate(&e1));
    Tiles best{};
    float best_ms = -1.f;
    int rc = MSAMD_OK;
    // coordinate search: each tile dimension in turn, the others at their best so far
    for (int dim = 0; dim < 3 && rc == MSAMD_OK; ++dim) {
        Tiles base = best;
        Tiles dim_best = best;
        float dim_ms = -1.f;
        for_each_candidate(base, [&](int d, const Tiles& t) {
            if (d != dim || rc != MSAMD_OK) return;
            msamd_pca_plan P;
            if (check_tiles(t) != MSAMD_OK) {
                clear_error();  // candidate outside the tile limits: skipped
                return;
            }
            rc = plan_build(&P, G, num_events_host, widths_host, T, num_features_per_channel,
                            num_features, max_pca_samples, t);
            if (rc == MSAMD_OK) rc = plan_upload(&P, workspace, workspace_bytes, st);
            float ms = 0.f;
            for (int rep = 0; rep < 2 && rc == MSAMD_OK; ++rep) {  // second run timed
                if (hipEventRecord(e0, st) != hipSuccess) { rc = MSAMD_ERR_HIP; break; }
                rc = plan_enqueue(&P, clips, reduced, features, st, gs);
                if (rc != MSAMD_OK) break;
                if (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                    hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                    set_error("pca_autotune: event timing failed");
                    rc = MSAMD_ERR_HIP;
                }
            }
            if (rc == MSAMD_OK && (dim_ms < 0.f || ms < dim_ms)) {
                dim_ms = ms;
                dim_best = t;
            }
        });
        best = dim_best;
        best_ms = dim_ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != MSAMD_OK) return rc;
    tiles_out[0] = best.gram_chunk;
    tiles_out[1] = best.pe1;
    tiles_out[2] = best.pe2;
    if (best_ms_out) *best_ms_out = best_ms;
    return MSAMD_OK;
}

}  // namespace

extern "C" int msamd_pca_autotune(const float* clips, int32_t G, const int64_t* num_events_host,
                                  const int32_t* widths_host, int32_t T,
                                  int32_t num_features_per_channel, int32_t num_features,
                                  int64_t max_pca_samples, float* reduced, float* features,
                                  void* workspace, int64_t workspace_bytes, void* stream,
                                  int32_t* tiles_out, float* best_ms_out) {
    clear_error();
    MSAMD_REQUIRE(clips, "pca_autotune: null clips");
    return autotune(clips, nullptr, G, num_events_host, widths_host, T, num_features_per_channel,
                    num_features, max_pca_samples, reduced, features, workspace,
                    workspace_bytes, stream, tiles_out, best_ms_out);
}

extern "C" int msamd_pca_autotune_timeseries(
    const float* X, int64_t M, int64_t N, const int64_t* times, const int32_t* nbhd_channels,
    const int64_t* nbhd_offsets, int32_t G, const int64_t* num_events_host,
    const int32_t* widths_host, int32_t T, int32_t num_features_per_channel,
    int32_t num_features, int64_t max_pca_samples, float* reduced, float* features,
    void* workspace, int64_t workspace_bytes, void* stream, int32_t* tiles_out,
    float* best_ms_out) {
    clear_error();
    GatherSrc gs;
    const int rc = gather_source(nullptr, X, M, N, times, nbhd_channels, nbhd_offsets, &gs);
    if (rc != MSAMD_OK) return rc;
    MSAMD_REQUIRE(T >= 1, "pca_autotune: bad clip_size");
    gs.tmid = (T + 1) / 2 - 1;
    return autotune(nullptr, &gs, G, num_events_host, widths_host, T, num_features_per_channel,
                    num_features, max_pca_samples, reduced, features, workspace,
                    workspace_bytes, stream, tiles_out, best_ms_out);
}

// --------------------------------------------------- PCA of index subsets
// The PCA of sort_clips_subset (packages/mountainsortalg/sort_clips.cpp:47-70)
// for many subsets at once: subset g takes the columns indices[off[g] ..
// off[g+1]) of an R x L array, in that order; each subset's columns are
// gathered into one packed R x L_g block and the stage-2 kernels run on all
// blocks (pca_subsampled column rule, power method, projection).
namespace {

__global__ void __launch_bounds__(256) subset_gather_kernel(const float* __restrict__ data, int64_t R,
                                                            const int64_t* __restrict__ indices,
                                                            int64_t total, float* __restrict__ out) {
    const int64_t n = total * R;
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t j = q / R;
        out[q] = data[indices[j] * R + (q - j * R)];
    }
}

int subsets_plan(msamd_pca_plan* P, int32_t G, const int64_t* offsets, int32_t R, int32_t nf,
                 int64_t max_s, std::vector<int64_t>& sizes) {
    MSAMD_REQUIRE(G >= 1 && offsets && offsets[0] == 0, "pca_subsets: bad subset offsets");
    MSAMD_REQUIRE(R >= 1 && R <= kMaxR, "pca_subsets: vector length %d outside 1..%d", R, kMaxR);
    sizes.resize(G);
    for (int32_t g = 0; g < G; ++g) {
        MSAMD_REQUIRE(offsets[g + 1] >= offsets[g], "pca_subsets: offsets not ascending");
        sizes[g] = offsets[g + 1] - offsets[g];
    }
    std::vector<int32_t> widths(G, R);
    return plan_build(P, G, sizes.data(), widths.data(), 1, 1, nf, max_s, Tiles{}, true);
}

}  // namespace

extern "C" int64_t msamd_pca_subsets_workspace_size(int32_t G, const int64_t* offsets_host,
                                                    int32_t R, int32_t num_features,
                                                    int64_t max_pca_samples) {
    msamd_pca_plan P;
    std::vector<int64_t> sizes;
    if (subsets_plan(&P, G, offsets_host, R, num_features, max_pca_samples, sizes) != MSAMD_OK) {
        clear_error();
        return -1;
    }
    Carver c(nullptr);
    c.take<float>(std::max<int64_t>(offsets_host[G] * R, 1));
    c.take<char>(P.l.bytes);
    return c.bytes();
}

extern "C" int msamd_pca_subsets(const float* data, int32_t R, int64_t L, const int64_t* indices,
                                 const int64_t* offsets_host, int32_t G, int32_t num_features,
                                 int64_t max_pca_samples, float* features, void* workspace,
                                 int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(L >= 0, "pca_subsets: bad L");
    msamd_pca_plan P;
    std::vector<int64_t> sizes;
    int rc = subsets_plan(&P, G, offsets_host, R, num_features, max_pca_samples, sizes);
    if (rc != MSAMD_OK) return rc;
    const int64_t total = offsets_host[G];
    Carver c(workspace);
    float* gathered = c.take<float>(std::max<int64_t>(total * R, 1));
    char* pws = c.take<char>(P.l.bytes);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("pca_subsets: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    if (total > 0) {
        int64_t blocks = ceil_div(total * R, 256);
        if (blocks > 65536) blocks = 65536;
        subset_gather_kernel<<<static_cast<unsigned>(blocks), 256, 0, st>>>(data, R, indices, total,
                                                                            gathered);
        MSAMD_LAUNCH_CHECK();
    }
    rc = plan_upload(&P, pws, P.l.bytes, st);
    if (rc != MSAMD_OK) return rc;
    return plan_enqueue(&P, nullptr, gathered, features, st);
}
