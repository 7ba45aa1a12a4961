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

// Streaming form for nB <= 24 (M <= 384): G workgroups of 8 waves share a
// time span and hold the output tiles of the upper triangle between them
// (up to 20 per wave); each streams the span once, staging the fragments of
// all channel blocks per 32 samples in a double-buffered LDS area while the
// next 32 samples load.  The G workgroups of one span run on the same XCD,
// so the span is read from HBM about once.
constexpr int kCovAllWaves = 8;
constexpr int kCovAllTPW = 20;
constexpr int kCovAllKSteps = 1200;  // 32-sample steps per workgroup (FP32 partial length)
constexpr int kCovAllItems = 3;      // (channel, 8-sample group) items per thread (4*Mp <= 1536)

struct CovPlan {
    int64_t M, N, span, nsplit;
    int32_t Mp, nB, nTiles, G;
    int32_t all;  // single-pass kernel
};

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
        for (int q = threadIdx.x; q < nitems; q += blockDim.x) {
            const int m = q % p.Mp;
            const int b = q / p.Mp;
            float v[8];
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t t = t0 + 8 * b + e;
                v[e] = (m < p.M && t < te) ? __builtin_nontemporal_load(X + t * p.M + m) : 0.0f;
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
    auto load = [&](int64_t t0) {
#pragma unroll
        for (int it = 0; it < kCovAllItems; ++it) {
            const int q = threadIdx.x + it * kCovAllWaves * 64;
            const int m = q % p.Mp;
            const int b = q / p.Mp;
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int64_t t = t0 + 8 * b + e;
                v[it][e] = (q < nitems && m < p.M && t < te)
                               ? __builtin_nontemporal_load(X + t * p.M + m)
                               : 0.0f;
            }
        }
    };
    load(tb);
    int buf = 0;
    for (int64_t t0 = tb; t0 < te; t0 += 32, buf ^= 1) {
        bf16x8* fh = lds + buf * 2 * plane;
        bf16x8* fl = fh + plane;
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

// FP64 sum of the block partials in split order, / (N-1), mirrored.
__global__ void __launch_bounds__(256) cov_reduce_kernel(const float* __restrict__ part,
                                                         CovPlan p, double* __restrict__ C) {
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
    double s = 0.0;
    for (int64_t sp = 0; sp < p.nsplit; ++sp)
        s = s + static_cast<double>(part[(sp * p.nTiles + tile) * 256 + e]);
    if (p.N > 1) s = s / static_cast<double>(p.N - 1);
    C[ii * p.M + jj] = s;
    C[jj * p.M + ii] = s;
}

// ------------------------------------------------------- whitening matrix
// The power method is a chain of M * 11 dependent matrix-vector products, so
// its speed is set by how fast one product and one exchange complete.  G is
// split by rows over NP workgroups and held in their LDS (24 rows of 384
// doubles = 74 KB each at M = 384); every power step each workgroup forms
// its rows of w = G v plus its partial |w|^2, publishes them, and all
// workgroups meet at a device-scope barrier before normalizing the full
// vector (partials summed in workgroup order, so every workgroup holds the
// same v).  Deflation G <- (I-vv')G(I-vv') is applied to each workgroup's
// rows in the rank-1 form G - v u' - u v' + (v'u) v v' with u = G v, equal
// to the reference's two M^3 products in exact arithmetic.
//
// The launch is an ordinary one (a cooperative launch is refused while
// another process on the GPU holds the cooperative queue); the NP groups
// are few enough to be co-resident on an otherwise idle device, and every
// barrier wait is bounded: on timeout the error flag is raised, every
// workgroup leaves and the call returns an error, so the kernel always
// terminates.
constexpr int kWThreads = 256;
constexpr int kWRowsPerThread = 4;   // M <= kWThreads * kWRowsPerThread = 1024
constexpr unsigned kBarrierSpinLimit = 1u << 22;

struct WPlan {
    int M, NP, RW;
};

inline WPlan w_plan(int64_t M) {
    WPlan p{};
    p.M = static_cast<int>(M);
    int rw = static_cast<int>(ceil_div(M, 16));
    // rows per workgroup so that rows + two vectors fit in ~150 KB of LDS
    const int cap = static_cast<int>((19200 - 2 * M) / M);
    if (rw > cap) rw = cap;
    if (rw < 1) rw = 1;
    p.RW = rw;
    p.NP = static_cast<int>(ceil_div(M, rw));
    return p;
}

__device__ __forceinline__ bool grid_barrier(unsigned* counter, unsigned target, int* err) {
    // every wave's stores complete (workgroup scope: a wait, no cache flush),
    // then one agent-scope release by thread 0 publishes the workgroup's data
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __syncthreads();
    __shared__ int failed;
    if (threadIdx.x == 0) {
        failed = 0;
        __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        unsigned spins = 0;
        // relaxed polling (coherent loads, no invalidate per poll)
        while (__hip_atomic_load(counter, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 ||
                ++spins > kBarrierSpinLimit) {
                __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                failed = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        // one acquire for the workgroup (the CU's caches are shared by its waves)
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    return failed == 0;
}

struct WBufs {
    double* wbuf;     // [2][M] published rows of G v
    double* psum;     // [2][NP] published partial sums
    unsigned* counter;
    int* err;
};

// sum over this workgroup's rows of (G v)_r * y_r into w_out (rows), returns
// the workgroup-local sum of w_r * z_r where z = w (norm) or v (v'u).
__device__ double rows_matvec(const double* Gl, int M, int nrows, const double* v, double* w_out,
                              const double* dotv, double* red) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    for (int r = wid; r < nrows; r += nw) {
        const double* row = Gl + static_cast<int64_t>(r) * M;
        double acc = 0.0;
        for (int n = lane; n < M; n += 64) acc += row[n] * v[n];
        for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
        if (lane == 0) w_out[r] = acc;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int r = 0; r < nrows; ++r) s = s + w_out[r] * (dotv ? dotv[r] : w_out[r]);
        red[0] = s;
    }
    __syncthreads();
    return red[0];
}

__global__ void __launch_bounds__(kWThreads) whitening_power_kernel(
    const double* __restrict__ XXt, WPlan p, const double* __restrict__ v0,
    double* __restrict__ comp, double* __restrict__ sig, WBufs bufs) {
    extern __shared__ double sh[];
    const int M = p.M;
    const int wg = blockIdx.x;
    const int r0 = wg * p.RW;
    const int nrows = (M - r0) < p.RW ? (M - r0) : p.RW;
    double* Gl = sh;                                    // [RW][M]
    double* v = Gl + static_cast<int64_t>(p.RW) * M;     // [M]
    double* u = v + M;                                  // [M]
    double* wl = u + M;                                 // [RW]
    double* red = wl + p.RW;                            // [2]
    for (int64_t q = threadIdx.x; q < static_cast<int64_t>(nrows) * M; q += blockDim.x) {
        const int r = static_cast<int>(q / M);
        const int n = static_cast<int>(q - static_cast<int64_t>(r) * M);
        Gl[q] = XXt[static_cast<int64_t>(r0 + r) * M + n];
    }
    unsigned epoch = 0;
    for (int k = 0; k < M; ++k) {
        for (int i = threadIdx.x; i < M; i += blockDim.x) v[i] = v0[i];
        __syncthreads();
        double sigma = 0.0;
        for (int it = 0; it < 11; ++it) {
            // it < 10: power step w = G v; it == 10: u = G v for the deflation
            const bool defl = (it == 10);
            const int par = epoch & 1;
            const double part = rows_matvec(Gl, M, nrows, v, wl, defl ? v + r0 : nullptr, red);
            for (int r = threadIdx.x; r < nrows; r += blockDim.x) bufs.wbuf[par * M + r0 + r] = wl[r];
            if (threadIdx.x == 0) bufs.psum[par * p.NP + wg] = part;
            ++epoch;
            if (!grid_barrier(bufs.counter, epoch * static_cast<unsigned>(p.NP), bufs.err)) return;
            // published rows and partials are loaded together (after the
            // acquire); the partials are summed in workgroup order
            double wreg[kWRowsPerThread];
#pragma unroll
            for (int k2 = 0; k2 < kWRowsPerThread; ++k2) {
                const int i = threadIdx.x + k2 * kWThreads;
                wreg[k2] = i < M ? bufs.wbuf[par * M + i] : 0.0;
            }
            for (int q = threadIdx.x; q < p.NP; q += blockDim.x) u[q] = bufs.psum[par * p.NP + q];
            __syncthreads();
            if (threadIdx.x == 0) {
                double s = 0.0;
                for (int q = 0; q < p.NP; ++q) s = s + u[q];
                red[1] = s;
            }
            __syncthreads();
            const double tot = red[1];
            if (!defl) {
                sigma = sqrt(tot);
#pragma unroll
                for (int k2 = 0; k2 < kWRowsPerThread; ++k2) {
                    const int i = threadIdx.x + k2 * kWThreads;
                    if (i < M) v[i] = sigma != 0.0 ? wreg[k2] / sigma : wreg[k2];
                }
                __syncthreads();
                if (it == 9) {
                    for (int r = threadIdx.x; r < nrows; r += blockDim.x)
                        comp[static_cast<int64_t>(r0 + r) * M + k] = v[r0 + r];
                    if (wg == 0 && threadIdx.x == 0) sig[k] = sigma;
                }
            } else {
                const double alpha = tot;
                __syncthreads();  // the partials in u are read before u is overwritten
#pragma unroll
                for (int k2 = 0; k2 < kWRowsPerThread; ++k2) {
                    const int i = threadIdx.x + k2 * kWThreads;
                    if (i < M) u[i] = wreg[k2];
                }
                __syncthreads();
                for (int64_t q = threadIdx.x; q < static_cast<int64_t>(nrows) * M; q += blockDim.x) {
                    const int r = static_cast<int>(q / M);
                    const int n = static_cast<int>(q - static_cast<int64_t>(r) * M);
                    const int i = r0 + r;
                    const double cross = v[i] * u[n] + u[i] * v[n];
                    const double vv = v[i] * v[n];
                    Gl[q] = (Gl[q] - cross) + alpha * vv;
                }
                __syncthreads();
            }
        }
    }
}

// W[m, n] = sum_l (C[m,l] d_l) C[n,l], l ascending (mult_AB then
// mult_ABtrans, pca.cpp:648-649), d_l = 1/sqrt(sigma_l) or 0.
// comp is row-major C[i*M + k]; W is written column-major like the Mda.
__global__ void __launch_bounds__(256) whitening_assemble_kernel(
    const double* __restrict__ comp, const double* __restrict__ sig, int M,
    double* __restrict__ W) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= static_cast<int64_t>(M) * M) return;
    const int n = static_cast<int>(q / M);
    const int m = static_cast<int>(q - static_cast<int64_t>(n) * M);
    const double* cm = comp + static_cast<int64_t>(m) * M;
    const double* cn = comp + static_cast<int64_t>(n) * M;
    double acc = 0.0;
    for (int l = 0; l < M; ++l) {
        const double s = sig[l];
        const double d = s != 0.0 ? 1.0 / sqrt(s) : 0.0;
        acc = acc + (cm[l] * d) * cn[l];
    }
    W[m + static_cast<int64_t>(M) * n] = acc;
}

// ------------------------------------------------------------------- apply
// Y[i, t] = sum_j X[j, t] W(j, i) with W(j, i) = Wcm[j + M*i] (the reference
// walks W column-major, p_whiten.cpp:174-182).  MFMA view: D(16 samples x
// 16 outputs) += A(16 samples x 32 channels) B(32 channels x 16 outputs).
//
// Precision: the 1e-4 output quantum of the reference needs FP32-level
// products, and bf16 hi+lo keeps only 16 significant bits (|x - hi - lo| <=
// 2^-16 |x|), which with cancellation inside W x reaches two quanta.  Both
// operands are therefore split into three bf16 pieces x = h + m + l (exact
// for an fp32 value) and the six products of order >= 2^-16 are accumulated:
// hh + hm + mh + hl + lh + mm; the dropped terms are O(2^-24) relative.
//
// Each workgroup keeps the three-piece fragments of CT column tiles (16*CT
// outputs) of W in LDS and streams 8 waves x 64 samples per pass over its
// 2048-sample time block.  Workgroups are numbered so that all column
// blocks of one time block run back to back on the same XCD (blockIdx % 8
// selects the XCD), and the time block is read once from HBM into that
// XCD's L2.
constexpr int kApplyWaves = 8;
constexpr int kApplyRowTiles = 4;    // 16-sample row tiles per wave per pass
constexpr int kApplyPasses = 4;      // passes per workgroup
constexpr int kApplyTB = kApplyWaves * kApplyRowTiles * 16 * kApplyPasses;  // 2048 samples
constexpr int kXcd = 8;

struct ApplyPlan {
    int64_t M, N, nTB, nWG;
    int32_t Mp, nK, nColTiles, nCB, CT;
    double qunit;
};

inline ApplyPlan apply_plan(int64_t M, int64_t N, double qunit) {
    ApplyPlan p{};
    p.M = M;
    p.N = N;
    p.Mp = static_cast<int32_t>(align_up(M, 64));
    p.nK = p.Mp / 32;
    p.nColTiles = p.Mp / 16;
    p.CT = 4;
    while (p.CT > 1 && 3 * p.CT * p.nK * 1024 > 150 * 1024) p.CT >>= 1;
    p.nCB = p.nColTiles / p.CT;
    p.qunit = qunit;
    p.nTB = ceil_div(N, kApplyTB);
    p.nWG = align_up(p.nTB, kXcd) * p.nCB;
    return p;
}

__device__ __forceinline__ void split3_8(const float* v, bf16x8& h, bf16x8& m, bf16x8& l) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const __bf16 hb = static_cast<__bf16>(v[e]);
        const float r1 = v[e] - static_cast<float>(hb);   // exact
        const __bf16 mb = static_cast<__bf16>(r1);
        const float r2 = r1 - static_cast<float>(mb);     // exact
        h[e] = hb;
        m[e] = mb;
        l[e] = static_cast<__bf16>(r2);
    }
}

__device__ __forceinline__ f32x4 mfma6(const bf16x8& ah, const bf16x8& am, const bf16x8& al,
                                       const bf16x8& bh, const bf16x8& bm, const bf16x8& bl,
                                       f32x4 acc) {
    // smallest terms first
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, acc, 0, 0, 0);
    return acc;
}

// Fragments of W: index ((ct * nK + kk) * 64 + lane) within each piece array;
// lane (i + 16b) of column tile ct holds W(32kk + 8b + e, 16ct + i), e = 0..7.
__global__ void __launch_bounds__(256) apply_prep_kernel(const double* __restrict__ Wcm,
                                                         ApplyPlan p, bf16x8* __restrict__ fr) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t total = static_cast<int64_t>(p.nColTiles) * p.nK * 64;
    if (q >= total) return;
    const int lane = static_cast<int>(q & 63);
    const int64_t tk = q >> 6;
    const int kk = static_cast<int>(tk % p.nK);
    const int ct = static_cast<int>(tk / p.nK);
    const int64_t i = ct * 16 + (lane & 15);
    float v[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int64_t j = kk * 32 + 8 * (lane >> 4) + e;
        v[e] = (i < p.M && j < p.M) ? static_cast<float>(Wcm[j + p.M * i]) : 0.0f;
    }
    bf16x8 h, m, l;
    split3_8(v, h, m, l);
    fr[q] = h;
    fr[total + q] = m;
    fr[2 * total + q] = l;
}

template <bool kVec, int CT>
__global__ void __launch_bounds__(kApplyWaves * 64) apply_kernel(
    const float* __restrict__ X, ApplyPlan p, const bf16x8* __restrict__ gfr,
    float* __restrict__ Y, int16_t* __restrict__ Yq) {
    extern __shared__ bf16x8 lds[];
    const int nfr = CT * p.nK * 64;  // fragments of this column block, per piece
    // XCD-aware numbering: all column blocks of a time block share blockIdx % 8
    const int64_t w = blockIdx.x;
    const int64_t xcd = w % kXcd;
    const int64_t jx = w / kXcd;
    const int cb = static_cast<int>(jx % p.nCB);
    const int64_t tb = (jx / p.nCB) * kXcd + xcd;
    if (tb >= p.nTB) return;
    bf16x8* fh = lds;
    bf16x8* fm = lds + nfr;
    bf16x8* fl = lds + 2 * nfr;
    {
        const int64_t total = static_cast<int64_t>(p.nColTiles) * p.nK * 64;
        const int64_t base = static_cast<int64_t>(cb) * nfr;
        for (int q = threadIdx.x; q < nfr; q += blockDim.x) {
            fh[q] = gfr[base + q];
            fm[q] = gfr[total + base + q];
            fl[q] = gfr[2 * total + base + q];
        }
    }
    __syncthreads();
    const int wid = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int li = lane & 15;
    const int lb = lane >> 4;
    const int64_t tblock = tb * kApplyTB;
    for (int pass = 0; pass < kApplyPasses; ++pass) {
        const int64_t row0 = tblock + (static_cast<int64_t>(pass) * kApplyWaves + wid) * (kApplyRowTiles * 16);
        if (row0 >= p.N) break;
        f32x4 acc[kApplyRowTiles][CT];
#pragma unroll
        for (int rt = 0; rt < kApplyRowTiles; ++rt)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[rt][ct] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int kk = 0; kk < p.nK; ++kk) {
            bf16x8 ah[kApplyRowTiles], am[kApplyRowTiles], al[kApplyRowTiles];
#pragma unroll
            for (int rt = 0; rt < kApplyRowTiles; ++rt) {
                const int64_t t = row0 + rt * 16 + li;
                const int64_t j0 = static_cast<int64_t>(kk) * 32 + 8 * lb;
                float v[8];
                if (kVec) {
                    if (t < p.N && j0 < p.M) {
                        const vf4* src = reinterpret_cast<const vf4*>(X + t * p.M + j0);
                        const vf4 a = src[0];
                        const vf4 b = src[1];
                        v[0] = a[0]; v[1] = a[1]; v[2] = a[2]; v[3] = a[3];
                        v[4] = b[0]; v[5] = b[1]; v[6] = b[2]; v[7] = b[3];
                    } else {
#pragma unroll
                        for (int e = 0; e < 8; ++e) v[e] = 0.0f;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < 8; ++e)
                        v[e] = (t < p.N && j0 + e < p.M) ? X[t * p.M + j0 + e] : 0.0f;
                }
                split3_8(v, ah[rt], am[rt], al[rt]);
            }
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int fi = (ct * p.nK + kk) * 64 + lane;
                const bf16x8 bh = fh[fi];
                const bf16x8 bm = fm[fi];
                const bf16x8 bl = fl[fi];
#pragma unroll
                for (int rt = 0; rt < kApplyRowTiles; ++rt)
                    acc[rt][ct] = mfma6(ah[rt], am[rt], al[rt], bh, bm, bl, acc[rt][ct]);
            }
        }
        // epilogue: quantize(x, 1e-4) = floor(x/1e-4 + 0.5)*1e-4 in double,
        // stored as float; optional int16 (int)(y/q + 0.5).
#pragma unroll
        for (int rt = 0; rt < kApplyRowTiles; ++rt) {
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) {
                const int64_t i = static_cast<int64_t>(cb) * (CT * 16) + ct * 16 + li;
                if (i >= p.M) continue;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t t = row0 + rt * 16 + 4 * lb + r;
                    if (t >= p.N) continue;
                    const double xq = floor(static_cast<double>(acc[rt][ct][r]) / 0.0001 + 0.5) * 0.0001;
                    const float yf = static_cast<float>(xq);
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

extern "C" int msamd_whiten_covariance(const float* X, int64_t M, int64_t N, double* XXt,
                                       void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= 1024, "whiten_covariance: M=%lld outside 1..1024", (long long)M);
    MSAMD_REQUIRE(N >= 0, "whiten_covariance: negative N");
    hipStream_t st = as_stream(stream);
    if (N == 0) {
        MSAMD_HIP(hipMemsetAsync(XXt, 0, sizeof(double) * M * M, st));
        return MSAMD_OK;
    }
    const CovPlan p = cov_plan(M, N);
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
    cov_reduce_kernel<<<static_cast<unsigned>(ceil_div(nred, 256)), 256, 0, st>>>(part, p, XXt);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int64_t msamd_whitening_matrix_workspace_size(int64_t M) {
    if (M < 1) return -1;
    const WPlan wp = w_plan(M);
    Carver c(nullptr);
    c.take<double>(M * M);  // components
    c.take<double>(M);      // sigma
    c.take<double>(M);      // v0
    c.take<double>(2 * M);  // published G v rows
    c.take<double>(2 * wp.NP);
    c.take<int>(2);         // barrier counter, error flag
    return c.bytes();
}

extern "C" int msamd_whitening_matrix(const double* XXt, int64_t M, double* W, void* workspace,
                                      int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= 1024, "whitening_matrix: M=%lld outside 1..1024", (long long)M);
    const WPlan wp0 = w_plan(M);
    Carver c(workspace);
    double* comp = c.take<double>(M * M);
    double* sig = c.take<double>(M);
    double* v0 = c.take<double>(M);
    double* wbuf = c.take<double>(2 * M);
    double* psum = c.take<double>(2 * wp0.NP);
    int* flags = c.take<int>(2);
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
    const WPlan wp = w_plan(M);
    WBufs wb{};
    wb.wbuf = wbuf;
    wb.psum = psum;
    wb.counter = reinterpret_cast<unsigned*>(flags);
    wb.err = flags + 1;
    MSAMD_HIP(hipMemsetAsync(flags, 0, 2 * sizeof(int), st));
    const size_t lds = sizeof(double) * (static_cast<size_t>(wp.RW) * M + 2 * M + wp.RW + 2);
    MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&whitening_power_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    whitening_power_kernel<<<static_cast<unsigned>(wp.NP), kWThreads, lds, st>>>(
        XXt, wp, v0, comp, sig, wb);
    MSAMD_LAUNCH_CHECK();
    int herr = 0;
    MSAMD_HIP(hipMemcpyAsync(&herr, flags + 1, sizeof(int), hipMemcpyDeviceToHost, st));
    MSAMD_HIP(hipStreamSynchronize(st));
    if (herr) {
        set_error("whitening_matrix: workgroup barrier timed out");
        return MSAMD_ERR_HIP;
    }
    whitening_assemble_kernel<<<static_cast<unsigned>(ceil_div(M * M, 256)), 256, 0, st>>>(
        comp, sig, static_cast<int>(M), W);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int64_t msamd_whiten_apply_workspace_size(int64_t M) {
    if (M < 1) return -1;
    const ApplyPlan p = apply_plan(M, 1, 0.0);
    Carver c(nullptr);
    c.take<bf16x8>(3 * static_cast<int64_t>(p.nColTiles) * p.nK * 64);
    return c.bytes();
}

extern "C" int msamd_whiten_apply(const float* X, int64_t M, int64_t N, const double* W, float* Y,
                                  double quantization_unit, int16_t* Yq, void* workspace,
                                  int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= 1024, "whiten_apply: M=%lld outside 1..1024", (long long)M);
    MSAMD_REQUIRE(N >= 0, "whiten_apply: negative N");
    MSAMD_REQUIRE(Y || Yq, "whiten_apply: no output");
    MSAMD_REQUIRE(!Yq || quantization_unit > 0, "whiten_apply: int16 output needs quantization_unit > 0");
    const ApplyPlan p = apply_plan(M, N, quantization_unit);
    Carver c(workspace);
    const int64_t nfr = static_cast<int64_t>(p.nColTiles) * p.nK * 64;
    bf16x8* fr = c.take<bf16x8>(3 * nfr);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("whiten_apply: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    apply_prep_kernel<<<static_cast<unsigned>(ceil_div(nfr, 256)), 256, 0, st>>>(W, p, fr);
    MSAMD_LAUNCH_CHECK();
    if (N == 0) return MSAMD_OK;
    const size_t lds = sizeof(bf16x8) * 3 * p.CT * p.nK * 64;
    const dim3 grid(static_cast<unsigned>(p.nWG));
    const bool vec = (M % 8 == 0);
#define MSAMD_APPLY_LAUNCH(V, C)                                                              \
    do {                                                                                      \
        MSAMD_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&apply_kernel<V, C>),      \
                                      hipFuncAttributeMaxDynamicSharedMemorySize,             \
                                      static_cast<int>(lds)));                                \
        apply_kernel<V, C><<<grid, kApplyWaves * 64, lds, st>>>(X, p, fr, Y, Yq);             \
    } while (0)
    if (p.CT == 4) {
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
