// Event detection (a10): per (central channel, padded time chunk) statistics
// and the detect_interval suppression stencil.
//
// Reference: detect_events (packages/mountainsortalg/detect_events.cpp:69-135)
// driven by NeighborhoodSorter::addTimeChunk (neighborhoodsorter.cpp:74-106)
// over the chunk plan of p_mountainsort3.cpp:170-216,507-533.
//
// The reference's sequential candidate/last-best state machine keeps sample n
// exactly when  v[n] > thr,  v[n] > v[u] for every u in [n-D, n-1]  and
// v[n] >= v[u] for every u in [n+1, n+D]  (windows cut at the padded-chunk
// ends), with v = x - mean (negated for sign < 0, |.| for sign == 0) and
// thr = detect_threshold * stdev.  That closed form is a per-sample stencil,
// evaluated here in parallel; only the two statistics need the reference's
// sequential order, and they are computed that way (one thread per
// (chunk, channel), double sums in time order) so that thresholds are
// bit-identical.  Contraction is off in this file: every double operation
// is rounded as on the reference's x86-64 build.
#include "common.h"
#include "tcn.h"

#pragma clang fp contract(off)

namespace msamd {
namespace {

constexpr int kSeg = 1024;   // samples per (segment, channel) scan (fits a uint16 offset)
constexpr int kB = 16;        // register batch of the stencil scan

struct DetectPlan {
    int64_t M, N_file, t_start, n_range, cs, pad, nchunks, spc, nseg;
    int32_t Mc;
    int D;         // floor(detect_interval), lags <= D are suppressed
    int64_t cap;   // slot capacity per (segment, channel)
};

inline DetectPlan make_plan(int64_t M, int64_t N_file, int64_t t_start, int64_t n_range,
                            int64_t cs, int64_t pad, int32_t Mc, double detect_interval) {
    DetectPlan p{};
    p.M = M;
    p.N_file = N_file;
    p.t_start = t_start;
    p.n_range = n_range;
    p.cs = cs;
    p.pad = pad;
    p.Mc = Mc;
    p.nchunks = n_range > 0 ? ceil_div(n_range, cs) : 0;
    p.spc = ceil_div(cs < n_range ? cs : n_range, kSeg);
    if (p.spc < 1) p.spc = 1;
    p.nseg = p.nchunks * p.spc;
    double d = detect_interval;
    p.D = d >= 0 ? static_cast<int>(d) : -1;  // reference compares lag > D in double
    if (p.D > (1 << 20)) p.D = 1 << 20;
    p.cap = p.D >= 1 ? (kSeg + p.D) / (p.D + 1) : kSeg;
    return p;
}

struct DetectBuffers {
    double* mean;      // [nchunks * Mc]
    double* thr;       // [nchunks * Mc]
    int32_t* counts;   // [Mc * nseg] (channel-major)
    int64_t* seg_off;  // [Mc * nseg] exclusive scan per channel
    int64_t* chan_off; // [Mc + 1]
    uint16_t* slots;   // [nseg * Mc * cap]
    DetectPlan* header;  // the plan, read back by msamd_detect_fetch
};

inline DetectBuffers carve(void* ws, const DetectPlan& p, int64_t* bytes) {
    Carver c(ws);
    DetectBuffers b{};
    b.header = c.take<DetectPlan>(1);
    b.mean = c.take<double>(p.nchunks * p.Mc);
    b.thr = c.take<double>(p.nchunks * p.Mc);
    b.counts = c.take<int32_t>(p.nseg * p.Mc);
    b.seg_off = c.take<int64_t>(p.nseg * p.Mc);
    b.chan_off = c.take<int64_t>(p.Mc + 1);
    b.slots = c.take<uint16_t>(p.nseg * p.Mc * p.cap);
    if (bytes) *bytes = c.bytes();
    return b;
}

// One thread per (chunk, channel): the sequential double sums of
// MLCompute::mean / MLCompute::stdev (common/mlcompute.cpp:62-85) over the
// padded chunk.  Zero-filled samples outside the file add exact zeros and are
// skipped.  Lanes map to consecutive channels, so each time step is one
// coalesced row read.
__global__ void __launch_bounds__(256) detect_stats_kernel(
    const float* __restrict__ X, DetectPlan p, const int32_t* __restrict__ channels,
    double k_thr, double* __restrict__ mean_out, double* __restrict__ thr_out) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= p.nchunks * p.Mc) return;
    const int j = static_cast<int>(g % p.Mc);
    const int64_t c = g / p.Mc;
    int64_t size = p.n_range - c * p.cs;
    if (size > p.cs) size = p.cs;
    const int64_t len = size + 2 * p.pad;
    const int64_t base = p.t_start + c * p.cs - p.pad;
    int64_t a = base < 0 ? 0 : base;
    int64_t b = base + len;
    if (b > p.N_file) b = p.N_file;
    const float* xp = X + channels[j];
    double s = 0.0, ss = 0.0;
    int64_t t = a;
    constexpr int U = 16;
    for (; t + U <= b; t += U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) v[u] = __builtin_nontemporal_load(xp + (t + u) * p.M);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double d = static_cast<double>(v[u]);
            s = s + d;
            ss = ss + d * d;  // d*d is exact (24-bit mantissa squared)
        }
    }
    for (; t < b; ++t) {
        const double d = static_cast<double>(xp[t * p.M]);
        s = s + d;
        ss = ss + d * d;
    }
    const double mean = s / static_cast<double>(len);
    const int ct = static_cast<int>(len);
    const double sd = (ct >= 2) ? sqrt((ss - s * s / ct) / (ct - 1)) : 0.0;
    mean_out[c * p.Mc + j] = mean;
    thr_out[c * p.Mc + j] = k_thr * sd;
}

__device__ __forceinline__ double detect_value(float x, double mean, int sign) {
    double v = static_cast<double>(x) - mean;
    if (sign < 0)
        v = -v;
    else if (sign == 0)
        v = fabs(v);
    return v;
}

// One thread per (segment, central channel), lanes over consecutive channels
// so every load is one coalesced row.  The thread streams its segment in
// register batches of kB samples, loading the next batch while it tests the
// current one.  Threshold crossings are rare; for D <= kB their +-D
// neighbours are in the previous, current and next batches held in
// registers, otherwise they are read again through the cache.
__device__ __forceinline__ float load_x(const float* xp, int64_t t, int64_t M, int64_t N) {
    return (t >= 0 && t < N) ? xp[t * M] : 0.0f;
}

// The +-D suppression test of sample t (windows cut at [wa, wb)): v must
// beat every earlier value and be >= every later value within D.  The
// neighbours were just streamed by the wavefront and come from the cache;
// walking them in order keeps the kernel small enough for high occupancy,
// which matters more here than the latency of the rare candidate.
__device__ __forceinline__ bool stencil_ok(const float* xp, int64_t t, double v, double mean,
                                           int sign, int D, int64_t wa, int64_t wb, int64_t M,
                                           int64_t N) {
    for (int d = 1; d <= D; ++d) {
        if (t - d < wa) break;
        if (detect_value(load_x(xp, t - d, M, N), mean, sign) >= v) return false;
    }
    for (int d = 1; d <= D; ++d) {
        if (t + d >= wb) break;
        if (detect_value(load_x(xp, t + d, M, N), mean, sign) > v) return false;
    }
    return true;
}

// detect_value with the sign fixed at compile time
template <int kSign>
__device__ __forceinline__ double detect_value_s(float x, double mean) {
    const double v = static_cast<double>(x) - mean;
    return kSign < 0 ? -v : (kSign == 0 ? fabs(v) : v);
}

// The +-D test of sample t for D <= kB, decided in float whenever float
// decides it exactly: 1 keep, 0 reject, -1 undecided (left to the FP64 test).
// The 2D neighbours are loaded together (one trip through the cache rather
// than one per neighbour).
//
// With s = kSign = +-1, v = s * RN(x - mean) is a non-decreasing function of
// the key s*x, strictly increasing except where the rounding of x - mean
// merges two inputs, which needs the keys to differ by less than
// 2^-51 * (|x| + |mean|).  So with bk / fk the largest key among the earlier
// / later neighbours inside D and the padded window:
//   bk >= kt                 some earlier v >= v_t: reject (exact)
//   fk - kt > tau            some later v > v_t: reject (no merge possible)
//   fk <= kt, kt - bk > tau  every earlier v < v_t, later v <= v_t: keep
// with tau = 2^-40 * (|x_t| + |mean|) far above the merge distance; keys
// within tau stay undecided.
template <int kSign>
__device__ __forceinline__ int stencil_float(const float* xp, int64_t t, float xt, float mabs,
                                             int D, int64_t wa, int64_t wb, int64_t M,
                                             int64_t N) {
    const int nbk = static_cast<int>(t - wa < D ? t - wa : D);
    const int nfw = static_cast<int>(wb - 1 - t < D ? wb - 1 - t : D);
    float xb[kB], xf[kB];
#pragma unroll
    for (int d = 1; d <= kB; ++d) {
        xb[d - 1] = d <= nbk ? load_x(xp, t - d, M, N) : 0.0f;
        xf[d - 1] = d <= nfw ? load_x(xp, t + d, M, N) : 0.0f;
    }
    const float kt = kSign < 0 ? -xt : xt;
    float bk = -INFINITY, fk = -INFINITY;
#pragma unroll
    for (int d = 1; d <= kB; ++d) {
        if (d <= nbk) bk = fmaxf(bk, kSign < 0 ? -xb[d - 1] : xb[d - 1]);
        if (d <= nfw) fk = fmaxf(fk, kSign < 0 ? -xf[d - 1] : xf[d - 1]);
    }
    const float tau = 0x1p-40f * (fabsf(xt) + mabs);
    if (bk >= kt || fk - kt > tau) return 0;
    if (fk <= kt && kt - bk > tau) return 1;
    return -1;
}

// The same test with the neighbours read from the wavefront's LDS window:
// lane column `win` (stride 64 floats per row) holds samples
// tb - kLdsD .. tb + kB + kLdsD - 1 of its channel (row kLdsD + i is sample
// tb + i, zero outside the recording, as load_x), so D <= kLdsD neighbours of
// batch index i need no load through the texture path.
constexpr int kLdsD = 10;
constexpr int kWinRows = kLdsD + kB + kLdsD;
template <int kSign>
__device__ __forceinline__ int stencil_win(const float* win, int i, int64_t t, float xt, float mabs,
                                           int D, int64_t wa, int64_t wb) {
    const int nbk = static_cast<int>(t - wa < D ? t - wa : D);
    const int nfw = static_cast<int>(wb - 1 - t < D ? wb - 1 - t : D);
    const float kt = kSign < 0 ? -xt : xt;
    float bk = -INFINITY, fk = -INFINITY;
#pragma unroll
    for (int d = 1; d <= kLdsD; ++d) {
        if (d <= nbk) {
            const float x = win[(kLdsD + i - d) * 64];
            bk = fmaxf(bk, kSign < 0 ? -x : x);
        }
        if (d <= nfw) {
            const float x = win[(kLdsD + i + d) * 64];
            fk = fmaxf(fk, kSign < 0 ? -x : x);
        }
    }
    const float tau = 0x1p-40f * (fabsf(xt) + mabs);
    if (bk >= kt || fk - kt > tau) return 0;
    if (fk <= kt && kt - bk > tau) return 1;
    return -1;
}

// kGated: the learned detector's gated scan (channels 0..M-1): a crossing is
// kept only when the network proposed its 32-sample window (class 1 or 2 in
// cls[w * M + channel]); the stencil itself still reads every neighbour, so
// every kept sample is an event of plain detection and the slot capacity
// holds.  Window classes are read only for batches with a crossing.
template <int kSign, bool kGated = false>
__global__ void __launch_bounds__(256) detect_scan_kernel(
    const float* __restrict__ X, DetectPlan p, const int32_t* __restrict__ channels,
    const double* __restrict__ mean_in, const double* __restrict__ thr_in,
    int32_t* __restrict__ counts, uint16_t* __restrict__ slots,
    const uint8_t* __restrict__ cls = nullptr) {
    // per wavefront: the window of each lane's channel around the batch (a
    // column of stride 64 floats), written only by lanes with a crossing
    __shared__ float win_all[4][kWinRows * 64];
    float* const win = win_all[threadIdx.x >> 6] + (threadIdx.x & 63);
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= p.nseg * p.Mc) return;
    const int j = static_cast<int>(g % p.Mc);
    const int64_t seg = g / p.Mc;
    const int64_t ck = seg / p.spc;
    const int64_t s = seg % p.spc;
    int64_t size = p.n_range - ck * p.cs;
    if (size > p.cs) size = p.cs;
    const int64_t r0 = s * kSeg;
    int64_t r1 = r0 + kSeg;
    if (r1 > size) r1 = size;
    int32_t cnt = 0;
    if (r0 < r1) {
        const double mean = mean_in[ck * p.Mc + j];
        const double thr = thr_in[ck * p.Mc + j];
        const int64_t core0 = p.t_start + ck * p.cs;
        const int64_t wa = core0 - p.pad;  // padded window [wa, wb)
        const int64_t wb = core0 + size + p.pad;
        const float* xp = X + channels[j];
        uint16_t* my_slots = slots + (seg * p.Mc + j) * p.cap;
        const int64_t T0 = core0 + r0;
        const int64_t T1 = core0 + r1;
        // one-sided sign and a window within kB: the float test first
        const bool fast = kSign != 0 && p.D <= kB;
        // D <= kLdsD: a crossing's neighbours come from the wave's LDS window
        const bool winpath = fast && p.D <= kLdsD;
        // |mean| rounded up to float
        const float mabs = static_cast<float>(fabs(mean)) * (1.0f + 0x1p-20f);
        float cur[kB], nxt[kB];
#pragma unroll
        for (int i = 0; i < kB; ++i) cur[i] = load_x(xp, T0 + i, p.M, p.N_file);
        float prv[kLdsD];  // the kLdsD samples before the batch
#pragma unroll
        for (int k = 0; k < kLdsD; ++k) prv[k] = load_x(xp, T0 - kLdsD + k, p.M, p.N_file);
        float prev = prv[kLdsD - 1];  // the sample before the batch
        for (int64_t tb = T0; tb < T1; tb += kB) {
#pragma unroll
            for (int i = 0; i < kB; ++i) nxt[i] = load_x(xp, tb + kB + i, p.M, p.N_file);
            uint32_t over = 0;  // threshold crossings of this batch
#pragma unroll
            for (int i = 0; i < kB; ++i)
                if (tb + i < T1 && detect_value_s<kSign>(cur[i], mean) > thr) over |= 1u << i;
            // A crossing that its immediate neighbour already beats (by the
            // float rules of stencil_float) is rejected from the registers:
            // only the local maxima of a run above threshold reach the
            // neighbour loads.  lo / hi: the first / last batch index whose
            // left / right neighbour lies inside the padded window.
            if (fast && p.D >= 1 && over) {
                const int64_t lo64 = wa - tb + 1, hi64 = wb - tb - 2;
                const int lo = static_cast<int>(lo64 < 0 ? 0 : (lo64 > kB ? kB : lo64));
                const int hi = static_cast<int>(hi64 < -1 ? -1 : (hi64 > kB ? kB : hi64));
                uint32_t rej = 0;
#pragma unroll
                for (int i = 0; i < kB; ++i) {
                    const float xm = i > 0 ? cur[i - 1] : prev;
                    const float xq = i < kB - 1 ? cur[i + 1] : nxt[0];
                    const float kc = kSign < 0 ? -cur[i] : cur[i];
                    const float km = kSign < 0 ? -xm : xm;
                    const float kq = kSign < 0 ? -xq : xq;
                    const float tau = 0x1p-40f * (fabsf(cur[i]) + mabs);
                    if ((i >= lo && km >= kc) || (i <= hi && kq - kc > tau)) rej |= 1u << i;
                }
                over &= ~rej;
            }
            if (kGated && over) {
                constexpr int64_t W = tcn::kWin;
                const int64_t nW = (p.N_file + W - 1) / W;
                const int64_t wa_ = tb / W;
                int64_t wb_ = (tb + kB - 1) / W;
                wb_ = wb_ < nW ? wb_ : nW - 1;
                const bool pa = cls[wa_ * p.M + j] != 0;
                const bool pb = cls[wb_ * p.M + j] != 0;
                const uint32_t all = (1u << kB) - 1;
                uint32_t gate;
                if (wb_ == wa_) {
                    gate = pa ? all : 0u;
                } else {
                    const uint32_t lowm = (1u << static_cast<int>(wb_ * W - tb)) - 1;  // batch indices in wa_
                    gate = (pa ? lowm : 0u) | (pb ? (all & ~lowm) : 0u);
                }
                over &= gate;
            }
            if (winpath && over) {
                // this lane's column of the window: rows tb - kLdsD .. tb + kB + kLdsD - 1
#pragma unroll
                for (int k = 0; k < kLdsD; ++k) win[k * 64] = prv[k];
#pragma unroll
                for (int k = 0; k < kB; ++k) win[(kLdsD + k) * 64] = cur[k];
#pragma unroll
                for (int k = 0; k < kLdsD; ++k) win[(kLdsD + kB + k) * 64] = nxt[k];
            }
            // in time order; kept events are > D apart, so cnt < cap
            while (over) {
                const int i = __builtin_ctz(over);
                over &= over - 1;
                const int64_t t = tb + i;
                const float xt = winpath ? win[(kLdsD + i) * 64] : load_x(xp, t, p.M, p.N_file);
                int keep = winpath ? stencil_win<kSign>(win, i, t, xt, mabs, p.D, wa, wb)
                           : fast  ? stencil_float<kSign>(xp, t, xt, mabs, p.D, wa, wb, p.M, p.N_file)
                                   : -1;
                if (keep < 0)
                    keep = stencil_ok(xp, t, detect_value(xt, mean, kSign), mean, kSign, p.D, wa, wb,
                                      p.M, p.N_file);
                if (keep && cnt < p.cap) my_slots[cnt++] = static_cast<uint16_t>(t - T0);
            }
            prev = cur[kB - 1];
#pragma unroll
            for (int k = 0; k < kLdsD; ++k) prv[k] = cur[kB - kLdsD + k];
#pragma unroll
            for (int i = 0; i < kB; ++i) cur[i] = nxt[i];
        }
    }
    counts[static_cast<int64_t>(j) * p.nseg + seg] = cnt;
}

// Class histogram over the windows that hold samples of [t_start, t_end].
__global__ void __launch_bounds__(256) detect_class_count_kernel(
    const uint8_t* __restrict__ cls, int64_t w_begin, int64_t w_end, int64_t M,
    unsigned long long* __restrict__ out) {
    __shared__ unsigned int h[3];
    if (threadIdx.x < 3) h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t n = (w_end - w_begin) * M;
    unsigned int loc[3] = {0, 0, 0};
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < n;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int c = cls[w_begin * M + q];
        loc[c < 3 ? c : 2]++;
    }
    for (int k = 0; k < 3; ++k) atomicAdd(&h[k], loc[k]);
    __syncthreads();
    if (threadIdx.x < 3) atomicAdd(&out[threadIdx.x], static_cast<unsigned long long>(h[threadIdx.x]));
}

__global__ void iota_kernel(int32_t* __restrict__ a, int32_t n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = i;
}

// One block per channel: exclusive scan of the per-segment counts.
__global__ void __launch_bounds__(1024) detect_seg_scan_kernel(
    const int32_t* __restrict__ counts, int64_t nseg, int64_t* __restrict__ seg_off,
    int64_t* __restrict__ chan_total) {
    __shared__ int64_t warp_sums[16];
    __shared__ int64_t carry_s;
    const int j = blockIdx.x;
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    const int32_t* cj = counts + static_cast<int64_t>(j) * nseg;
    int64_t* oj = seg_off + static_cast<int64_t>(j) * nseg;
    if (tid == 0) carry_s = 0;
    __syncthreads();
    for (int64_t base = 0; base < nseg; base += blockDim.x) {
        const int64_t i = base + tid;
        const int64_t v = (i < nseg) ? cj[i] : 0;
        int64_t x = v;
        for (int o = 1; o < 64; o <<= 1) {
            const int64_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) warp_sums[wid] = x;
        __syncthreads();
        if (tid == 0) {
            int64_t acc = 0;
            for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
                const int64_t t = warp_sums[w];
                warp_sums[w] = acc;
                acc += t;
            }
        }
        __syncthreads();
        const int64_t carry = carry_s;
        if (i < nseg) oj[i] = carry + warp_sums[wid] + x - v;
        __syncthreads();
        if (tid == static_cast<int>(blockDim.x) - 1) carry_s = carry + warp_sums[wid] + x;
        __syncthreads();
    }
    if (tid == 0) chan_total[j] = carry_s;
}

// Single block: exclusive scan over channels -> chan_off[0..Mc].
__global__ void detect_chan_scan_kernel(int64_t* __restrict__ chan_off, int32_t Mc) {
    if (threadIdx.x != 0) return;
    int64_t acc = 0;
    for (int32_t j = 0; j < Mc; ++j) {
        const int64_t t = chan_off[j];
        chan_off[j] = acc;
        acc += t;
    }
    chan_off[Mc] = acc;
}

__global__ void __launch_bounds__(256) detect_compact_kernel(
    DetectPlan p, const int32_t* __restrict__ counts, const int64_t* __restrict__ seg_off,
    const int64_t* __restrict__ chan_off, const uint16_t* __restrict__ slots,
    int64_t* __restrict__ times) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= p.nseg * p.Mc) return;
    const int j = static_cast<int>(g % p.Mc);
    const int64_t seg = g / p.Mc;
    const int64_t idx = static_cast<int64_t>(j) * p.nseg + seg;
    const int32_t n = counts[idx];
    if (n == 0) return;
    const int64_t c = seg / p.spc;
    const int64_t s = seg % p.spc;
    const int64_t t0 = p.t_start + c * p.cs + s * kSeg;
    const int64_t o = chan_off[j] + seg_off[idx];
    const uint16_t* my_slots = slots + (seg * p.Mc + j) * p.cap;
    for (int32_t k = 0; k < n; ++k) times[o + k] = t0 + my_slots[k];
}

__global__ void detect_offsets_kernel(const int64_t* __restrict__ chan_off, int32_t Mc,
                                      int64_t* __restrict__ offsets) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= Mc) offsets[i] = chan_off[i];
}

}  // namespace
}  // namespace msamd

using namespace msamd;

extern "C" int64_t msamd_detect_chunk_size(int64_t M, int64_t N, int64_t num_threads) {
    // p_mountainsort3.cpp:170-172 (simultaneous neighborhoods, batch count)
    // and get_time_chunk_infos :507-533 with min_num_chunks = 1.
    if (M < 1) M = 1;
    int64_t nsim = num_threads < M ? num_threads : M;
    if (nsim < 1) nsim = 1;
    const int64_t num_batches = ceil_div(M, nsim);
    const int64_t ram = 1000000000LL;
    int64_t chunk = ram / (M * num_batches * 4);
    if (chunk > N) chunk = N;
    if (chunk < 20000) chunk = 20000;
    return chunk;
}

extern "C" int64_t msamd_detect_workspace_size(int64_t n_range, int64_t chunk_size,
                                               int32_t num_channels, double detect_interval) {
    if (n_range < 0 || chunk_size < 1 || num_channels < 0) return -1;
    const DetectPlan p = make_plan(1, n_range, 0, n_range, chunk_size, 0, num_channels,
                                   detect_interval);
    int64_t bytes = 0;
    carve(nullptr, p, &bytes);
    return bytes;
}

extern "C" int msamd_detect_events(const float* X, int64_t M, int64_t N_file, int64_t t1,
                                   int64_t t2, int64_t chunk_size, int64_t padding,
                                   const int32_t* channels, int32_t num_channels,
                                   double detect_threshold, double detect_interval,
                                   int detect_sign, void* workspace, int64_t workspace_bytes,
                                   int64_t* num_events_host, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N_file >= 0, "detect: bad shape M=%lld N=%lld", (long long)M,
                  (long long)N_file);
    MSAMD_REQUIRE(chunk_size >= 1 && padding >= 0, "detect: bad chunk plan");
    MSAMD_REQUIRE(num_channels >= 0 && num_events_host, "detect: bad channel list");
    const int64_t t_start = t1 < 0 ? 0 : t1;
    const int64_t t_end = t2 < 0 ? N_file - 1 : t2;
    const int64_t n_range = t_end - t_start + 1;
    MSAMD_REQUIRE(n_range >= 0, "detect: empty or inverted time range");
    const DetectPlan p = make_plan(M, N_file, t_start, n_range, chunk_size, padding,
                                   num_channels, detect_interval);
    int64_t need = 0;
    DetectBuffers b = carve(workspace, p, &need);
    if (workspace_bytes < need || !workspace) {
        set_error("detect: workspace too small (%lld < %lld bytes)", (long long)workspace_bytes,
                  (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    MSAMD_HIP(hipMemcpyAsync(b.header, &p, sizeof(DetectPlan), hipMemcpyHostToDevice, st));
    if (num_channels == 0 || n_range == 0) {
        MSAMD_HIP(hipMemsetAsync(b.chan_off, 0, sizeof(int64_t) * (num_channels + 1), st));
        MSAMD_HIP(hipStreamSynchronize(st));
        *num_events_host = 0;
        return MSAMD_OK;
    }
    {
        const int64_t n = p.nchunks * p.Mc;
        detect_stats_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, 256))), 256, 0, st>>>(
            X, p, channels, detect_threshold, b.mean, b.thr);
        MSAMD_LAUNCH_CHECK();
    }
    {
        const int64_t n = p.nseg * p.Mc;
        const dim3 grid(static_cast<unsigned>(ceil_div(n, 256)));
        if (detect_sign < 0)
            detect_scan_kernel<-1><<<grid, 256, 0, st>>>(X, p, channels, b.mean, b.thr, b.counts, b.slots);
        else if (detect_sign == 0)
            detect_scan_kernel<0><<<grid, 256, 0, st>>>(X, p, channels, b.mean, b.thr, b.counts, b.slots);
        else
            detect_scan_kernel<1><<<grid, 256, 0, st>>>(X, p, channels, b.mean, b.thr, b.counts, b.slots);
        MSAMD_LAUNCH_CHECK();
    }
    detect_seg_scan_kernel<<<dim3(p.Mc), 1024, 0, st>>>(b.counts, p.nseg, b.seg_off,
                                                         b.chan_off);
    MSAMD_LAUNCH_CHECK();
    detect_chan_scan_kernel<<<1, 64, 0, st>>>(b.chan_off, p.Mc);
    MSAMD_LAUNCH_CHECK();
    int64_t total = 0;
    MSAMD_HIP(hipMemcpyAsync(&total, b.chan_off + p.Mc, sizeof(int64_t), hipMemcpyDeviceToHost,
                             st));
    MSAMD_HIP(hipStreamSynchronize(st));
    *num_events_host = total;
    return MSAMD_OK;
}

extern "C" int msamd_detect_fetch(const void* workspace, int32_t num_channels, int64_t* times,
                                  int64_t* offsets, void* stream) {
    clear_error();
    MSAMD_REQUIRE(workspace != nullptr, "detect_fetch: null workspace");
    hipStream_t st = as_stream(stream);
    // The plan was stored at the head of the workspace by msamd_detect_events.
    DetectPlan p{};
    MSAMD_HIP(hipMemcpyAsync(&p, workspace, sizeof(DetectPlan), hipMemcpyDeviceToHost, st));
    MSAMD_HIP(hipStreamSynchronize(st));
    MSAMD_REQUIRE(p.Mc == num_channels, "detect_fetch: channel count %d does not match %d",
                  num_channels, p.Mc);
    DetectBuffers b = carve(const_cast<void*>(workspace), p, nullptr);
    if (offsets) {
        detect_offsets_kernel<<<dim3(static_cast<unsigned>(ceil_div(p.Mc + 1, 256))), 256, 0,
                                st>>>(b.chan_off, p.Mc, offsets);
        MSAMD_LAUNCH_CHECK();
    }
    const int64_t n = p.nseg * p.Mc;
    if (times && n > 0) {
        detect_compact_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, 256))), 256, 0, st>>>(
            p, b.counts, b.seg_off, b.chan_off, b.slots, times);
        MSAMD_LAUNCH_CHECK();
    }
    return MSAMD_OK;
}

// ------------------------------------------------------- learned detector
// Workspace: the plain detection buffers (same offsets, so msamd_detect_fetch
// reads the result), then the channel list, the window classes [nW][M], the
// network fragments and the class histogram.
namespace {
struct LearnedBuffers {
    int32_t* channels;
    uint8_t* classes;
    void* frag;
    unsigned long long* hist;
};

inline LearnedBuffers carve_learned(void* ws, const DetectPlan& p, int64_t* bytes) {
    int64_t prefix = 0;
    carve(nullptr, p, &prefix);
    Carver c(ws ? static_cast<char*>(ws) + prefix : nullptr);
    LearnedBuffers b{};
    b.channels = c.take<int32_t>(p.M);
    b.classes = c.take<uint8_t>(tcn::num_windows(p.N_file) * p.M);
    b.frag = c.take<char>(tcn::frag_bytes());
    b.hist = c.take<unsigned long long>(3);
    if (bytes) *bytes = prefix + c.bytes();
    return b;
}
}  // namespace

extern "C" int64_t msamd_tcn_detect_workspace_size(int64_t N_file, int64_t n_range,
                                                   int64_t chunk_size, int32_t M,
                                                   double detect_interval) {
    if (N_file < 0 || n_range < 0 || chunk_size < 1 || M < 1) return -1;
    const DetectPlan p = make_plan(M, N_file, 0, n_range, chunk_size, 0, M, detect_interval);
    int64_t bytes = 0;
    carve_learned(nullptr, p, &bytes);
    return bytes;
}

extern "C" int msamd_tcn_detect(const float* X, int64_t M, int64_t N_file, int64_t t1,
                                int64_t t2, int64_t chunk_size, int64_t padding,
                                const float* params, double detect_threshold,
                                double detect_interval, int detect_sign, double score_lo,
                                double score_hi, void* workspace, int64_t workspace_bytes,
                                int64_t* num_events_host, int64_t* class_counts_host,
                                void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && M <= INT32_MAX && N_file >= 0, "tcn_detect: bad shape");
    MSAMD_REQUIRE(chunk_size >= 1 && padding >= 0, "tcn_detect: bad chunk plan");
    MSAMD_REQUIRE(params && num_events_host, "tcn_detect: null argument");
    MSAMD_REQUIRE(score_lo > 0.0 && score_hi < 1.0 && score_lo <= score_hi,
                  "tcn_detect: score band must satisfy 0 < lo <= hi < 1");
    const int64_t t_start = t1 < 0 ? 0 : t1;
    const int64_t t_end = t2 < 0 ? N_file - 1 : t2;
    const int64_t n_range = t_end - t_start + 1;
    MSAMD_REQUIRE(n_range >= 0 && t_end < N_file + (n_range == 0 ? 1 : 0),
                  "tcn_detect: time range outside the recording");
    const int32_t Mc = static_cast<int32_t>(M);
    const DetectPlan p = make_plan(M, N_file, t_start, n_range, chunk_size, padding, Mc,
                                   detect_interval);
    int64_t need = 0;
    DetectBuffers b = carve(workspace, p, nullptr);
    LearnedBuffers lb = carve_learned(workspace, p, &need);
    if (!workspace || workspace_bytes < need) {
        set_error("tcn_detect: workspace too small (%lld < %lld bytes)",
                  (long long)workspace_bytes, (long long)need);
        return MSAMD_ERR_WORKSPACE;
    }
    hipStream_t st = as_stream(stream);
    MSAMD_HIP(hipMemcpyAsync(b.header, &p, sizeof(DetectPlan), hipMemcpyHostToDevice, st));
    if (n_range == 0) {
        MSAMD_HIP(hipMemsetAsync(b.chan_off, 0, sizeof(int64_t) * (Mc + 1), st));
        MSAMD_HIP(hipStreamSynchronize(st));
        *num_events_host = 0;
        if (class_counts_host) class_counts_host[0] = class_counts_host[1] = class_counts_host[2] = 0;
        return MSAMD_OK;
    }
    iota_kernel<<<static_cast<unsigned>(ceil_div(M, 256)), 256, 0, st>>>(lb.channels, Mc);
    MSAMD_LAUNCH_CHECK();
    {
        const int64_t n = p.nchunks * p.Mc;
        detect_stats_kernel<<<dim3(static_cast<unsigned>(ceil_div(n, 256))), 256, 0, st>>>(
            X, p, lb.channels, detect_threshold, b.mean, b.thr);
        MSAMD_LAUNCH_CHECK();
    }
    const float lo = static_cast<float>(log(score_lo / (1.0 - score_lo)));
    const float hi = static_cast<float>(log(score_hi / (1.0 - score_hi)));
    int rc = tcn::prepare(params, lb.frag, st);
    if (rc != MSAMD_OK) return rc;
    rc = tcn::classify(X, M, N_file, lb.frag, lo, hi, nullptr, lb.classes, st);
    if (rc != MSAMD_OK) return rc;
    {
        const dim3 grid(static_cast<unsigned>(ceil_div(p.nseg * p.Mc, 256)));
        if (detect_sign < 0)
            detect_scan_kernel<-1, true><<<grid, 256, 0, st>>>(X, p, lb.channels, b.mean, b.thr,
                                                                b.counts, b.slots, lb.classes);
        else if (detect_sign == 0)
            detect_scan_kernel<0, true><<<grid, 256, 0, st>>>(X, p, lb.channels, b.mean, b.thr,
                                                               b.counts, b.slots, lb.classes);
        else
            detect_scan_kernel<1, true><<<grid, 256, 0, st>>>(X, p, lb.channels, b.mean, b.thr,
                                                               b.counts, b.slots, lb.classes);
        MSAMD_LAUNCH_CHECK();
    }
    detect_seg_scan_kernel<<<dim3(p.Mc), 1024, 0, st>>>(b.counts, p.nseg, b.seg_off,
                                                         b.chan_off);
    MSAMD_LAUNCH_CHECK();
    detect_chan_scan_kernel<<<1, 64, 0, st>>>(b.chan_off, p.Mc);
    MSAMD_LAUNCH_CHECK();
    unsigned long long hist[3] = {0, 0, 0};
    if (class_counts_host) {
        MSAMD_HIP(hipMemsetAsync(lb.hist, 0, sizeof(unsigned long long) * 3, st));
        const int64_t wb0 = t_start / tcn::kWin, wb1 = t_end / tcn::kWin + 1;
        detect_class_count_kernel<<<1024, 256, 0, st>>>(lb.classes, wb0, wb1, M, lb.hist);
        MSAMD_LAUNCH_CHECK();
        MSAMD_HIP(hipMemcpyAsync(hist, lb.hist, sizeof(hist), hipMemcpyDeviceToHost, st));
    }
    int64_t total = 0;
    MSAMD_HIP(hipMemcpyAsync(&total, b.chan_off + p.Mc, sizeof(int64_t), hipMemcpyDeviceToHost,
                             st));
    MSAMD_HIP(hipStreamSynchronize(st));
    *num_events_host = total;
    if (class_counts_host)
        for (int k = 0; k < 3; ++k) class_counts_host[k] = static_cast<int64_t>(hist[k]);
    return MSAMD_OK;
}
