// Templates and the fit stage (template-subtraction event pruning).
//
// Reference: p_compute_templates (packages/ms3/p_compute_templates.cpp:24-118),
// GlobalTemplateComputer::processTimeChunk / templates
// (packages/mountainsortalg/globaltemplatecomputer.cpp:63-135),
// FitStageComputer::processTimeChunk (fitstagecomputer.cpp:55-104) and
// fit_stage (fit_stage.cpp:27-275), driven by p_mountainsort3.cpp:270-356.
//
// Templates: template k = the mean of the M x T clips of the events labeled
// k, summed in FP64 in event order (the order of the reference's loop), so
// the float templates equal the reference's.  One thread per (k, m, t)
// entry walks the label's events.
//
// Fit stage: each time chunk (padded by `padding` zero-filled samples) is
// fitted on its own copy of the data by one persistent workgroup that runs
// the reference's passes until nothing changes:
//   score    every considered event whose masked samples changed in the
//            previous pass (all of them on the first pass) gets
//            sum(x^2) - sum((x - template)^2) over its template's mask
//            |template| > thresh (FP64; a wave per event, a fixed lane
//            order, so equal inputs give equal scores and ties break as in
//            the reference); the others keep their score
//   select   candidates (score > 0, in time order) survive unless an
//            earlier one within T samples has a score >= theirs or a later
//            one within T a score > theirs
//   subtract each survivor's template is subtracted over its mask
//            ((float)((double)x - (double)template), the reference's
//            rounding; survivors are > T apart, so no two overlap) and the
//            samples are stamped with the pass number
//   the candidates that did not survive are considered again.
// "Changed in the previous pass" is a 16-bit stamp per sample equal to the
// previous pass number, so nothing is cleared between passes.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

#pragma clang fp contract(off)

namespace msamd {
namespace {

constexpr int kFitThreads = 1024;

struct FitTables {
    float* xc;                 // chunk copies, chunk c at xoff[c] (M x N2_c)
    uint16_t* stamp;           // one per element of xc
    const int64_t* xoff;       // [nch]
    const int32_t* n2;         // [nch] padded chunk lengths
    const int64_t* eoff;       // [nch + 1] chunk-event offsets
    const int32_t* ltime;      // [ne] chunk-local times
    const int32_t* lab;        // [ne] labels
    const float* tmpl;         // M x T x K
    const int32_t* mask;       // [K][M*T] mask indices m + M*t, ascending
    const int32_t* mlen;       // [K]
    double* score;             // [ne]
    int32_t* cons;             // [ne] considered local ids
    double* cand_s;            // [ne] score of consider[j] this pass
    int32_t* tri;              // [ne] candidates: local id
    double* trs;               // [ne] candidates: score
    int32_t* trt;              // [ne] candidates: time
    uint8_t* use;              // [ne]
    uint8_t* keep;             // [ne]
    int32_t* passes;           // [nch]
    int32_t* status;           // 1 when a chunk ran out of pass stamps
    int32_t M, T, K;
};

// Block-wide ordered compaction helper: exclusive prefix of `flag` over the
// workgroup (sh: kFitThreads / 64 + 1 ints); returns this thread's position
// and writes the total to *total (all threads).
__device__ int block_excl_scan(int flag, int* sh, int* total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(flag);
    const int before = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) sh[wid] = __popcll(bal);
    __syncthreads();
    if (threadIdx.x == 0) {
        int run = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            const int c = sh[w];
            sh[w] = run;
            run += c;
        }
        sh[kFitThreads / 64] = run;
    }
    __syncthreads();
    const int pos = sh[wid] + before;
    *total = sh[kFitThreads / 64];
    __syncthreads();
    return pos;
}

// ------------------------------------------------------------- templates
__global__ void __launch_bounds__(256) templates_kernel(const float* __restrict__ X, int64_t M,
                                                        int64_t N, const int64_t* __restrict__ gt,
                                                        const int64_t* __restrict__ goff, int32_t T,
                                                        float* __restrict__ out) {
    const int k = blockIdx.y;
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= M * T) return;
    const int64_t m = e % M;
    const int64_t t = e / M;
    const int64_t tmid = (T + 1) / 2 - 1;
    const int64_t a = goff[k], b = goff[k + 1];
    double s = 0.0;
    int64_t i = a;
    // four loads in flight, added in event order
    for (; i + 4 <= b; i += 4) {
        float v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t tt = gt[i + u] - tmid + t;
            v[u] = (tt >= 0 && tt < N) ? X[tt * M + m] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) s = s + static_cast<double>(v[u]);
    }
    for (; i < b; ++i) {
        const int64_t tt = gt[i] - tmid + t;
        s = s + static_cast<double>((tt >= 0 && tt < N) ? X[tt * M + m] : 0.0f);
    }
    const int64_t n = b - a;
    out[static_cast<int64_t>(k) * M * T + e] = static_cast<float>(n ? s / static_cast<double>(n) : s);
}

// ------------------------------------------------------------- fit stage
// Chunk copies: sample (m, j) of chunk c is X[m, t1_c - pad + j], zero
// outside [0, N).
__global__ void __launch_bounds__(256) fit_copy_kernel(const float* __restrict__ X, int64_t M,
                                                       int64_t N, const int64_t* __restrict__ xoff,
                                                       const int64_t* __restrict__ start,
                                                       const int32_t* __restrict__ n2,
                                                       float* __restrict__ xc) {
    const int c = blockIdx.y;
    const int64_t total = M * n2[c];
    for (int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; q < total;
         q += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t j = q / M;
        const int64_t tt = start[c] + j;
        xc[xoff[c] + q] = (tt >= 0 && tt < N) ? X[tt * M + (q - j * M)] : 0.0f;
    }
}

// Time-channel masks (fit_stage.cpp:176-192): one workgroup per template,
// ordered compaction of the entries with |template| > thresh.
__global__ void __launch_bounds__(kFitThreads) fit_mask_kernel(const float* __restrict__ tmpl,
                                                               int32_t MT, double thresh,
                                                               int32_t* __restrict__ mask,
                                                               int32_t* __restrict__ mlen) {
    __shared__ int sh[kFitThreads / 64 + 1];
    const int k = blockIdx.x;
    const float* tk = tmpl + static_cast<int64_t>(k) * MT;
    int32_t* mk = mask + static_cast<int64_t>(k) * MT;
    int base = 0;
    for (int e0 = 0; e0 < MT; e0 += blockDim.x) {
        const int e = e0 + threadIdx.x;
        const int f = e < MT && fabs(static_cast<double>(tk[e])) > thresh;
        int tot;
        const int pos = block_excl_scan(f, sh, &tot);
        if (f) mk[base + pos] = e;
        base += tot;
    }
    if (threadIdx.x == 0) mlen[k] = base;
}

__global__ void __launch_bounds__(kFitThreads) fit_chunk_kernel(FitTables tb) {
    __shared__ int sh[kFitThreads / 64 + 1];
    __shared__ int n_used_sh;
    const int c = blockIdx.x;
    const int64_t e0 = tb.eoff[c];
    const int ne = static_cast<int>(tb.eoff[c + 1] - e0);
    const int N2 = tb.n2[c];
    const int M = tb.M, T = tb.T, K = tb.K;
    const int tmid = (T + 1) / 2 - 1;
    float* X = tb.xc + tb.xoff[c];
    uint16_t* st = tb.stamp + tb.xoff[c];
    const int32_t* ltime = tb.ltime + e0;
    const int32_t* lab = tb.lab + e0;
    double* score = tb.score + e0;
    int32_t* cons = tb.cons + e0;
    double* cand_s = tb.cand_s + e0;
    int32_t* tri = tb.tri + e0;
    double* trs = tb.trs + e0;
    int32_t* trt = tb.trt + e0;
    uint8_t* use = tb.use + e0;
    uint8_t* keep = tb.keep + e0;
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
    for (int j = threadIdx.x; j < ne; j += blockDim.x) {
        cons[j] = j;
        keep[j] = 0;
    }
    int n_cons = ne;
    int pass = 0;
    __syncthreads();
    for (;;) {
        ++pass;
        if (pass > 65535) {  // the 16-bit stamps would wrap
            if (threadIdx.x == 0) atomicExch(tb.status, 1);
            break;
        }
        // score (fit_stage.cpp:86-133)
        for (int j = wid; j < n_cons; j += nw) {
            const int i = cons[j];
            const int k0 = lab[i];
            double s = 0.0;
            if (k0 > 0 && k0 <= K) {
                // (bigint)(t0 - Tmid + 0.5): C truncation of the double
                const int64_t tt = static_cast<int64_t>(static_cast<double>(ltime[i]) -
                                                        static_cast<double>(tmid) + 0.5);
                if (tt >= 0 && tt + T <= N2) {
                    const int32_t* mk = tb.mask + static_cast<int64_t>(k0 - 1) * M * T;
                    const int n = tb.mlen[k0 - 1];
                    const int64_t base = tt * M;
                    bool dirty = pass == 1;
                    if (!dirty) {
                        int d = 0;
                        for (int q = lane; q < n && !d; q += 64)
                            d = st[base + mk[q]] == static_cast<uint16_t>(pass - 1);
                        dirty = __ballot(d) != 0;
                    }
                    if (!dirty) {
                        s = score[i];
                    } else {
                        const float* tp = tb.tmpl + static_cast<int64_t>(k0 - 1) * M * T;
                        double bsum = 0.0, asum = 0.0;
                        for (int q = lane; q < n; q += 64) {
                            const int idx = mk[q];
                            double v = static_cast<double>(X[base + idx]);
                            bsum = bsum + v * v;
                            v = v - static_cast<double>(tp[idx]);
                            asum = asum + v * v;
                        }
                        for (int o = 32; o > 0; o >>= 1) {
                            bsum = bsum + __shfl_xor(bsum, o, 64);
                            asum = asum + __shfl_xor(asum, o, 64);
                        }
                        // lane 0's association order for every lane
                        const double b0 = __shfl(bsum, 0, 64), a0 = __shfl(asum, 0, 64);
                        s = b0 - a0;
                        if (lane == 0) score[i] = s;
                    }
                }
            }
            if (lane == 0) cand_s[j] = s;
        }
        __syncthreads();
        // candidates: score > 0, in consider order = time order (:120-125)
        int nt = 0;
        for (int j0 = 0; j0 < n_cons; j0 += blockDim.x) {
            const int j = j0 + threadIdx.x;
            const int f = j < n_cons && cand_s[j] > 0.0;
            int tot;
            const int pos = block_excl_scan(f, sh, &tot);
            if (f) {
                const int i = cons[j];
                tri[nt + pos] = i;
                trs[nt + pos] = cand_s[j];
                trt[nt + pos] = ltime[i];
            }
            nt += tot;
        }
        __syncthreads();
        // find_events_to_use (:221-252)
        for (int a = threadIdx.x; a < nt; a += blockDim.x) {
            const double sa = trs[a];
            const int ta = trt[a];
            int u = 1;
            for (int j = a - 1; j >= 0 && trt[j] >= ta - T; --j)
                if (trs[j] >= sa) u = 0;
            for (int j = a + 1; j < nt && trt[j] <= ta + T; ++j)
                if (trs[j] > sa) u = 0;
            use[a] = static_cast<uint8_t>(u);
        }
        __syncthreads();
        // subtract the survivors' templates (:148-167), stamp with the pass
        for (int a = wid; a < nt; a += nw) {
            if (!use[a]) continue;
            const int i = tri[a];
            const int k0 = lab[i];
            const int32_t* mk = tb.mask + static_cast<int64_t>(k0 - 1) * M * T;
            const int n = tb.mlen[k0 - 1];
            const float* tp = tb.tmpl + static_cast<int64_t>(k0 - 1) * M * T;
            const int64_t base =
                static_cast<int64_t>(static_cast<double>(trt[a]) - static_cast<double>(tmid) + 0.5) * M;
            for (int q = lane; q < n; q += 64) {
                const int idx = mk[q];
                X[base + idx] = static_cast<float>(static_cast<double>(X[base + idx]) -
                                                   static_cast<double>(tp[idx]));
                st[base + idx] = static_cast<uint16_t>(pass);
            }
            if (lane == 0) keep[i] = 1;
        }
        // the others are considered again, in order
        int nn = 0, nu = 0;
        for (int a0 = 0; a0 < nt; a0 += blockDim.x) {
            const int a = a0 + threadIdx.x;
            const int f = a < nt && !use[a];
            int tot;
            const int pos = block_excl_scan(f, sh, &tot);
            if (f) cons[nn + pos] = tri[a];
            nn += tot;
            nu += (a0 + static_cast<int>(blockDim.x) < nt ? static_cast<int>(blockDim.x) : nt - a0) - tot;
        }
        if (threadIdx.x == 0) n_used_sh = nu;
        __syncthreads();
        n_cons = nn;
        const int changed = n_used_sh;
        __syncthreads();
        if (changed == 0) break;
    }
    if (threadIdx.x == 0) tb.passes[c] = pass;
}

}  // namespace
}  // namespace msamd

using namespace msamd;

// ------------------------------------------------------------ templates ABI
extern "C" int64_t msamd_templates_workspace_size(int64_t L, int32_t K) {
    if (L < 0 || K < 0) return -1;
    Carver c(nullptr);
    c.take<int64_t>(L > 0 ? L : 1);
    c.take<int64_t>(static_cast<int64_t>(K) + 1);
    return c.bytes();
}

extern "C" int msamd_compute_templates(const float* X, int64_t M, int64_t N,
                                       const int64_t* times_host, const int32_t* labels_host,
                                       int64_t L, int32_t K, int32_t T, float* templates,
                                       void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 0 && L >= 0 && K >= 0 && T >= 1, "compute_templates: bad shape");
    MSAMD_REQUIRE(L == 0 || (times_host && labels_host), "compute_templates: null events");
    if (K == 0) return MSAMD_OK;
    Carver c(workspace);
    int64_t* gt = c.take<int64_t>(L > 0 ? L : 1);
    int64_t* goff = c.take<int64_t>(static_cast<int64_t>(K) + 1);
    if (!workspace || workspace_bytes < c.bytes()) {
        set_error("compute_templates: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)c.bytes());
        return MSAMD_ERR_WORKSPACE;
    }
    // events grouped by label, event order kept within a label
    std::vector<int64_t> off(static_cast<size_t>(K) + 1, 0);
    for (int64_t i = 0; i < L; ++i)
        if (labels_host[i] >= 1 && labels_host[i] <= K) off[labels_host[i]]++;
    for (int32_t k = 0; k < K; ++k) off[k + 1] += off[k];
    std::vector<int64_t> pos(off.begin(), off.end() - 1), grouped(off[K] > 0 ? off[K] : 1);
    for (int64_t i = 0; i < L; ++i) {
        const int32_t k = labels_host[i];
        if (k >= 1 && k <= K) grouped[pos[k - 1]++] = times_host[i];
    }
    hipStream_t st = as_stream(stream);
    MSAMD_HIP(hipMemcpyAsync(goff, off.data(), off.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    if (off[K] > 0)
        MSAMD_HIP(hipMemcpyAsync(gt, grouped.data(), off[K] * sizeof(int64_t), hipMemcpyHostToDevice, st));
    MSAMD_HIP(hipStreamSynchronize(st));  // pageable sources
    const dim3 grid(static_cast<unsigned>(ceil_div(M * T, 256)), static_cast<unsigned>(K));
    templates_kernel<<<grid, 256, 0, st>>>(X, M, N, gt, goff, T, templates);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

// ------------------------------------------------------------ fit stage ABI
namespace {

struct FitPlan {
    std::vector<int64_t> start, xoff, eoff;
    std::vector<int32_t> n2, ltime, lab;
    std::vector<int64_t> gidx;
    int64_t xtotal = 0;
};

int fit_plan(int64_t N, const int64_t* times, const int32_t* labels, int64_t L, int64_t chunk,
             int64_t pad, int64_t M, FitPlan& P) {
    MSAMD_REQUIRE(chunk >= 1 && pad >= 0, "fit_stage: bad chunk plan");
    for (int64_t i = 1; i < L; ++i)
        MSAMD_REQUIRE(times[i] >= times[i - 1], "fit_stage: event times must be ascending");
    int64_t lo = 0;
    for (int64_t t1 = 0; t1 < N; t1 += chunk) {
        const int64_t size = t1 + chunk > N ? N - t1 : chunk;
        const int64_t n2 = size + 2 * pad;
        MSAMD_REQUIRE(n2 < (1LL << 31) && M * n2 < (1LL << 31),
                      "fit_stage: chunk of %lld samples too long", (long long)n2);
        P.start.push_back(t1 - pad);
        P.n2.push_back(static_cast<int32_t>(n2));
        P.xoff.push_back(P.xtotal);
        P.xtotal += M * n2;
        P.eoff.push_back(static_cast<int64_t>(P.ltime.size()));
        while (lo < L && times[lo] < t1 - pad) ++lo;
        for (int64_t i = lo; i < L && times[i] < t1 + size + pad; ++i) {
            P.ltime.push_back(static_cast<int32_t>(times[i] - t1 + pad));
            P.lab.push_back(labels[i]);
            P.gidx.push_back(i);
        }
    }
    P.eoff.push_back(static_cast<int64_t>(P.ltime.size()));
    return MSAMD_OK;
}

struct FitLayout {
    FitTables tb{};
    int64_t* d_start = nullptr;
    int64_t bytes = 0;
};

FitLayout fit_layout(void* ws, const FitPlan& P, int64_t M, int32_t K, int32_t T) {
    FitLayout l;
    Carver c(ws);
    const int64_t nch = static_cast<int64_t>(P.n2.size());
    const int64_t ne = std::max<int64_t>(static_cast<int64_t>(P.ltime.size()), 1);
    l.tb.xc = c.take<float>(std::max<int64_t>(P.xtotal, 1));
    l.tb.stamp = c.take<uint16_t>(std::max<int64_t>(P.xtotal, 1));
    l.tb.xoff = c.take<int64_t>(std::max<int64_t>(nch, 1));
    l.d_start = c.take<int64_t>(std::max<int64_t>(nch, 1));
    l.tb.n2 = c.take<int32_t>(std::max<int64_t>(nch, 1));
    l.tb.eoff = c.take<int64_t>(nch + 1);
    l.tb.ltime = c.take<int32_t>(ne);
    l.tb.lab = c.take<int32_t>(ne);
    l.tb.mask = c.take<int32_t>(std::max<int64_t>(M * T * K, 1));
    l.tb.mlen = c.take<int32_t>(std::max<int32_t>(K, 1));
    l.tb.score = c.take<double>(ne);
    l.tb.cons = c.take<int32_t>(ne);
    l.tb.cand_s = c.take<double>(ne);
    l.tb.tri = c.take<int32_t>(ne);
    l.tb.trs = c.take<double>(ne);
    l.tb.trt = c.take<int32_t>(ne);
    l.tb.use = c.take<uint8_t>(ne);
    l.tb.keep = c.take<uint8_t>(ne);
    l.tb.passes = c.take<int32_t>(std::max<int64_t>(nch, 1));
    l.tb.status = c.take<int32_t>(1);
    l.tb.M = static_cast<int32_t>(M);
    l.tb.T = T;
    l.tb.K = K;
    l.bytes = c.bytes();
    return l;
}

}  // namespace

extern "C" int64_t msamd_fit_stage_workspace_size(int64_t M, int64_t N, const int64_t* times_host,
                                                  int64_t L, int32_t K, int32_t T,
                                                  int64_t chunk_size, int64_t padding) {
    if (M < 1 || N < 0 || L < 0 || K < 0 || T < 1 || (L > 0 && !times_host)) return -1;
    FitPlan P;
    std::vector<int32_t> dummy(static_cast<size_t>(L > 0 ? L : 1), 1);
    if (fit_plan(N, times_host, dummy.data(), L, chunk_size, padding, M, P) != MSAMD_OK) {
        clear_error();
        return -1;
    }
    return fit_layout(nullptr, P, M, K, T).bytes;
}

extern "C" int msamd_fit_stage(const float* X, int64_t M, int64_t N, const int64_t* times_host,
                               const int32_t* labels_host, int64_t L, const float* templates,
                               int32_t K, int32_t T, double mask_threshold, int64_t chunk_size,
                               int64_t padding, uint8_t* keep_host, int32_t* max_passes_host,
                               void* workspace, int64_t workspace_bytes, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 0 && L >= 0 && K >= 0 && T >= 1, "fit_stage: bad shape");
    MSAMD_REQUIRE(L == 0 || (times_host && labels_host && keep_host), "fit_stage: null events");
    MSAMD_REQUIRE(M * T < (1LL << 31), "fit_stage: template too large");
    FitPlan P;
    int rc = fit_plan(N, times_host, labels_host, L, chunk_size, padding, M, P);
    if (rc != MSAMD_OK) return rc;
    FitLayout l = fit_layout(workspace, P, M, K, T);
    if (!workspace || workspace_bytes < l.bytes) {
        set_error("fit_stage: workspace too small (%lld < %lld)", (long long)workspace_bytes,
                  (long long)l.bytes);
        return MSAMD_ERR_WORKSPACE;
    }
    for (int64_t i = 0; i < L; ++i) keep_host[i] = 0;
    if (max_passes_host) *max_passes_host = 0;
    const int64_t nch = static_cast<int64_t>(P.n2.size());
    if (nch == 0 || L == 0 || K == 0) return MSAMD_OK;
    hipStream_t st = as_stream(stream);
    FitTables& tb = l.tb;
    tb.tmpl = templates;
    const int64_t ne = static_cast<int64_t>(P.ltime.size());
    auto up = [&](void* d, const void* h, size_t n) {
        return n ? hipMemcpyAsync(d, h, n, hipMemcpyHostToDevice, st) : hipSuccess;
    };
    MSAMD_HIP(up(const_cast<int64_t*>(tb.xoff), P.xoff.data(), nch * 8));
    MSAMD_HIP(up(l.d_start, P.start.data(), nch * 8));
    MSAMD_HIP(up(const_cast<int32_t*>(tb.n2), P.n2.data(), nch * 4));
    MSAMD_HIP(up(const_cast<int64_t*>(tb.eoff), P.eoff.data(), (nch + 1) * 8));
    MSAMD_HIP(up(const_cast<int32_t*>(tb.ltime), P.ltime.data(), ne * 4));
    MSAMD_HIP(up(const_cast<int32_t*>(tb.lab), P.lab.data(), ne * 4));
    MSAMD_HIP(hipMemsetAsync(tb.stamp, 0, sizeof(uint16_t) * P.xtotal, st));
    MSAMD_HIP(hipMemsetAsync(tb.status, 0, sizeof(int32_t), st));
    MSAMD_HIP(hipStreamSynchronize(st));  // pageable sources
    {
        int64_t maxn2 = 0;
        for (int32_t v : P.n2) maxn2 = std::max<int64_t>(maxn2, v);
        const dim3 grid(static_cast<unsigned>(std::min<int64_t>(ceil_div(M * maxn2, 256), 4096)),
                        static_cast<unsigned>(nch));
        fit_copy_kernel<<<grid, 256, 0, st>>>(X, M, N, tb.xoff, l.d_start, tb.n2, tb.xc);
        MSAMD_LAUNCH_CHECK();
    }
    fit_mask_kernel<<<static_cast<unsigned>(K), kFitThreads, 0, st>>>(
        templates, static_cast<int32_t>(M * T), mask_threshold, const_cast<int32_t*>(tb.mask),
        const_cast<int32_t*>(tb.mlen));
    MSAMD_LAUNCH_CHECK();
    fit_chunk_kernel<<<static_cast<unsigned>(nch), kFitThreads, 0, st>>>(tb);
    MSAMD_LAUNCH_CHECK();
    std::vector<uint8_t> keep(static_cast<size_t>(ne));
    std::vector<int32_t> passes(static_cast<size_t>(nch));
    int32_t status = 0;
    MSAMD_HIP(hipMemcpyAsync(keep.data(), tb.keep, ne, hipMemcpyDeviceToHost, st));
    MSAMD_HIP(hipMemcpyAsync(passes.data(), tb.passes, nch * 4, hipMemcpyDeviceToHost, st));
    MSAMD_HIP(hipMemcpyAsync(&status, tb.status, 4, hipMemcpyDeviceToHost, st));
    MSAMD_HIP(hipStreamSynchronize(st));
    if (status) {
        set_error("fit_stage: a chunk needed more than 65535 passes");
        return MSAMD_ERR_CAPACITY;
    }
    // an event is used when the chunk that fitted it has it in its core:
    // padding <= t1 < size, the reference's test (fitstagecomputer.cpp:92-93)
    for (int64_t c = 0; c < nch; ++c) {
        const int64_t size = P.n2[c] - 2 * padding;
        for (int64_t e = P.eoff[c]; e < P.eoff[c + 1]; ++e)
            if (keep[e] && P.ltime[e] >= padding && P.ltime[e] < size) keep_host[P.gidx[e]] = 1;
        if (max_passes_host) *max_passes_host = std::max(*max_passes_host, passes[c]);
    }
    return MSAMD_OK;
}

extern "C" int64_t msamd_time_chunk_size(int64_t M, int64_t N, int64_t num_simultaneous,
                                         int64_t min_num_chunks) {
    if (M < 1 || N < 0 || num_simultaneous < 1 || min_num_chunks < 1) return -1;
    const int64_t ram = 1000000000;  // p_mountainsort3.cpp:509
    int64_t chunk = ram / (M * num_simultaneous * 4);
    if (chunk > N) chunk = N;
    if (N < chunk * min_num_chunks) chunk = N / min_num_chunks;
    if (chunk < 20000) chunk = 20000;
    return chunk;
}
