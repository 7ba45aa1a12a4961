// Clip extraction (a11): gather W x T windows of the timeseries at event times.
//
// Reference: NeighborhoodSorterPrivate::get_clip
// (packages/mountainsortalg/neighborhoodsorter.cpp:409-419) and the clip loop
// of p_extract_clips (packages/ms3/p_extract_clips.cpp:53-69), whose reads
// are zero-filled outside the file (mda/diskreadmda32.cpp:631-650).
//
// One wavefront per clip: lanes walk the W*T outputs in storage order
// (w fastest), so each clip is written as one contiguous run and each time
// row of the source is read as a short run of nearby channels.  Pure copy:
// results are bit-identical to the reference.
#include "common.h"

namespace msamd {
namespace {

__global__ void __launch_bounds__(256) clips_kernel(
    const float* __restrict__ X, int64_t M, int64_t N, const int64_t* __restrict__ starts,
    int64_t L, const int32_t* __restrict__ channels, int32_t W, int32_t T,
    float* __restrict__ clips) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int64_t nwaves = (static_cast<int64_t>(gridDim.x) * blockDim.x) >> 6;
    const int32_t WT = W * T;
    for (int64_t i = wave; i < L; i += nwaves) {
        const int64_t t0 = starts[i];
        float* out = clips + i * WT;
        for (int32_t r = lane; r < WT; r += 64) {
            const int32_t t = r / W;
            const int32_t w = r - t * W;
            const int64_t tt = t0 + t;
            out[r] = (tt >= 0 && tt < N) ? X[tt * M + channels[w]] : 0.0f;
        }
    }
}

// Batched form: events of neighborhood g are times[eo[g]..eo[g+1]) (centers),
// its channels nc[no[g]..no[g+1]), its clips start at clips + co[g].
// Work item (g, j) takes the j-th of P equal slices of g's events, one
// wavefront per clip.  Events are time-sorted per neighborhood, so slice j of
// every neighborhood lies near time j/P of the recording.  Workgroups are
// dispatched round-robin over the eight XCDs; XCD k is given the contiguous
// neighborhoods [k*gx, (k+1)*gx) and walks them slice by slice, so the
// overlapping source rows of neighboring channels' clips meet in its L2.  For W <= 64 each lane owns one channel
// (w = lane % W) and a row offset (lane / W), so the channel index is loaded
// once and the W x T clip is covered by rows of 64 / W time samples; the
// loads of up to 16 rows per lane (a whole clip of T = 50 samples for
// W <= 16) are issued before their stores.
constexpr int kClipRows = 16;

__global__ void __launch_bounds__(256) clips_batched_kernel(
    const float* __restrict__ X, int64_t M, int64_t N, const int64_t* __restrict__ times,
    const int64_t* __restrict__ eo, const int32_t* __restrict__ nc,
    const int64_t* __restrict__ no, int32_t G, int32_t P, int32_t T, int32_t tmid,
    const int64_t* __restrict__ co, float* __restrict__ clips) {
    const int64_t b = blockIdx.x;
    const int gx = (G + 7) / 8;                 // neighborhoods per XCD
    const int64_t sx = b / 8;                   // position in this XCD's sequence
    const int g = static_cast<int>((b % 8) * gx + sx % gx);
    const int64_t j = sx / gx;                  // slice
    if (g >= G || j >= P) return;
    const int lane = threadIdx.x & 63;
    const int64_t eg = eo[g];
    const int64_t Lg = eo[g + 1] - eg;
    const int64_t s0 = Lg * j / P;              // this slice: events [s0, s1) of g
    const int64_t s1 = Lg * (j + 1) / P;
    const int64_t e0 = eg + s0;
    const int64_t L = s1 - s0;
    const int32_t* ch = nc + no[g];
    const int32_t W = static_cast<int32_t>(no[g + 1] - no[g]);
    const int32_t WT = W * T;
    float* base = clips + co[g] + s0 * WT;
    const int64_t wave = threadIdx.x >> 6;
    const int64_t nwaves = blockDim.x >> 6;
    if (W <= 64) {
        const int rpi = 64 / W;           // rows per pass
        const int w = lane % W;
        const int tl = lane / W;
        const bool active = tl < rpi;
        const int64_t chw = active ? ch[w] : 0;
        for (int64_t i = wave; i < L; i += nwaves) {
            const int64_t t0 = times[e0 + i] - tmid;
            float* out = base + i * WT;
            if (!active) continue;
            // up to kClipRows rows per lane in flight before their stores
            for (int tb = tl; tb < T; tb += kClipRows * rpi) {
                float v[kClipRows];
#pragma unroll
                for (int k = 0; k < kClipRows; ++k) {
                    const int t = tb + k * rpi;
                    const int64_t tt = t0 + t;
                    v[k] = (t < T && tt >= 0 && tt < N) ? X[tt * M + chw] : 0.0f;
                }
#pragma unroll
                for (int k = 0; k < kClipRows; ++k) {
                    const int t = tb + k * rpi;
                    if (t < T) out[t * W + w] = v[k];
                }
            }
        }
        return;
    }
    for (int64_t i = wave; i < L; i += nwaves) {
        const int64_t t0 = times[e0 + i] - tmid;
        float* out = base + i * WT;
        for (int32_t r = lane; r < WT; r += 64) {
            const int32_t t = r / W;
            const int32_t w = r - t * W;
            const int64_t tt = t0 + t;
            out[r] = (tt >= 0 && tt < N) ? X[tt * M + ch[w]] : 0.0f;
        }
    }
}

inline unsigned grid_for_waves(int64_t nwaves) {
    int64_t blocks = ceil_div(nwaves, 4);
    if (blocks > 65536) blocks = 65536;
    if (blocks < 1) blocks = 1;
    return static_cast<unsigned>(blocks);
}

}  // namespace
}  // namespace msamd

using namespace msamd;

extern "C" int msamd_extract_clips(const float* X, int64_t M, int64_t N, const int64_t* starts,
                                   int64_t L, const int32_t* channels, int32_t W, int32_t T,
                                   float* clips, void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 0 && L >= 0, "extract_clips: bad shape");
    MSAMD_REQUIRE(T >= 1, "extract_clips: clip_size is zero");
    MSAMD_REQUIRE(W >= 1, "extract_clips: empty channel list");
    if (L == 0) return MSAMD_OK;
    clips_kernel<<<grid_for_waves(L), 256, 0, as_stream(stream)>>>(X, M, N, starts, L, channels,
                                                                   W, T, clips);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}

extern "C" int msamd_extract_clips_batched(const float* X, int64_t M, int64_t N,
                                           const int64_t* times, const int64_t* event_offsets,
                                           int64_t total_events, const int32_t* nbhd_channels,
                                           const int64_t* nbhd_offsets, int32_t G, int32_t T,
                                           const int64_t* clip_offsets, float* clips,
                                           void* stream) {
    clear_error();
    MSAMD_REQUIRE(M >= 1 && N >= 0 && G >= 1 && total_events >= 0,
                  "extract_clips_batched: bad shape");
    MSAMD_REQUIRE(T >= 1, "extract_clips_batched: clip_size is zero");
    if (total_events == 0) return MSAMD_OK;
    const int32_t tmid = (T + 1) / 2 - 1;
    // slices per neighborhood sized for ~16 clips (4 per wavefront) at the mean event count
    int64_t P = ceil_div(ceil_div(total_events, G), 16);
    if (P < 1) P = 1;
    const int64_t gx = ceil_div(G, 8);
    const int64_t blocks = 8 * gx * P;
    MSAMD_REQUIRE(blocks < (1LL << 31), "extract_clips_batched: too many work items");
    clips_batched_kernel<<<static_cast<unsigned>(blocks), 256, 0, as_stream(stream)>>>(
        X, M, N, times, event_offsets, nbhd_channels, nbhd_offsets, G, static_cast<int32_t>(P), T,
        tmid, clip_offsets, clips);
    MSAMD_LAUNCH_CHECK();
    return MSAMD_OK;
}
