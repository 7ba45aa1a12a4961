/*
 * msamd.h -- C ABI of the MI355X (gfx950) MountainSort front end.
 *
 * This is the drop-in seam beneath the MountainLab processors
 * ms3.bandpass_filter, ms3.mask_out_artifacts, ms3.whiten, ms3.extract_clips
 * and the detect / clip / PCA phase of mountainsortalg.ms3alg.  Each entry
 * point names the reference function it replaces (paths relative to the
 * MountainSort repository root).
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer unless its name ends in _host.
 *  - Timeseries use the .mda column-major layout of the reference: an M x N
 *    array stores sample (m, t) at index m + M*t (channels interleaved per
 *    time sample, packages/mountainsortalg/mda/mda32.h:91-96).
 *  - Clips W x T x L store (w, t, i) at w + W*t + W*T*i.
 *  - Buffers are caller-owned.  Scratch memory comes in through a workspace
 *    argument whose size is reported by the matching *_workspace_size call.
 *  - stream is a hipStream_t (NULL = default stream).  Calls are
 *    asynchronous unless documented otherwise.
 *  - Return value: MSAMD_OK (0) or an MSAMD_ERR_* code; msamd_last_error()
 *    returns the message of the last failure on the calling thread.  No C++
 *    exception crosses this boundary.
 */
#ifndef MSAMD_H
#define MSAMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSAMD_OK 0
#define MSAMD_ERR_ARG 1
#define MSAMD_ERR_HIP 2
#define MSAMD_ERR_WORKSPACE 3
#define MSAMD_ERR_CAPACITY 4
#define MSAMD_ERR_UNSUPPORTED 5
#define MSAMD_ERR_TIMEOUT 6   /* a bounded wait between workgroups gave up */

/* .mda element type codes (packages/ms3/mda/mdaio.h:55-62) accepted as input. */
#define MSAMD_DTYPE_FLOAT32 (-3)
#define MSAMD_DTYPE_INT16 (-4)

/* Library identity. */
int msamd_version(void);
const char* msamd_last_error(void);

/* ------------------------------------------------------------------------
 * Bandpass filter (a1-a3) -- replaces p_bandpass_filter with its
 * Kernel_runner::init/apply and define_kernel
 * (packages/ms3/p_bandpass_filter.cpp:47-101,115-229,421-447).
 * ------------------------------------------------------------------------ */

/* Core chunk length of the filter windows for M channels and the reference
 * thread count (p_bandpass_filter.cpp:147-155); the FFT length is
 * chunk + 2*overlap with overlap = 60000. */
int64_t msamd_bandpass_chunk_size(int64_t M, int64_t num_threads);

/* Gain H[i], i < n_fft, of define_kernel (FP64, host memory). */
int msamd_bandpass_kernel(int64_t n_fft, double samplerate, double freq_min, double freq_max,
                          double freq_wid, double* kernel_host);

/* 1 when the FFT length chunk_size + 2*overlap has a compiled factorization. */
int msamd_bandpass_supported(int64_t chunk_size, int64_t overlap);

/* Workspace bytes for msamd_bandpass_filter (-1 when unsupported). */
int64_t msamd_bandpass_workspace_size(int64_t M, int64_t N, int64_t chunk_size, int64_t overlap);

/* Filters X (M x N, dtype MSAMD_DTYPE_FLOAT32 or MSAMD_DTYPE_INT16) window by
 * window exactly as the reference: window [t-overlap, t+chunk+overlap),
 * zero outside [0, N), DFT, times H/n, inverse DFT, core kept.  FP32
 * arithmetic.  subsample_factor k > 1 keeps every k-th core sample
 * (p_bandpass_filter.cpp:480-491); Y has N/k time points.  Y is float32, or
 * int16 when quantization_unit > 0 (y * (1/q) rounded to float, then
 * truncated, :207-209).  freq_max == 0 (a plain copy in the reference) is an
 * error here. */
int msamd_bandpass_filter(const void* X, int dtype, int64_t M, int64_t N, double samplerate,
                          double freq_min, double freq_max, double freq_wid, int64_t chunk_size,
                          int64_t overlap, int subsample_factor, double quantization_unit,
                          void* Y, void* workspace, int64_t workspace_bytes, void* stream);

/* Number of filter windows of an N-timepoint recording: ceil(N / chunk_size)
 * (the window loop of p_bandpass_filter.cpp:179), -1 on bad arguments. */
int64_t msamd_bandpass_windows(int64_t N, int64_t chunk_size);

/* Windows [w0, w1) of msamd_bandpass_filter, for streaming a recording
 * through the device a segment at a time (the reference reads each window
 * from disk, p_bandpass_filter.cpp:179-214).  X holds recording rows
 * [x_t0, x_t0 + x_rows) of the N-timepoint recording and must cover the
 * windows' input [max(0, w0*cs - overlap), min(N, w1*cs + overlap)); Y row 0
 * is output row y_t0 and Y must cover output rows [(w0*cs)/k, min(N/k,
 * (w1*cs)/k)).  The output rows equal those of msamd_bandpass_filter on the
 * whole recording bit for bit.  The workspace of
 * msamd_bandpass_workspace_size(M, N, ...) suffices for any window range. */
int msamd_bandpass_filter_windows(const void* X, int dtype, int64_t M, int64_t N, int64_t x_t0,
                                  int64_t x_rows, int64_t w0, int64_t w1, double samplerate,
                                  double freq_min, double freq_max, double freq_wid,
                                  int64_t chunk_size, int64_t overlap, int subsample_factor,
                                  double quantization_unit, void* Y, int64_t y_t0, int64_t y_rows,
                                  void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Event detection (a10) -- replaces detect_events()
 * (packages/mountainsortalg/detect_events.cpp:69-135) as driven by
 * NeighborhoodSorter::addTimeChunk (neighborhoodsorter.cpp:74-106) over the
 * time-chunk plan of p_mountainsort3 (p_mountainsort3.cpp:170-216,507-533).
 * ------------------------------------------------------------------------ */

/* Chunk length of the detection plan for M channels, N timepoints and the
 * reference thread count (p_mountainsort3.cpp:170-172,191,507-533). */
int64_t msamd_detect_chunk_size(int64_t M, int64_t N, int64_t num_threads);

/* Workspace bytes for msamd_detect_events on the time range of n_range
 * samples and num_channels central channels. */
int64_t msamd_detect_workspace_size(int64_t n_range, int64_t chunk_size, int32_t num_channels,
                                    double detect_interval);

/* Detects events on every central channel: per (channel, padded chunk) the
 * mean and stdev are the reference's sequential double sums over the padded
 * chunk (zero-filled outside the file), events are threshold crossings that
 * win detect_interval suppression, and only core events are kept.
 * X: M x N_file float32.  t1/t2: inclusive time range (t1 < 0 -> 0, t2 < 0 ->
 * N_file-1), as the ms3alg t1/t2 parameters.  channels: num_channels 0-based
 * channel indices (int32, device).  padding: 1000 in the reference.
 * Synchronizes the stream once and returns the total event count in
 * *num_events_host; the events stay in the workspace until
 * msamd_detect_fetch. */
int msamd_detect_events(const float* X, int64_t M, int64_t N_file, int64_t t1, int64_t t2,
                        int64_t chunk_size, int64_t padding, const int32_t* channels,
                        int32_t num_channels, double detect_threshold, double detect_interval,
                        int detect_sign, void* workspace, int64_t workspace_bytes,
                        int64_t* num_events_host, void* stream);

/* Writes the events found by the last msamd_detect_events on this workspace:
 * times (int64, absolute 0-based sample indices, ascending per channel,
 * channels in list order) and offsets (num_channels+1 int64: events of
 * channel c are times[offsets[c]..offsets[c+1])). */
int msamd_detect_fetch(const void* workspace, int32_t num_channels, int64_t* times,
                       int64_t* offsets, void* stream);

/* ------------------------------------------------------------------------
 * Clip extraction (a11) -- replaces NeighborhoodSorterPrivate::get_clip
 * (neighborhoodsorter.cpp:409-419) and the gather of p_extract_clips
 * (packages/ms3/p_extract_clips.cpp:53-69).
 * ------------------------------------------------------------------------ */

/* clips[w + W*t + W*T*i] = X[channels[w] + M*(starts[i] + t)], 0 outside
 * [0, N).  starts: L int64 first-sample indices (the processor form uses
 * (int64)(event_time - Tmid), Tmid = (T+1)/2 - 1).  channels: W 0-based. */
int msamd_extract_clips(const float* X, int64_t M, int64_t N, const int64_t* starts, int64_t L,
                        const int32_t* channels, int32_t W, int32_t T, float* clips,
                        void* stream);

/* Clip gather for G neighborhoods in one launch.  Neighborhood g owns events
 * times[event_offsets[g] .. event_offsets[g+1]) (centers; start = time -
 * Tmid) and channels nbhd_channels[nbhd_offsets[g] .. nbhd_offsets[g+1]);
 * its W_g x T x L_g clips are written at clips + clip_offsets[g].
 * event_offsets, nbhd_offsets, clip_offsets: G+1 int64 (device). */
int msamd_extract_clips_batched(const float* X, int64_t M, int64_t N, const int64_t* times,
                                const int64_t* event_offsets, int64_t total_events,
                                const int32_t* nbhd_channels, const int64_t* nbhd_offsets,
                                int32_t G, int32_t T, const int64_t* clip_offsets, float* clips,
                                void* stream);

/* ------------------------------------------------------------------------
 * Two-stage PCA features (a12-a15) -- replaces dimension_reduce_clips
 * (packages/mountainsortalg/neighborhoodsorter.cpp:236-276), pca_subsampled /
 * pca (packages/mountainsortalg/pca.cpp:68-127) and the stage-2 PCA of
 * sort_clips_subset (sort_clips.cpp:58-70) that feeds isosplit5.
 * ------------------------------------------------------------------------ */

/* Workspace bytes for msamd_pca_features (-1 on invalid arguments). */
int64_t msamd_pca_workspace_size(int32_t G, const int64_t* num_events_host,
                                 const int32_t* widths_host, int32_t T,
                                 int32_t num_features_per_channel, int32_t num_features,
                                 int64_t max_pca_samples);

/* Features of G neighborhoods in one batched pass.  Neighborhood g has
 * num_events_host[g] clips of widths_host[g] channels (1..128, with
 * num_features_per_channel * width <= 2048) and T samples (<= 64), packed
 * one after another in clips (W x T x L each, the clip layout above).
 * Stage 1: per channel, PCA of the T x L slice with num_features_per_channel
 * components; reduced receives W x nf1 x L per neighborhood, packed.
 * Stage 2: PCA of reduced viewed as (nf1*W) x L with num_features
 * components; features receives nf2 x L per neighborhood, packed (the
 * isosplit5 input).  Both PCAs use the pca_subsampled column rule with
 * max_pca_samples, the 10-step power method with deflation from
 * normalize(sin(1..n)) (FP64), Gram matrices from two-piece FP16 MFMA
 * products after a power-of-two scale per 32-column step (hh + hl + lh, 22
 * significant bits; FP32 within a step, FP64 across steps), and FP32
 * projections over all columns. */
int msamd_pca_features(const float* clips, int32_t G, const int64_t* num_events_host,
                       const int32_t* widths_host, int32_t T, int32_t num_features_per_channel,
                       int32_t num_features, int64_t max_pca_samples, float* reduced,
                       float* features, void* workspace, int64_t workspace_bytes, void* stream);

/* The same computation as a reusable plan: the host work (neighborhood
 * tables, work items, their upload into the workspace) is done once by
 * msamd_pca_plan_create, and msamd_pca_plan_run only enqueues the nine
 * batched kernels, so it may be captured into a HIP graph.
 * msamd_pca_plan_launch_graph captures that sequence once per (clips,
 * reduced, features) triple on a private stream and replays it with one
 * hipGraphLaunch on the given stream.  tiles (nullable) = {Gram columns per
 * work item (multiple of 32), stage-1 events per projection item, stage-2
 * events per projection item}; null selects {2560, 64, 64}.  The plan keeps
 * pointers into the workspace, which must outlive it.  With workspace NULL
 * and workspace_bytes 0, msamd_pca_plan_create builds the host tables only
 * (the tables are built once, not once for the size query and once for the
 * plan): msamd_pca_plan_workspace_bytes gives the workspace it needs and
 * msamd_pca_plan_attach uploads the tables into one (it may be called again
 * with another workspace); runs are refused until then.  No reference
 * counterpart: the reference computes each neighborhood's PCA serially on
 * the CPU (neighborhoodsorter.cpp:236-276); callers that process many
 * recordings of one probe reuse a plan. */
typedef struct msamd_pca_plan msamd_pca_plan;
int64_t msamd_pca_plan_workspace_size(int32_t G, const int64_t* num_events_host,
                                      const int32_t* widths_host, int32_t T,
                                      int32_t num_features_per_channel, int32_t num_features,
                                      int64_t max_pca_samples, const int32_t* tiles);
int msamd_pca_plan_create(int32_t G, const int64_t* num_events_host, const int32_t* widths_host,
                          int32_t T, int32_t num_features_per_channel, int32_t num_features,
                          int64_t max_pca_samples, const int32_t* tiles, void* workspace,
                          int64_t workspace_bytes, void* stream, msamd_pca_plan** plan);
int msamd_pca_plan_run(msamd_pca_plan* plan, const float* clips, float* reduced, float* features,
                       void* stream);
int msamd_pca_plan_launch_graph(msamd_pca_plan* plan, const float* clips, float* reduced,
                                float* features, void* stream);
void msamd_pca_plan_tiles(const msamd_pca_plan* plan, int32_t* tiles);
void msamd_pca_plan_destroy(msamd_pca_plan* plan);
int64_t msamd_pca_plan_workspace_bytes(const msamd_pca_plan* plan);
int msamd_pca_plan_attach(msamd_pca_plan* plan, void* workspace, int64_t workspace_bytes,
                          void* stream);

/* The plan's computation with the clip gather fused into the stage-1
 * kernels: no clip array is written or read.  The stage-1 Gram and
 * projection read each clip's samples from the timeseries X (M x N,
 * x[m + M*t]) as msamd_extract_clips_batched would write them: neighborhood
 * g's events are times[ev_g .. ev_g + num_events_host[g]) (centers, packed
 * neighborhood after neighborhood, ev_g the sum of the earlier counts), its
 * channels nbhd_channels[nbhd_offsets[g] .. nbhd_offsets[g+1]) (0-based,
 * central first; the counts must equal the plan's widths), clip start time -
 * Tmid, zero outside [0, N).  reduced and features are bit-identical to
 * msamd_pca_plan_run on the gathered clips.  The replaced pair is get_clip
 * (neighborhoodsorter.cpp:409-419) feeding dimension_reduce_clips
 * (neighborhoodsorter.cpp:236-276). */
int msamd_pca_plan_run_timeseries(msamd_pca_plan* plan, const float* X, int64_t M, int64_t N,
                                  const int64_t* times, const int32_t* nbhd_channels,
                                  const int64_t* nbhd_offsets, float* reduced, float* features,
                                  void* stream);

/* Autotuner of the plan's tiles for one workload: a coordinate search over
 * Gram columns per item {1280, 2560, 5120, 10240}, stage-1 events per item
 * {32, 64, 128, 256} and stage-2 events per item {16, 32, 64, 128}, each
 * candidate timed with HIP events on the stream (second of two runs).
 * Writes the best tiles to tiles_out[3] and their time to best_ms_out
 * (nullable).  reduced/features receive the last candidate's results.  The
 * workspace must hold msamd_pca_autotune_workspace_size bytes.  Synchronizes
 * the stream. */
int64_t msamd_pca_autotune_workspace_size(int32_t G, const int64_t* num_events_host,
                                          const int32_t* widths_host, int32_t T,
                                          int32_t num_features_per_channel, int32_t num_features,
                                          int64_t max_pca_samples);
int msamd_pca_autotune(const float* clips, int32_t G, const int64_t* num_events_host,
                       const int32_t* widths_host, int32_t T, int32_t num_features_per_channel,
                       int32_t num_features, int64_t max_pca_samples, float* reduced,
                       float* features, void* workspace, int64_t workspace_bytes, void* stream,
                       int32_t* tiles_out, float* best_ms_out);
/* The same search for the fused form (msamd_pca_plan_run_timeseries). */
int msamd_pca_autotune_timeseries(const float* X, int64_t M, int64_t N, const int64_t* times,
                                  const int32_t* nbhd_channels, const int64_t* nbhd_offsets,
                                  int32_t G, const int64_t* num_events_host,
                                  const int32_t* widths_host, int32_t T,
                                  int32_t num_features_per_channel, int32_t num_features,
                                  int64_t max_pca_samples, float* reduced, float* features,
                                  void* workspace, int64_t workspace_bytes, void* stream,
                                  int32_t* tiles_out, float* best_ms_out);

/* PCA of index subsets, the PCA of sort_clips_subset
 * (packages/mountainsortalg/sort_clips.cpp:47-70) that the branch method
 * (:97-123) repeats on every cluster isosplit5 returns: subset g is the
 * columns indices[offsets_host[g] .. offsets_host[g+1]) (device int64, in
 * that order) of data, an R x L float32 array (R <= 2048; the reduced
 * clips W x nf1 x L viewed as (nf1*W) x L).  features receives
 * num_features x L_g per subset, packed.  Same arithmetic as the stage-2
 * PCA of msamd_pca_features (pca_subsampled rule with max_pca_samples). */
int64_t msamd_pca_subsets_workspace_size(int32_t G, const int64_t* offsets_host, int32_t R,
                                         int32_t num_features, int64_t max_pca_samples);
int msamd_pca_subsets(const float* data, int32_t R, int64_t L, const int64_t* indices,
                      const int64_t* offsets_host, int32_t G, int32_t num_features,
                      int64_t max_pca_samples, float* features, void* workspace,
                      int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Whitening (a4-a6) -- replaces p_whiten (packages/ms3/p_whiten.cpp:45-206)
 * and whitening_matrix_from_XXt (packages/mountainsortalg/pca.cpp:634-650).
 * ------------------------------------------------------------------------ */

/* Workspace bytes for msamd_whiten_covariance. */
int64_t msamd_whiten_covariance_workspace_size(int64_t M, int64_t N);

/* XXt = X X^T / (N-1) (no mean subtraction; p_whiten.cpp:64-118).
 * Split-BF16 3-term MFMA products (x = hi + lo; hi*hi + hi*lo + lo*hi)
 * accumulated in FP32 per block of samples and in FP64 across blocks.
 * XXt: M x M float64, symmetric.  M <= 1024. */
int msamd_whiten_covariance(const float* X, int64_t M, int64_t N, double* XXt, void* workspace,
                            int64_t workspace_bytes, void* stream);

/* The covariance over time segments, for streaming a recording through the
 * device (the reference sums per-chunk partials, p_whiten.cpp:75-111): the
 * unnormalized sum of x x^T over the N rows of X (row_keep may be NULL, as
 * in msamd_whiten_covariance_rows) is written to XXt_sum (accumulate == 0)
 * or added to it (accumulate != 0).  msamd_whiten_covariance_finish then
 * divides by N_total - 1 (when N_total > 1, p_whiten.cpp:114-118).  Same
 * workspace as msamd_whiten_covariance for this N. */
int msamd_whiten_covariance_accumulate(const float* X, int64_t M, int64_t N,
                                       const uint32_t* row_keep, double* XXt_sum, int accumulate,
                                       void* workspace, int64_t workspace_bytes, void* stream);
int msamd_whiten_covariance_finish(double* XXt, int64_t M, int64_t N_total, void* stream);

/* msamd_whiten_covariance of the masked timeseries, with the mask given as
 * msamd_mask_rows' row bitmap: rows whose bit is clear read as zero.  The
 * result equals msamd_whiten_covariance of the msamd_mask_out_artifacts
 * output bit for bit.  Same workspace as msamd_whiten_covariance. */
int msamd_whiten_covariance_rows(const float* X, int64_t M, int64_t N, const uint32_t* row_keep,
                                 double* XXt, void* workspace, int64_t workspace_bytes,
                                 void* stream);

/* Workspace bytes for msamd_whitening_matrix. */
int64_t msamd_whitening_matrix_workspace_size(int64_t M);

/* W from XXt by the reference's method: per component 10 power steps from
 * normalize(sin(1..M)), sigma = |G v|, deflation G <- (I-vv')G(I-vv'), then
 * W = C diag(sigma^-1/2) C' (zero where sigma == 0).  FP64 on the GPU.
 * XXt and W: M x M float64 (device).  The power steps exchange values
 * between workgroups that must run at the same time; when other work on the
 * device keeps some from starting, the bounded waits give up and the call
 * returns MSAMD_ERR_TIMEOUT (W is then not written; callers may retry or use
 * msamd_whitening_matrix_exact). */
int msamd_whitening_matrix(const double* XXt, int64_t M, double* W, void* workspace,
                           int64_t workspace_bytes, void* stream);

/* Exact mode of the whitening matrix (no reference counterpart; the
 * reference's W in whitening_matrix_from_XXt, packages/mountainsortalg/
 * pca.cpp:634-650, approximates it with unconverged power steps):
 * W = C^{-1/2}, the symmetric inverse square root of XXt, by the coupled
 * Newton-Schulz iteration in FP64 on the GPU (tiled FP64 GEMMs).  Channels
 * with a zero diagonal get zero rows and columns.  Stops when
 * ||(3I - ZY)/2 - I||_F < 1e-12 sqrt(M) or at rounding level; returns
 * MSAMD_ERR_UNSUPPORTED if XXt is not positive definite on its nonzero
 * channels or does not converge in 100 iterations.  iterations_out and
 * residual_out (nullable) receive the iteration count and final residual.
 * Synchronizes the stream every six iterations.  M <= 4096. */
int64_t msamd_whitening_matrix_exact_workspace_size(int64_t M);
int msamd_whitening_matrix_exact(const double* XXt, int64_t M, double* W, void* workspace,
                                 int64_t workspace_bytes, void* stream, int32_t* iterations_out,
                                 double* residual_out);

/* Workspace bytes for msamd_whiten_apply. */
int64_t msamd_whiten_apply_workspace_size(int64_t M);

/* Y = quantize(W X, 1e-4) with quantize(x,u) = floor(x/u + 0.5)*u
 * (p_whiten.cpp:31-34,153-206).  Split-FP16 MFMA with FP32 accumulation:
 * both operands are scaled by powers of two into fp16's range (W by its
 * maximum, X by its maximum, found by a pass over X) and split into two
 * fp16 pieces (22 significant bits); the three products of order >= 2^-11
 * are accumulated, so products carry error O(2^-22), below the reference's
 * own float accumulation (the 1e-4 output quantum needs this; bf16 hi+lo
 * would not).
 * Y: M x N float32 (may be NULL when Yq is given).  When quantization_unit
 * > 0, Yq receives the int16 values (int)(y/q + 0.5) (p_whiten.cpp:445-451). */
int msamd_whiten_apply(const float* X, int64_t M, int64_t N, const double* W, float* Y,
                       double quantization_unit, int16_t* Yq, void* workspace,
                       int64_t workspace_bytes, void* stream);

/* msamd_whiten_apply of the masked timeseries (row bitmap as for
 * msamd_whiten_covariance_rows; masked rows give zero output rows), equal
 * bit for bit to msamd_whiten_apply of the masked array.  Same workspace as
 * msamd_whiten_apply. */
int msamd_whiten_apply_rows(const float* X, int64_t M, int64_t N, const double* W,
                            const uint32_t* row_keep, float* Y, double quantization_unit,
                            int16_t* Yq, void* workspace, int64_t workspace_bytes, void* stream);

/* msamd_whiten_apply (or _rows, row_keep non-NULL) with the covariance XXt
 * of this X (M x M float64, as msamd_whiten_covariance / _rows returned it
 * for the same rows): its diagonal bounds |x| (x_tj^2 <= (N-1) C_jj), which
 * sets the fp16 scaling of the samples without the pass over X that the
 * other forms make.  No reference counterpart; results equal those forms'
 * up to that scale choice. */
int msamd_whiten_apply_cov(const float* X, int64_t M, int64_t N, const double* W,
                           const double* XXt, const uint32_t* row_keep, float* Y,
                           double quantization_unit, int16_t* Yq, void* workspace,
                           int64_t workspace_bytes, void* stream);

/* msamd_whiten_apply_cov without a workspace and without any allocation: the
 * W fragments and scale words live in a device array owned by the library
 * (about 1 MiB per device, held until the process ends), so M <= 512
 * (MSAMD_ERR_CAPACITY beyond).  Results equal msamd_whiten_apply_cov's bit
 * for bit (the same kernels).  Calls on one device share that array: calls
 * that overlap in time on one device race; calls ordered on one stream are
 * safe.  No counterpart in the reference; it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops,
 * harnesses that forbid hipMalloc). */
int msamd_whiten_apply_cov_noalloc(const float* X, int64_t M, int64_t N, const double* W,
                                   const double* XXt, const uint32_t* row_keep, float* Y,
                                   double quantization_unit, int16_t* Yq, void* stream);

/* ms3.whiten_clips (packages/ms3/p_whiten.cpp:295-345): every time sample of
 * every clip times W, without the 1e-4 quantize of the timeseries form.
 * clips, clips_out: M x T x L float32 (the clip layout); with
 * quantization_unit > 0, clips_out_q receives the int16 file values
 * (int)(y/q + 0.5) and clips_out may be NULL.  Same workspace as
 * msamd_whiten_apply. */
int msamd_whiten_clips(const float* clips, int64_t M, int64_t T, int64_t L, const double* W,
                       float* clips_out, double quantization_unit, int16_t* clips_out_q,
                       void* workspace, int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Artifact masking (a7) -- replaces p_mask_out_artifacts
 * (packages/ms3/p_mask_out_artifacts.cpp:24-104).
 * ------------------------------------------------------------------------ */

/* Workspace bytes for msamd_mask_out_artifacts. */
int64_t msamd_mask_workspace_size(int64_t M, int64_t N, int64_t interval_size);

/* Per (channel, interval) norms sqrt(sum x^2) (sequential double sums),
 * per-channel mean/stdev of the norms, intervals whose norm exceeds
 * mean + stdev*threshold on any channel are cleared together with both
 * neighbours; kept intervals are copied and everything else (including the
 * tail N mod interval_size) is zero.  X, Y: M x N float32.  use_it (may be
 * NULL): N/interval_size + 1 int32 flags.  threshold == 0 or
 * interval_size == 0 is an error, as in the reference. */
int msamd_mask_out_artifacts(const float* X, int64_t M, int64_t N, double threshold,
                             int64_t interval_size, float* Y, int32_t* use_it, void* workspace,
                             int64_t workspace_bytes, void* stream);

/* The mask's decisions without the copy: row_keep receives ceil(N/32) words,
 * bit t%32 of word t/32 set when row t lies in a kept interval (so the tail
 * N mod interval_size is clear), for the masked-input forms of the
 * covariance and the apply below.  use_it as above (may be NULL).  Same
 * workspace as msamd_mask_out_artifacts.  No reference counterpart: it
 * splits p_mask_out_artifacts so that the masked timeseries need not be
 * written when only the whitening consumes it. */
int msamd_mask_rows(const float* X, int64_t M, int64_t N, double threshold,
                    int64_t interval_size, uint32_t* row_keep, int32_t* use_it, void* workspace,
                    int64_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Templates and the fit stage -- replace p_compute_templates
 * (packages/ms3/p_compute_templates.cpp:24-118), the global templates of
 * ms3alg (packages/mountainsortalg/globaltemplatecomputer.cpp:63-135) and
 * its fit stage (fitstagecomputer.cpp:55-104, fit_stage.cpp:27-275, driven
 * by p_mountainsort3.cpp:270-356).  Events come in host arrays (times as
 * 0-based sample indices, labels 1..K; other labels are skipped).
 * ------------------------------------------------------------------------ */

/* Workspace bytes for msamd_compute_templates. */
int64_t msamd_templates_workspace_size(int64_t L, int32_t K);

/* templates (M x T x K float32, device): template k = the mean of the
 * M x T clips starting at time - Tmid (Tmid = (T+1)/2 - 1; samples outside
 * [0, N) read as zero) of the events labeled k, summed in FP64 in event
 * order and divided by their count (zero when there are none): the
 * reference's arithmetic and order.  Synchronizes the stream (the grouped
 * event lists are uploaded from the host). */
int msamd_compute_templates(const float* X, int64_t M, int64_t N, const int64_t* times_host,
                            const int32_t* labels_host, int64_t L, int32_t K, int32_t T,
                            float* templates, void* workspace, int64_t workspace_bytes,
                            void* stream);

/* Workspace bytes for msamd_fit_stage (-1 on invalid arguments). */
int64_t msamd_fit_stage_workspace_size(int64_t M, int64_t N, const int64_t* times_host, int64_t L,
                                       int32_t K, int32_t T, int64_t chunk_size,
                                       int64_t padding);

/* The fit stage over the time chunks [t1, t1 + chunk_size) of X (M x N
 * float32, device), each fitted on its own copy padded by `padding`
 * zero-filled samples with the events (times ascending) of that padded
 * window: passes of score (sum x^2 - sum (x - template)^2 over the
 * template's mask |template| > mask_threshold), local-maximum selection
 * within T samples and template subtraction, until a pass uses no event.
 * keep_host[i] = 1 for event i when the chunk whose core holds it used it
 * (the reference's chunk-local test padding <= t1 < size).  In ms3alg:
 * chunk_size from the plan get_time_chunk_infos(M, N, threads, threads),
 * padding 1000, mask_threshold 0.1.  max_passes_host (nullable): the most
 * passes any chunk took.  Synchronizes the stream. */
int msamd_fit_stage(const float* X, int64_t M, int64_t N, const int64_t* times_host,
                    const int32_t* labels_host, int64_t L, const float* templates, int32_t K,
                    int32_t T, double mask_threshold, int64_t chunk_size, int64_t padding,
                    uint8_t* keep_host, int32_t* max_passes_host, void* workspace,
                    int64_t workspace_bytes, void* stream);

/* Chunk length of get_time_chunk_infos(M, N, num_simultaneous,
 * min_num_chunks) (p_mountainsort3.cpp:507-533): 1e9 / (4 M
 * num_simultaneous) samples, at most N, at most N / min_num_chunks when
 * that many chunks would not fit, at least 20000. */
int64_t msamd_time_chunk_size(int64_t M, int64_t N, int64_t num_simultaneous,
                              int64_t min_num_chunks);

/* ------------------------------------------------------------------------
 * Learned spike-presence detector.  No counterpart in the reference: a
 * 3-layer temporal convolution network (32 hidden channels, kernel 9, ReLU,
 * strides 8 and 4) over the 7-channel window c-3..c+3 scores every
 * 32-sample window [32w, 32w+32) of every channel; detection then only
 * looks where the network proposes, with the reference's detection rule
 * (detect_events.cpp:69-135, time-chunk plan p_mountainsort3.cpp:170-216)
 * deciding uncertain windows.  Parameters: one float vector of
 * msamd_tcn_num_params() values, W1 [32][7][9] | b1 [32] | W2 [32][32][9] |
 * b2 [32] | W3 [32][9] | b3; weights are used as bf16, sums are fp32.
 * Windows: nW = ceil(N / 32); window-indexed arrays are [nW][M].
 * ------------------------------------------------------------------------ */

int32_t msamd_tcn_num_params(void);

/* He-normal weights from a counter-based generator, zero biases, b3 = -3. */
int msamd_tcn_init(float* params, uint64_t seed, void* stream);

/* Training labels from plain detection output (times/offsets as returned by
 * msamd_detect_fetch for channels 0..M-1): labels[w][c] = 1 when window w of
 * channel c holds an event, else 0.  positives (may be NULL): one entry
 * w*M + c per event, in event order. */
int msamd_tcn_labels(const int64_t* times, const int64_t* offsets, int32_t M, int64_t N,
                     uint8_t* labels, int64_t* positives, void* stream);

/* Workspace bytes for msamd_tcn_grad / msamd_tcn_train (-1 if invalid);
 * seg_windows (windows per training segment) is 1..64. */
int64_t msamd_tcn_train_workspace_size(int32_t batch, int32_t seg_windows);

/* Mean weighted binary cross-entropy over the windows of `batch` segments
 * (segment b: channel seg_channel[b], windows seg_window[b] ..
 * + seg_windows - 1; windows outside the recording have label 0) and its
 * gradient with respect to every parameter (grad: msamd_tcn_num_params()
 * floats; loss: one float).  Positive windows weigh pos_weight. */
int msamd_tcn_grad(const float* X, int64_t M, int64_t N, const uint8_t* labels,
                   const float* params, const int32_t* seg_channel, const int64_t* seg_window,
                   int32_t batch, int32_t seg_windows, double pos_weight, float* grad,
                   float* loss, void* workspace, int64_t workspace_bytes, void* stream);

/* `steps` Adam steps (lr, beta 0.9 / 0.999, eps 1e-8) on params (fp32
 * master copy) from adam_state (2 * num_params floats, zero before the first
 * step; step0 = steps already taken).  Each step samples `batch` segments
 * (half at random windows, half around a random entry of positives) with a
 * counter-based generator seeded by seed.  loss_trace (may be NULL): the
 * loss of every step. */
int msamd_tcn_train(const float* X, int64_t M, int64_t N, const uint8_t* labels,
                    const int64_t* positives, int64_t num_positives, float* params,
                    float* adam_state, int32_t steps, int32_t step0, int32_t batch,
                    int32_t seg_windows, double lr, double pos_weight, uint64_t seed,
                    float* loss_trace, void* workspace, int64_t workspace_bytes, void* stream);

/* Inference logits of every window of every channel (logits: [nW][M]),
 * computed by the bf16 MFMA inference kernel. */
int64_t msamd_tcn_score_workspace_size(void);
int msamd_tcn_score(const float* X, int64_t M, int64_t N, const float* params, float* logits,
                    void* workspace, int64_t workspace_bytes, void* stream);

/* Workspace bytes for msamd_tcn_detect. */
int64_t msamd_tcn_detect_workspace_size(int64_t N_file, int64_t n_range, int64_t chunk_size,
                                        int32_t M, double detect_interval);

/* Learned detection on every channel 0..M-1.  Statistics and chunk plan are
 * those of msamd_detect_events.  Each window is classified by its score
 * s = sigmoid(z): s < score_lo: skipped (no event); score_lo <= s <= score_hi:
 * uncertain; s > score_hi: present.  In uncertain and present windows the
 * plain rule decides every sample, so every reported event is an event of
 * plain detection and only skipped windows can lose events.
 * Fetch the result with msamd_detect_fetch(workspace, M, ...).
 * class_counts_host (may be NULL): windows per class {skipped, uncertain,
 * present} over the detection range. */
int msamd_tcn_detect(const float* X, int64_t M, int64_t N_file, int64_t t1, int64_t t2,
                     int64_t chunk_size, int64_t padding, const float* params,
                     double detect_threshold, double detect_interval, int detect_sign,
                     double score_lo, double score_hi, void* workspace, int64_t workspace_bytes,
                     int64_t* num_events_host, int64_t* class_counts_host, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MSAMD_H */
