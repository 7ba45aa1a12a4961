/*
 * CPU oracle for the MountainSort front end: a plain-C restatement of the
 * reference algorithms with the reference's float/double semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP path is
 * compared against.  Only tests/, the smoke check and bench.py's CPU-baseline
 * leg may load it; the product path (mountainsort_amd) never does.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the MountainSort repository root).  Rounding notes:
 *  - float*float products inside std::inner_product / "+=" on a double
 *    accumulator are rounded to float before the double add (C++ usual
 *    arithmetic conversions on x86-64 SSE, FLT_EVAL_METHOD == 0).  The helper
 *    fprod() makes that rounding explicit.
 *  - the reference is built without FMA contraction; this file is compiled
 *    with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

static inline double fprod(float a, float b)
{
    volatile float p = a * b; /* rounded to float, as in the reference */
    return (double)p;
}

/* ------------------------------------------------------------------------ */
/* a1: bandpass gain, packages/ms3/p_bandpass_filter.cpp:421-447            */
/* ------------------------------------------------------------------------ */
EXPORT void orc_define_kernel(int64_t N, double* kernel, double samplefreq,
    double freq_min, double freq_max, double freq_wid)
{
    double T = N / samplefreq; /* :425 */
    double df = 1 / T; /* :426 */
    double relwid = 3.0; /* :427 */
    for (int64_t i = 0; i < N; i++) {
        double fgrid = (i <= (N + 1) / 2) ? df * i : df * (i - N); /* :433 */
        double absf = fabs(fgrid);
        double val = 1.0;
        if (freq_min != 0) { /* :436-441 */
            if (i == 0)
                val = 0.0;
            else
                val *= (1 + erf(relwid * (absf - freq_min) / freq_min)) / 2;
        }
        if (freq_max != 0) /* :442-444 */
            val *= (1 - erf((absf - freq_max) / freq_wid)) / 2;
        kernel[i] = sqrt(val); /* :445 */
    }
}

/* a3: bandpass chunk plan, p_bandpass_filter.cpp:147-155 */
EXPORT int64_t orc_bandpass_chunk_size(int64_t M, int64_t num_threads)
{
    int overhead = 6;
    int64_t overlap_size = 60000;
    int64_t min_chunk_size = overlap_size * 2;
    double target_ram_bytes = 1.0 * (1024 * 1024 * 1024);
    double target_chunk_size = target_ram_bytes / (overhead * sizeof(float) * M * num_threads) - 2 * overlap_size;
    int64_t chunk_size = (int64_t)(target_chunk_size);
    if (chunk_size < min_chunk_size)
        chunk_size = min_chunk_size;
    return chunk_size;
}

/* ------------------------------------------------------------------------ */
/* a8: detection chunk plan, packages/mountainsortalg/p_mountainsort3.cpp    */
/* ------------------------------------------------------------------------ */
/* :170-172 (num simultaneous neighborhoods, batches) and :507-533 */
EXPORT int64_t orc_detect_chunk_size(int64_t M, int64_t N, int64_t num_threads)
{
    int64_t nsim = num_threads < M ? num_threads : M; /* :170 */
    if (nsim < 1)
        nsim = 1;
    int64_t num_batches = (M + nsim - 1) / nsim; /* get_neighborhood_batches :47-57 */
    int64_t ram = (int64_t)1e9; /* :509 */
    int64_t chunk_size = ram / (M * num_batches * 4); /* :510 */
    if (chunk_size > N)
        chunk_size = N;
    if (N < chunk_size * 1) /* min_num_chunks = 1, :191 */
        chunk_size = N / 1;
    if (chunk_size < 20000)
        chunk_size = 20000;
    return chunk_size;
}

/* ------------------------------------------------------------------------ */
/* MLCompute::mean / stdev on QVector<double>,                              */
/* packages/mountainsortalg/common/mlcompute.cpp:62-85                     */
/* ------------------------------------------------------------------------ */
EXPORT void orc_mean_stdev(const double* X, int64_t n, double* mean_out, double* stdev_out)
{
    double s = 0, ss = 0;
    for (int64_t i = 0; i < n; i++)
        s = s + X[i]; /* std::accumulate :64 */
    for (int64_t i = 0; i < n; i++)
        ss = ss + X[i] * X[i]; /* std::inner_product :77 */
    *mean_out = n ? s / (double)n : 0; /* :67-73 (count is int) */
    int ct = (int)n;
    *stdev_out = (ct >= 2) ? sqrt((ss - s * s / ct) / (ct - 1)) : 0; /* :80-84 */
}

/* ------------------------------------------------------------------------ */
/* a10: detect_events, packages/mountainsortalg/detect_events.cpp:69-135.   */
/* The literal candidate / last-best state machine.  Returns the number of  */
/* events written to times_out (indices into X, ascending).                 */
/* ------------------------------------------------------------------------ */
EXPORT int64_t orc_detect_events(const double* X, int64_t n, double detect_threshold,
    double detect_interval, int sign, int64_t* times_out)
{
    double mean, stdev;
    orc_mean_stdev(X, n, &mean, &stdev); /* :71-72 */
    double threshold2 = detect_threshold * stdev; /* :73 */
    int64_t* tp = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    double* vals = (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
    char* to_use = (char*)malloc(n > 0 ? n : 1);
    int64_t nc = 0;
    for (int64_t i = 0; i < n; i++) { /* :78-89 */
        double val = X[i] - mean;
        if (sign < 0)
            val = -val;
        else if (sign == 0)
            val = fabs(val);
        if (val > threshold2) {
            tp[nc] = i;
            vals[nc] = val;
            nc++;
        }
    }
    for (int64_t i = 0; i < nc; i++)
        to_use[i] = 0;
    int64_t last_best_ind = -1;
    for (int64_t i = 0; i < nc; i++) { /* :92-126 */
        if (last_best_ind >= 0) {
            if ((double)tp[last_best_ind] < (double)tp[i] - detect_interval) {
                last_best_ind = -1;
                if ((i > 0) && ((double)tp[i - 1] >= (double)tp[i] - detect_interval)) {
                    last_best_ind = i - 1;
                    int64_t jj = last_best_ind;
                    while ((jj - 1 >= 0) && ((double)tp[jj - 1] >= (double)tp[i] - detect_interval)) {
                        if (vals[jj - 1] > vals[last_best_ind])
                            last_best_ind = jj - 1;
                        jj--;
                    }
                } else {
                    last_best_ind = -1;
                }
            }
        }
        if (last_best_ind >= 0) {
            if (vals[i] > vals[last_best_ind]) {
                to_use[i] = 1;
                to_use[last_best_ind] = 0;
                last_best_ind = i;
            } else {
                to_use[i] = 0;
            }
        } else {
            to_use[i] = 1;
            last_best_ind = i;
        }
    }
    int64_t ne = 0;
    for (int64_t i = 0; i < nc; i++) /* :128-133 */
        if (to_use[i])
            times_out[ne++] = tp[i];
    free(tp);
    free(vals);
    free(to_use);
    return ne;
}

/*
 * Phase 1 of mountainsortalg.ms3alg restricted to detection
 * (p_mountainsort3.cpp:145-216 with NeighborhoodSorter::addTimeChunk,
 * neighborhoodsorter.cpp:74-106): for every time chunk (padding 1000, read
 * zero-filled outside [0, N_file)), for every central channel, detect on the
 * padded chunk and keep the core events.  Times are returned as absolute
 * 0-based sample indices (t_start + chunk-relative time).
 * Output is per channel: counts[c] events for channels[c] (0-based), times in
 * ascending order, concatenated in channel order.  Returns the total, or -1
 * when cap is exceeded.
 */
EXPORT int64_t orc_ms3alg_detect(const float* X, int64_t M, int64_t N_file, int64_t t1, int64_t t2,
    int64_t chunk_size, int64_t padding, const int32_t* channels, int32_t num_channels,
    double detect_threshold, double detect_interval, int sign,
    int64_t* counts, int64_t* times_out, int64_t cap)
{
    int64_t t_start = t1 < 0 ? 0 : t1; /* :149-155 */
    int64_t t_end = t2 < 0 ? N_file - 1 : t2;
    int64_t N = t_end - t_start + 1;
    int64_t maxlen = chunk_size + 2 * padding;
    double* X0 = (double*)malloc(sizeof(double) * maxlen);
    int64_t* ev = (int64_t*)malloc(sizeof(int64_t) * maxlen);
    /* per channel we need time order over chunks; collect per channel */
    int64_t total = 0;
    for (int32_t c = 0; c < num_channels; c++)
        counts[c] = 0;
    /* two passes: count, then fill in channel-major order */
    int64_t* chan_off = (int64_t*)calloc(num_channels + 1, sizeof(int64_t));
    for (int pass = 0; pass < 2; pass++) {
        int64_t* fill = (int64_t*)calloc(num_channels, sizeof(int64_t));
        for (int64_t t = 0; t < N; t += chunk_size) { /* get_time_chunk_infos :516-528 */
            int64_t size = chunk_size;
            if (t + size > N)
                size = N - t;
            int64_t len = size + 2 * padding;
            int64_t base = t_start + t - padding; /* :195 */
            for (int32_t c = 0; c < num_channels; c++) {
                int64_t ch = channels[c];
                for (int64_t i = 0; i < len; i++) { /* :83-87, zero fill diskreadmda32.cpp:631-650 */
                    int64_t tt = base + i;
                    X0[i] = (tt >= 0 && tt < N_file) ? (double)X[ch + M * tt] : 0.0;
                }
                int64_t ne = orc_detect_events(X0, len, detect_threshold, detect_interval, sign, ev);
                for (int64_t k = 0; k < ne; k++) {
                    int64_t t0 = ev[k];
                    if ((0 <= t0 - padding) && (t0 - padding < len - padding - padding)) { /* :94 */
                        if (pass == 0) {
                            counts[c]++;
                        } else {
                            times_out[chan_off[c] + fill[c]] = t0 - padding + t + t_start; /* :105, :458 */
                            fill[c]++;
                        }
                    }
                }
            }
        }
        free(fill);
        if (pass == 0) {
            total = 0;
            for (int32_t c = 0; c < num_channels; c++) {
                chan_off[c] = total;
                total += counts[c];
            }
            chan_off[num_channels] = total;
            if (total > cap) {
                free(chan_off);
                free(X0);
                free(ev);
                return -1;
            }
        }
    }
    free(chan_off);
    free(X0);
    free(ev);
    return total;
}

/* ------------------------------------------------------------------------ */
/* a4: whitening covariance, packages/ms3/p_whiten.cpp:64-118.              */
/* Chunk partials (1e5 samples) are merged in ascending chunk order (the     */
/* reference merges in OpenMP completion order, :97-110).                   */
/* ------------------------------------------------------------------------ */
EXPORT void orc_whiten_covariance(const float* X, int64_t M, int64_t N, double* XXt)
{
    int64_t chunk_size = 100000; /* :53 */
    if (N < chunk_size)
        chunk_size = N;
    double* part = (double*)malloc(sizeof(double) * M * M);
    memset(XXt, 0, sizeof(double) * M * M);
    for (int64_t tp = 0; tp < N; tp += chunk_size) {
        int64_t n = chunk_size < N - tp ? chunk_size : N - tp;
        memset(part, 0, sizeof(double) * M * M);
        for (int64_t i = 0; i < n; i++) { /* :87-96 */
            const float* x = X + M * (tp + i);
            int64_t bb = 0;
            for (int64_t m1 = 0; m1 < M; m1++)
                for (int64_t m2 = 0; m2 < M; m2++)
                    part[bb++] += fprod(x[m1], x[m2]);
        }
        for (int64_t bb = 0; bb < M * M; bb++) /* :99-105 */
            XXt[bb] += part[bb];
    }
    if (N > 1) /* :114-118 */
        for (int64_t ii = 0; ii < M * M; ii++)
            XXt[ii] /= (N - 1);
    free(part);
}

/* ------------------------------------------------------------------------ */
/* a5: whitening matrix, packages/mountainsortalg/pca.cpp (Mda, double      */
/* forms): whitening_matrix_from_XXt :634-650, pca_from_XXt :171-193,       */
/* iterate_XXt_to_get_top_component :541-557, subtract_out_rank_1_from_XXt */
/* :408-430, normalize_vector :456-465, matvec :478-490, mult_AB :219-242,  */
/* mult_ABtrans :318-341.  All matrices column-major (Mda order).          */
/* ------------------------------------------------------------------------ */
static double norm_d(int64_t n, const double* x) /* mlcompute.cpp:131-139 */
{
    double s = 0;
    for (int64_t i = 0; i < n; i++)
        s = s + x[i] * x[i];
    return sqrt(s);
}

static void normalize_d(int64_t n, double* v) /* pca.cpp:456-465 */
{
    double nrm = norm_d(n, v);
    if (!nrm)
        return;
    for (int64_t i = 0; i < n; i++)
        v[i] /= nrm;
}

static void matvec_d(int64_t M, int64_t N, double* ret, const double* A, const double* x) /* :478-490 */
{
    for (int64_t m = 0; m < M; m++)
        ret[m] = 0;
    int64_t iA = 0;
    for (int64_t n = 0; n < N; n++) {
        double xval = x[n];
        for (int64_t m = 0; m < M; m++) {
            ret[m] += A[iA] * xval;
            iA++;
        }
    }
}

static void mult_AB_d(int64_t M, int64_t L, int64_t N, const double* A, const double* B, double* C) /* :219-242 */
{
    int64_t iC = 0;
    for (int64_t n = 0; n < N; n++) {
        for (int64_t m = 0; m < M; m++) {
            double acc = 0;
            for (int64_t l = 0; l < L; l++)
                acc += A[m + l * M] * B[l + L * n];
            C[iC++] = acc;
        }
    }
}

EXPORT void orc_pca_from_XXt(const double* XXt, int64_t M, int64_t K, double* C, double* sigma)
{
    double* G = (double*)malloc(sizeof(double) * M * M);
    double* B = (double*)malloc(sizeof(double) * M * M);
    double* tmp = (double*)malloc(sizeof(double) * M * M);
    double* v = (double*)malloc(sizeof(double) * M);
    double* w = (double*)malloc(sizeof(double) * M);
    memcpy(G, XXt, sizeof(double) * M * M);
    for (int64_t k = 0; k < K; k++) { /* :182-192 */
        double sig = 0;
        for (int64_t i = 0; i < M; i++) /* :545-547 */
            v[i] = sin((double)(i + 1));
        normalize_d(M, v);
        for (int it = 0; it < 10; it++) { /* :549-556 */
            matvec_d(M, M, w, G, v);
            memcpy(v, w, sizeof(double) * M);
            sig = norm_d(M, v);
            normalize_d(M, v);
        }
        for (int64_t m = 0; m < M; m++)
            C[m + M * k] = v[m];
        sigma[k] = sig;
        /* deflate: G <- B G B with B = I - v v' (:419-429) */
        for (int64_t j = 0; j < M; j++)
            for (int64_t i = 0; i < M; i++)
                B[i + M * j] = (i == j) ? 1 - v[i] * v[j] : -v[i] * v[j];
        mult_AB_d(M, M, M, G, B, tmp);
        mult_AB_d(M, M, M, B, tmp, G);
    }
    free(G);
    free(B);
    free(tmp);
    free(v);
    free(w);
}

EXPORT void orc_whitening_matrix(const double* XXt, int64_t M, double* W)
{
    double* C = (double*)malloc(sizeof(double) * M * M);
    double* sigma = (double*)malloc(sizeof(double) * M);
    double* CD = (double*)malloc(sizeof(double) * M * M);
    orc_pca_from_XXt(XXt, M, M, C, sigma); /* :638 */
    for (int64_t j = 0; j < M; j++) { /* :640-648, tmp = C * D */
        double val = 0;
        if (sigma[j])
            val = 1 / sqrt(sigma[j]);
        for (int64_t i = 0; i < M; i++) {
            /* mult_AB with diagonal D: sum over l of C[i,l] D[l,j]; only l == j is nonzero,
               but the reference sums every term; zeros do not change a double sum */
            CD[i + M * j] = C[i + M * j] * val;
        }
    }
    /* W = CD * C' (mult_ABtrans :318-341) */
    for (int64_t n = 0; n < M; n++) {
        for (int64_t m = 0; m < M; m++) {
            double acc = 0;
            for (int64_t l = 0; l < M; l++)
                acc += CD[m + l * M] * C[n + l * M];
            W[m + M * n] = acc;
        }
    }
    free(C);
    free(sigma);
    free(CD);
}

/* ------------------------------------------------------------------------ */
/* a6: whitening apply + quantize, packages/ms3/p_whiten.cpp:153-206         */
/* (also apply_whitening_matrix :357-428).                                 */
/* out_f32 receives the quantized float output; when quantization_unit > 0 */
/* out_i16 (may be NULL otherwise) receives the int16 file values.         */
/* ------------------------------------------------------------------------ */
static inline double quantize_d(float X, double unit) /* :31-34 */
{
    return (floor(X / unit + 0.5)) * unit;
}

EXPORT void orc_whiten_apply(const float* X, int64_t M, int64_t N, const double* W,
    float* out_f32, double quantization_unit, int16_t* out_i16)
{
    for (int64_t i = 0; i < N; i++) { /* :174-183 */
        const float* x = X + M * i;
        float* y = out_f32 + M * i;
        int64_t bb = 0;
        for (int64_t m1 = 0; m1 < M; m1++) {
            float acc = 0;
            for (int64_t m2 = 0; m2 < M; m2++) {
                acc = (float)((double)acc + (double)x[m2] * W[bb]);
                bb++;
            }
            y[m1] = acc;
        }
        for (int64_t m = 0; m < M; m++) { /* :188 */
            y[m] = (float)quantize_d(y[m], 0.0001);
            if (quantization_unit > 0 && out_i16) { /* :189-191, :445-451 */
                float s = (float)(int)((y[m] / quantization_unit) + 0.5);
                out_i16[M * i + m] = (int16_t)s;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* f1: p_whiten_clips, packages/ms3/p_whiten.cpp:295-345.  Per clip and time */
/* sample the same float accumulation over m2 as pass 2 (:326-334), with no */
/* 1e-4 quantize (it is commented out there, :347-352); with              */
/* quantization_unit > 0, scale_for_quantization (:445-451) gives the int16 */
/* file values.  Clips: M x T x L, i.e. an M x (T*L) array.                 */
/* ------------------------------------------------------------------------ */
EXPORT void orc_whiten_clips(const float* X, int64_t M, int64_t TL, const double* W,
    float* out_f32, double quantization_unit, int16_t* out_i16)
{
    for (int64_t i = 0; i < TL; i++) {
        const float* x = X + M * i;
        float* y = out_f32 + M * i;
        int64_t aa = 0;
        for (int64_t m1 = 0; m1 < M; m1++) {
            float acc = 0;
            for (int64_t m2 = 0; m2 < M; m2++) {
                acc = (float)((double)acc + (double)x[m2] * W[aa]);
                aa++;
            }
            y[m1] = acc;
            if (quantization_unit > 0 && out_i16) {
                float s = (float)(int)((y[m1] / quantization_unit) + 0.5);
                out_i16[M * i + m1] = (int16_t)s;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* a7: mask_out_artifacts, packages/ms3/p_mask_out_artifacts.cpp:24-104.     */
/* norms_out (M x N/interval, column-major) and use_it_out (N/interval + 1) */
/* may be NULL.  Returns 0, or -1 on the parameter check of :26-29.         */
/* ------------------------------------------------------------------------ */
EXPORT int orc_mask_out_artifacts(const float* X, int64_t M, int64_t N, double threshold,
    int64_t interval_size, float* Y, double* norms_out, int32_t* use_it_out)
{
    if ((!threshold) || (!interval_size))
        return -1;
    int64_t nint = N / interval_size;
    double* norms = (double*)malloc(sizeof(double) * (M * nint > 0 ? M * nint : 1));
    for (int64_t i = 0; i < nint; i++) { /* :39-56 */
        int64_t tp = i * interval_size;
        for (int64_t m = 0; m < M; m++) {
            double sumsqr = 0;
            for (int64_t aa = 0; aa < interval_size; aa++) {
                double v = (double)X[m + M * (tp + aa)];
                sumsqr += v * v;
            }
            norms[m + M * i] = sqrt(sumsqr);
        }
    }
    int32_t* use_it = (int32_t*)malloc(sizeof(int32_t) * (nint + 1)); /* :59-61 */
    for (int64_t i = 0; i < nint + 1; i++)
        use_it[i] = 1;
    double* vals = (double*)malloc(sizeof(double) * (nint > 0 ? nint : 1));
    for (int64_t m = 0; m < M; m++) { /* :62-76 */
        for (int64_t i = 0; i < nint; i++)
            vals[i] = norms[m + M * i];
        double mean0, sigma0;
        orc_mean_stdev(vals, nint, &mean0, &sigma0);
        for (int64_t i = 0; i < nint; i++) {
            if (norms[m + M * i] > mean0 + sigma0 * threshold) {
                if (i - 1 >= 0) /* i == 0 writes use_it[-1] in the reference (undefined); no-op here */
                    use_it[i - 1] = 0;
                use_it[i] = 0;
                use_it[i + 1] = 0;
            }
        }
    }
    memset(Y, 0, sizeof(float) * M * N); /* pre-zeroed output, diskwritemda.cpp:114-116 */
    for (int64_t i = 0; i < nint; i++) { /* :80-99 */
        if (use_it[i])
            memcpy(Y + M * i * interval_size, X + M * i * interval_size, sizeof(float) * M * interval_size);
    }
    if (norms_out)
        memcpy(norms_out, norms, sizeof(double) * M * nint);
    if (use_it_out)
        memcpy(use_it_out, use_it, sizeof(int32_t) * (nint + 1));
    free(norms);
    free(use_it);
    free(vals);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* a13/a14: PCA on Mda32 data, packages/mountainsortalg/pca.cpp.            */
/* pca :68-97, iterate_to_get_top_component :523-539, mult_AtransB          */
/* :293-316, matvec :492-503, subtract_out_rank_1 :388-406,                 */
/* normalize_vector :467-476, MLCompute::dotProduct(float) mlcompute.cpp:141 */
/* X is M x N column-major; C is M x K, F is K x N, sigma is K (float).     */
/* ------------------------------------------------------------------------ */
static double dot_f(int64_t n, const float* a, const float* b) /* mlcompute.cpp:141-145 */
{
    double s = 0.0;
    for (int64_t i = 0; i < n; i++)
        s = s + fprod(a[i], b[i]);
    return s;
}

static void normalize_f(int64_t n, float* v) /* pca.cpp:467-476 */
{
    double nrm = sqrt(dot_f(n, v, v));
    if (!nrm)
        return;
    for (int64_t i = 0; i < n; i++)
        v[i] = (float)(v[i] / nrm);
}

static void top_component_f(int64_t M, int64_t N, const float* Xw, float* C, double* sigma, float* tmp)
{
    for (int64_t i = 0; i < M; i++) /* :528-530 */
        C[i] = (float)sin((double)(i + 1));
    normalize_f(M, C);
    for (int it = 0; it < 10; it++) {
        for (int64_t n = 0; n < N; n++) /* mult_AtransB(C, X) :293-316 */
            tmp[n] = (float)dot_f(M, C, Xw + M * n);
        for (int64_t m = 0; m < M; m++) /* matvec :492-503 */
            C[m] = 0;
        int64_t iA = 0;
        for (int64_t n = 0; n < N; n++) {
            double xval = tmp[n];
            for (int64_t m = 0; m < M; m++) {
                C[m] = (float)((double)C[m] + (double)Xw[iA] * xval);
                iA++;
            }
        }
        *sigma = sqrt(dot_f(M, C, C)); /* :536 */
        normalize_f(M, C);
    }
}

EXPORT void orc_pca(const float* X, int64_t M, int64_t N, int64_t K, float* C, float* F, float* sigma)
{
    float* Xw = (float*)malloc(sizeof(float) * (M * N > 0 ? M * N : 1));
    float* C0 = (float*)malloc(sizeof(float) * M);
    float* tmp = (float*)malloc(sizeof(float) * (N > 0 ? N : 1));
    memcpy(Xw, X, sizeof(float) * M * N);
    for (int64_t k = 0; k < K; k++) { /* :83-94 */
        double sig = 0;
        top_component_f(M, N, Xw, C0, &sig, tmp);
        for (int64_t m = 0; m < M; m++)
            C[m + M * k] = C0[m];
        sigma[k] = (float)sig;
        for (int64_t n = 0; n < N; n++) { /* subtract_out_rank_1 :388-406 */
            float* col = Xw + M * n;
            double dp = dot_f(M, col, C0);
            for (int64_t m = 0; m < M; m++)
                col[m] = (float)((double)col[m] - dp * (double)C0[m]);
        }
    }
    if (F) {
        for (int64_t n = 0; n < N; n++) /* F = mult_AtransB(C, X) :96 */
            for (int64_t k = 0; k < K; k++)
                F[k + K * n] = (float)dot_f(M, C + M * k, X + M * n);
    }
    free(Xw);
    free(C0);
    free(tmp);
}

/* pca_subsampled, pca.cpp:99-127.  Returns the number of subsampled columns. */
EXPORT int64_t orc_pca_subsample_indices(int64_t N, int64_t max_samples, int64_t* idx_out)
{
    if (N <= max_samples) {
        for (int64_t i = 0; i < N; i++)
            idx_out[i] = i;
        return N;
    }
    double increment = N * 1.0 / max_samples; /* :110 */
    int64_t c = 0;
    for (double i = 0; i < N; i += increment) /* :112-114 */
        idx_out[c++] = (int64_t)i;
    return c;
}

EXPORT void orc_pca_subsampled(const float* X, int64_t M, int64_t N, int64_t K, int64_t max_samples,
    float* C, float* F, float* sigma)
{
    if (N <= max_samples) { /* :106-109 */
        orc_pca(X, M, N, K, C, F, sigma);
        return;
    }
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * (max_samples + 2));
    int64_t N2 = orc_pca_subsample_indices(N, max_samples, idx);
    float* X2 = (float*)malloc(sizeof(float) * M * N2);
    for (int64_t j = 0; j < N2; j++) /* :116-121 */
        for (int64_t m = 0; m < M; m++)
            X2[m + M * j] = X[m + M * idx[j]];
    orc_pca(X2, M, N2, K, C, NULL, sigma); /* :123 */
    for (int64_t n = 0; n < N; n++) /* :126 */
        for (int64_t k = 0; k < K; k++)
            F[k + K * n] = (float)dot_f(M, C + M * k, X + M * n);
    free(idx);
    free(X2);
}

/*
 * a12: dimension_reduce_clips, packages/mountainsortalg/neighborhoodsorter.cpp:236-276.
 * clips W x T x L (column-major); out W x nf x L.
 */
EXPORT void orc_dimension_reduce_clips(const float* clips, int64_t W, int64_t T, int64_t L,
    int64_t nf, int64_t max_samples, float* out)
{
    float* A = (float*)malloc(sizeof(float) * (T * L > 0 ? T * L : 1));
    float* CC = (float*)malloc(sizeof(float) * T * nf);
    float* FF = (float*)malloc(sizeof(float) * (nf * L > 0 ? nf * L : 1));
    float* sg = (float*)malloc(sizeof(float) * nf);
    for (int64_t m = 0; m < W; m++) {
        for (int64_t i = 0; i < L; i++) /* :247-260 */
            for (int64_t t = 0; t < T; t++)
                A[t + T * i] = clips[m + W * t + W * T * i];
        orc_pca_subsampled(A, T, L, nf, max_samples, CC, FF, sg); /* :262 */
        for (int64_t i = 0; i < L; i++) /* :263-274 */
            for (int64_t a = 0; a < nf; a++)
                out[m + W * a + W * nf * i] = FF[a + nf * i];
    }
    free(A);
    free(CC);
    free(FF);
    free(sg);
}

/* ------------------------------------------------------------------------
 * Templates and the fit stage (test infrastructure: the oracle of
 * msamd_compute_templates / msamd_fit_stage).
 * ------------------------------------------------------------------------ */

/*
 * ms3.compute_templates, packages/ms3/p_compute_templates.cpp:24-118.
 * X M x N (column-major), events (times as doubles, labels 1..K; others are
 * skipped).  Clip start t1 = (int64)(time - Tmid) (:88), samples outside
 * [0, N) read as zero (DiskReadMda32::readChunk zero-fills,
 * diskreadmda32.cpp:631-650).  Double sums in event order (:95-99), divided
 * by the count when nonzero (:101-108), stored as float (write32).
 * templates M x T x K; counts K.
 */
EXPORT void orc_compute_templates(const float* X, int64_t M, int64_t N, const double* times,
    const int32_t* labels, int64_t L, int32_t K, int32_t T, float* templates, double* counts)
{
    const int64_t Tmid = (T + 1) / 2 - 1;
    double* sum = (double*)calloc((size_t)(M * T * (K > 0 ? K : 1)), sizeof(double));
    for (int32_t k = 0; k < K; k++)
        counts[k] = 0;
    for (int64_t i = 0; i < L; i++) {
        const int32_t k = labels[i] - 1;
        if (k < 0 || k >= K)
            continue;
        const int64_t t1 = (int64_t)(times[i] - (double)Tmid);
        double* s = sum + M * T * k;
        for (int64_t t = 0; t < T; t++) {
            const int64_t tt = t1 + t;
            for (int64_t m = 0; m < M; m++) {
                const float v = (tt >= 0 && tt < N) ? X[m + M * tt] : 0.0f;
                s[m + M * t] = s[m + M * t] + (double)v;
            }
        }
        counts[k] = counts[k] + 1.0;
    }
    for (int32_t k = 0; k < K; k++)
        for (int64_t e = 0; e < M * T; e++) {
            double v = sum[M * T * k + e];
            if (counts[k] != 0)
                v = v / counts[k];
            templates[M * T * k + e] = (float)v;
        }
    free(sum);
}

/* fit_stage.cpp:176-192: indices m + M*t of |template| > thresh, ascending. */
static int64_t tchmask(const float* tmpl, int64_t M, int64_t T, double thresh, int64_t* out)
{
    int64_t n = 0;
    for (int64_t t = 0; t < T; t++)
        for (int64_t m = 0; m < M; m++) {
            const double v = fabs((double)tmpl[m + M * t]);
            if (v > thresh)
                out[n++] = m + M * t;
        }
    return n;
}

/*
 * fit_stage, packages/mountainsortalg/fit_stage.cpp:27-174 (helpers :194-275),
 * on one (padded) chunk X (M x N2, modified in place as the reference does).
 * times: chunk-local times (ascending), labels 1..K.  keep[i] = 1 for the
 * events the fit uses.  Returns the number of passes.
 */
EXPORT int64_t orc_fit_stage(float* X, int64_t M, int64_t N2, const double* times,
    const int32_t* labels, int64_t L, const float* templates, int32_t K, int32_t T,
    double thresh, uint8_t* keep)
{
    const int64_t Tmid = (T + 1) / 2 - 1;
    int64_t* mask = (int64_t*)malloc(sizeof(int64_t) * (size_t)(M * T * (K > 0 ? K : 1)));
    int64_t* msz = (int64_t*)malloc(sizeof(int64_t) * (size_t)(K > 0 ? K : 1));
    for (int32_t k = 0; k < K; k++)
        msz[k] = tchmask(templates + M * T * k, M, T, thresh, mask + M * T * k);
    double* scores = (double*)calloc((size_t)(L > 0 ? L : 1), sizeof(double));
    float* dirty = (float*)malloc(sizeof(float) * (size_t)(M * N2 > 0 ? M * N2 : 1));
    for (int64_t i = 0; i < M * N2; i++)
        dirty[i] = 1;
    int64_t* consider = (int64_t*)malloc(sizeof(int64_t) * (size_t)(L > 0 ? L : 1));
    int64_t* tri = (int64_t*)malloc(sizeof(int64_t) * (size_t)(L > 0 ? L : 1));
    double* trs = (double*)malloc(sizeof(double) * (size_t)(L > 0 ? L : 1));
    double* trt = (double*)malloc(sizeof(double) * (size_t)(L > 0 ? L : 1));
    int32_t* trl = (int32_t*)malloc(sizeof(int32_t) * (size_t)(L > 0 ? L : 1));
    uint8_t* use = (uint8_t*)malloc((size_t)(L > 0 ? L : 1));
    int64_t nc = L;
    for (int64_t i = 0; i < L; i++) {
        consider[i] = i;
        keep[i] = 0;
    }
    int changed = 1;
    int64_t passes = 0;
    while (changed) { /* :74 */
        passes++;
        int64_t nt = 0;
        for (int64_t kk = 0; kk < nc; kk++) { /* :86-133 */
            const int64_t i = consider[kk];
            const double t0 = times[i];
            const int32_t k0 = labels[i];
            if (k0 > 0 && k0 <= K) {
                const int64_t tt = (int64_t)(t0 - (double)Tmid + 0.5); /* :92 */
                double score0 = 0;
                if (tt >= 0 && tt + T <= N2) {
                    const int64_t* mk = mask + M * T * (k0 - 1);
                    const int64_t n = msz[k0 - 1];
                    int dz = 0; /* is_dirty, :194-201 */
                    for (int64_t j = 0; j < n; j++)
                        if (dirty[M * tt + mk[j]] != 0) {
                            dz = 1;
                            break;
                        }
                    if (!dz) {
                        score0 = scores[i];
                    } else { /* compute_score, :203-219 */
                        const float* xp = X + M * tt;
                        const float* tp = templates + M * T * (k0 - 1);
                        double before = 0, after = 0;
                        for (int64_t j = 0; j < n; j++) {
                            double val = xp[mk[j]];
                            before += val * val;
                            val -= tp[mk[j]];
                            after += val * val;
                        }
                        score0 = before - after;
                        scores[i] = score0;
                    }
                }
                if (score0 > 0) { /* neglogprior = 0, :120-125 */
                    trs[nt] = score0;
                    trt[nt] = t0;
                    trl[nt] = k0;
                    tri[nt] = i;
                    nt++;
                }
            }
        }
        /* find_events_to_use, :221-252 */
        for (int64_t i = 0; i < nt; i++) {
            use[i] = 0;
            if (trs[i] > 0) {
                use[i] = 1;
                for (int64_t j = i; j >= 0 && trt[j] >= trt[i] - T; j--)
                    if (i != j && trs[j] >= trs[i])
                        use[i] = 0;
                for (int64_t j = i; j < nt && trt[j] <= trt[i] + T; j++)
                    if (trs[j] > trs[i])
                        use[i] = 0;
            }
        }
        for (int64_t i = 0; i < M * N2; i++) /* :138-141 */
            dirty[i] = 0;
        changed = 0;
        int64_t nn = 0;
        for (int64_t a = 0; a < nt; a++) { /* :148-167 */
            if (use[a]) {
                const int32_t k0 = trl[a];
                const int64_t* mk = mask + M * T * (k0 - 1);
                const int64_t n = msz[k0 - 1];
                const int64_t tt = (int64_t)(trt[a] - (double)Tmid + 0.5);
                float* xp = X + M * tt;
                float* dp = dirty + M * tt;
                const float* tp = templates + M * T * (k0 - 1);
                /* subtract_scaled_template with scale_min = scale_max = 1, :254-275:
                 * alpha is clamped to 1 whatever S12/S22 are */
                for (int64_t j = 0; j < n; j++) {
                    xp[mk[j]] -= 1.0 * tp[mk[j]];
                    dp[mk[j]] = 1;
                }
                changed = 1;
                keep[tri[a]] = 1;
            } else {
                consider[nn++] = tri[a];
            }
        }
        nc = nn;
    }
    free(mask);
    free(msz);
    free(scores);
    free(dirty);
    free(consider);
    free(tri);
    free(trs);
    free(trt);
    free(trl);
    free(use);
    return passes;
}

/* ------------------------------------------------------------------------
 * OpenMP forms of the reference's threaded loops, for timing the CPU path
 * (the CPU-baseline leg of bench.py).  Per element they do the arithmetic of
 * the sequential forms above; the threads split the work as the reference's
 * OpenMP loops do, with the reference's critical sections around its reads
 * and writes (here in-memory copies in place of its .mda chunk I/O).
 * ------------------------------------------------------------------------ */

/* p_whiten.cpp:75-111: one OpenMP iteration per chunk of chunk_size samples
 * (1e5 in the reference, :53), a double M x M partial per chunk, merged in a
 * critical section in completion order.  Divided by N-1 (:114-118). */
EXPORT void orc_whiten_covariance_mt(const float* X, int64_t M, int64_t N, int64_t chunk_size,
    int nthreads, double* XXt)
{
    if (N < chunk_size)
        chunk_size = N;
    memset(XXt, 0, sizeof(double) * M * M);
    int64_t nch = chunk_size > 0 ? (N + chunk_size - 1) / chunk_size : 0;
#pragma omp parallel for num_threads(nthreads)
    for (int64_t c = 0; c < nch; c++) {
        int64_t tp = c * chunk_size;
        int64_t n = chunk_size < N - tp ? chunk_size : N - tp;
        float* chunk = (float*)malloc(sizeof(float) * M * n);
        double* part = (double*)calloc(M * M, sizeof(double));
#pragma omp critical(orc_lock1)
        memcpy(chunk, X + M * tp, sizeof(float) * M * n);
        for (int64_t i = 0; i < n; i++) { /* :87-96 */
            const float* x = chunk + M * i;
            int64_t bb = 0;
            for (int64_t m1 = 0; m1 < M; m1++)
                for (int64_t m2 = 0; m2 < M; m2++) {
                    part[bb] += x[m1] * x[m2]; /* float product, double sum */
                    bb++;
                }
        }
#pragma omp critical(orc_lock2)
        for (int64_t bb = 0; bb < M * M; bb++)
            XXt[bb] += part[bb];
        free(chunk);
        free(part);
    }
    if (N > 1)
        for (int64_t ii = 0; ii < M * M; ii++)
            XXt[ii] /= (N - 1);
}

/* p_whiten.cpp:160-200: per chunk the float accumulation into the output
 * chunk (:174-183), then the 1e-4 quantize and the write in a critical
 * section (:185-195). */
EXPORT void orc_whiten_apply_mt(const float* X, int64_t M, int64_t N, const double* W,
    int64_t chunk_size, int nthreads, float* Y)
{
    if (N < chunk_size)
        chunk_size = N;
    int64_t nch = chunk_size > 0 ? (N + chunk_size - 1) / chunk_size : 0;
#pragma omp parallel for num_threads(nthreads)
    for (int64_t c = 0; c < nch; c++) {
        int64_t tp = c * chunk_size;
        int64_t n = chunk_size < N - tp ? chunk_size : N - tp;
        float* in = (float*)malloc(sizeof(float) * M * n);
        float* out = (float*)calloc(M * n, sizeof(float));
#pragma omp critical(orc_lock1)
        memcpy(in, X + M * tp, sizeof(float) * M * n);
        for (int64_t i = 0; i < n; i++) {
            int64_t aa = M * i;
            int64_t bb = 0;
            for (int64_t m1 = 0; m1 < M; m1++)
                for (int64_t m2 = 0; m2 < M; m2++) {
                    out[aa + m1] += in[aa + m2] * W[bb];
                    bb++;
                }
        }
#pragma omp critical(orc_lock2)
        {
            for (int64_t i = 0; i < M * n; i++)
                Y[M * tp + i] = (float)quantize_d(out[i], 0.0001);
        }
        free(in);
        free(out);
    }
}

/* Clips of the ms3alg read phase, per neighborhood. */
typedef struct {
    int64_t G, T;
    int64_t* W;
    int64_t* L;
    int64_t* cap;
    float** clips; /* neighborhood g: W[g] x T x L[g] (column-major) */
} orc_clipset;

EXPORT void orc_clipset_free(orc_clipset* cs)
{
    if (!cs)
        return;
    for (int64_t g = 0; g < cs->G; g++)
        free(cs->clips[g]);
    free(cs->clips);
    free(cs->W);
    free(cs->L);
    free(cs->cap);
    free(cs);
}

EXPORT void orc_clipset_counts(const orc_clipset* cs, int64_t* L)
{
    for (int64_t g = 0; g < cs->G; g++)
        L[g] = cs->L[g];
}

EXPORT void orc_clipset_get(const orc_clipset* cs, int64_t g, float* out)
{
    memcpy(out, cs->clips[g], sizeof(float) * cs->W[g] * cs->T * cs->L[g]);
}

/*
 * The read phase of mountainsortalg.ms3alg (p_mountainsort3.cpp:170-216):
 * neighborhoods in batches of `batch` (get_neighborhood_batches :47-57); for
 * each time chunk (chunk_size, padding; zero-filled outside [0, N)), for each
 * batch, an OpenMP loop over the batch's neighborhoods running
 * NeighborhoodSorter::addTimeChunk (neighborhoodsorter.cpp:74-110): detect on
 * the central channel (nb_ch[nb_off[g]]), keep the core events, gather their
 * clips (get_clip :409-419).  X is N x M row-major (the M x N .mda memory).
 */
EXPORT orc_clipset* orc_ms3alg_clips_mt(const float* X, int64_t M, int64_t N, int64_t chunk_size,
    int64_t padding, const int32_t* nb_ch, const int64_t* nb_off, int64_t G, int64_t batch,
    double detect_threshold, double detect_interval, int sign, int64_t T, int nthreads)
{
    orc_clipset* cs = (orc_clipset*)calloc(1, sizeof(orc_clipset));
    cs->G = G;
    cs->T = T;
    cs->W = (int64_t*)calloc(G > 0 ? G : 1, sizeof(int64_t));
    cs->L = (int64_t*)calloc(G > 0 ? G : 1, sizeof(int64_t));
    cs->cap = (int64_t*)calloc(G > 0 ? G : 1, sizeof(int64_t));
    cs->clips = (float**)calloc(G > 0 ? G : 1, sizeof(float*));
    for (int64_t g = 0; g < G; g++)
        cs->W[g] = nb_off[g + 1] - nb_off[g];
    const int64_t Tmid = (T + 1) / 2 - 1;
    float* chunk = (float*)malloc(sizeof(float) * M * (chunk_size + 2 * padding));
    for (int64_t t = 0; t < N; t += chunk_size) { /* get_time_chunk_infos */
        int64_t size = chunk_size < N - t ? chunk_size : N - t;
        int64_t len = size + 2 * padding;
        for (int64_t i = 0; i < len; i++) { /* readChunk, zero fill outside [0, N) */
            int64_t tt = t - padding + i;
            if (tt >= 0 && tt < N)
                memcpy(chunk + M * i, X + M * tt, sizeof(float) * M);
            else
                memset(chunk + M * i, 0, sizeof(float) * M);
        }
        for (int64_t b0 = 0; b0 < G; b0 += batch) {
            int64_t b1 = b0 + batch < G ? b0 + batch : G;
#pragma omp parallel for num_threads(nthreads)
            for (int64_t g = b0; g < b1; g++) {
                const int64_t W = cs->W[g];
                if (W == 0)
                    continue;
                const int32_t* ch = nb_ch + nb_off[g];
                double* X0 = (double*)malloc(sizeof(double) * len);
                int64_t* ev = (int64_t*)malloc(sizeof(int64_t) * len);
                for (int64_t i = 0; i < len; i++) /* :83-87 */
                    X0[i] = chunk[ch[0] + M * i];
                int64_t ne = orc_detect_events(X0, len, detect_threshold, detect_interval, sign, ev);
                int64_t nk = 0;
                for (int64_t k = 0; k < ne; k++) /* :92-97 */
                    if ((0 <= ev[k] - padding) && (ev[k] - padding < len - padding - padding))
                        ev[nk++] = ev[k];
                if (cs->L[g] + nk > cs->cap[g]) {
                    int64_t nc = 2 * (cs->L[g] + nk) + 1024;
                    cs->clips[g] = (float*)realloc(cs->clips[g], sizeof(float) * W * T * nc);
                    cs->cap[g] = nc;
                }
                float* out = cs->clips[g] + W * T * cs->L[g];
                for (int64_t k = 0; k < nk; k++) /* :99-106, get_clip :409-419 */
                    for (int64_t tt = 0; tt < T; tt++) {
                        int64_t s = ev[k] + tt - Tmid;
                        for (int64_t w = 0; w < W; w++)
                            out[w + W * tt + W * T * k] = (s >= 0 && s < len) ? chunk[ch[w] + M * s] : 0.0f;
                    }
                cs->L[g] += nk;
                free(X0);
                free(ev);
            }
        }
    }
    free(chunk);
    return cs;
}

/*
 * The PCA part of NeighborhoodSorter::sort (neighborhoodsorter.cpp:153-160),
 * neighborhoods in batches of `batch`, an OpenMP loop over each batch
 * (p_mountainsort3.cpp:222-241): dimension_reduce_clips (stage 1) and the
 * first sort_clips_subset PCA (sort_clips.cpp:58-70, stage 2) on the
 * neighborhood's (W*nf) x L reduced clips.  clips[g]: W[g] x T x L[g];
 * reduced[g]: W[g] x nf x L[g]; features[g]: nf2 x L[g] (either may be NULL).
 */
EXPORT void orc_pca_neighborhoods_mt(const float* const* clips, const int64_t* W, const int64_t* L,
    int64_t G, int64_t T, int64_t nf, int64_t nf2, int64_t max_samples, int64_t batch,
    int nthreads, float* const* reduced, float* const* features)
{
    for (int64_t b0 = 0; b0 < G; b0 += batch) {
        int64_t b1 = b0 + batch < G ? b0 + batch : G;
#pragma omp parallel for num_threads(nthreads)
        for (int64_t g = b0; g < b1; g++) {
            int64_t Wg = W[g], Lg = L[g];
            if (Wg == 0 || Lg == 0)
                continue;
            float* red = reduced && reduced[g] ? reduced[g] : (float*)malloc(sizeof(float) * Wg * nf * Lg);
            orc_dimension_reduce_clips(clips[g], Wg, T, Lg, nf, max_samples, red);
            int64_t R = Wg * nf;
            float* C = (float*)malloc(sizeof(float) * R * nf2);
            float* F = features && features[g] ? features[g] : (float*)malloc(sizeof(float) * nf2 * Lg);
            float* s = (float*)malloc(sizeof(float) * nf2);
            orc_pca_subsampled(red, R, Lg, nf2, max_samples, C, F, s);
            if (!(features && features[g]))
                free(F);
            if (!(reduced && reduced[g]))
                free(red);
            free(C);
            free(s);
        }
    }
}
