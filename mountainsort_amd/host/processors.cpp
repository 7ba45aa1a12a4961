// The processors of msamd.mp: the reference's hot-path processors with the
// same names, versions, parameters and defaults
// (packages/ms3/mountainsort3_main.cpp:145-198,343-350) and stage processors
// for detection, neighborhood clips and PCA features that use the parameter
// names of mountainsortalg.ms3alg (mountainsortalg_main.cpp:31-46,87-102;
// p_mountainsort3.h:20-40).  Arrays are read into device memory, processed
// by libmsamd (include/msamd.h) and written back.
#include "processors.h"

#include <sched.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <thread>

#include "device.h"
#include "mda.h"
#include "msamd.h"

namespace msamd_host {

namespace {

constexpr int64_t kBandpassOverlap = 60000;  // p_bandpass_filter.cpp:149
constexpr int64_t kDetectPadding = 1000;     // p_mountainsort3.cpp:515

// omp_get_max_threads() of the reference process: _request_num_threads when
// given and nonzero (mountainsort3_main.cpp:428-434), else OMP_NUM_THREADS,
// else the CPUs this process may run on.  It fixes the chunk plans.
int reference_threads(const Ctx& c) {
    const long long explicit_t = c.clp.get("reference_num_threads").to_long();
    if (explicit_t > 0) return static_cast<int>(explicit_t);
    const long long req = c.clp.get("_request_num_threads").to_long();
    if (req > 0) return static_cast<int>(req);
    if (const char* e = getenv("OMP_NUM_THREADS")) {
        const int v = atoi(e);
        if (v > 0) return v;
    }
    cpu_set_t set;
    if (sched_getaffinity(0, sizeof set, &set) == 0) return std::max(1, CPU_COUNT(&set));
    return std::max(1u, std::thread::hardware_concurrency());
}

bool lib_ok(Ctx& c, int status, const char* what) {
    if (status == MSAMD_OK) return true;
    const char* msg = msamd_last_error();
    return c.fail(std::string(what) + ": " + (msg ? msg : "error"));
}

bool open_timeseries(Ctx& c, const std::string& param, TimeseriesSource& src) {
    const std::vector<std::string> paths = c.param(param).to_string_list();
    if (paths.empty()) return c.fail("missing parameter " + param);
    return src.open(paths, c.err);
}

// Device-resident float32 copy of a timeseries.
bool load_f32(Ctx& c, const TimeseriesSource& src, DeviceBuffer& X, hipStream_t s) {
    return X.alloc(src.M * src.N * 4, c.err) && upload_timeseries(src, MDA_FLOAT32, X.get(), s, c.err);
}

bool write_device(Ctx& c, const std::string& path, int32_t type, const std::vector<int64_t>& dims,
                  const void* dev, hipStream_t s) {
    int64_t bytes = mda_bytes_per_entry(type);
    for (int64_t d : dims) bytes *= d;
    MdaWriter w;
    return w.open(path, type, dims, c.err) && download_to_writer(dev, bytes, w, s, c.err) &&
           w.close(c.err);
}

bool copy_file(Ctx& c, const std::string& from, const std::string& to) {
    FILE* in = fopen(from.c_str(), "rb");
    if (!in) return c.fail("unable to open " + from);
    FILE* out = fopen((to + ".tmp").c_str(), "wb");
    if (!out) {
        fclose(in);
        return c.fail("unable to open " + to + ".tmp");
    }
    std::vector<char> buf(1 << 20);
    bool ok = true;
    size_t n;
    while ((n = fread(buf.data(), 1, buf.size(), in)) > 0) ok = ok && fwrite(buf.data(), 1, n, out) == n;
    fclose(in);
    ok = (fclose(out) == 0) && ok;
    if (!ok || std::rename((to + ".tmp").c_str(), to.c_str()) != 0) {
        std::remove((to + ".tmp").c_str());
        return c.fail("unable to copy " + from + " to " + to);
    }
    return true;
}

// 1-based channel list -> 0-based, checked against M.
bool channel_list(Ctx& c, const std::string& param, int64_t M, std::vector<int32_t>& out) {
    out.clear();
    for (int ch : int_list(c.str(param))) {
        if (ch < 1 || ch > M)
            return c.fail("channel " + std::to_string(ch) + " outside 1.." + std::to_string(M));
        out.push_back(ch - 1);
    }
    return true;
}

// Neighborhood channel lists (0-based, central channel first) of every
// channel: get_channels_from_geom (p_mountainsort3.cpp:473-491) with the
// geometry of p_mountainsort3.cpp:159-168 (no geom file: all coordinates 0).
bool neighborhoods(Ctx& c, int64_t M, std::vector<std::vector<int32_t>>& nb) {
    std::vector<std::vector<double>> geom;
    const std::string gpath = c.str("geom");
    if (gpath.empty()) {
        geom.assign(M, std::vector<double>(2, 0.0));
    } else {
        if (!read_geom_csv(gpath, geom, c.err)) return false;
        if (static_cast<int64_t>(geom.size()) != M)
            return c.fail("Error: Inconsistent number of channels in geometry file: " +
                          std::to_string(geom.size()) + " <> " + std::to_string(M));
    }
    const double r = c.num("adjacency_radius");
    nb.assign(M, {});
    for (int64_t m = 0; m < M; ++m) {
        nb[m].push_back(static_cast<int32_t>(m));
        for (int64_t m2 = 0; m2 < M; ++m2) {
            if (m2 == m) continue;
            double s = 0.0;
            for (size_t i = 0; i < geom[m].size(); ++i) {
                const double d = geom[m][i] - geom[m2][i];
                s += d * d;
            }
            const double dist = std::sqrt(s);
            if ((r > 0 && dist <= r) || r < 0) nb[m].push_back(static_cast<int32_t>(m2));
        }
    }
    return true;
}

// Events of a 2 x L (or R x L, R >= 2) firings-like array: row 0 central
// channel (1-based), row 1 time.
bool read_events(Ctx& c, int64_t M, std::vector<int32_t>& chan, std::vector<double>& time) {
    std::vector<double> a;
    MdaHeader h;
    if (!mda_read_f64(c.str("event_times"), a, h, c.err)) return false;
    if (h.dim(0) < 2) return c.fail("event_times must have rows (central channel, time)");
    const int64_t R = h.dim(0), L = h.size() / R;
    chan.resize(L);
    time.resize(L);
    for (int64_t i = 0; i < L; ++i) {
        const double ch = a[R * i];
        if (!(ch >= 1 && ch <= M) || ch != std::floor(ch))
            return c.fail("event " + std::to_string(i) + ": central channel outside 1.." +
                          std::to_string(M));
        chan[i] = static_cast<int32_t>(ch) - 1;
        time[i] = a[R * i + 1];
    }
    return true;
}

int clip_tmid(int64_t T) { return static_cast<int>((T + 1) / 2) - 1; }

// ---------------------------------------------------------------- processors

// Segments of about this many input bytes stream through the device
// (MSAMD_SEGMENT_BYTES overrides it).
constexpr int64_t kSegmentBytes = 256ll << 20;

// Pinned staging bytes of a segment pipeline (two input and two output buffers).
double pipeline_ram_mb(int64_t in_bytes, int64_t out_bytes) {
    return 2.0 * static_cast<double>(in_bytes + out_bytes) / (1024.0 * 1024.0);
}

// ms3.bandpass_filter, streamed: segments of whole windows of the reference's
// chunk plan (p_bandpass_filter.cpp:147-155,179-214); segment k reads the
// input rows of its windows (with the 60000-sample halos, zero outside the
// recording), filters them on the device and writes its output rows, while
// the next segment is read and the previous one written.
struct BandpassStream {
    RowReader* rd;
    MdaWriter* wr;
    int64_t M, N, cs, ov, nwin, S, k;
    int32_t in_type, out_type;
    double fs, fmin, fmax, fwid, q;
    void* ws;
    int64_t wsb;
    int64_t w0(int64_t s) const { return s * S; }
    int64_t w1(int64_t s) const { return std::min(nwin, (s + 1) * S); }
    int64_t xlo(int64_t s) const { return std::max<int64_t>(0, w0(s) * cs - ov); }
    int64_t xhi(int64_t s) const { return std::min<int64_t>(N, w1(s) * cs + ov); }
    int64_t ylo(int64_t s) const { return (w0(s) * cs) / k; }
    int64_t yhi(int64_t s) const { return std::min<int64_t>(N / k, (w1(s) * cs) / k); }

    static bool prepare(void* ctx, int64_t s, void* host, int64_t* nb, void**, std::string& err) {
        auto* b = static_cast<BandpassStream*>(ctx);
        *nb = (b->xhi(s) - b->xlo(s)) * b->M * mda_bytes_per_entry(b->in_type);
        return b->rd->read(b->in_type, b->xlo(s), b->xhi(s), host, err);
    }
    static bool compute(void* ctx, int64_t s, const void* din, void* dout, int64_t* ob,
                        hipStream_t st, std::string& err) {
        auto* b = static_cast<BandpassStream*>(ctx);
        const int rc = msamd_bandpass_filter_windows(
            din, b->in_type == MDA_INT16 ? MSAMD_DTYPE_INT16 : MSAMD_DTYPE_FLOAT32, b->M, b->N,
            b->xlo(s), b->xhi(s) - b->xlo(s), b->w0(s), b->w1(s), b->fs, b->fmin, b->fmax, b->fwid,
            b->cs, b->ov, static_cast<int>(b->k), b->q, dout, b->ylo(s), b->yhi(s) - b->ylo(s),
            b->ws, b->wsb, st);
        if (rc != MSAMD_OK) {
            const char* m = msamd_last_error();
            err = std::string("bandpass_filter: ") + (m ? m : "error");
            return false;
        }
        *ob = (b->yhi(s) - b->ylo(s)) * b->M * mda_bytes_per_entry(b->out_type);
        return true;
    }
    static bool consume(void* ctx, int64_t, const void* host, int64_t nb, std::string& err) {
        return static_cast<BandpassStream*>(ctx)->wr->write(host, nb, err);
    }
};

bool run_bandpass(Ctx& c) {
    const double fs = c.num("samplerate"), fmin = c.num("freq_min"), fmax = c.num("freq_max");
    const double fwid = c.num("freq_wid"), q = c.num("quantization_unit");
    long long k = c.integer("subsample_factor");
    if (k == 0) k = 1;  // p_bandpass_filter.cpp:121-122
    if (k < 0) return c.fail("subsample_factor must be positive");
    if (fmax == 0) {
        if (c.requirements_only) return true;
        return copy_file(c, c.str("timeseries"), c.str("timeseries_out"));  // :117-119
    }
    TimeseriesSource src;
    if (!open_timeseries(c, "timeseries", src)) {
        if (!c.requirements_only) return false;
        // inputs not there yet: the staging of the largest segments (float32
        // in, float32 out)
        c.peak_ram_mb = pipeline_ram_mb(segment_bytes(kSegmentBytes), segment_bytes(kSegmentBytes));
        return true;
    }
    BandpassStream b{};
    b.M = src.M;
    b.N = src.N;
    b.k = k;
    b.cs = msamd_bandpass_chunk_size(b.M, reference_threads(c));
    b.ov = kBandpassOverlap;
    if (!msamd_bandpass_supported(b.cs, b.ov))
        return c.fail("bandpass window length " + std::to_string(b.cs + 2 * b.ov) +
                      " (chunk plan for M=" + std::to_string(b.M) + ", " +
                      std::to_string(reference_threads(c)) + " threads) is not supported");
    b.in_type = src.common_type == MDA_INT16 ? MDA_INT16 : MDA_FLOAT32;
    b.out_type = q != 0 ? MDA_INT16 : MDA_FLOAT32;
    b.nwin = msamd_bandpass_windows(b.N, b.cs);
    const int64_t in_row = b.M * mda_bytes_per_entry(b.in_type);
    b.S = std::max<int64_t>(1, segment_bytes(kSegmentBytes) / (b.cs * in_row));
    b.S = std::min<int64_t>(b.S, std::max<int64_t>(1, b.nwin));
    const int64_t max_in = (b.S * b.cs + 2 * b.ov) * in_row;
    const int64_t max_out = (b.S * b.cs / k + 1) * b.M * mda_bytes_per_entry(b.out_type);
    if (c.requirements_only) {
        c.peak_ram_mb = pipeline_ram_mb(max_in, max_out);
        return true;
    }
    b.fs = fs;
    b.fmin = fmin;
    b.fmax = fmax;
    b.fwid = fwid;
    b.q = q;
    DeviceBuffer ws;
    b.wsb = msamd_bandpass_workspace_size(b.M, std::max<int64_t>(b.N, 1), b.cs, b.ov);
    if (b.wsb < 0) return c.fail("bandpass: unsupported window length");
    RowReader rd(src);
    MdaWriter wr;
    SegmentPipeline pipe;
    if (!ws.alloc(b.wsb, c.err) || !pipe.init(max_in, max_out, true, c.err) ||
        !wr.open(c.str("timeseries_out"), b.out_type, {b.M, b.N / k}, c.err))
        return false;
    b.ws = ws.get();
    b.rd = &rd;
    b.wr = &wr;
    const int64_t K = (b.nwin + b.S - 1) / b.S;
    SegmentPipeline::Stage st{&BandpassStream::prepare, &BandpassStream::compute,
                              &BandpassStream::consume, &b};
    if (!pipe.run(K, st, c.err)) {
        wr.abort();
        return false;
    }
    return wr.close(c.err);
}

// ms3.whiten and ms3.apply_whitening_matrix, streamed over row segments
// (p_whiten.cpp:64-206 reads the recording in 1e5-sample chunks, twice).
// Pass 1 accumulates the covariance segment by segment; pass 2 applies W to
// each segment and writes it.  When the float32 recording fits in device
// memory it is kept there after pass 1 and read from disk once; otherwise
// pass 2 reads it again.  The covariance sums are the same either way.
struct WhitenStream {
    RowReader* rd;
    MdaWriter* wr;
    int64_t M, N, R;
    float* resident;  // the device copy of the recording, or nullptr
    double* C;        // device M x M: covariance sums, then the covariance
    const double* W;  // device M x M (column-major), pass 2
    bool with_cov;    // pass 2 bounds |x| by the covariance diagonal
    double q;
    int32_t out_type;
    void* ws;
    int64_t wsb;
    int64_t r0(int64_t s) const { return s * R; }
    int64_t r1(int64_t s) const { return std::min(N, (s + 1) * R); }

    static bool lib(int rc, const char* what, std::string& err) {
        if (rc == MSAMD_OK) return true;
        const char* m = msamd_last_error();
        err = std::string(what) + ": " + (m ? m : "error");
        return false;
    }
    static bool read(void* ctx, int64_t s, void* host, int64_t* nb, void** dst, std::string& err) {
        auto* w = static_cast<WhitenStream*>(ctx);
        *nb = (w->r1(s) - w->r0(s)) * w->M * 4;
        if (w->resident) *dst = w->resident + w->r0(s) * w->M;
        return w->rd->read(MDA_FLOAT32, w->r0(s), w->r1(s), host, err);
    }
    static bool reread(void* ctx, int64_t s, void* host, int64_t* nb, void** dst, std::string& err) {
        auto* w = static_cast<WhitenStream*>(ctx);
        if (w->resident) {
            *nb = 0;
            return true;
        }
        return read(ctx, s, host, nb, dst, err);
    }
    static bool covariance(void* ctx, int64_t s, const void* din, void*, int64_t* ob,
                           hipStream_t st, std::string& err) {
        auto* w = static_cast<WhitenStream*>(ctx);
        *ob = 0;
        return lib(msamd_whiten_covariance_accumulate(static_cast<const float*>(din), w->M,
                                                      w->r1(s) - w->r0(s), nullptr, w->C,
                                                      s > 0 ? 1 : 0, w->ws, w->wsb, st),
                   "whiten_covariance", err);
    }
    static bool apply(void* ctx, int64_t s, const void* din, void* dout, int64_t* ob,
                      hipStream_t st, std::string& err) {
        auto* w = static_cast<WhitenStream*>(ctx);
        const float* X = w->resident ? w->resident + w->r0(s) * w->M : static_cast<const float*>(din);
        const int64_t n = w->r1(s) - w->r0(s);
        float* Yf = w->q > 0 ? nullptr : static_cast<float*>(dout);
        int16_t* Yq = w->q > 0 ? static_cast<int16_t*>(dout) : nullptr;
        const int rc = w->with_cov
                           ? msamd_whiten_apply_cov(X, w->M, n, w->W, w->C, nullptr, Yf, w->q, Yq,
                                                    w->ws, w->wsb, st)
                           : msamd_whiten_apply(X, w->M, n, w->W, Yf, w->q, Yq, w->ws, w->wsb, st);
        *ob = n * w->M * mda_bytes_per_entry(w->out_type);
        return lib(rc, "whiten_apply", err);
    }
    static bool write(void* ctx, int64_t, const void* host, int64_t nb, std::string& err) {
        return static_cast<WhitenStream*>(ctx)->wr->write(host, nb, err);
    }
};

// The reference's power-method W; if its workgroups could not run together
// (MSAMD_ERR_TIMEOUT: other work held the device), one retry, then the
// exact inverse square root with a warning.
bool whitening_matrix_or_exact(Ctx& c, const double* C, int64_t M, double* W, hipStream_t s) {
    DeviceBuffer ws;
    if (!ws.alloc(msamd_whitening_matrix_workspace_size(M), c.err)) return false;
    int rc = MSAMD_ERR_TIMEOUT;
    for (int attempt = 0; attempt < 2 && rc == MSAMD_ERR_TIMEOUT; ++attempt)
        rc = msamd_whitening_matrix(C, M, W, ws.get(), ws.bytes(), s);
    if (rc != MSAMD_ERR_TIMEOUT) return lib_ok(c, rc, "whitening_matrix");
    const char* m = msamd_last_error();
    fprintf(stderr, "warning: %s; using the exact whitening matrix instead\n", m ? m : "timeout");
    DeviceBuffer wx;
    int32_t iters = 0;
    double resid = 0.0;
    if (!wx.alloc(msamd_whitening_matrix_exact_workspace_size(M), c.err)) return false;
    return lib_ok(c, msamd_whitening_matrix_exact(C, M, W, wx.get(), wx.bytes(), s, &iters, &resid),
                  "whitening_matrix_exact");
}

// W given (apply_whitening_matrix, device column-major) or computed
// (whiten).  Returns false with c.err on failure.
bool whiten_streamed(Ctx& c, const TimeseriesSource& src, const std::vector<double>* Wgiven) {
    WhitenStream w{};
    w.M = src.M;
    w.N = src.N;
    w.q = c.num("quantization_unit");
    w.out_type = w.q > 0 ? MDA_INT16 : MDA_FLOAT32;
    const int64_t M = w.M;
    w.R = std::max<int64_t>(1, segment_bytes(kSegmentBytes) / (M * 4));
    w.R = std::min<int64_t>(w.R, std::max<int64_t>(1, w.N));
    const int64_t max_in = w.R * M * 4;
    const int64_t max_out = w.R * M * mda_bytes_per_entry(w.out_type);
    if (c.requirements_only) {
        c.peak_ram_mb = pipeline_ram_mb(max_in, max_out);
        return true;
    }
    const int64_t K = (w.N + w.R - 1) / w.R;
    DeviceBuffer Wd, Cd, ws, res;
    // the covariance workspace of a full segment and of the last, shorter one
    const int64_t last = w.N - (K > 0 ? K - 1 : 0) * w.R;
    const int64_t wsb = std::max({msamd_whiten_covariance_workspace_size(M, w.R),
                                  msamd_whiten_covariance_workspace_size(M, std::max<int64_t>(last, 1)),
                                  msamd_whiten_apply_workspace_size(M)});
    if (!Wd.alloc(M * M * 8, c.err) || !Cd.alloc(M * M * 8, c.err) || !ws.alloc(wsb, c.err))
        return false;
    w.ws = ws.get();
    w.wsb = wsb;
    w.C = Cd.as<double>();
    w.W = Wd.as<double>();
    // keep the recording on the device when it fits with room to spare
    const char* e = getenv("MSAMD_STREAM_RESIDENT");
    const bool allow_resident = !(e && std::string(e) == "0");
    const int64_t need = M * std::max<int64_t>(w.N, 1) * 4;
    const bool resident = Wgiven == nullptr && allow_resident &&
                          need + 4 * (max_in + max_out) < device_free_bytes() * 4 / 5;
    if (resident) {
        if (!res.alloc(need, c.err)) return false;
        w.resident = res.as<float>();
    }
    RowReader rd(src);
    MdaWriter wr;
    w.rd = &rd;
    w.wr = &wr;
    SegmentPipeline pipe;
    if (!pipe.init(max_in, max_out, !resident, c.err)) return false;
    Stream s;
    if (!s.create(c.err)) return false;
    if (Wgiven) {
        if (!to_device(Wd.get(), Wgiven->data(), M * M * 8, s.get(), c.err)) return false;
        w.with_cov = false;
    } else {
        // pass 1: covariance sums, then / (N-1) and the whitening matrix
        SegmentPipeline::Stage p1{&WhitenStream::read, &WhitenStream::covariance, nullptr, &w};
        if (w.N == 0) {
            if (!lib_ok(c, msamd_whiten_covariance_accumulate(nullptr, M, 0, nullptr, w.C, 0, w.ws,
                                                              w.wsb, s.get()),
                        "whiten_covariance"))
                return false;
        } else if (!pipe.run(K, p1, c.err)) {
            return false;
        }
        if (!lib_ok(c, msamd_whiten_covariance_finish(w.C, M, w.N, s.get()), "whiten_covariance") ||
            !whitening_matrix_or_exact(c, w.C, M, Wd.as<double>(), s.get()) || !s.sync(c.err))
            return false;
        w.with_cov = true;
    }
    // pass 2: apply and write
    if (!wr.open(c.str("timeseries_out"), w.out_type, {M, w.N}, c.err)) return false;
    SegmentPipeline::Stage p2{&WhitenStream::reread, &WhitenStream::apply, &WhitenStream::write, &w};
    if (w.N > 0 && !pipe.run(K, p2, c.err)) {
        wr.abort();
        return false;
    }
    return wr.close(c.err);
}

bool run_whiten(Ctx& c) {
    TimeseriesSource src;
    if (!open_timeseries(c, "timeseries", src)) {
        if (!c.requirements_only) return false;
        c.peak_ram_mb = pipeline_ram_mb(segment_bytes(kSegmentBytes), segment_bytes(kSegmentBytes));
        return true;
    }
    return whiten_streamed(c, src, nullptr);
}

// Covariance and whitening matrix of a device timeseries (M x N float32) into
// a device M x M float64 column-major buffer.
bool whitening_matrix_of(Ctx& c, const float* X, int64_t M, int64_t N, DeviceBuffer& Wd,
                         hipStream_t s, DeviceBuffer* Cout = nullptr) {
    DeviceBuffer Cl, ws1;
    DeviceBuffer& C = Cout ? *Cout : Cl;
    if (!C.alloc(M * M * 8, c.err) || !Wd.alloc(M * M * 8, c.err) ||
        !ws1.alloc(msamd_whiten_covariance_workspace_size(M, N), c.err))
        return false;
    return lib_ok(c, msamd_whiten_covariance(X, M, N, C.as<double>(), ws1.get(), ws1.bytes(), s),
                  "whiten_covariance") &&
           whitening_matrix_or_exact(c, C.as<double>(), M, Wd.as<double>(), s);
}

bool run_compute_whitening_matrix(Ctx& c) {
    TimeseriesSource src;
    Stream s;
    DeviceBuffer X, W;
    if (!open_timeseries(c, "timeseries_list", src) || !s.create(c.err) ||
        !load_f32(c, src, X, s.get()))
        return false;
    std::vector<int32_t> ch;
    if (!channel_list(c, "channels", src.M, ch)) return false;
    const float* Xs = X.as<float>();
    int64_t M2 = src.M;
    DeviceBuffer sub, dstarts, dch;
    if (!ch.empty()) {
        // The channel subset as an M2 x N timeseries: consecutive "clips" of
        // Tb samples laid end to end are exactly that array.
        M2 = static_cast<int64_t>(ch.size());
        const int32_t Tb = 4096;
        const int64_t L = (src.N + Tb - 1) / Tb;
        std::vector<int64_t> starts(L);
        for (int64_t i = 0; i < L; ++i) starts[i] = i * Tb;
        if (!sub.alloc(std::max<int64_t>(1, M2 * Tb * L) * 4, c.err) ||
            !dstarts.alloc(L * 8, c.err) || !dch.alloc(M2 * 4, c.err) ||
            !to_device(dstarts.get(), starts.data(), L * 8, s.get(), c.err) ||
            !to_device(dch.get(), ch.data(), M2 * 4, s.get(), c.err))
            return false;
        if (!lib_ok(c, msamd_extract_clips(Xs, src.M, src.N, dstarts.as<int64_t>(), L,
                                           dch.as<int32_t>(), static_cast<int32_t>(M2), Tb,
                                           sub.as<float>(), s.get()),
                    "extract_clips"))
            return false;
        Xs = sub.as<float>();
    }
    if (!whitening_matrix_of(c, Xs, M2, src.N, W, s.get())) return false;
    return write_device(c, c.str("whitening_matrix_out"), MDA_FLOAT64, {M2, M2}, W.get(), s.get());
}

bool run_apply_whitening_matrix(Ctx& c) {
    TimeseriesSource src;
    if (!open_timeseries(c, "timeseries", src)) return false;
    std::vector<double> w;
    MdaHeader h;
    if (!mda_read_f64(c.str("whitening_matrix"), w, h, c.err)) return false;
    if (h.dim(0) != src.M || h.size() != src.M * src.M)
        return c.fail("whitening matrix is not " + std::to_string(src.M) + " x " +
                      std::to_string(src.M));
    return whiten_streamed(c, src, &w);
}

// ms3.whiten_clips (p_whiten.cpp:295-345): clips M x T x L, W M x M.
bool run_whiten_clips(Ctx& c) {
    std::vector<float> clips;
    std::vector<double> w;
    MdaHeader hc, hw;
    if (!mda_read_f32(c.str("clips"), clips, hc, c.err) ||
        !mda_read_f64(c.str("whitening_matrix"), w, hw, c.err))
        return false;
    const int64_t M = hc.dim(0), T = hc.dim(1), L = hc.dim(2);
    if (hc.size() != M * T * L) return c.fail("clips must be three-dimensional (M x T x L)");
    if (hw.dim(0) != M || hw.size() != M * M)
        return c.fail("whitening matrix is not " + std::to_string(M) + " x " + std::to_string(M));
    const double q = c.num("quantization_unit");
    const int32_t out_type = q > 0 ? MDA_INT16 : MDA_FLOAT32;
    Stream s;
    DeviceBuffer X, W, Y, ws;
    const int64_t n = std::max<int64_t>(1, M * T * L);
    if (!s.create(c.err) || !X.alloc(n * 4, c.err) || !W.alloc(M * M * 8, c.err) ||
        !Y.alloc(n * mda_bytes_per_entry(out_type), c.err) ||
        !ws.alloc(msamd_whiten_apply_workspace_size(M), c.err) ||
        !to_device(X.get(), clips.data(), M * T * L * 4, s.get(), c.err) ||
        !to_device(W.get(), w.data(), M * M * 8, s.get(), c.err))
        return false;
    const int st = q > 0 ? msamd_whiten_clips(X.as<float>(), M, T, L, W.as<double>(), nullptr, q,
                                              Y.as<int16_t>(), ws.get(), ws.bytes(), s.get())
                         : msamd_whiten_clips(X.as<float>(), M, T, L, W.as<double>(),
                                              Y.as<float>(), 0.0, nullptr, ws.get(), ws.bytes(),
                                              s.get());
    if (!lib_ok(c, st, "whiten_clips")) return false;
    return write_device(c, c.str("clips_out"), out_type, {M, T, L}, Y.get(), s.get());
}

bool run_mask(Ctx& c) {
    const double thr = c.num("threshold");
    const int64_t isz = static_cast<int64_t>(c.num("interval_size"));
    if (thr == 0 || isz == 0)  // p_mask_out_artifacts.cpp:26-29
        return c.fail("Problem with input parameters. Either threshold or interval_size is zero.");
    if (isz < 0) return c.fail("interval_size must be positive");
    TimeseriesSource src;
    Stream s;
    DeviceBuffer X, Y, ws, use;
    if (!open_timeseries(c, "timeseries", src) || !s.create(c.err) || !load_f32(c, src, X, s.get()))
        return false;
    const int64_t M = src.M, N = src.N;
    if (!Y.alloc(M * N * 4, c.err) || !ws.alloc(msamd_mask_workspace_size(M, N, isz), c.err) ||
        !use.alloc((N / isz + 1) * 4, c.err))
        return false;
    if (!lib_ok(c, msamd_mask_out_artifacts(X.as<float>(), M, N, thr, isz, Y.as<float>(),
                                            use.as<int32_t>(), ws.get(), ws.bytes(), s.get()),
                "mask_out_artifacts"))
        return false;
    return write_device(c, c.str("timeseries_out"), MDA_FLOAT32, {M, N}, Y.get(), s.get());
}

bool run_extract_clips(Ctx& c) {
    std::vector<double> et;
    MdaHeader h;
    if (!mda_read_f64(c.str("event_times"), et, h, c.err)) return false;
    if (h.dim(0) > 1 && h.dim(1) > 1)  // p_extract_clips.cpp:39-42
        return c.fail("Error: the input (event times) must be a vector.");
    const int64_t T = c.integer("clip_size");
    if (!T) return c.fail("Unexpected: Clip size is zero.");
    if (T < 0) return c.fail("clip_size must be positive");
    TimeseriesSource src;
    Stream s;
    DeviceBuffer X, clips, dstarts, dch;
    if (!open_timeseries(c, "timeseries", src) || !s.create(c.err)) return false;
    std::vector<int32_t> ch;
    if (!channel_list(c, "channels", src.M, ch)) return false;
    if (ch.empty()) {
        ch.resize(src.M);
        std::iota(ch.begin(), ch.end(), 0);
    }
    const int64_t L = static_cast<int64_t>(et.size()), W = static_cast<int64_t>(ch.size());
    std::vector<int64_t> starts(L);
    const int tmid = clip_tmid(T);
    for (int64_t i = 0; i < L; ++i) starts[i] = static_cast<int64_t>(et[i] - tmid);  // :59
    if (!load_f32(c, src, X, s.get()) || !clips.alloc(W * T * L * 4, c.err) ||
        !dstarts.alloc(L * 8, c.err) || !dch.alloc(W * 4, c.err) ||
        !to_device(dstarts.get(), starts.data(), L * 8, s.get(), c.err) ||
        !to_device(dch.get(), ch.data(), W * 4, s.get(), c.err))
        return false;
    if (!lib_ok(c, msamd_extract_clips(X.as<float>(), src.M, src.N, dstarts.as<int64_t>(), L,
                                       dch.as<int32_t>(), static_cast<int32_t>(W),
                                       static_cast<int32_t>(T), clips.as<float>(), s.get()),
                "extract_clips"))
        return false;
    return write_device(c, c.str("clips_out"), MDA_FLOAT32, {W, T, L}, clips.get(), s.get());
}

// Detection on every channel; returns (channel, time) pairs in the order of
// the ms3alg firings: by time, ties in channel order (p_mountainsort3.cpp:245-263).
bool detect_all(Ctx& c, const float* X, int64_t M, int64_t N, hipStream_t s,
                std::vector<int32_t>& chan, std::vector<int64_t>& time) {
    const double thr = c.num("detect_threshold");
    const double interval = static_cast<double>(static_cast<int>(c.num("detect_interval")));
    const int sign = c.param("detect_sign").to_int();
    const int64_t t1 = static_cast<int64_t>(c.num("t1")), t2 = static_cast<int64_t>(c.num("t2"));
    const int64_t t_start = t1 < 0 ? 0 : t1, t_end = t2 < 0 ? N - 1 : t2;
    if (t_end >= N || t_start > t_end + 1) return c.fail("t1/t2 outside the timeseries");
    const int64_t n_range = t_end - t_start + 1;
    const int64_t chunk = msamd_detect_chunk_size(M, n_range, reference_threads(c));
    std::vector<int32_t> all(M);
    std::iota(all.begin(), all.end(), 0);
    DeviceBuffer dch, ws, dtimes, doff;
    const int64_t wsb = msamd_detect_workspace_size(n_range, chunk, static_cast<int32_t>(M), interval);
    if (!dch.alloc(M * 4, c.err) || !to_device(dch.get(), all.data(), M * 4, s, c.err) ||
        !ws.alloc(wsb, c.err))
        return false;
    int64_t total = 0;
    if (!lib_ok(c, msamd_detect_events(X, M, N, t1, t2, chunk, kDetectPadding, dch.as<int32_t>(),
                                       static_cast<int32_t>(M), thr, interval, sign, ws.get(),
                                       ws.bytes(), &total, s),
                "detect_events"))
        return false;
    if (!dtimes.alloc(std::max<int64_t>(total, 1) * 8, c.err) || !doff.alloc((M + 1) * 8, c.err))
        return false;
    if (!lib_ok(c, msamd_detect_fetch(ws.get(), static_cast<int32_t>(M), dtimes.as<int64_t>(),
                                      doff.as<int64_t>(), s),
                "detect_fetch"))
        return false;
    std::vector<int64_t> t(total), off(M + 1);
    if (!to_host(t.data(), dtimes.get(), total * 8, s, c.err) ||
        !to_host(off.data(), doff.get(), (M + 1) * 8, s, c.err))
        return false;
    std::vector<int32_t> ch(total);
    for (int64_t m = 0; m < M; ++m)
        for (int64_t i = off[m]; i < off[m + 1]; ++i) ch[i] = static_cast<int32_t>(m);
    std::vector<int64_t> order(total);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return t[a] < t[b]; });
    chan.resize(total);
    time.resize(total);
    for (int64_t i = 0; i < total; ++i) {
        chan[i] = ch[order[i]];
        time[i] = t[order[i]];
    }
    return true;
}

bool run_detect(Ctx& c) {
    TimeseriesSource src;
    Stream s;
    DeviceBuffer X;
    if (!open_timeseries(c, "timeseries", src) || !s.create(c.err) || !load_f32(c, src, X, s.get()))
        return false;
    std::vector<int32_t> chan;
    std::vector<int64_t> time;
    if (!detect_all(c, X.as<float>(), src.M, src.N, s.get(), chan, time)) return false;
    const int64_t L = static_cast<int64_t>(time.size());
    std::vector<double> out(2 * L);
    for (int64_t i = 0; i < L; ++i) {
        out[2 * i] = chan[i] + 1;
        out[2 * i + 1] = static_cast<double>(time[i]);
    }
    return mda_write(c.str("event_times_out"), MDA_FLOAT64, {2, L}, out.data(), c.err);
}

bool run_neighborhood_clips(Ctx& c) {
    TimeseriesSource src;
    if (!open_timeseries(c, "timeseries", src)) return false;
    const int64_t M = src.M;
    const long long cc = c.integer("central_channel");
    if (cc < 1 || cc > M) return c.fail("central_channel outside 1.." + std::to_string(M));
    const int64_t T = c.integer("clip_size");
    if (T < 1) return c.fail("clip_size must be positive");
    std::vector<std::vector<int32_t>> nb;
    std::vector<int32_t> chan;
    std::vector<double> time;
    if (!neighborhoods(c, M, nb) || !read_events(c, M, chan, time)) return false;
    const std::vector<int32_t>& ch = nb[cc - 1];
    std::vector<int64_t> starts;
    const int tmid = clip_tmid(T);
    for (size_t i = 0; i < chan.size(); ++i)
        if (chan[i] == cc - 1) starts.push_back(static_cast<int64_t>(time[i] - tmid));
    const int64_t L = static_cast<int64_t>(starts.size()), W = static_cast<int64_t>(ch.size());
    Stream s;
    DeviceBuffer X, clips, dstarts, dch;
    if (!s.create(c.err) || !load_f32(c, src, X, s.get()) || !clips.alloc(W * T * L * 4, c.err) ||
        !dstarts.alloc(L * 8, c.err) || !dch.alloc(W * 4, c.err) ||
        !to_device(dstarts.get(), starts.data(), L * 8, s.get(), c.err) ||
        !to_device(dch.get(), ch.data(), W * 4, s.get(), c.err))
        return false;
    if (!lib_ok(c, msamd_extract_clips(X.as<float>(), M, src.N, dstarts.as<int64_t>(), L,
                                       dch.as<int32_t>(), static_cast<int32_t>(W),
                                       static_cast<int32_t>(T), clips.as<float>(), s.get()),
                "extract_clips"))
        return false;
    return write_device(c, c.str("clips_out"), MDA_FLOAT32, {W, T, L}, clips.get(), s.get());
}

// Clips and two-stage PCA features of events grouped by central channel.
// chan/time: events in any order; features come back in that order
// (nf2 x L, column i for event i).
bool features_of(Ctx& c, const float* X, int64_t M, int64_t N, const std::vector<int32_t>& chan,
                 const std::vector<int64_t>& center, hipStream_t s, std::vector<float>& feat) {
    const int64_t T = c.integer("clip_size");
    const int64_t nf1 = c.integer("num_features_per_channel"), nf2 = c.integer("num_features");
    const int64_t maxs = c.integer("max_pca_samples");
    if (T < 1) return c.fail("clip_size must be positive");
    std::vector<std::vector<int32_t>> nb;
    if (!neighborhoods(c, M, nb)) return false;
    const int32_t G = static_cast<int32_t>(M);
    const int64_t L = static_cast<int64_t>(chan.size());
    // events grouped by central channel, input order within a group
    std::vector<int64_t> eo(G + 1, 0);
    for (int32_t ch : chan) eo[ch + 1]++;
    for (int32_t g = 0; g < G; ++g) eo[g + 1] += eo[g];
    std::vector<int64_t> slot(eo.begin(), eo.end() - 1), pos(L), times(L);
    for (int64_t i = 0; i < L; ++i) {
        pos[i] = slot[chan[i]]++;
        times[pos[i]] = center[i];
    }
    std::vector<int64_t> num(G), no(G + 1, 0);
    std::vector<int32_t> width(G), nc;
    for (int32_t g = 0; g < G; ++g) {
        num[g] = eo[g + 1] - eo[g];
        width[g] = static_cast<int32_t>(nb[g].size());
        if (width[g] > 128)
            return c.fail("neighborhood of channel " + std::to_string(g + 1) + " has " +
                          std::to_string(width[g]) +
                          " channels; the PCA kernels take at most 128 (reduce adjacency_radius)");
        nc.insert(nc.end(), nb[g].begin(), nb[g].end());
        no[g + 1] = no[g] + width[g];
    }
    // the clip gather fused into the stage-1 PCA: no clip array on the device
    struct Plan {
        msamd_pca_plan* p = nullptr;
        ~Plan() { msamd_pca_plan_destroy(p); }
    } plan;
    if (!lib_ok(c, msamd_pca_plan_create(G, num.data(), width.data(), static_cast<int32_t>(T),
                                         static_cast<int32_t>(nf1), static_cast<int32_t>(nf2),
                                         maxs, nullptr, nullptr, 0, s, &plan.p),
                "pca_plan_create"))
        return false;
    const int64_t wsb = msamd_pca_plan_workspace_bytes(plan.p);
    int64_t red = 0;
    for (int32_t g = 0; g < G; ++g) red += width[g] * nf1 * num[g];
    DeviceBuffer dtimes, dnc, dno, reduced, features, ws;
    if (!dtimes.alloc(L * 8, c.err) || !dnc.alloc(static_cast<int64_t>(nc.size()) * 4, c.err) ||
        !dno.alloc((G + 1) * 8, c.err) || !reduced.alloc(red * 4, c.err) ||
        !features.alloc(nf2 * L * 4, c.err) || !ws.alloc(wsb, c.err))
        return false;
    if (!to_device(dtimes.get(), times.data(), L * 8, s, c.err) ||
        !to_device(dnc.get(), nc.data(), static_cast<int64_t>(nc.size()) * 4, s, c.err) ||
        !to_device(dno.get(), no.data(), (G + 1) * 8, s, c.err))
        return false;
    if (!lib_ok(c, msamd_pca_plan_attach(plan.p, ws.get(), ws.bytes(), s), "pca_plan_attach") ||
        !lib_ok(c, msamd_pca_plan_run_timeseries(plan.p, X, M, N, dtimes.as<int64_t>(),
                                                 dnc.as<int32_t>(), dno.as<int64_t>(),
                                                 reduced.as<float>(), features.as<float>(), s),
                "pca_plan_run_timeseries"))
        return false;
    std::vector<float> grouped(nf2 * L);
    if (!to_host(grouped.data(), features.get(), nf2 * L * 4, s, c.err)) return false;
    feat.assign(nf2 * L, 0.0f);
    for (int64_t i = 0; i < L; ++i)
        std::copy_n(&grouped[nf2 * pos[i]], nf2, &feat[nf2 * i]);
    return true;
}

bool run_pca_features(Ctx& c) {
    TimeseriesSource src;
    if (!open_timeseries(c, "timeseries", src)) return false;
    std::vector<int32_t> chan;
    std::vector<double> time;
    if (!read_events(c, src.M, chan, time)) return false;
    std::vector<int64_t> center(time.size());
    for (size_t i = 0; i < time.size(); ++i) center[i] = static_cast<int64_t>(time[i]);
    Stream s;
    DeviceBuffer X;
    if (!s.create(c.err) || !load_f32(c, src, X, s.get())) return false;
    std::vector<float> feat;
    if (!features_of(c, X.as<float>(), src.M, src.N, chan, center, s.get(), feat)) return false;
    const int64_t nf2 = c.integer("num_features");
    return mda_write(c.str("features_out"), MDA_FLOAT32, {nf2, static_cast<int64_t>(time.size())},
                     feat.data(), c.err);
}

// Firings (R x L, R >= 3: central channel, time, label) as times and labels.
bool read_firings(Ctx& c, std::vector<double>& time, std::vector<int32_t>& label,
                  std::vector<double>& all, int64_t& R) {
    MdaHeader h;
    if (!mda_read_f64(c.str("firings"), all, h, c.err)) return false;
    if (h.dim(0) < 3) return c.fail("firings must have rows (central channel, time, label)");
    R = h.dim(0);
    const int64_t L = h.size() / R;
    time.resize(L);
    label.resize(L);
    for (int64_t i = 0; i < L; ++i) {
        time[i] = all[R * i + 1];
        label[i] = static_cast<int32_t>(all[R * i + 2]);
    }
    return true;
}

// ms3.compute_templates (p_compute_templates.cpp:24-118): the mean clip of
// every listed cluster (all labels 1..max when `clusters` is empty), in list
// order, clips starting at (int64)(time - Tmid).
bool run_compute_templates(Ctx& c) {
    TimeseriesSource src;
    Stream s;
    DeviceBuffer X;
    const int64_t T = c.integer("clip_size");
    if (T < 1) return c.fail("Unexpected: Clip size is zero.");
    std::vector<double> time, all;
    std::vector<int32_t> label;
    int64_t R = 0;
    if (!open_timeseries(c, "timeseries", src) || !read_firings(c, time, label, all, R)) return false;
    std::vector<int> clusters = int_list(c.str("clusters"));
    if (clusters.empty()) {
        int kmax = 0;
        for (int32_t k : label) kmax = std::max(kmax, static_cast<int>(k));
        for (int k = 1; k <= kmax; ++k) clusters.push_back(k);
    }
    const int64_t K = static_cast<int64_t>(clusters.size());
    if (K == 0) return mda_write(c.str("templates_out"), MDA_FLOAT32, {src.M, T, 0}, nullptr, c.err);
    int kmax = *std::max_element(clusters.begin(), clusters.end());
    std::vector<int32_t> map(static_cast<size_t>(std::max(kmax, 0)) + 1, 0);
    for (int64_t q = 0; q < K; ++q)
        if (clusters[q] >= 0) map[clusters[q]] = static_cast<int32_t>(q + 1);
    const int tmid = clip_tmid(T);
    std::vector<int64_t> t;
    std::vector<int32_t> k;
    for (size_t i = 0; i < time.size(); ++i) {
        const int32_t lab = label[i];
        if (lab < 0 || lab > kmax || map[lab] == 0) continue;
        // (bigint)(time - Tmid) is the first sample (:88); the kernel takes time = start + Tmid
        t.push_back(static_cast<int64_t>(time[i] - tmid) + tmid);
        k.push_back(map[lab]);
    }
    DeviceBuffer tm, ws;
    const int64_t L = static_cast<int64_t>(t.size());
    if (!s.create(c.err) || !load_f32(c, src, X, s.get()) || !tm.alloc(src.M * T * K * 4, c.err) ||
        !ws.alloc(msamd_templates_workspace_size(L, static_cast<int32_t>(K)), c.err))
        return false;
    if (!lib_ok(c, msamd_compute_templates(X.as<float>(), src.M, src.N, t.data(), k.data(), L,
                                           static_cast<int32_t>(K), static_cast<int32_t>(T),
                                           tm.as<float>(), ws.get(), ws.bytes(), s.get()),
                "compute_templates"))
        return false;
    return write_device(c, c.str("templates_out"), MDA_FLOAT32, {src.M, T, K}, tm.get(), s.get());
}

// The template and fit phase of mountainsortalg.ms3alg
// (p_mountainsort3.cpp:270-289,313-344): global templates of every label
// over the chunks of get_time_chunk_infos(M, N, 1, 1) (an event counts when
// it lies in its chunk's first size - 1000 samples), then the fit stage over
// the chunks of get_time_chunk_infos(M, N, threads, threads).  Writes the
// firings columns of the events kept, in time order.
bool run_fit_stage(Ctx& c) {
    TimeseriesSource src;
    Stream s;
    DeviceBuffer X;
    const int64_t T = c.integer("clip_size");
    if (T < 1) return c.fail("clip_size must be positive");
    std::vector<double> time, all;
    std::vector<int32_t> label;
    int64_t R = 0;
    if (!open_timeseries(c, "timeseries", src) || !read_firings(c, time, label, all, R)) return false;
    const int64_t M = src.M, N = src.N, L = static_cast<int64_t>(time.size());
    std::vector<int64_t> order(L);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return time[a] < time[b]; });
    std::vector<int64_t> t(L);
    std::vector<int32_t> k(L);
    int32_t K = 0;
    for (int64_t i = 0; i < L; ++i) {
        t[i] = static_cast<int64_t>(time[order[i]]);
        k[i] = label[order[i]];
        K = std::max(K, k[i]);
    }
    const int threads = reference_threads(c);
    const int64_t pad = kDetectPadding;
    // global templates: the events of each chunk's first size - padding samples
    const int64_t gcs = msamd_time_chunk_size(M, N, 1, 1);
    std::vector<int64_t> gt;
    std::vector<int32_t> gk;
    for (int64_t i = 0; i < L; ++i) {
        const int64_t t1 = (t[i] / gcs) * gcs;
        const int64_t size = std::min(gcs, N - t1);
        if (t[i] >= 0 && t[i] < N && t[i] - t1 < size - pad) {
            gt.push_back(t[i]);
            gk.push_back(k[i]);
        }
    }
    const int64_t fcs = msamd_time_chunk_size(M, N, threads, threads);
    DeviceBuffer tm, tws, fws;
    const int64_t fwsb = msamd_fit_stage_workspace_size(M, N, t.data(), L, K, static_cast<int32_t>(T),
                                                        fcs, pad);
    if (fwsb < 0) return c.fail("fit_stage: invalid events");
    if (!s.create(c.err) || !load_f32(c, src, X, s.get()) ||
        !tm.alloc(std::max<int64_t>(M * T * K, 1) * 4, c.err) ||
        !tws.alloc(msamd_templates_workspace_size(static_cast<int64_t>(gt.size()), K), c.err) ||
        !fws.alloc(fwsb, c.err))
        return false;
    std::vector<uint8_t> keep(static_cast<size_t>(std::max<int64_t>(L, 1)), 0);
    if (K > 0) {
        if (!lib_ok(c, msamd_compute_templates(X.as<float>(), M, N, gt.data(), gk.data(),
                                               static_cast<int64_t>(gt.size()), K,
                                               static_cast<int32_t>(T), tm.as<float>(), tws.get(),
                                               tws.bytes(), s.get()),
                    "compute_templates") ||
            !lib_ok(c, msamd_fit_stage(X.as<float>(), M, N, t.data(), k.data(), L, tm.as<float>(), K,
                                       static_cast<int32_t>(T), 0.1, fcs, pad, keep.data(), nullptr,
                                       fws.get(), fws.bytes(), s.get()),
                    "fit_stage"))
            return false;
    }
    std::vector<double> out;
    int64_t n = 0;
    for (int64_t i = 0; i < L; ++i)
        if (keep[i]) {
            out.insert(out.end(), all.begin() + R * order[i], all.begin() + R * (order[i] + 1));
            ++n;
        }
    return mda_write(c.str("firings_out"), MDA_FLOAT64, {R, n}, out.data(), c.err);
}

// --------------------------------------------------------------- the registry

std::vector<ProcessorSpec> make_registry() {
    std::vector<ProcessorSpec> v;
    {
        ProcessorSpec p;
        p.name = "ms3.bandpass_filter";
        p.version = "0.20";
        p.input("timeseries").output("timeseries_out");
        p.required("samplerate").required("freq_min").required("freq_max");
        p.optional("freq_wid", "", 1000LL).optional("quantization_unit", "", 0LL);
        p.optional("subsample_factor", "", 1LL);
        p.can_return_requirements = true;
        p.run = run_bandpass;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.whiten";
        p.version = "0.1";
        p.input("timeseries").output("timeseries_out");
        p.optional("quantization_unit", "", 0LL);
        p.can_return_requirements = true;
        p.run = run_whiten;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.compute_whitening_matrix";
        p.version = "0.11";
        p.input("timeseries_list").output("whitening_matrix_out");
        p.optional("channels");
        p.run = run_compute_whitening_matrix;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.whiten_clips";
        p.version = "0.1";
        p.input("clips").input("whitening_matrix").output("clips_out");
        p.optional("quantization_unit", "", 0LL);
        p.run = run_whiten_clips;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.apply_whitening_matrix";
        p.version = "0.1";
        p.input("timeseries").input("whitening_matrix").output("timeseries_out");
        p.optional("quantization_unit", "", 0LL);
        p.run = run_apply_whitening_matrix;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.extract_clips";
        p.version = "0.11";
        p.input("timeseries").input("event_times", "1xL array of event timestamps");
        p.output("clips_out");
        p.required("clip_size").optional("channels");
        p.run = run_extract_clips;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.mask_out_artifacts";
        p.version = "0.1";
        p.input("timeseries").output("timeseries_out");
        p.optional("threshold", "", 6LL).optional("interval_size", "", 2000LL);
        p.run = run_mask;
        v.push_back(p);
    }
    const char* kRefThreads =
        "Thread count of the reference run whose time-chunk plan is reproduced (0: "
        "_request_num_threads, else OMP_NUM_THREADS, else the available CPUs)";
    {
        ProcessorSpec p;
        p.name = "msamd.detect_events";
        p.version = "0.1";
        p.description =
            "Detection phase of mountainsortalg.ms3alg on every channel: rows 0-1 of its "
            "firings (central channel, time) in firings order";
        p.input("timeseries", "Preprocessed timeseries");
        p.output("event_times_out", "2xL array: central channel (1-based), time (0-based)");
        p.optional("detect_interval", "", 10LL).optional("detect_threshold", "", 3LL);
        p.optional("detect_sign", "", 0LL);
        p.optional("t1", "Start timepoint (default 0 means start at beginning)", 0LL);
        p.optional("t2", "End timepoint (default -1 means go to end of timeseries)", -1LL);
        p.optional("reference_num_threads", kRefThreads, 0LL);
        p.run = run_detect;
        v.push_back(p);
    }
    const char* kRadius =
        "Radius of local sorting neighborhood, corresponding to the geometry file. 0 means "
        "each channel is sorted independently. -1 means all channels are included in every "
        "neighborhood.";
    {
        ProcessorSpec p;
        p.name = "msamd.extract_neighborhood_clips";
        p.version = "0.1";
        p.description = "Clips of the events of one central channel over its neighborhood "
                        "channels (central channel first, then ascending)";
        p.input("timeseries").input("event_times", "2xL array (central channel, time)");
        p.input("geom", "", true);
        p.output("clips_out", "WxTxL float32");
        p.required("central_channel");
        p.optional("adjacency_radius", kRadius, -1LL).optional("clip_size", "", 50LL);
        p.run = run_neighborhood_clips;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "msamd.pca_features";
        p.version = "0.1";
        p.description = "Two-stage per-neighborhood PCA features (the isosplit5 input of "
                        "ms3alg) of every event";
        p.input("timeseries").input("event_times", "2xL array (central channel, time)");
        p.input("geom", "", true);
        p.output("features_out", "num_features x L float32, column i for event i");
        p.optional("adjacency_radius", kRadius, -1LL).optional("clip_size", "", 50LL);
        p.optional("num_features", "", 10LL).optional("num_features_per_channel", "", 10LL);
        p.optional("max_pca_samples", "", 10000LL);
        p.run = run_pca_features;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "ms3.compute_templates";
        p.version = "0.11";
        p.input("timeseries").input("firings");
        p.output("templates_out");
        p.required("clip_size");
        p.optional("clusters", "Comma-separated list of clusters to inclue", "");
        p.run = run_compute_templates;
        v.push_back(p);
    }
    {
        ProcessorSpec p;
        p.name = "msamd.fit_stage";
        p.version = "0.1";
        p.description = "Global templates and fit stage of mountainsortalg.ms3alg: the firings "
                        "(central channel, time, label) the template-subtraction fit keeps";
        p.input("timeseries", "Preprocessed timeseries");
        p.input("firings", "RxL firings, R >= 3 (central channel, time, label)");
        p.output("firings_out", "The kept columns of the firings, in time order");
        p.optional("clip_size", "", 50LL);
        p.optional("reference_num_threads", kRefThreads, 0LL);
        p.run = run_fit_stage;
        v.push_back(p);
    }
    return v;
}

}  // namespace

Value Ctx::param(const std::string& name) const {
    if (clp.has(name)) return clp.get(name);
    for (const auto& p : spec.params)
        if (p.name == name && p.has_default) return p.default_value;
    return Value();
}

const std::vector<ProcessorSpec>& processors() {
    static const std::vector<ProcessorSpec> reg = make_registry();
    return reg;
}

const ProcessorSpec* find_processor(const std::string& name) {
    for (const auto& p : processors())
        if (p.name == name) return &p;
    return nullptr;
}

Json processor_json(const ProcessorSpec& p, const std::string& exe) {
    Json j = Json::object();
    j["name"] = p.name;
    j["version"] = p.version;
    j["description"] = p.description;
    auto files = [](const std::vector<FileSpec>& fs) {
        Json a = Json::array();
        for (const auto& f : fs) {
            Json o = Json::object();
            o["name"] = f.name;
            o["description"] = f.description;
            o["optional"] = f.optional;
            a.push_back(o);
        }
        return a;
    };
    j["inputs"] = files(p.inputs);
    j["outputs"] = files(p.outputs);
    Json ps = Json::array();
    for (const auto& q : p.params) {
        Json o = Json::object();
        o["name"] = q.name;
        o["description"] = q.description;
        o["optional"] = q.optional;
        if (q.has_default) o["default_value"] = q.default_json;
        ps.push_back(o);
    }
    j["parameters"] = ps;
    j["exe_command"] = exe + " " + p.name + " $(arguments)";
    if (p.can_return_requirements)
        j["requirements_command"] = exe + " requirements " + p.name + " $(arguments)";
    return j;
}

Json spec_json(const std::string& exe) {
    Json arr = Json::array();
    for (const auto& p : processors()) arr.push_back(processor_json(p, exe));
    Json j = Json::object();
    j["processors"] = arr;
    return j;
}

}  // namespace msamd_host
