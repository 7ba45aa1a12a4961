// Device memory and streaming I/O (see device.h).
#include "device.h"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <thread>

#include <unistd.h>

namespace msamd_host {

namespace {

constexpr int64_t kStage = 64ll << 20;  // bytes per pinned staging buffer

struct Pinned {
    void* p = nullptr;
    ~Pinned() {
        if (p) (void)hipHostFree(p);
    }
};

struct Event {
    hipEvent_t e = nullptr;
    ~Event() {
        if (e) (void)hipEventDestroy(e);
    }
};

}  // namespace

bool hip_ok(hipError_t e, const char* what, std::string& err) {
    if (e == hipSuccess) return true;
    err = std::string(what) + " failed: " + hipGetErrorString(e);
    return false;
}

bool DeviceBuffer::alloc(int64_t bytes, std::string& err) {
    reset();
    if (bytes <= 0) bytes = 256;
    if (!hip_ok(hipMalloc(&p_, static_cast<size_t>(bytes)), "hipMalloc", err)) {
        p_ = nullptr;
        return false;
    }
    n_ = bytes;
    return true;
}

void DeviceBuffer::reset() {
    if (p_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
}

Stream::~Stream() {
    if (s_) (void)hipStreamDestroy(s_);
}

bool Stream::create(std::string& err) { return hip_ok(hipStreamCreate(&s_), "hipStreamCreate", err); }

bool Stream::sync(std::string& err) {
    return hip_ok(hipStreamSynchronize(s_), "hipStreamSynchronize", err);
}

int64_t staging_bytes() { return 2 * kStage + kStage; }

bool upload_timeseries(const TimeseriesSource& src, int32_t target, void* dev, hipStream_t s,
                       std::string& err) {
    const int tb = mda_bytes_per_entry(target);
    if (target == MDA_INT16 && src.common_type != MDA_INT16) {
        err = "internal: int16 upload of a non-int16 timeseries";
        return false;
    }
    Pinned buf[2];
    Event done[2];
    for (int b = 0; b < 2; ++b) {
        if (!hip_ok(hipHostMalloc(&buf[b].p, kStage, hipHostMallocDefault), "hipHostMalloc", err))
            return false;
        if (!hip_ok(hipEventCreateWithFlags(&done[b].e, hipEventDisableTiming), "hipEventCreate",
                    err))
            return false;
    }
    RowReader rd(src);
    const int64_t row = std::max<int64_t>(1, src.M * tb);
    const int64_t per = std::max<int64_t>(1, kStage / row);
    char* out = static_cast<char*>(dev);
    int cur = 0;
    bool used[2] = {false, false};
    // read block k while block k-1 is copied
    for (int64_t t = 0; t < src.N; t += per) {
        const int64_t t1 = std::min(src.N, t + per);
        if (used[cur] && !hip_ok(hipEventSynchronize(done[cur].e), "hipEventSynchronize", err))
            return false;
        if (!rd.read(target, t, t1, buf[cur].p, err)) return false;
        const int64_t n = (t1 - t) * row;
        if (!hip_ok(hipMemcpyAsync(out, buf[cur].p, static_cast<size_t>(n), hipMemcpyHostToDevice, s),
                    "hipMemcpyAsync", err) ||
            !hip_ok(hipEventRecord(done[cur].e, s), "hipEventRecord", err))
            return false;
        used[cur] = true;
        out += n;
        cur ^= 1;
    }
    return hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize", err);
}

bool download_to_writer(const void* dev, int64_t n_bytes, MdaWriter& w, hipStream_t s,
                        std::string& err) {
    Pinned buf[2];
    Event done[2];
    for (int b = 0; b < 2; ++b) {
        if (!hip_ok(hipHostMalloc(&buf[b].p, kStage, hipHostMallocDefault), "hipHostMalloc", err))
            return false;
        if (!hip_ok(hipEventCreateWithFlags(&done[b].e, hipEventDisableTiming), "hipEventCreate",
                    err))
            return false;
    }
    const char* in = static_cast<const char*>(dev);
    int64_t off = 0;
    int64_t pending[2] = {0, 0};
    int cur = 0;
    // copy block k while block k-1 is written to the file
    while (off < n_bytes || pending[cur ^ 1] > 0) {
        if (off < n_bytes) {
            const int64_t n = std::min(kStage, n_bytes - off);
            if (!hip_ok(hipMemcpyAsync(buf[cur].p, in + off, n, hipMemcpyDeviceToHost, s),
                        "hipMemcpyAsync", err) ||
                !hip_ok(hipEventRecord(done[cur].e, s), "hipEventRecord", err))
                return false;
            pending[cur] = n;
            off += n;
        }
        const int prev = cur ^ 1;
        if (pending[prev] > 0) {
            if (!hip_ok(hipEventSynchronize(done[prev].e), "hipEventSynchronize", err)) return false;
            if (!w.write(buf[prev].p, pending[prev], err)) return false;
            pending[prev] = 0;
        }
        cur ^= 1;
    }
    return true;
}

bool to_device(void* dev, const void* host, int64_t bytes, hipStream_t s, std::string& err) {
    if (bytes <= 0) return true;
    return hip_ok(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync",
                  err) &&
           hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize", err);
}

bool to_host(void* host, const void* dev, int64_t bytes, hipStream_t s, std::string& err) {
    if (bytes <= 0) return true;
    return hip_ok(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, s), "hipMemcpyAsync",
                  err) &&
           hip_ok(hipStreamSynchronize(s), "hipStreamSynchronize", err);
}


RowReader::~RowReader() {
    for (FILE* f : f_)
        if (f) fclose(f);
}

bool RowReader::read(int32_t target, int64_t t0, int64_t t1, void* dst, std::string& err) {
    const int tb = mda_bytes_per_entry(target);
    const int64_t M = src_.M;
    if (target == MDA_INT16 && src_.common_type != MDA_INT16) {
        err = "internal: int16 read of a non-int16 timeseries";
        return false;
    }
    if (f_.empty()) f_.assign(src_.files.size(), nullptr);
    char* out = static_cast<char*>(dst);
    // rows before 0 and at or past N are zero
    if (t0 < 0) {
        const int64_t z = std::min<int64_t>(t1, 0) - t0;
        std::memset(out, 0, static_cast<size_t>(z * M * tb));
        out += z * M * tb;
        t0 += z;
    }
    int64_t file_t0 = 0;
    for (size_t fi = 0; fi < src_.files.size() && t0 < t1; ++fi) {
        const MdaHeader& h = src_.headers[fi];
        const int64_t n = h.dim(1);
        const int64_t lo = std::max(t0, file_t0), hi = std::min(t1, file_t0 + n);
        if (hi > lo) {
            if (!f_[fi]) {
                f_[fi] = fopen(src_.files[fi].c_str(), "rb");
                if (!f_[fi]) {
                    err = "unable to open " + src_.files[fi];
                    return false;
                }
            }
            const int fd = fileno(f_[fi]);
            const int sb = h.bytes_per_entry;
            const int64_t cnt = (hi - lo) * M;
            const off_t base = h.header_size + (lo - file_t0) * M * sb;
            // positioned reads, split over threads for large ranges (each
            // thread converts its own part)
            const int P = cnt * sb >= kParallelIoBytes ? kParallelIoThreads : 1;
            const int64_t part = (cnt + P - 1) / P;
            if (h.type != target) raw_.resize(static_cast<size_t>(cnt * sb));
            std::vector<char> ok(P, 1);
            auto work = [&](int i) {
                const int64_t a = i * part, b = std::min(cnt, a + part);
                if (b <= a) return;
                char* dst = h.type == target ? out + a * sb : raw_.data() + a * sb;
                const int64_t want = (b - a) * sb;
                int64_t done = 0;
                while (done < want) {
                    const ssize_t r = pread(fd, dst + done, static_cast<size_t>(want - done),
                                            base + a * sb + done);
                    if (r <= 0) {
                        ok[i] = 0;
                        return;
                    }
                    done += r;
                }
                if (h.type != target)
                    mda_convert_to_f32(h.type, dst, reinterpret_cast<float*>(out) + a, b - a);
            };
            if (P == 1) {
                work(0);
            } else {
                std::vector<std::thread> th;
                for (int i = 0; i < P; ++i) th.emplace_back(work, i);
                for (auto& t : th) t.join();
            }
            if (std::find(ok.begin(), ok.end(), 0) != ok.end()) {
                err = "truncated data in " + src_.files[fi];
                return false;
            }
            out += cnt * tb;
            t0 = hi;
        }
        file_t0 += n;
    }
    if (t1 > t0) std::memset(out, 0, static_cast<size_t>((t1 - t0) * M * tb));
    return true;
}

SegmentPipeline::~SegmentPipeline() {
    for (int b = 0; b < 2; ++b) {
        if (host_in_[b]) (void)hipHostFree(host_in_[b]);
        if (host_out_[b]) (void)hipHostFree(host_out_[b]);
        if (up_[b]) (void)hipEventDestroy(up_[b]);
        if (comp_[b]) (void)hipEventDestroy(comp_[b]);
        if (down_[b]) (void)hipEventDestroy(down_[b]);
    }
    if (su_) (void)hipStreamDestroy(su_);
    if (sc_) (void)hipStreamDestroy(sc_);
    if (sd_) (void)hipStreamDestroy(sd_);
}

bool SegmentPipeline::init(int64_t max_in, int64_t max_out, bool dev_in_buffers,
                           std::string& err) {
    max_in_ = std::max<int64_t>(max_in, 256);
    max_out_ = std::max<int64_t>(max_out, 256);
    for (int b = 0; b < 2; ++b) {
        if (!hip_ok(hipHostMalloc(&host_in_[b], static_cast<size_t>(max_in_), hipHostMallocDefault),
                    "hipHostMalloc", err) ||
            !hip_ok(hipHostMalloc(&host_out_[b], static_cast<size_t>(max_out_), hipHostMallocDefault),
                    "hipHostMalloc", err) ||
            (dev_in_buffers && !dev_in_[b].alloc(max_in_, err)) || !dev_out_[b].alloc(max_out_, err) ||
            !hip_ok(hipEventCreateWithFlags(&up_[b], hipEventDisableTiming), "hipEventCreate", err) ||
            !hip_ok(hipEventCreateWithFlags(&comp_[b], hipEventDisableTiming), "hipEventCreate", err) ||
            !hip_ok(hipEventCreateWithFlags(&down_[b], hipEventDisableTiming), "hipEventCreate", err))
            return false;
    }
    return hip_ok(hipStreamCreateWithFlags(&su_, hipStreamNonBlocking), "hipStreamCreate", err) &&
           hip_ok(hipStreamCreateWithFlags(&sc_, hipStreamNonBlocking), "hipStreamCreate", err) &&
           hip_ok(hipStreamCreateWithFlags(&sd_, hipStreamNonBlocking), "hipStreamCreate", err);
}

bool SegmentPipeline::run(int64_t K, const Stage& st, std::string& err) {
    const void* dev_src[2] = {nullptr, nullptr};
    int64_t out_n[2] = {0, 0};
    bool used_up[2] = {false, false}, used_comp[2] = {false, false}, used_down[2] = {false, false};
    // consume (the file write) of segment j runs on its own thread while the
    // main thread reads segment j+2; it is joined before host_out[j & 1] is
    // filled again
    std::thread writer;
    bool wrote_ok = true;
    std::string werr;
    auto join_writer = [&]() {
        if (writer.joinable()) writer.join();
        if (!wrote_ok && err.empty()) err = werr;
        return wrote_ok;
    };
    auto drain = [&]() {
        join_writer();
        (void)hipStreamSynchronize(su_);
        (void)hipStreamSynchronize(sc_);
        (void)hipStreamSynchronize(sd_);
    };
    // iteration k: read + upload segment k, compute + download segment k-1,
    // write segment k-2
    for (int64_t k = 0; k < K + 2; ++k) {
        if (k < K) {
            const int b = static_cast<int>(k & 1);
            // the pinned buffer's previous upload (segment k-2) is complete
            if (used_up[b] && !hip_ok(hipEventSynchronize(up_[b]), "hipEventSynchronize", err)) {
                drain();
                return false;
            }
            int64_t in_bytes = 0;
            void* dst = nullptr;
            if (!st.prepare(st.ctx, k, host_in_[b], &in_bytes, &dst, err)) {
                drain();
                return false;
            }
            if (in_bytes > max_in_) {
                drain();
                err = "internal: segment input larger than its buffer";
                return false;
            }
            dev_src[b] = nullptr;
            if (in_bytes > 0) {
                if (!dst) dst = dev_in_[b].get();
                // the device input buffer's previous reader (segment k-2) is done
                if ((used_comp[b] && !hip_ok(hipStreamWaitEvent(su_, comp_[b], 0), "hipStreamWaitEvent", err)) ||
                    !hip_ok(hipMemcpyAsync(dst, host_in_[b], static_cast<size_t>(in_bytes),
                                           hipMemcpyHostToDevice, su_),
                            "hipMemcpyAsync", err) ||
                    !hip_ok(hipEventRecord(up_[b], su_), "hipEventRecord", err)) {
                    drain();
                    return false;
                }
                used_up[b] = true;
                dev_src[b] = dst;
            }
        }
        if (k >= 1 && k - 1 < K) {
            const int64_t j = k - 1;
            const int b = static_cast<int>(j & 1);
            int64_t ob = 0;
            if ((dev_src[b] && !hip_ok(hipStreamWaitEvent(sc_, up_[b], 0), "hipStreamWaitEvent", err)) ||
                (used_down[b] && !hip_ok(hipStreamWaitEvent(sc_, down_[b], 0), "hipStreamWaitEvent", err)) ||
                !st.compute(st.ctx, j, dev_src[b], dev_out_[b].get(), &ob, sc_, err) ||
                !hip_ok(hipEventRecord(comp_[b], sc_), "hipEventRecord", err)) {
                drain();
                return false;
            }
            used_comp[b] = true;
            if (ob > max_out_) {
                drain();
                err = "internal: segment output larger than its buffer";
                return false;
            }
            out_n[b] = ob;
            if (ob > 0) {
                // the writer of segment j-2 (the same host buffer) has finished
                if (!join_writer()) {
                    drain();
                    return false;
                }
                if (!hip_ok(hipStreamWaitEvent(sd_, comp_[b], 0), "hipStreamWaitEvent", err) ||
                    !hip_ok(hipMemcpyAsync(host_out_[b], dev_out_[b].get(), static_cast<size_t>(ob),
                                           hipMemcpyDeviceToHost, sd_),
                            "hipMemcpyAsync", err) ||
                    !hip_ok(hipEventRecord(down_[b], sd_), "hipEventRecord", err)) {
                    drain();
                    return false;
                }
                used_down[b] = true;
            }
        }
        if (k >= 2) {
            const int64_t j = k - 2;
            const int b = static_cast<int>(j & 1);
            const hipEvent_t done = out_n[b] > 0 ? down_[b] : comp_[b];
            // segments are written in order: the previous writer finishes first
            if (!join_writer() || !hip_ok(hipEventSynchronize(done), "hipEventSynchronize", err)) {
                drain();
                return false;
            }
            if (st.consume) {
                const void* buf = host_out_[b];
                const int64_t nb = out_n[b];
                writer = std::thread([&st, &wrote_ok, &werr, j, buf, nb]() {
                    wrote_ok = st.consume(st.ctx, j, buf, nb, werr);
                });
            }
        }
    }
    if (!join_writer()) {
        drain();
        return false;
    }
    return hip_ok(hipStreamSynchronize(sc_), "hipStreamSynchronize", err);
}

int64_t segment_bytes(int64_t dflt) {
    if (const char* e = getenv("MSAMD_SEGMENT_BYTES")) {
        const long long v = atoll(e);
        if (v > 0) return v;
    }
    return dflt;
}

int64_t device_free_bytes() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return 0;
    return static_cast<int64_t>(fr);
}

}  // namespace msamd_host
