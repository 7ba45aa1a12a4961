// Device memory and streaming I/O (see device.h).
#include "device.h"

#include <algorithm>
#include <cstdio>
#include <cstring>

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
    std::vector<char> raw;
    char* out = static_cast<char*>(dev);
    int cur = 0;
    bool used[2] = {false, false};
    for (size_t fi = 0; fi < src.files.size(); ++fi) {
        const MdaHeader& h = src.headers[fi];
        FILE* f = fopen(src.files[fi].c_str(), "rb");
        if (!f) {
            err = "unable to open " + src.files[fi];
            return false;
        }
        if (fseeko(f, h.header_size, SEEK_SET) != 0) {
            fclose(f);
            err = "seek failed in " + src.files[fi];
            return false;
        }
        const int sb = h.bytes_per_entry;
        const int64_t per = kStage / std::max(sb, tb);
        int64_t left = h.dim(0) * h.dim(1);
        while (left > 0) {
            const int64_t n = std::min(per, left);
            if (used[cur] && !hip_ok(hipEventSynchronize(done[cur].e), "hipEventSynchronize", err)) {
                fclose(f);
                return false;
            }
            void* dst = buf[cur].p;
            size_t got;
            if (h.type == target) {
                got = fread(dst, sb, static_cast<size_t>(n), f);
            } else {
                raw.resize(static_cast<size_t>(n * sb));
                got = fread(raw.data(), sb, static_cast<size_t>(n), f);
                mda_convert_to_f32(h.type, raw.data(), static_cast<float*>(dst), n);
            }
            if (static_cast<int64_t>(got) != n) {
                fclose(f);
                err = "truncated data in " + src.files[fi];
                return false;
            }
            if (!hip_ok(hipMemcpyAsync(out, dst, n * tb, hipMemcpyHostToDevice, s),
                        "hipMemcpyAsync", err) ||
                !hip_ok(hipEventRecord(done[cur].e, s), "hipEventRecord", err)) {
                fclose(f);
                return false;
            }
            used[cur] = true;
            out += n * tb;
            left -= n;
            cur ^= 1;
        }
        fclose(f);
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

}  // namespace msamd_host
