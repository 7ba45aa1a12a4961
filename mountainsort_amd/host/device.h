// Device memory, streams and host<->device streaming of .mda arrays for the
// processor executable.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <string>
#include <vector>

#include "mda.h"

namespace msamd_host {

// Owning device allocation (hipMalloc / hipFree).
class DeviceBuffer {
public:
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { reset(); }
    bool alloc(int64_t bytes, std::string& err);
    void reset();
    void* get() const { return p_; }
    int64_t bytes() const { return n_; }
    template <typename T>
    T* as() const { return static_cast<T*>(p_); }

private:
    void* p_ = nullptr;
    int64_t n_ = 0;
};

// A stream at the default priority, destroyed with the object.
class Stream {
public:
    Stream() = default;
    ~Stream();
    bool create(std::string& err);
    hipStream_t get() const { return s_; }
    bool sync(std::string& err);

private:
    hipStream_t s_ = nullptr;
};

bool hip_ok(hipError_t e, const char* what, std::string& err);

// Streams a timeseries (all files in order) into device memory as float32
// (target MDA_FLOAT32) or int16 (target MDA_INT16, only when every file is
// int16).  Reads go through two pinned staging buffers so that the copy of
// one block overlaps the read of the next.
bool upload_timeseries(const TimeseriesSource& src, int32_t target, void* dev, hipStream_t s,
                       std::string& err);

// Streams n_bytes of device memory into an open writer, block by block.
bool download_to_writer(const void* dev, int64_t n_bytes, MdaWriter& w, hipStream_t s,
                        std::string& err);

// Small transfers.
bool to_device(void* dev, const void* host, int64_t bytes, hipStream_t s, std::string& err);
bool to_host(void* host, const void* dev, int64_t bytes, hipStream_t s, std::string& err);

// Host bytes held by the staging buffers of the two streaming helpers.
int64_t staging_bytes();

// Reads time rows [t0, t1) of a timeseries (its files concatenated along
// time) as `target` entries (MDA_FLOAT32, converted as mda_convert_to_f32;
// or MDA_INT16 when every file is int16) into dst, M entries per row.  Rows
// outside [0, N) read as zero (DiskReadMda32::readChunk,
// packages/ms3/mda/diskreadmda32.cpp:631-650).  Files stay open between
// calls.
class RowReader {
public:
    explicit RowReader(const TimeseriesSource& src) : src_(src) {}
    ~RowReader();
    RowReader(const RowReader&) = delete;
    RowReader& operator=(const RowReader&) = delete;
    bool read(int32_t target, int64_t t0, int64_t t1, void* dst, std::string& err);

private:
    const TimeseriesSource& src_;
    std::vector<FILE*> f_;
    std::vector<char> raw_;
};

// A recording streamed through the device one segment at a time.  For
// segment k the host fills a pinned input buffer (prepare), the pipeline
// copies it to the device on an upload stream, runs compute on a compute
// stream, copies the result back on a download stream, and hands it to the
// host (consume).  Two buffers of each kind let the file read of segment k+1
// and the file write of segment k-1 run while the device works on segment k.
class SegmentPipeline {
public:
    // prepare: fill host_in with *in_bytes bytes (0: nothing to upload); it
    // may set *dev_dst to upload into caller-owned device memory instead of
    // the pipeline's input buffer.  compute: dev_in is where the input went
    // (nullptr when nothing was uploaded); set *out_bytes of dev_out to
    // download.  consume: the downloaded bytes.
    struct Stage {
        bool (*prepare)(void* ctx, int64_t k, void* host_in, int64_t* in_bytes, void** dev_dst,
                        std::string& err);
        bool (*compute)(void* ctx, int64_t k, const void* dev_in, void* dev_out,
                        int64_t* out_bytes, hipStream_t s, std::string& err);
        bool (*consume)(void* ctx, int64_t k, const void* host_out, int64_t out_bytes,
                        std::string& err);
        void* ctx;
    };
    ~SegmentPipeline();
    // Buffers for segments of at most max_in input and max_out output bytes;
    // dev_in_buffers: false when every upload goes to caller memory.
    bool init(int64_t max_in, int64_t max_out, bool dev_in_buffers, std::string& err);
    bool run(int64_t K, const Stage& st, std::string& err);

private:
    int64_t max_in_ = 0, max_out_ = 0;
    void* host_in_[2] = {nullptr, nullptr};
    void* host_out_[2] = {nullptr, nullptr};
    DeviceBuffer dev_in_[2], dev_out_[2];
    hipStream_t su_ = nullptr, sc_ = nullptr, sd_ = nullptr;
    hipEvent_t up_[2] = {nullptr, nullptr}, comp_[2] = {nullptr, nullptr},
               down_[2] = {nullptr, nullptr};
};

// Bytes of one streamed segment: MSAMD_SEGMENT_BYTES in the environment, or
// dflt.
int64_t segment_bytes(int64_t dflt);

// Device memory free for this process now.
int64_t device_free_bytes();

}  // namespace msamd_host
