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

}  // namespace msamd_host
