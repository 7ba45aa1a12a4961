// .mda array files: header, whole-array reads, streamed writes with
// tmp-then-rename, concatenation along time, and the CSV geometry reader.
//
// Format (packages/ms3/mda/mdaio.cpp:50-199): int32 type code, int32 bytes
// per entry, int32 number of dimensions (negative: the dimensions follow as
// uint64, else int32), then the data in column-major order (first index
// fastest).  Type codes: -2 uint8, -3 float32, -4 int16, -5 int32, -6 uint16,
// -7 float64, -8 uint32 (mdaio.h).
#pragma once

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

namespace msamd_host {

enum MdaType : int32_t {
    MDA_BYTE = -2,
    MDA_FLOAT32 = -3,
    MDA_INT16 = -4,
    MDA_INT32 = -5,
    MDA_UINT16 = -6,
    MDA_FLOAT64 = -7,
    MDA_UINT32 = -8,
};

int mda_bytes_per_entry(int32_t type);

struct MdaHeader {
    int32_t type = 0;
    int32_t bytes_per_entry = 0;
    std::vector<int64_t> dims;
    int64_t header_size = 0;
    int64_t dim(int i) const { return i < static_cast<int>(dims.size()) ? dims[i] : 1; }
    int64_t size() const {
        int64_t n = 1;
        for (int64_t d : dims) n *= d;
        return n;
    }
};

// Reads the header; on failure returns false with a message in err.
bool mda_read_header(FILE* f, MdaHeader& h, std::string& err);
bool mda_read_header(const std::string& path, MdaHeader& h, std::string& err);

// Converts n entries of type `type` at src to float32 / float64 (the
// std::copy conversion of mdaReadData_impl, mdaio.cpp:201-214).
void mda_convert_to_f32(int32_t type, const void* src, float* dst, int64_t n);
void mda_convert_to_f64(int32_t type, const void* src, double* dst, int64_t n);

// Whole-array reads, converted (small arrays: event times, matrices, geometry).
bool mda_read_f64(const std::string& path, std::vector<double>& out, MdaHeader& h,
                  std::string& err);
bool mda_read_f32(const std::string& path, std::vector<float>& out, MdaHeader& h,
                  std::string& err);

// Reads and writes of at least this many bytes are split over
// kParallelIoThreads threads (positioned I/O on the same file).
constexpr int64_t kParallelIoBytes = 32ll << 20;
constexpr int kParallelIoThreads = 4;

// Streamed writer: writes path + ".tmp", renamed to path by close()
// (packages/ms3/mda/diskwritemda.cpp:84,142-146).  The number of dimensions
// written drops trailing 1s (minimum 2); 64-bit dimensions are used when any
// exceeds 2e9 (mdaio.cpp:160-165).
class MdaWriter {
public:
    ~MdaWriter();
    bool open(const std::string& path, int32_t type, const std::vector<int64_t>& dims,
              std::string& err);
    bool write(const void* data, int64_t bytes, std::string& err);
    bool close(std::string& err);
    void abort();

private:
    FILE* f_ = nullptr;
    std::string path_;
    int64_t expected_ = 0, written_ = 0;
};

// Writes a whole small array.
bool mda_write(const std::string& path, int32_t type, const std::vector<int64_t>& dims,
               const void* data, std::string& err);

// An M x N timeseries that may be one file, a directory of .mda files, or a
// list of files concatenated along time (DiskReadMda32 with concatenation,
// packages/ms3/mda/diskreadmda32.cpp:235-288,896-913).  A directory
// contributes its *.mda files in name order; the reference's recursion into
// subdirectories looks under a fixed wrong path and adds nothing, so none is
// done here either.
struct TimeseriesSource {
    std::vector<std::string> files;
    std::vector<MdaHeader> headers;
    int64_t M = 0, N = 0;
    int32_t common_type = 0;  // the type shared by every file, 0 if mixed

    bool open(const std::vector<std::string>& paths, std::string& err);
};

// Mda::readCsv (packages/mountainsortalg/mda/mda_p.h:179-215): one line per
// channel; commas or whitespace separate values; a first line whose first
// value is not a number is a header and skipped.  Returns coordinates
// coords[channel][d] (D = values on the first data line).
bool read_geom_csv(const std::string& path, std::vector<std::vector<double>>& coords,
                   std::string& err);

bool is_directory(const std::string& path);

}  // namespace msamd_host
