// .mda file I/O (see mda.h).
#include "mda.h"

#include <dirent.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <thread>
#include <vector>

namespace msamd_host {

int mda_bytes_per_entry(int32_t type) {
    switch (type) {
        case MDA_BYTE: return 1;
        case MDA_FLOAT32: return 4;
        case MDA_INT16: return 2;
        case MDA_INT32: return 4;
        case MDA_UINT16: return 2;
        case MDA_FLOAT64: return 8;
        case MDA_UINT32: return 4;
        default: return 0;
    }
}

bool mda_read_header(FILE* f, MdaHeader& h, std::string& err) {
    h = MdaHeader();
    int32_t code, nbytes, ndims;
    if (fread(&code, 4, 1, f) != 1) {
        err = "problem reading .mda header";
        return false;
    }
    // The reference accepts codes -7..-1 (mdaio.cpp:88-91); of those, -1
    // (complex) has no reader, so it is refused here as well.
    if (code < -7 || code >= -1) {
        err = "unsupported .mda data type " + std::to_string(code);
        return false;
    }
    if (fread(&nbytes, 4, 1, f) != 1 || nbytes <= 0 || nbytes > 8) {
        err = "bad .mda bytes-per-entry field";
        return false;
    }
    if (fread(&ndims, 4, 1, f) != 1 || ndims == 0 || ndims > 50 || ndims < -50) {
        err = "bad .mda number of dimensions";
        return false;
    }
    const bool d64 = ndims < 0;
    if (d64) ndims = -ndims;
    for (int i = 0; i < ndims; ++i) {
        int64_t d;
        if (d64) {
            uint64_t v;
            if (fread(&v, 8, 1, f) != 1) {
                err = "truncated .mda header";
                return false;
            }
            d = static_cast<int64_t>(v);
        } else {
            int32_t v;
            if (fread(&v, 4, 1, f) != 1) {
                err = "truncated .mda header";
                return false;
            }
            d = v;
        }
        if (d < 0) {
            err = "negative .mda dimension";
            return false;
        }
        h.dims.push_back(d);
    }
    h.type = code;
    h.bytes_per_entry = mda_bytes_per_entry(code);  // the proper value, as mdaio.cpp:100-103
    h.header_size = 12 + static_cast<int64_t>(ndims) * (d64 ? 8 : 4);
    return true;
}

bool mda_read_header(const std::string& path, MdaHeader& h, std::string& err) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        err = "unable to open " + path;
        return false;
    }
    const bool ok = mda_read_header(f, h, err);
    fclose(f);
    if (!ok) err = path + ": " + err;
    return ok;
}

template <typename D>
static void convert(int32_t type, const void* src, D* dst, int64_t n) {
    switch (type) {
        case MDA_BYTE: std::copy_n(static_cast<const uint8_t*>(src), n, dst); break;
        case MDA_FLOAT32: std::copy_n(static_cast<const float*>(src), n, dst); break;
        case MDA_INT16: std::copy_n(static_cast<const int16_t*>(src), n, dst); break;
        case MDA_INT32: std::copy_n(static_cast<const int32_t*>(src), n, dst); break;
        case MDA_UINT16: std::copy_n(static_cast<const uint16_t*>(src), n, dst); break;
        case MDA_FLOAT64: std::copy_n(static_cast<const double*>(src), n, dst); break;
        case MDA_UINT32: std::copy_n(static_cast<const uint32_t*>(src), n, dst); break;
        default: std::fill_n(dst, n, D(0)); break;
    }
}

void mda_convert_to_f32(int32_t type, const void* src, float* dst, int64_t n) {
    convert<float>(type, src, dst, n);
}

void mda_convert_to_f64(int32_t type, const void* src, double* dst, int64_t n) {
    convert<double>(type, src, dst, n);
}

bool mda_read_f64(const std::string& path, std::vector<double>& out, MdaHeader& h,
                  std::string& err) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        err = "unable to open " + path;
        return false;
    }
    if (!mda_read_header(f, h, err)) {
        fclose(f);
        err = path + ": " + err;
        return false;
    }
    const int64_t n = h.size();
    std::vector<char> raw(static_cast<size_t>(n * h.bytes_per_entry));
    const size_t got = n ? fread(raw.data(), h.bytes_per_entry, n, f) : 0;
    fclose(f);
    if (static_cast<int64_t>(got) != n) {
        err = path + ": truncated data";
        return false;
    }
    out.resize(n);
    mda_convert_to_f64(h.type, raw.data(), out.data(), n);
    return true;
}

bool mda_read_f32(const std::string& path, std::vector<float>& out, MdaHeader& h,
                  std::string& err) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) {
        err = "unable to open " + path;
        return false;
    }
    if (!mda_read_header(f, h, err)) {
        fclose(f);
        err = path + ": " + err;
        return false;
    }
    const int64_t n = h.size();
    std::vector<char> raw(static_cast<size_t>(n * h.bytes_per_entry));
    const size_t got = n ? fread(raw.data(), h.bytes_per_entry, n, f) : 0;
    fclose(f);
    if (static_cast<int64_t>(got) != n) {
        err = path + ": truncated data";
        return false;
    }
    out.resize(n);
    mda_convert_to_f32(h.type, raw.data(), out.data(), n);
    return true;
}

MdaWriter::~MdaWriter() { abort(); }

bool MdaWriter::open(const std::string& path, int32_t type, const std::vector<int64_t>& dims_in,
                     std::string& err) {
    std::vector<int64_t> dims = dims_in;
    while (dims.size() > 2 && dims.back() == 1) dims.pop_back();
    while (dims.size() < 2) dims.push_back(1);
    bool d64 = false;
    for (int64_t d : dims) d64 = d64 || d > 2000000000LL;
    path_ = path;
    f_ = fopen((path + ".tmp").c_str(), "wb");
    if (!f_) {
        err = "unable to open " + path + ".tmp for writing";
        return false;
    }
    const int32_t nb = mda_bytes_per_entry(type);
    const int32_t nd = d64 ? -static_cast<int32_t>(dims.size()) : static_cast<int32_t>(dims.size());
    bool ok = fwrite(&type, 4, 1, f_) == 1 && fwrite(&nb, 4, 1, f_) == 1 && fwrite(&nd, 4, 1, f_) == 1;
    for (int64_t d : dims) {
        if (d64) {
            const uint64_t v = static_cast<uint64_t>(d);
            ok = ok && fwrite(&v, 8, 1, f_) == 1;
        } else {
            const int32_t v = static_cast<int32_t>(d);
            ok = ok && fwrite(&v, 4, 1, f_) == 1;
        }
    }
    if (!ok) {
        err = "problem writing header of " + path;
        abort();
        return false;
    }
    expected_ = nb;
    for (int64_t d : dims) expected_ *= d;
    written_ = 0;
    return true;
}

bool MdaWriter::write(const void* data, int64_t bytes, std::string& err) {
    if (!f_) {
        err = "writer is not open";
        return false;
    }
    if (bytes >= kParallelIoBytes) {
        // large blocks: positioned writes from several threads at once (the
        // page-cache copy is the cost, and one thread moves ~5 GB/s)
        if (fflush(f_) != 0) {
            err = "problem writing " + path_ + ".tmp";
            return false;
        }
        const off_t off = ftello(f_);
        const int fd = fileno(f_);
        const int P = kParallelIoThreads;
        const int64_t part = (bytes + P - 1) / P;
        std::vector<char> ok(P, 1);
        std::vector<std::thread> th;
        for (int i = 0; i < P; ++i) {
            const int64_t a = i * part, b = std::min(bytes, a + part);
            if (b <= a) break;
            th.emplace_back([&, i, a, b]() {
                const char* src = static_cast<const char*>(data) + a;
                int64_t done = 0;
                while (done < b - a) {
                    const ssize_t w = pwrite(fd, src + done, static_cast<size_t>(b - a - done),
                                             off + a + done);
                    if (w <= 0) {
                        ok[i] = 0;
                        return;
                    }
                    done += w;
                }
            });
        }
        for (auto& t : th) t.join();
        if (std::find(ok.begin(), ok.end(), 0) != ok.end() || fseeko(f_, off + bytes, SEEK_SET) != 0) {
            err = "problem writing " + path_ + ".tmp";
            return false;
        }
    } else if (bytes > 0 &&
               fwrite(data, 1, static_cast<size_t>(bytes), f_) != static_cast<size_t>(bytes)) {
        err = "problem writing " + path_ + ".tmp";
        return false;
    }
    written_ += bytes;
    return true;
}

bool MdaWriter::close(std::string& err) {
    if (!f_) {
        err = "writer is not open";
        return false;
    }
    if (written_ != expected_) {
        err = "internal: wrote " + std::to_string(written_) + " of " + std::to_string(expected_) +
              " bytes of " + path_;
        abort();
        return false;
    }
    const bool ok = fclose(f_) == 0;
    f_ = nullptr;
    if (!ok || std::rename((path_ + ".tmp").c_str(), path_.c_str()) != 0) {
        err = "unable to rename " + path_ + ".tmp to " + path_;
        std::remove((path_ + ".tmp").c_str());
        return false;
    }
    return true;
}

void MdaWriter::abort() {
    if (f_) {
        fclose(f_);
        f_ = nullptr;
        std::remove((path_ + ".tmp").c_str());
    }
}

bool mda_write(const std::string& path, int32_t type, const std::vector<int64_t>& dims,
               const void* data, std::string& err) {
    MdaWriter w;
    int64_t n = mda_bytes_per_entry(type);
    for (int64_t d : dims) n *= d;
    return w.open(path, type, dims, err) && w.write(data, n, err) && w.close(err);
}

bool is_directory(const std::string& path) {
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

static std::vector<std::string> mda_files_in(const std::string& dir) {
    std::vector<std::string> names;
    if (DIR* d = opendir(dir.c_str())) {
        while (dirent* e = readdir(d)) {
            const std::string n = e->d_name;
            if (n.size() >= 4 && n.compare(n.size() - 4, 4, ".mda") == 0 &&
                !is_directory(dir + "/" + n))
                names.push_back(n);
        }
        closedir(d);
    }
    std::sort(names.begin(), names.end());
    for (auto& n : names) n = dir + "/" + n;
    return names;
}

bool TimeseriesSource::open(const std::vector<std::string>& paths, std::string& err) {
    files.clear();
    headers.clear();
    for (const auto& p : paths) {
        if (is_directory(p)) {
            for (const auto& f : mda_files_in(p)) files.push_back(f);
        } else {
            files.push_back(p);
        }
    }
    if (files.empty()) {
        err = "no timeseries input";
        return false;
    }
    M = -1;
    N = 0;
    common_type = 0;
    for (size_t i = 0; i < files.size(); ++i) {
        MdaHeader h;
        if (!mda_read_header(files[i], h, err)) return false;
        if (h.dims.size() > 2 && h.size() != h.dim(0) * h.dim(1)) {
            err = files[i] + ": a timeseries must be two-dimensional";
            return false;
        }
        if (M < 0) M = h.dim(0);
        if (h.dim(0) != M) {
            err = "inconsistent number of channels in " + files[i];
            return false;
        }
        N += h.dim(1);
        if (i == 0) common_type = h.type;
        else if (h.type != common_type) common_type = 0;
        headers.push_back(h);
    }
    return true;
}

bool read_geom_csv(const std::string& path, std::vector<std::vector<double>>& coords,
                   std::string& err) {
    std::ifstream in(path);
    if (!in) {
        err = "unable to open geometry file " + path;
        return false;
    }
    std::vector<std::vector<std::string>> rows;
    std::string line;
    bool first = true;
    while (std::getline(in, line)) {
        for (char& c : line)
            if (c == ',') c = ' ';
        std::istringstream ss(line);
        std::vector<std::string> vals;
        std::string tok;
        while (ss >> tok) vals.push_back(tok);
        if (vals.empty()) continue;
        if (first) {
            first = false;
            char* end = nullptr;
            strtod(vals[0].c_str(), &end);
            if (end != vals[0].c_str() + vals[0].size()) continue;  // header line
        }
        rows.push_back(vals);
    }
    if (rows.empty()) {
        err = "empty geometry file " + path;
        return false;
    }
    const size_t D = rows[0].size();
    coords.assign(rows.size(), std::vector<double>(D, 0.0));
    for (size_t i = 0; i < rows.size(); ++i)
        for (size_t j = 0; j < D && j < rows[i].size(); ++j) {
            char* end = nullptr;
            const double v = strtod(rows[i][j].c_str(), &end);
            coords[i][j] = (end == rows[i][j].c_str() + rows[i][j].size()) ? v : 0.0;
        }
    return true;
}

}  // namespace msamd_host
