// ddj: command-line compressor over the C ABI (the reference's main.cpp:22-124).
//
//   ddj -c -h HEADER -i INPUT -o OUTPUT [-f bin|csv] [-p PADDING] [-b ROWS] [--device N]
//   ddj -d -h HEADER -i INPUT -o OUTPUT [-f bin|csv] [-p PADDING] [-b ROWS] [--device N]
//   ddj -g ROWS [-o DIR]
//
// -c reads batches of ROWS records (ParallelTSCompressor's batch size,
// default 1,000,000, parallel_ts_compressor.cpp:17-23) from INPUT and
// appends every column's [size_t][blob] to OUTPUT through the column API
// (one stateful planner per column, CompressionOptimizer::CompressData);
// -d decodes such a file back into records; -g writes DIR/generated.header,
// generated.csv and generated.inf with the reference's six sample columns
// (main.cpp:127-180).  INPUT and OUTPUT may also be given positionally.
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "ddj.h"

namespace {

struct Options {
    bool compress = false, decompress = false;
    long long generate = -1;
    std::string header, input, output;
    int format = DDJ_FORMAT_BIN;
    int padding = 0;
    size_t batch = 1000000;
    int device = 0;
};

void usage() {
    fprintf(stderr,
            "usage: ddj -c|-d -h HEADER -i INPUT -o OUTPUT [-f bin|csv] [-p PADDING] [-b ROWS] [--device N]\n"
            "       ddj -g ROWS [-o DIR]\n");
}

int parse(int argc, char** argv, Options& o) {
    std::vector<std::string> pos;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto value = [&](const char* what) -> std::string {
            if (i + 1 >= argc) {
                fprintf(stderr, "ddj: %s needs a value\n", what);
                exit(2);
            }
            return argv[++i];
        };
        if (a == "-c" || a == "--compress") o.compress = true;
        else if (a == "-d" || a == "--decompress") o.decompress = true;
        else if (a == "-h" || a == "--header-file") o.header = value("-h");
        else if (a == "-i" || a == "--input-file") o.input = value("-i");
        else if (a == "-o" || a == "--output-file") o.output = value("-o");
        else if (a == "-g" || a == "--generate") o.generate = atoll(value("-g").c_str());
        else if (a == "-p" || a == "--padding") o.padding = atoi(value("-p").c_str());
        else if (a == "-b" || a == "--batch") o.batch = (size_t)atoll(value("-b").c_str());
        else if (a == "--device") o.device = atoi(value("--device").c_str());
        else if (a == "-f" || a == "--format") {
            const std::string f = value("-f");
            if (f == "bin") o.format = DDJ_FORMAT_BIN;
            else if (f == "csv") o.format = DDJ_FORMAT_CSV;
            else {
                fprintf(stderr, "ddj: unknown format %s\n", f.c_str());
                return 2;
            }
        } else if (!a.empty() && a[0] == '-') {
            fprintf(stderr, "ddj: unknown option %s\n", a.c_str());
            return 2;
        } else {
            pos.push_back(a);
        }
    }
    if (!pos.empty() && o.input.empty()) o.input = pos[0];
    if (pos.size() > 1 && o.output.empty()) o.output = pos[1];
    if (o.batch == 0) o.batch = 1;
    return 0;
}

struct Columns {
    std::vector<std::vector<uint8_t>> data;
    std::vector<void*> ptrs;
    std::vector<size_t> es;
    Columns(const ddj_file_def* def, size_t rows) {
        const uint32_t n = ddj_file_def_columns(def);
        static const size_t sizes[] = {4, 8, 4, 8, 2, 1};
        for (uint32_t i = 0; i < n; i++) {
            int dt = 0;
            ddj_file_def_column(def, i, nullptr, &dt, nullptr);
            es.push_back(sizes[dt]);
            data.emplace_back(rows * sizes[dt]);
        }
        for (auto& d : data) ptrs.push_back(d.data());
    }
    void grow(size_t rows) {
        for (size_t i = 0; i < data.size(); i++) {
            data[i].resize(rows * es[i]);
            ptrs[i] = data[i].data();
        }
    }
};

std::vector<int> dtypes_of(const ddj_file_def* def) {
    std::vector<int> dt(ddj_file_def_columns(def));
    for (uint32_t i = 0; i < dt.size(); i++) ddj_file_def_column(def, i, nullptr, &dt[i], nullptr);
    return dt;
}

int io_error(const char* what) {
    fprintf(stderr, "ddj: %s: %s\n", what, ddj_io_last_error());
    return 1;
}

int ctx_error(ddj_ctx* ctx, const char* what, ddj_status st) {
    fprintf(stderr, "ddj: %s failed (%d): %s\n", what, (int)st, ctx ? ddj_last_error(ctx) : "");
    return 1;
}

// ParallelTSCompressor::Compress (parallel_ts_compressor.cpp:39-61)
int compress(const Options& o, const ddj_file_def* def) {
    ddj_series_file* src = nullptr;
    if (ddj_series_open(def, o.input.c_str(), o.format, o.padding, 0, &src) != DDJ_OK) return io_error("open input");
    ddj_ctx* ctx = nullptr;
    ddj_status st = ddj_ctx_create(&ctx, o.device);
    if (st != DDJ_OK) return ctx_error(nullptr, "context", st);
    const std::vector<int> dt = dtypes_of(def);
    ddj_ts_compressor* ts = nullptr;
    st = ddj_ts_create(ctx, (uint32_t)dt.size(), dt.data(), o.batch, &ts);
    int rc = 0;
    if (st != DDJ_OK) {
        rc = ctx_error(ctx, "column compressor", st);
    } else {
        Columns cols(def, o.batch);
        size_t rows = 0, total = 0;
        bool first = true;
        do {
            if (ddj_series_read(src, o.batch, cols.ptrs.data(), &rows) != DDJ_OK) {
                rc = io_error("read input");
                break;
            }
            if (rows == 0 && !first) break;
            st = ddj_ts_compress(ts, cols.ptrs.data(), rows, o.output.c_str(), first ? 0 : 1);
            if (st != DDJ_OK) {
                rc = ctx_error(ctx, "compress", st);
                break;
            }
            first = false;
            total += rows;
        } while (rows == o.batch);
        if (!rc) printf("compressed %zu records of %zu columns into %s\n", total, dt.size(), o.output.c_str());
        ddj_ts_destroy(ts);
    }
    ddj_series_close(src);
    ddj_ctx_destroy(ctx);
    return rc;
}

// ParallelTSCompressor::Decompress (parallel_ts_compressor.cpp:63-100)
int decompress(const Options& o, const ddj_file_def* def) {
    ddj_ctx* ctx = nullptr;
    ddj_status st = ddj_ctx_create(&ctx, o.device);
    if (st != DDJ_OK) return ctx_error(nullptr, "context", st);
    const std::vector<int> dt = dtypes_of(def);
    ddj_ts_reader* rd = nullptr;
    st = ddj_ts_reader_open(ctx, o.input.c_str(), (uint32_t)dt.size(), dt.data(), &rd);
    if (st != DDJ_OK) {
        int rc = ctx_error(ctx, "open input", st);
        ddj_ctx_destroy(ctx);
        return rc;
    }
    ddj_series_file* dst = nullptr;
    int rc = 0;
    if (ddj_series_open(def, o.output.c_str(), o.format, o.padding, 1, &dst) != DDJ_OK) {
        rc = io_error("open output");
    } else {
        size_t cap = o.batch, rows = 0, total = 0;
        Columns cols(def, cap);
        for (;;) {
            st = ddj_ts_reader_next(rd, cols.ptrs.data(), cap, &rows);
            if (st == DDJ_EINVAL && strstr(ddj_last_error(ctx), "capacity")) {  // a batch larger than -b: grow
                cap *= 2;
                cols.grow(cap);
                continue;
            }
            if (st != DDJ_OK) {
                rc = ctx_error(ctx, "decompress", st);
                break;
            }
            if (rows == 0) break;
            if (ddj_series_write(dst, cols.ptrs.data(), rows) != DDJ_OK) {
                rc = io_error("write output");
                break;
            }
            total += rows;
        }
        if (!rc) printf("decompressed %zu records of %zu columns into %s\n", total, dt.size(), o.output.c_str());
        ddj_series_close(dst);
    }
    ddj_ts_reader_close(rd);
    ddj_ctx_destroy(ctx);
    return rc;
}

// The reference's sample data (main.cpp:127-180 with
// CudaArrayGenerator::GetFakeDataForTime / GetFakeDataWithPatternA / B,
// util/generator/cuda_array_generator.cpp:69-166), from a fixed seed.
template <typename T> std::vector<T> pattern_a(size_t len, T step, T mn, T mx, size_t n) {
    std::vector<T> out(n);
    const T old_min = mn;
    for (size_t i = 0; i < n; i++) {
        T v = mn < mx ? (mn > old_min ? mn : old_min) : mx;
        if (i % len == 0) {
            v = mx;
            mn += step;
            if (mn > mx || mn < old_min) step = -step;
        }
        out[i] = v;
    }
    return out;
}
template <typename T> std::vector<T> pattern_b(size_t len, T mn, T mx, size_t n, std::mt19937_64& rng) {
    std::vector<T> out(n);
    const T step = (mx - mn) / (T)(len / 2);
    for (size_t i = 0; i < n; i++) {
        if ((i / len) % 2 == 0) {
            const T r = (T)(rng() % 6);
            out[i] = i % 2 ? mn + r : mx - r;
        } else {
            const size_t x = i % len;
            out[i] = x < len / 2 ? mx - (T)x * step : mx - (T)(len - x) * step;
        }
    }
    return out;
}

int generate(const Options& o) {
    const size_t n = (size_t)o.generate;
    const std::string dir = o.output.empty() ? "sample_data" : o.output;
    const char* names[] = {"TIME", "PATTERN_A_INT", "PATTERN_B_INT", "FLOAT_PREC_4", "PATTERN_A_FLOAT",
                           "PATTERN_B_FLOAT"};
    const int dts[] = {DDJ_D_TIME, DDJ_D_INT, DDJ_D_INT, DDJ_D_FLOAT, DDJ_D_FLOAT, DDJ_D_FLOAT};
    const int decs[] = {0, 0, 0, 4, 6, 6};
    ddj_file_def* def = nullptr;
    if (ddj_file_def_create(6, names, dts, decs, &def) != DDJ_OK) return io_error("definition");
    std::mt19937_64 rng(12345);
    std::vector<int64_t> t(n);
    int64_t v = 1450000000;
    for (size_t i = 0; i < n; i++) {  // flatness 0.05: the time advances on 95% of rows by 1 or 2
        const int64_t step = 1 + (int64_t)(rng() % 2);
        if ((int)(rng() % 100) > 5) v += step;
        t[i] = v;
    }
    auto a = pattern_a<int>(10, 10, 0, 1000000, n);
    auto b = pattern_b<int>(100000, 0, 1000000, n, rng);
    std::vector<float> f4(n);
    for (size_t i = 0; i < n; i++) f4[i] = (float)((double)(rng() % 20000000) / 1e4);  // 4 decimals
    auto fa = pattern_a<float>(10, 0.01f, 0.f, 1e3f, n);
    auto fb = pattern_b<float>(1000000, 0.f, 1e7f, n, rng);
    const void* cols[] = {t.data(), a.data(), b.data(), f4.data(), fa.data(), fb.data()};
    int rc = 0;
    for (int format : {DDJ_FORMAT_CSV, DDJ_FORMAT_BIN}) {
        const std::string path = dir + (format == DDJ_FORMAT_CSV ? "/generated.csv" : "/generated.inf");
        ddj_series_file* f = nullptr;
        if (ddj_series_open(def, path.c_str(), format, 0, 1, &f) != DDJ_OK || ddj_series_write(f, cols, n) != DDJ_OK) {
            rc = io_error(path.c_str());
        }
        ddj_series_close(f);
        if (rc) break;
    }
    if (!rc && ddj_file_def_write(def, (dir + "/generated.header").c_str()) != DDJ_OK) rc = io_error("header");
    if (!rc) printf("generated %zu records in %s/generated.{header,csv,inf}\n", n, dir.c_str());
    ddj_file_def_free(def);
    return rc;
}

}  // namespace

int main(int argc, char** argv) {
    Options o;
    if (parse(argc, argv, o) != 0) {
        usage();
        return 2;
    }
    if (o.generate >= 0 && !o.compress && !o.decompress) return generate(o);
    if (o.compress == o.decompress || o.header.empty() || o.input.empty() || o.output.empty()) {
        usage();
        return 2;
    }
    ddj_file_def* def = nullptr;
    if (ddj_file_def_read(o.header.c_str(), &def) != DDJ_OK) return io_error("header file");
    printf("Input file: %s\nOutput file: %s\nHeader file: %s\n", o.input.c_str(), o.output.c_str(), o.header.c_str());
    const int rc = o.compress ? compress(o, def) : decompress(o, def);
    ddj_file_def_free(def);
    return rc;
}
