// Time-series files: the column definitions of a header file and the
// binary / CSV record files the column API reads and writes (host code).
//
// The reference takes these from its data-manager submodule (File,
// FileDefinition, TimeSeriesReaderBinary / TimeSeriesReaderCSV, used by
// main.cpp:59-76 and parallel_ts_compressor.cpp:39-102), which is absent
// from its tree; the formats below are this library's, built on what the
// reference's sample headers (sample_data/*.header) and options
// (main.cpp:22-51) show:
//
// * header file: one column per line, "name,type,decimals", type one of
//   int, time, float, double, short, char; blank lines are skipped;
// * binary file: records of every column's value in order, each in its
//   native little-endian width (time 8 bytes, char 1), followed by
//   `padding` bytes (skipped on read, zero on write);
// * CSV file: one record per line, values separated by commas; integers in
//   decimal, floats written with the column's decimals ("%.*f") and read
//   with strtod.
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "dispatch.h"
#include "ddj_internal.h"

namespace {

thread_local std::string g_io_error;

ddj_status io_fail(ddj_status st, const std::string& msg) {
    g_io_error = msg;
    return st;
}

const char* kTypeNames[] = {"int", "time", "float", "double", "short", "char"};

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && (s[a] == ' ' || s[a] == '\t' || s[a] == '\r' || s[a] == '\n')) a++;
    while (b > a && (s[b - 1] == ' ' || s[b - 1] == '\t' || s[b - 1] == '\r' || s[b - 1] == '\n')) b--;
    return s.substr(a, b - a);
}

}  // namespace

struct ddj_file_def {
    std::vector<std::string> names;
    std::vector<int> dtypes;
    std::vector<int> decimals;
};

struct ddj_series_file {
    ddj_file_def def;
    FILE* f = nullptr;
    int format = DDJ_FORMAT_BIN;
    int padding = 0;
    bool write = false;
    size_t record_bytes = 0;
    std::vector<uint8_t> rec;
    std::string line;
};

extern "C" {

const char* ddj_io_last_error(void) { return g_io_error.c_str(); }

ddj_status ddj_file_def_create(uint32_t n, const char* const* names, const int* dtypes, const int* decimals,
                               ddj_file_def** out) {
    if (!out || !dtypes || n == 0) return io_fail(DDJ_EINVAL, "bad argument");
    auto d = std::make_unique<ddj_file_def>();
    for (uint32_t i = 0; i < n; i++) {
        if (!ddj::dtype_ok(dtypes[i])) return io_fail(DDJ_EINVAL, "bad column type");
        d->names.push_back(names && names[i] ? names[i] : ("column" + std::to_string(i)));
        d->dtypes.push_back(dtypes[i]);
        d->decimals.push_back(decimals ? decimals[i] : 0);
    }
    *out = d.release();
    return DDJ_OK;
}

ddj_status ddj_file_def_read(const char* header_path, ddj_file_def** out) {
    if (!out || !header_path) return io_fail(DDJ_EINVAL, "bad argument");
    FILE* f = fopen(header_path, "r");
    if (!f) return io_fail(DDJ_EINVAL, std::string("cannot open ") + header_path);
    std::unique_ptr<FILE, int (*)(FILE*)> guard(f, fclose);
    auto d = std::make_unique<ddj_file_def>();
    char buf[4096];
    int line_no = 0;
    while (fgets(buf, sizeof buf, f)) {
        line_no++;
        const std::string line = trim(buf);
        if (line.empty()) continue;
        // name,type,decimals: the name may hold spaces, not commas
        const size_t c2 = line.rfind(','), c1 = c2 == std::string::npos ? c2 : line.rfind(',', c2 - 1);
        if (c1 == std::string::npos || c2 == std::string::npos || c1 == 0)
            return io_fail(DDJ_EINVAL, "header line " + std::to_string(line_no) + ": expected name,type,decimals");
        const std::string name = trim(line.substr(0, c1)), type = trim(line.substr(c1 + 1, c2 - c1 - 1));
        const std::string dec = trim(line.substr(c2 + 1));
        int dt = -1;
        for (int k = 0; k < 6; k++)
            if (type == kTypeNames[k]) dt = k;
        if (dt < 0) return io_fail(DDJ_EINVAL, "header line " + std::to_string(line_no) + ": unknown type " + type);
        char* end = nullptr;
        const long dv = strtol(dec.c_str(), &end, 10);
        if (dec.empty() || *end || dv < 0 || dv > 30)
            return io_fail(DDJ_EINVAL, "header line " + std::to_string(line_no) + ": bad decimals " + dec);
        d->names.push_back(name);
        d->dtypes.push_back(dt);
        d->decimals.push_back((int)dv);
    }
    if (d->names.empty()) return io_fail(DDJ_EINVAL, std::string("no columns in ") + header_path);
    *out = d.release();
    return DDJ_OK;
}

ddj_status ddj_file_def_write(const ddj_file_def* def, const char* header_path) {
    if (!def || !header_path) return io_fail(DDJ_EINVAL, "bad argument");
    FILE* f = fopen(header_path, "w");
    if (!f) return io_fail(DDJ_EINVAL, std::string("cannot open ") + header_path);
    for (size_t i = 0; i < def->names.size(); i++)
        fprintf(f, "%s,%s,%d\n", def->names[i].c_str(), kTypeNames[def->dtypes[i]], def->decimals[i]);
    return fclose(f) == 0 ? DDJ_OK : io_fail(DDJ_EINVAL, "write failed");
}

uint32_t ddj_file_def_columns(const ddj_file_def* def) { return def ? (uint32_t)def->names.size() : 0; }

ddj_status ddj_file_def_column(const ddj_file_def* def, uint32_t i, const char** name, int* dtype, int* decimals) {
    if (!def || i >= def->names.size()) return io_fail(DDJ_EINVAL, "no such column");
    if (name) *name = def->names[i].c_str();
    if (dtype) *dtype = def->dtypes[i];
    if (decimals) *decimals = def->decimals[i];
    return DDJ_OK;
}

void ddj_file_def_free(ddj_file_def* def) { delete def; }

ddj_status ddj_series_open(const ddj_file_def* def, const char* path, int format, int padding, int write,
                           ddj_series_file** out) {
    if (!def || !path || !out) return io_fail(DDJ_EINVAL, "bad argument");
    if (format != DDJ_FORMAT_BIN && format != DDJ_FORMAT_CSV) return io_fail(DDJ_EINVAL, "unknown file format");
    if (padding < 0) return io_fail(DDJ_EINVAL, "negative padding");
    FILE* f = fopen(path, write ? "wb" : "rb");
    if (!f) return io_fail(DDJ_EINVAL, std::string("cannot open ") + path);
    auto s = std::make_unique<ddj_series_file>();
    s->def = *def;
    s->f = f;
    s->format = format;
    s->padding = padding;
    s->write = write != 0;
    for (int dt : def->dtypes) s->record_bytes += ddj::dtype_size(dt);
    s->rec.assign(s->record_bytes + (size_t)padding, 0);
    *out = s.release();
    return DDJ_OK;
}

void ddj_series_close(ddj_series_file* s) {
    if (!s) return;
    if (s->f) fclose(s->f);
    delete s;
}

ddj_status ddj_series_read(ddj_series_file* s, size_t max_rows, void* const* host_columns, size_t* rows) {
    if (!s || !rows || (max_rows && !host_columns) || s->write) return io_fail(DDJ_EINVAL, "bad argument");
    *rows = 0;
    const size_t nc = s->def.dtypes.size();
    for (size_t r = 0; r < max_rows;) {
        if (s->format == DDJ_FORMAT_BIN) {
            const size_t got = fread(s->rec.data(), 1, s->rec.size(), s->f);
            if (got == 0 && feof(s->f)) return DDJ_OK;
            if (got < s->record_bytes) return io_fail(DDJ_ECORRUPT, "truncated record");
            size_t off = 0;
            for (size_t k = 0; k < nc; k++) {
                const size_t es = ddj::dtype_size(s->def.dtypes[k]);
                memcpy(static_cast<uint8_t*>(host_columns[k]) + r * es, s->rec.data() + off, es);
                off += es;
            }
        } else {
            s->line.clear();
            int ch;
            while ((ch = fgetc(s->f)) != EOF && ch != '\n') s->line.push_back((char)ch);
            if (ch == EOF && s->line.empty()) return DDJ_OK;
            if (trim(s->line).empty()) continue;  // blank line
            const char* p = s->line.c_str();
            for (size_t k = 0; k < nc; k++) {
                while (*p == ' ') p++;
                char* end = nullptr;
                errno = 0;
                const int dt = s->def.dtypes[k];
                uint8_t* dst = static_cast<uint8_t*>(host_columns[k]) + r * ddj::dtype_size(dt);
                if (dt == DDJ_D_FLOAT || dt == DDJ_D_DOUBLE) {
                    const double v = strtod(p, &end);
                    if (dt == DDJ_D_FLOAT) {
                        const float fv = (float)v;
                        memcpy(dst, &fv, 4);
                    } else {
                        memcpy(dst, &v, 8);
                    }
                } else {
                    const long long v = strtoll(p, &end, 10);
                    ddj::dispatch_int((int)ddj::dtype_size(dt), [&](auto tag) {
                        using T = decltype(tag);
                        const T t = (T)v;
                        memcpy(dst, &t, sizeof t);
                    });
                }
                if (end == p || errno == ERANGE)
                    return io_fail(DDJ_ECORRUPT, "bad value in column " + s->def.names[k] + ": " + s->line);
                p = end;
                while (*p == ' ' || *p == '\r') p++;
                if (k + 1 < nc) {
                    if (*p != ',') return io_fail(DDJ_ECORRUPT, "too few values in record: " + s->line);
                    p++;
                } else if (*p) {
                    return io_fail(DDJ_ECORRUPT, "too many values in record: " + s->line);
                }
            }
        }
        *rows = ++r;
    }
    return DDJ_OK;
}

ddj_status ddj_series_write(ddj_series_file* s, const void* const* host_columns, size_t rows) {
    if (!s || (rows && !host_columns) || !s->write) return io_fail(DDJ_EINVAL, "bad argument");
    const size_t nc = s->def.dtypes.size();
    char num[512];
    for (size_t r = 0; r < rows; r++) {
        if (s->format == DDJ_FORMAT_BIN) {
            size_t off = 0;
            for (size_t k = 0; k < nc; k++) {
                const size_t es = ddj::dtype_size(s->def.dtypes[k]);
                memcpy(s->rec.data() + off, static_cast<const uint8_t*>(host_columns[k]) + r * es, es);
                off += es;
            }
            if (fwrite(s->rec.data(), 1, s->rec.size(), s->f) != s->rec.size())
                return io_fail(DDJ_EINVAL, "write failed");
        } else {
            s->line.clear();
            for (size_t k = 0; k < nc; k++) {
                const int dt = s->def.dtypes[k];
                const uint8_t* src = static_cast<const uint8_t*>(host_columns[k]) + r * ddj::dtype_size(dt);
                if (dt == DDJ_D_FLOAT) {
                    float v;
                    memcpy(&v, src, 4);
                    snprintf(num, sizeof num, "%.*f", s->def.decimals[k], (double)v);
                } else if (dt == DDJ_D_DOUBLE) {
                    double v;
                    memcpy(&v, src, 8);
                    snprintf(num, sizeof num, "%.*f", s->def.decimals[k], v);
                } else {
                    long long v = ddj::dispatch_int((int)ddj::dtype_size(dt), [&](auto tag) -> long long {
                        using T = decltype(tag);
                        T t;
                        memcpy(&t, src, sizeof t);
                        return (long long)t;
                    });
                    snprintf(num, sizeof num, "%lld", v);
                }
                if (k) s->line.push_back(',');
                s->line += num;
            }
            s->line.push_back('\n');
            if (fwrite(s->line.data(), 1, s->line.size(), s->f) != s->line.size())
                return io_fail(DDJ_EINVAL, "write failed");
        }
    }
    return DDJ_OK;
}

}  // extern "C"
