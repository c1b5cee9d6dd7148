// Edge-file reader and output writers of the drop-in (host side).
//
// Reader: read_TO_edges_from_file<double> (cpu_curvelet/data_io.hpp:20-78).
// The reference splits the file with std::getline, rejects blank lines
// (only " \t\r\n"), extracts exactly values_per_edge doubles per line with
// istream >> double and rejects any non-blank remainder.  istream >> double
// (libstdc++ num_get) skips isspace characters, collects the longest prefix
// of the form [sign] digits [. digits] [(e|E) [sign] digits] (an exponent is
// taken only after at least one mantissa digit), converts it with strtod and
// fails if strtod does not consume all of it or overflows to infinity.  The
// scanner below applies that grammar and the same strtod, so values and
// accept/reject decisions are identical; the messages are the reference's.
//
// Writers: write_int_array_to_file / write_double_array_to_file
// (data_io.hpp:80-133): every value followed by a tab, a newline per row,
// doubles in default ostream notation, which is printf's %.6g; std::to_chars
// with chars_format::general and precision 6 produces the same text.
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/curvelet_gpu.h"

namespace {

bool is_blank_line(const char *b, const char *e)
{
    for (const char *p = b; p < e; ++p)
        if (!(*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) return false;
    return true;
}

bool is_stream_space(char c)
{
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

bool is_digit(char c) { return c >= '0' && c <= '9'; }

// One istream >> double on [*p, e).  Advances *p past what the extraction
// consumed.  Returns false on failure (as the stream's failbit).
bool extract_double(const char **p, const char *e, double *out)
{
    const char *s = *p;
    while (s < e && is_stream_space(*s)) ++s;
    const char *start = s;
    const char *q = s;
    if (q < e && (*q == '+' || *q == '-')) ++q;
    bool mantissa_digit = false;
    bool seen_point = false;
    bool seen_exp = false;
    while (q < e) {
        const char c = *q;
        if (is_digit(c)) {
            if (!seen_exp) mantissa_digit = true;
            ++q;
        } else if (c == '.' && !seen_point && !seen_exp) {
            seen_point = true;
            ++q;
        } else if ((c == 'e' || c == 'E') && mantissa_digit && !seen_exp) {
            seen_exp = true;
            ++q;
            if (q < e && (*q == '+' || *q == '-')) ++q;
        } else {
            break;
        }
    }
    *p = q;
    if (q == start) return false;
    std::string tok(start, q);
    char *end = nullptr;
    errno = 0;
    const double v = std::strtod(tok.c_str(), &end);
    if (end == tok.c_str() || *end != '\0') return false;
    if (std::isinf(v)) return false;
    *out = v;
    return true;
}

void set_err(char *err, size_t err_len, const std::string &msg)
{
    if (err && err_len > 0) {
        std::strncpy(err, msg.c_str(), err_len - 1);
        err[err_len - 1] = '\0';
    }
}

}  // namespace

extern "C" {

int cvg_read_edge_file(const char *path, int values_per_edge, double **edges, int *num_edges, char *err,
                       size_t err_len)
{
    if (!path || !edges || !num_edges || values_per_edge <= 0) {
        set_err(err, err_len, "Error: invalid arguments to cvg_read_edge_file");
        return CVG_ERR_INVALID_ARGUMENT;
    }
    *edges = nullptr;
    *num_edges = 0;
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        set_err(err, err_len, std::string("Error: input file does not exist: ") + path);
        return CVG_ERR_IO;
    }
    std::vector<char> buf;
    {
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    }
    std::fclose(f);

    std::vector<double> data;
    data.reserve(buf.size() / 16);
    const char *p = buf.data();
    const char *end = p + buf.size();
    long line_num = 0;
    long edge_count = 0;
    while (p < end) {
        const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(end - p)));
        const char *le = nl ? nl : end;
        ++line_num;
        if (is_blank_line(p, le)) {
            set_err(err, err_len,
                    "Error: blank line at line " + std::to_string(line_num) + " in " + path +
                        " (each line must contain exactly " + std::to_string(values_per_edge) + " numeric values)");
            return CVG_ERR_BAD_INPUT;
        }
        const char *q = p;
        for (int j = 0; j < values_per_edge; ++j) {
            double v;
            if (!extract_double(&q, le, &v)) {
                set_err(err, err_len,
                        std::string("Error: invalid edge file format in ") + path + " at line " +
                            std::to_string(line_num) + ": expected " + std::to_string(values_per_edge) +
                            " numeric values per line, found " + std::to_string(j) + ".");
                return CVG_ERR_BAD_INPUT;
            }
            data.push_back(v);
        }
        if (!is_blank_line(q, le)) {
            set_err(err, err_len,
                    std::string("Error: invalid edge file format in ") + path + " at line " +
                        std::to_string(line_num) + ": expected exactly " + std::to_string(values_per_edge) +
                        " numeric values per line, found extra data.");
            return CVG_ERR_BAD_INPUT;
        }
        ++edge_count;
        p = nl ? nl + 1 : end;
    }
    if (edge_count == 0) {
        set_err(err, err_len, std::string("Error: input file contains no valid edges: ") + path);
        return CVG_ERR_BAD_INPUT;
    }
    if (edge_count > 0x7fffffffL / values_per_edge) {
        set_err(err, err_len, "Error: too many edges");
        return CVG_ERR_BAD_INPUT;
    }
    double *out = static_cast<double *>(std::malloc(data.size() * sizeof(double)));
    if (!out) {
        set_err(err, err_len, "Error: out of host memory");
        return CVG_ERR_OUT_OF_MEMORY;
    }
    std::memcpy(out, data.data(), data.size() * sizeof(double));
    *edges = out;
    *num_edges = static_cast<int>(edge_count);
    return CVG_OK;
}

void cvg_free(void *ptr) { std::free(ptr); }

int cvg_write_chain_file(const char *path, const unsigned *chains, unsigned rows, unsigned width)
{
    if (!path || (!chains && rows > 0 && width > 0)) return CVG_ERR_INVALID_ARGUMENT;
    FILE *f = std::fopen(path, "wb");
    if (!f) return CVG_ERR_IO;
    std::vector<char> out;
    out.reserve(static_cast<size_t>(rows) * width * 8 + 16);
    char tmp[32];
    for (unsigned i = 0; i < rows; ++i) {
        for (unsigned j = 0; j < width; ++j) {
            // main.cpp:76 casts each unsigned entry to int before printing.
            const int v = static_cast<int>(chains[static_cast<size_t>(i) * width + j]);
            const auto r = std::to_chars(tmp, tmp + sizeof(tmp), v);
            out.insert(out.end(), tmp, r.ptr);
            out.push_back('\t');
        }
        out.push_back('\n');
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? CVG_OK : CVG_ERR_IO;
}

int cvg_write_info_file(const char *path, const double *info, unsigned rows, unsigned width)
{
    if (!path || (!info && rows > 0 && width > 0)) return CVG_ERR_INVALID_ARGUMENT;
    FILE *f = std::fopen(path, "wb");
    if (!f) return CVG_ERR_IO;
    std::vector<char> out;
    out.reserve(static_cast<size_t>(rows) * width * 12 + 16);
    char tmp[64];
    for (unsigned i = 0; i < rows; ++i) {
        for (unsigned j = 0; j < width; ++j) {
            const double v = info[static_cast<size_t>(i) * width + j];
            const auto r = std::to_chars(tmp, tmp + sizeof(tmp), v, std::chars_format::general, 6);
            out.insert(out.end(), tmp, r.ptr);
            out.push_back('\t');
        }
        out.push_back('\n');
    }
    const bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
    return (std::fclose(f) == 0 && ok) ? CVG_OK : CVG_ERR_IO;
}

}  // extern "C"
