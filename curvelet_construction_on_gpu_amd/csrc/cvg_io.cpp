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
// doubles in default ostream notation, which is printf's %.6g (format_g6
// below: a scaled-integer fast path, std::to_chars for the cases it cannot
// decide).
//
// Both directions run on several host threads for large files: the reader
// splits the text at line boundaries, parses the pieces independently and
// reports the first error in file order (its line number counted over the
// pieces before it); the writers format row ranges independently and write
// them in order.  Numbers are converted by the Eisel-Lemire algorithm (up to
// 19 significant digits, correctly rounded like strtod), else by strtod_l in
// the "C" locale from a stack copy of the token; no allocation per number.
#include <locale.h>

#include <algorithm>
#include <cerrno>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/curvelet_gpu.h"
#include "cvg_pow10_table.hpp"
#include "cvg_pow5_table.hpp"

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

// The "C" locale, created once (istream >> double uses the classic locale;
// the process's LC_NUMERIC may differ).
locale_t c_locale()
{
    static const locale_t loc = newlocale(LC_ALL_MASK, "C", static_cast<locale_t>(0));
    return loc;
}

// Host threads for a job of `bytes` (at least 1 MiB per thread, at most 16).
int io_threads(size_t bytes)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t by_size = bytes / (1u << 20) + 1;
    return static_cast<int>(std::min<size_t>({static_cast<size_t>(hw), by_size, static_cast<size_t>(16)}));
}

// Runs f(0) .. f(n-1) on up to n threads (f(0) on the caller's); the parts
// whose thread cannot be started run on the caller's thread.
template <typename F>
void run_parallel(int n, F f)
{
    std::vector<std::thread> th;
    int started = 1;
    try {
        for (; started < n; ++started) th.emplace_back(f, started);
    } catch (...) {
    }
    f(0);
    for (int i = started; i < n; ++i) f(i);
    for (auto &t : th) t.join();
}

// Decimal to double for w * 10^q (w < 10^19): the Eisel-Lemire algorithm
// (D. Lemire, "Number Parsing at a Gigabyte per Second", Software: Practice
// and Experience 51(8), 2021; N. Mushtak and D. Lemire, "Fast number parsing
// without fallback", 2023): w, normalized, times a 128-bit approximation of
// 5^q (cvg_pow5_table.hpp) gives the correctly rounded significand, ties to
// even.  Returns false for a subnormal result, which is left to strtod.
bool eisel_lemire(uint64_t w, int q, bool neg, double *out)
{
    uint64_t bits;
    if (w == 0 || q < cvg::kPow5Min) {
        bits = 0;
    } else if (q > cvg::kPow5Max) {
        bits = uint64_t(0x7ff) << 52;  // infinity
    } else {
        const int lz = __builtin_clzll(w);
        w <<= lz;
        const int idx = 2 * (q - cvg::kPow5Min);
        const unsigned __int128 p = static_cast<unsigned __int128>(w) * cvg::kPow5[idx];
        uint64_t hi = static_cast<uint64_t>(p >> 64), lo = static_cast<uint64_t>(p);
        if ((hi & 0x1FF) == 0x1FF) {  // the low 9 bits could carry: add the next 64 bits of 5^q
            const uint64_t h2 = static_cast<uint64_t>((static_cast<unsigned __int128>(w) * cvg::kPow5[idx + 1]) >> 64);
            lo += h2;
            if (h2 > lo) ++hi;
        }
        const int upper = static_cast<int>(hi >> 63);
        const int shift = upper + 9;  // keep 52 + 3 bits
        uint64_t m = hi >> shift;
        int e2 = ((((152170 + 65536) * q) >> 16) + 63) + upper - lz + 1023;
        if (e2 <= 0) return false;
        // exactly halfway (only possible for small |q|): round to even
        if (lo <= 1 && q >= -4 && q <= 23 && (m & 3) == 1 && (m << shift) == hi) m &= ~uint64_t(1);
        m += m & 1;
        m >>= 1;
        if (m >= (uint64_t(2) << 52)) {
            m = uint64_t(1) << 52;
            ++e2;
        }
        m &= ~(uint64_t(1) << 52);
        bits = e2 >= 0x7ff ? uint64_t(0x7ff) << 52 : (static_cast<uint64_t>(e2) << 52) | m;
    }
    if (neg) bits |= uint64_t(1) << 63;
    std::memcpy(out, &bits, sizeof bits);
    return true;
}

// The token [s, e) has the scanner's form [sign] digits [. digits] [(e|E)
// [sign] digits].  Returns false when it is outside the fast path (no
// mantissa digit, an exponent without digits, more than 19 significant
// digits, a huge exponent, a subnormal result); the caller then uses strtod.
bool parse_fast(const char *s, const char *e, double *out)
{
    bool neg = false;
    if (s < e && (*s == '+' || *s == '-')) {
        neg = *s == '-';
        ++s;
    }
    uint64_t w = 0;
    int nd = 0, q = 0;
    bool any = false, nonzero = false;
    for (; s < e && is_digit(*s); ++s) {
        any = true;
        const unsigned dg = static_cast<unsigned>(*s - '0');
        if (nonzero || dg != 0) {
            if (nd == 19) return false;
            w = w * 10 + dg;
            ++nd;
            nonzero = true;
        }
    }
    if (s < e && *s == '.') {
        for (++s; s < e && is_digit(*s); ++s) {
            any = true;
            const unsigned dg = static_cast<unsigned>(*s - '0');
            if (nonzero || dg != 0) {
                if (nd == 19) return false;
                w = w * 10 + dg;
                ++nd;
                nonzero = true;
            }
            --q;
        }
    }
    if (!any) return false;
    if (s < e && (*s == 'e' || *s == 'E')) {
        ++s;
        bool eneg = false;
        if (s < e && (*s == '+' || *s == '-')) {
            eneg = *s == '-';
            ++s;
        }
        if (s == e) return false;
        int x = 0;
        for (; s < e && is_digit(*s); ++s)
            if (x < 100000) x = x * 10 + (*s - '0');
        if (x >= 100000) return false;
        q += eneg ? -x : x;
    }
    if (s != e) return false;
    return eisel_lemire(w, q, neg, out);
}

// One istream >> double (f32: istream >> float) on [*p, e).  Advances *p past
// what the extraction consumed.  Returns false on failure (as the stream's
// failbit).  A float is converted from the decimal text directly (strtof),
// never through a double, whose rounding could land on a float midpoint.
bool extract_double(const char **p, const char *e, double *out, bool f32 = false)
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
    double fast;
    if (!f32 && parse_fast(start, q, &fast)) {
        if (std::isinf(fast)) return false;  // overflow: the stream's failbit
        *out = fast;
        return true;
    }
    // strtod must see exactly the collected characters (it would read
    // further, e.g. a hexadecimal form, on the raw buffer)
    const size_t len = static_cast<size_t>(q - start);
    char small[128];
    std::string big;
    const char *tok;
    if (len < sizeof(small)) {
        std::memcpy(small, start, len);
        small[len] = '\0';
        tok = small;
    } else {
        big.assign(start, q);
        tok = big.c_str();
    }
    char *end = nullptr;
    const double v = f32 ? static_cast<double>(strtof_l(tok, &end, c_locale())) : strtod_l(tok, &end, c_locale());
    if (end == tok || *end != '\0') return false;
    if (std::isinf(v)) return false;
    *out = v;
    return true;
}

// printf("%.6g") of a double, as std::ostream writes it.  Fast path: with
// X = floor(log10 |x|), s = |x| * RN(10^(5 - X)) lies in [99999.5, 999999.5)
// and carries a relative error below 2^-52 (one rounding of the table value,
// one of the product), so it is within 2.3e-10 of the exact scaled value; when
// its fraction is more than 1e-9 away from one half, rounding s to an integer
// rounds the exact value the same way and gives the 6 significant digits.
// Everything else (near a tie or a power of ten, zero's sign aside tiny or
// huge magnitudes, inf, nan) goes through std::to_chars, which is exact.
char *format_g6(char *p, double x)
{
    auto slow = [&]() { return std::to_chars(p, p + 32, x, std::chars_format::general, 6).ptr; };
    uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    const int be = static_cast<int>((u >> 52) & 0x7ff);
    if (be < 60 || be > 1990) {
        if ((u << 1) == 0) {  // +-0
            if (u >> 63) *p++ = '-';
            *p++ = '0';
            return p;
        }
        return slow();
    }
    const double v = std::fabs(x);
    int X = ((be - 1023) * 78913) >> 18;  // floor(e2 log10 2) up to one
    if (v >= cvg::kPow10[X + 1 - cvg::kPow10Min]) ++X;
    double s = v * cvg::kPow10[5 - X - cvg::kPow10Min];
    if (s < 99999.5) {
        --X;
        s = v * cvg::kPow10[5 - X - cvg::kPow10Min];
    }
    const long n0 = static_cast<long>(s);
    const double f = s - static_cast<double>(n0);
    if (!(s >= 99999.5 && s < 999999.4999) || std::fabs(f - 0.5) < 1e-9) return slow();
    unsigned n = static_cast<unsigned>(n0) + (f > 0.5 ? 1u : 0u);
    if (n >= 1000000u) {
        n /= 10u;
        ++X;
    }
    char d[6];
    const unsigned hi = n / 1000u, lo = n - hi * 1000u;
    d[0] = static_cast<char>('0' + hi / 100u);
    d[1] = static_cast<char>('0' + (hi / 10u) % 10u);
    d[2] = static_cast<char>('0' + hi % 10u);
    d[3] = static_cast<char>('0' + lo / 100u);
    d[4] = static_cast<char>('0' + (lo / 10u) % 10u);
    d[5] = static_cast<char>('0' + lo % 10u);
    int nd = 6;  // significant digits after trailing zeros are dropped
    while (nd > 1 && d[nd - 1] == '0') --nd;
    if (u >> 63) *p++ = '-';
    if (X >= -4 && X < 6) {  // fixed notation
        if (X >= 0) {
            for (int i = 0; i <= X; ++i) *p++ = i < nd ? d[i] : '0';
            if (nd > X + 1) {
                *p++ = '.';
                for (int i = X + 1; i < nd; ++i) *p++ = d[i];
            }
        } else {
            *p++ = '0';
            *p++ = '.';
            for (int i = 0; i < -X - 1; ++i) *p++ = '0';
            for (int i = 0; i < nd; ++i) *p++ = d[i];
        }
    } else {  // d.ddddde+XX
        *p++ = d[0];
        if (nd > 1) {
            *p++ = '.';
            for (int i = 1; i < nd; ++i) *p++ = d[i];
        }
        *p++ = 'e';
        int ex = X;
        if (ex < 0) {
            *p++ = '-';
            ex = -ex;
        } else {
            *p++ = '+';
        }
        if (ex >= 100) {
            *p++ = static_cast<char>('0' + ex / 100);
            ex %= 100;
        }
        *p++ = static_cast<char>('0' + ex / 10);
        *p++ = static_cast<char>('0' + ex % 10);
    }
    return p;
}

enum ParseError { kNone = 0, kBlank, kShort, kExtra };

// Parses the lines of [b, e) (whole lines).  Stops at the first error.
struct Piece {
    std::vector<double> values;
    long lines = 0;   // lines parsed (the failing one included)
    long edges = 0;
    ParseError error = kNone;
    int found = 0;    // values read on the failing line
};

void parse_piece(const char *b, const char *e, int values_per_edge, Piece &out, bool f32 = false)
{
    out.values.reserve(static_cast<size_t>(e - b) / 16 + 16);
    const char *p = b;
    while (p < e) {
        const char *nl = static_cast<const char *>(std::memchr(p, '\n', static_cast<size_t>(e - p)));
        const char *le = nl ? nl : e;
        ++out.lines;
        if (is_blank_line(p, le)) {
            out.error = kBlank;
            return;
        }
        const char *q = p;
        for (int j = 0; j < values_per_edge; ++j) {
            double v;
            if (!extract_double(&q, le, &v, f32)) {
                out.error = kShort;
                out.found = j;
                return;
            }
            out.values.push_back(v);
        }
        if (!is_blank_line(q, le)) {
            out.error = kExtra;
            return;
        }
        ++out.edges;
        p = nl ? nl + 1 : e;
    }
}

// Formats rows [r0, r1) of each thread's share with fmt(row, p), which writes
// at most max_row bytes at p and returns the end.  Rows go through a small
// local buffer appended to the thread's output in bulk (per-value vector
// appends do not scale across threads); the outputs are written in order.
template <typename RowFmt>
int write_rows(const char *path, unsigned rows, size_t max_row, size_t bytes_per_row_guess, RowFmt fmt)
{
    FILE *f = std::fopen(path, "wb");
    if (!f) return CVG_ERR_IO;
    const int nt = io_threads(static_cast<size_t>(rows) * bytes_per_row_guess);
    std::vector<std::vector<char>> out(nt);
    run_parallel(nt, [&](int t) {
        const unsigned r0 = static_cast<unsigned>(static_cast<unsigned long long>(rows) * t / nt);
        const unsigned r1 = static_cast<unsigned>(static_cast<unsigned long long>(rows) * (t + 1) / nt);
        std::vector<char> &o = out[t];
        o.reserve(static_cast<size_t>(r1 - r0) * bytes_per_row_guess + 16);
        const size_t cap = std::max<size_t>(size_t(1) << 18, 2 * max_row);
        std::vector<char> local(cap);
        size_t used = 0;
        for (unsigned i = r0; i < r1; ++i) {
            if (cap - used < max_row) {
                o.insert(o.end(), local.data(), local.data() + used);
                used = 0;
            }
            used = static_cast<size_t>(fmt(i, local.data() + used) - local.data());
        }
        o.insert(o.end(), local.data(), local.data() + used);
    });
    bool ok = true;
    for (const auto &o : out)
        if (!o.empty()) ok = ok && std::fwrite(o.data(), 1, o.size(), f) == o.size();
    return (std::fclose(f) == 0 && ok) ? CVG_OK : CVG_ERR_IO;
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

namespace {

// read_TO_edges_from_file<T> (data_io.hpp:20-78): the values as doubles
// (for T = float, the float values), malloc'd.
int read_edges(const char *path, int values_per_edge, double **edges, int *num_edges, char *err, size_t err_len,
               bool f32)
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
        // one read of the whole file when its size is known, else chunks
        long size = -1;
        if (std::fseek(f, 0, SEEK_END) == 0) size = std::ftell(f);
        if (size >= 0 && std::fseek(f, 0, SEEK_SET) == 0) {
            buf.resize(static_cast<size_t>(size));
            buf.resize(std::fread(buf.data(), 1, buf.size(), f));
        }
        char chunk[1 << 16];
        size_t got;
        while ((got = std::fread(chunk, 1, sizeof(chunk), f)) > 0) buf.insert(buf.end(), chunk, chunk + got);
    }
    std::fclose(f);

    // pieces end at line boundaries
    const char *begin = buf.data();
    const char *end = begin + buf.size();
    const int nt = io_threads(buf.size());
    std::vector<const char *> cut(nt + 1, end);
    cut[0] = begin;
    for (int t = 1; t < nt; ++t) {
        const char *c = std::max(cut[t - 1], begin + buf.size() * t / nt);
        if (c > begin && c < end && c[-1] != '\n') {
            const char *nl = static_cast<const char *>(std::memchr(c, '\n', static_cast<size_t>(end - c)));
            c = nl ? nl + 1 : end;
        }
        cut[t] = c;
    }
    std::vector<Piece> pieces(nt);
    run_parallel(nt, [&](int t) { parse_piece(cut[t], cut[t + 1], values_per_edge, pieces[t], f32); });

    long line_num = 0;
    long edge_count = 0;
    for (const Piece &pc : pieces) {
        line_num += pc.lines;
        edge_count += pc.edges;
        if (pc.error == kNone) continue;
        if (pc.error == kBlank)
            set_err(err, err_len,
                    "Error: blank line at line " + std::to_string(line_num) + " in " + path +
                        " (each line must contain exactly " + std::to_string(values_per_edge) + " numeric values)");
        else if (pc.error == kShort)
            set_err(err, err_len,
                    std::string("Error: invalid edge file format in ") + path + " at line " +
                        std::to_string(line_num) + ": expected " + std::to_string(values_per_edge) +
                        " numeric values per line, found " + std::to_string(pc.found) + ".");
        else
            set_err(err, err_len,
                    std::string("Error: invalid edge file format in ") + path + " at line " +
                        std::to_string(line_num) + ": expected exactly " + std::to_string(values_per_edge) +
                        " numeric values per line, found extra data.");
        return CVG_ERR_BAD_INPUT;
    }
    if (edge_count == 0) {
        set_err(err, err_len, std::string("Error: input file contains no valid edges: ") + path);
        return CVG_ERR_BAD_INPUT;
    }
    if (edge_count > 0x7fffffffL / values_per_edge) {
        set_err(err, err_len, "Error: too many edges");
        return CVG_ERR_BAD_INPUT;
    }
    const size_t total = static_cast<size_t>(edge_count) * values_per_edge;
    double *out = static_cast<double *>(std::malloc(total * sizeof(double)));
    if (!out) {
        set_err(err, err_len, "Error: out of host memory");
        return CVG_ERR_OUT_OF_MEMORY;
    }
    size_t o = 0;
    for (const Piece &pc : pieces) {
        if (!pc.values.empty()) std::memcpy(out + o, pc.values.data(), pc.values.size() * sizeof(double));
        o += pc.values.size();
    }
    *edges = out;
    *num_edges = static_cast<int>(edge_count);
    return CVG_OK;
}

}  // namespace

int cvg_read_edge_file(const char *path, int values_per_edge, double **edges, int *num_edges, char *err,
                       size_t err_len)
{
    return read_edges(path, values_per_edge, edges, num_edges, err, err_len, false);
}

int cvg_read_edge_file_f32(const char *path, int values_per_edge, float **edges, int *num_edges, char *err,
                           size_t err_len)
{
    if (!edges) {
        set_err(err, err_len, "Error: invalid arguments to cvg_read_edge_file_f32");
        return CVG_ERR_INVALID_ARGUMENT;
    }
    *edges = nullptr;
    double *wide = nullptr;
    const int rc = read_edges(path, values_per_edge, &wide, num_edges, err, err_len, true);
    if (rc != CVG_OK) return rc;
    const size_t total = static_cast<size_t>(*num_edges) * values_per_edge;
    float *out = static_cast<float *>(std::malloc(total * sizeof(float)));
    if (!out) {
        std::free(wide);
        set_err(err, err_len, "Error: out of host memory");
        return CVG_ERR_OUT_OF_MEMORY;
    }
    for (size_t i = 0; i < total; ++i) out[i] = static_cast<float>(wide[i]);  // exact: float values
    std::free(wide);
    *edges = out;
    return CVG_OK;
}

void cvg_free(void *ptr) { std::free(ptr); }

int cvg_write_chain_file(const char *path, const unsigned *chains, unsigned rows, unsigned width)
{
    if (!path || (!chains && rows > 0 && width > 0)) return CVG_ERR_INVALID_ARGUMENT;
    return write_rows(path, rows, static_cast<size_t>(width) * 12 + 1, static_cast<size_t>(width) * 8 + 1,
                      [&](unsigned i, char *p) {
                          for (unsigned j = 0; j < width; ++j) {
                              // main.cpp:76 casts each unsigned entry to int before printing.
                              const int v = static_cast<int>(chains[static_cast<size_t>(i) * width + j]);
                              p = std::to_chars(p, p + 11, v).ptr;
                              *p++ = '\t';
                          }
                          *p++ = '\n';
                          return p;
                      });
}

int cvg_write_info_file(const char *path, const double *info, unsigned rows, unsigned width)
{
    if (!path || (!info && rows > 0 && width > 0)) return CVG_ERR_INVALID_ARGUMENT;
    // %.6g of a double is at most 13 characters ("-1.23457e-308"); 24 leaves room
    return write_rows(path, rows, static_cast<size_t>(width) * 25 + 1, static_cast<size_t>(width) * 12 + 1,
                      [&](unsigned i, char *p) {
                          for (unsigned j = 0; j < width; ++j) {
                              const double v = info[static_cast<size_t>(i) * width + j];
                              p = format_g6(p, v);
                              *p++ = '\t';
                          }
                          *p++ = '\n';
                          return p;
                      });
}

}  // extern "C"
