// CURVELET: drop-in command line for cpu_curvelet (cpu_curvelet/main.cpp and
// param_settings.hpp), running the curvelet-construction path on an MI355X
// through libcurvelet_gpu.so.
//
// Same flags, defaults, argument-order rules and messages as the reference's
// parse_args (param_settings.hpp:86-155); input from "../test_files/<file>"
// (data_io.hpp:10-13); chain and info files written to "../outputs/" in the
// reference's text format (data_io.hpp:80-145).  Additive: --device <N>
// selects the HIP device.  --nthreads is accepted and has no effect (there is
// no CPU path).  --float runs the reference's float instantiation
// (run_curvelet<float>): float edges, float arithmetic and its double
// promotions, info widened to double.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "curvelet_gpu.h"

namespace {

struct Options {
    cvg_params p{};
    std::string edge_file = "eth3d_cables2.txt";
    std::string out_file = "chain_cpu.txt";
    bool use_double = true;
    int nthreads = 1;
    int device = 0;
};

void usage(const char *prog)
{
    std::cerr << "Usage: " << prog << " [options]\n"
              << "\n"
              << "Options:\n"
              << "  --float, --double          Scalar type (default: double)\n"
              << "  --output <file>            Output chain file (default: chain_cpu.txt)\n"
              << "  --nthreads <N>             OpenMP thread count (default: 1)\n"
              << "  --edge-file <file>         Input edge file in (default: ../test_files/eth3d_cables2.txt)\n"
              << "  --nrad <val>               Neighbor search radius (default: 3.5)\n"
              << "  --dx <val>                 Edge position tolerance in pixels (default: 0.4)\n"
              << "  --dt-deg <val>             Edge angle tolerance in degrees (default: 15)\n"
              << "  --max-k <val>              Maximum curvature (default: 0.3)\n"
              << "  --curvelet-style <N>       Curvelet style (default: 2)\n"
              << "  --group-max-sz <N>         Maximum group size (default: 4)\n"
              << "  --out-type <N>             Output type (default: 0)\n"
              << "  --sx <val>                 Position sampling step (default: 0.1)\n"
              << "  --st <val>                 Angle sampling step in radians (default: 0.08)\n"
              << "  --edge-data-sz <N>         Values per edge in input file (default: 4)\n"
              << "  --max-candidates <N>       Max neighbors per anchor in CSR build (default: 128)\n"
              << "  --device <N>               HIP device (default: 0)\n"
              << "  --help                     Show this help message\n";
}

bool to_unsigned(const char *arg, const char *name, unsigned &v)
{
    char *end = nullptr;
    const unsigned long x = std::strtoul(arg, &end, 10);
    if (end == arg || *end != '\0') {
        std::cerr << "Error: invalid value for " << name << ": " << arg << std::endl;
        return false;
    }
    v = static_cast<unsigned>(x);
    return true;
}

bool to_int(const char *arg, const char *name, int &v)
{
    char *end = nullptr;
    const long x = std::strtol(arg, &end, 10);
    if (end == arg || *end != '\0') {
        std::cerr << "Error: invalid value for " << name << ": " << arg << std::endl;
        return false;
    }
    v = static_cast<int>(x);
    return true;
}

bool to_double(const char *arg, const char *name, double &v)
{
    char *end = nullptr;
    const double x = std::strtod(arg, &end);
    if (end == arg || *end != '\0') {
        std::cerr << "Error: invalid value for " << name << ": " << arg << std::endl;
        return false;
    }
    v = x;
    return true;
}

// parse_args (param_settings.hpp:86-155): a flag that needs a value and has
// none falls through to "Unknown argument"; --float/--double also reset the
// output name, so their position relative to --output matters.
bool parse(int argc, char **argv, Options &o, bool &help)
{
    help = false;
    for (int i = 1; i < argc; ++i) {
        const char *a = argv[i];
        const bool more = i + 1 < argc;
        auto is = [&](const char *s) { return std::strcmp(a, s) == 0; };
        if (is("--help") || is("-h")) {
            usage(argv[0]);
            help = true;
            return false;
        }
        if (is("--float")) {
            o.use_double = false;
            o.out_file = "chain_cpu_float.txt";
        } else if (is("--double")) {
            o.use_double = true;
            o.out_file = "chain_cpu_double.txt";
        } else if (is("--output") && more) {
            o.out_file = argv[++i];
        } else if (is("--nthreads") && more) {
            o.nthreads = std::atoi(argv[++i]);
        } else if ((is("--input") || is("--edge-file")) && more) {
            o.edge_file = argv[++i];
        } else if (is("--nrad") && more) {
            if (!to_double(argv[++i], "--nrad", o.p.nrad)) return false;
        } else if (is("--dx") && more) {
            if (!to_double(argv[++i], "--dx", o.p.dx)) return false;
        } else if (is("--dt-deg") && more) {
            if (!to_double(argv[++i], "--dt-deg", o.p.dt_deg)) return false;
        } else if (is("--max-k") && more) {
            if (!to_double(argv[++i], "--max-k", o.p.max_k)) return false;
        } else if (is("--curvelet-style") && more) {
            if (!to_unsigned(argv[++i], "--curvelet-style", o.p.curvelet_style)) return false;
        } else if (is("--group-max-sz") && more) {
            if (!to_unsigned(argv[++i], "--group-max-sz", o.p.group_max_sz)) return false;
        } else if (is("--out-type") && more) {
            if (!to_unsigned(argv[++i], "--out-type", o.p.out_type)) return false;
        } else if (is("--sx") && more) {
            if (!to_double(argv[++i], "--sx", o.p.sx)) return false;
        } else if (is("--st") && more) {
            if (!to_double(argv[++i], "--st", o.p.st)) return false;
        } else if (is("--edge-data-sz") && more) {
            if (!to_int(argv[++i], "--edge-data-sz", o.p.edge_data_sz)) return false;
        } else if (is("--max-candidates") && more) {
            if (!to_unsigned(argv[++i], "--max-candidates", o.p.max_candidates)) return false;
        } else if (is("--device") && more) {
            if (!to_int(argv[++i], "--device", o.device)) return false;
        } else {
            std::cerr << "Unknown argument: " << a << std::endl;
            usage(argv[0]);
            return false;
        }
    }
    return true;
}

// Lap timer with the reference's output format (timer.hpp:10-71).
class Laps {
public:
    void start() { origin_ = last_ = Clock::now(); }
    void lap(const std::string &name)
    {
        const auto now = Clock::now();
        const double s = std::chrono::duration<double>(now - last_).count();
        recs_.push_back({name, s});
        std::cout << "[timer] " << name << ": " << fmt(s) << std::endl;
        last_ = now;
    }
    void summary() const
    {
        const double total = std::chrono::duration<double>(Clock::now() - origin_).count();
        std::cout << "\n========== timing summary ==========\n";
        for (const auto &r : recs_) {
            const double pct = total > 0.0 ? 100.0 * r.second / total : 0.0;
            std::cout << "  " << std::left << std::setw(40) << r.first << std::right << std::setw(12) << fmt(r.second)
                      << "  (" << std::fixed << std::setprecision(1) << pct << "%)\n";
        }
        std::cout << "  " << std::left << std::setw(40) << "total" << std::right << std::setw(12) << fmt(total) << "\n";
        std::cout << "====================================\n";
        std::cout.unsetf(std::ios::fixed);
        std::cout << std::setprecision(6);
    }

private:
    using Clock = std::chrono::steady_clock;
    static std::string fmt(double s)
    {
        std::ostringstream o;
        o << std::fixed;
        if (s < 1.0)
            o << std::setprecision(2) << s * 1000.0 << " ms";
        else
            o << std::setprecision(4) << s << " s";
        return o.str();
    }
    Clock::time_point origin_, last_;
    std::vector<std::pair<std::string, double>> recs_;
};

std::string info_name(const std::string &chain_file)
{
    const size_t pos = chain_file.find("chain");
    if (pos != std::string::npos) {
        std::string s = chain_file;
        s.replace(pos, 5, "info");
        return s;
    }
    return "info_" + chain_file;
}

// The reference writers announce the file, create ../outputs and keep going
// when the file cannot be opened (data_io.hpp:80-133).
template <typename Writer>
void write_output(const std::string &name, Writer w)
{
    std::cout << "writing data to a file " << name << " ..." << std::endl;
    const std::filesystem::path dir = "../outputs";
    std::error_code ec;
    std::filesystem::create_directories(dir, ec);
    if (ec) {
        std::cerr << "Failed to create output directory " << dir << ": " << ec.message() << std::endl;
        return;
    }
    const std::string path = dir.string() + "/" + name;
    if (w(path.c_str()) != CVG_OK) std::cout << "write data file cannot be opened!" << std::endl;
}

int fail_ctx(cvg_context *ctx, const char *what)
{
    std::cerr << "Error: " << what << ": " << cvg_last_error(ctx) << std::endl;
    cvg_destroy(ctx);
    return 1;
}

int run(const Options &o)
{
    std::cout << "Using scalar type: " << (o.use_double ? "double" : "float") << std::endl;
    const std::string path = "../test_files/" + o.edge_file;  // data_io.hpp:10-13
    if (FILE *f = std::fopen(path.c_str(), "rb")) {
        std::fclose(f);
    } else {
        std::cerr << "Error: input file does not exist: " << path << std::endl;
        return 1;
    }
    std::cout << "reading data from file " << o.edge_file << std::endl;
    double *edges = nullptr;
    float *edges_f = nullptr;
    int n = 0;
    char err[1024];
    const int rrc = o.use_double ? cvg_read_edge_file(path.c_str(), o.p.edge_data_sz, &edges, &n, err, sizeof(err))
                                 : cvg_read_edge_file_f32(path.c_str(), o.p.edge_data_sz, &edges_f, &n, err,
                                                          sizeof(err));
    if (rrc != CVG_OK) {
        std::cerr << err << std::endl;
        return 1;
    }
    std::cout << "read " << n << " edges from " << o.edge_file << std::endl;

    cvg_context *ctx = nullptr;
    const int rc = cvg_create(o.device, &ctx);
    if (rc != CVG_OK) {
        std::cerr << "Error: cannot open HIP device " << o.device << ": " << cvg_status_string(rc) << std::endl;
        cvg_free(edges);
        cvg_free(edges_f);
        return 1;
    }
    Laps timer;
    timer.start();
    const int grc = o.use_double
                        ? cvg_build_neighbor_csr_graph(ctx, n, o.p.edge_data_sz, edges, o.p.nrad, 3u, o.p.max_candidates)
                        : cvg_build_neighbor_csr_graph_f32(ctx, n, o.p.edge_data_sz, edges_f, o.p.nrad, 3u,
                                                           o.p.max_candidates);
    cvg_free(edges);
    cvg_free(edges_f);
    if (grc != CVG_OK) return fail_ctx(ctx, "build_neighbor_csr_graph");
    cvg_graph_info gi{};
    cvg_get_graph_info(ctx, &gi);
    if (gi.truncated_anchors > 0)
        std::cerr << "Warning: " << gi.truncated_anchors << " anchors hit max_candidates=" << o.p.max_candidates
                  << "; neighbor lists may be truncated\n";
    std::cout << "CSR neighbor graph: " << gi.num_edges << " anchors, " << gi.total_neighbor_pairs
              << " total pairs, max number of neighbors per anchor: " << gi.max_num_of_neighbors << std::endl;
    timer.lap("build CSR neighbor graph");

    unsigned valid_pairs = 0;
    if (cvg_form_pairwise_bundles(ctx, &o.p, &valid_pairs) != CVG_OK) return fail_ctx(ctx, "form_pairwise_bundles");
    timer.lap("form_pairwise_bundles");
    std::cout << "Pairwise curve bundles formed: " << valid_pairs << " valid pairs" << std::endl;

    unsigned num = 0, width = 0;
    if (cvg_build_curvelets_greedy(ctx, &o.p, &num) != CVG_OK) return fail_ctx(ctx, "build_curvelets_greedy");
    cvg_curvelet_shape(ctx, &num, &width);
    std::cout << "- Number of curvelets formed: " << num << std::endl;
    timer.lap("build_curvelets_greedy (pairwise+growth+dedup; see phase breakdown above)");
    timer.summary();

    std::cout << "(out_h, out_w) = (" << num << ", " << width << ")" << std::endl;
    std::vector<unsigned> chains(static_cast<size_t>(num) * width);
    std::vector<double> info(static_cast<size_t>(num) * 10);
    if (cvg_get_curvelets(ctx, chains.data(), info.data()) != CVG_OK) return fail_ctx(ctx, "get_curvelets");
    cvg_destroy(ctx);
    write_output(o.out_file, [&](const char *p) { return cvg_write_chain_file(p, chains.data(), num, width); });
    write_output(info_name(o.out_file),
                 [&](const char *p) { return cvg_write_info_file(p, info.data(), num, 10); });
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    Options o;
    cvg_params_default(&o.p);
    bool help = false;
    if (!parse(argc, argv, o, help)) return help ? 0 : 1;
    return run(o);
}
