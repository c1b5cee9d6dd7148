// Test infrastructure only: never linked into, or called by, the product path.
//
// A thin extern "C" harness around the reference's own header-only cpu_curvelet
// code, compiled in place from the reference tree by oracle/Makefile into
// oracle/_ref/libcurvelet_ref.so.  The reference CLI prints curvelet info with
// only 6 significant digits (cpu_curvelet/data_io.hpp:128); this harness hands
// the same computation back at full precision so the oracle restatement and the
// GPU path can be pinned to it.  It mirrors run_curvelet<double>
// (cpu_curvelet/main.cpp:19-96): the same parameter casts (:22-30), the same
// call sequence (:47-64) and the same column-major → row-major info transpose
// (:81-88).
// main.cpp includes <math.h> (libstdc++'s wrapper, which puts std::sin,
// std::cos, ... into the global namespace) before the reference headers, so
// their unqualified sin/cos calls resolve to the float overloads in the
// float instantiation; the harness includes it first for the same meaning.
#include <math.h>

#include <chrono>
#include <iostream>
#include <sstream>
#include <vector>

#include "preprocess.hpp"
#include "cpu_curvelet.hpp"

extern "C" {
typedef struct ref_params {
    double nrad;
    double dx;
    double dt_deg;
    double max_k;
    double sx;
    double st;
    unsigned group_max_sz;
    unsigned max_candidates;
} ref_params;
}

namespace {

struct RefState {
    CPUNeighborGraph<double> graph;
    std::vector<double> edges;
    int n = 0;
    int sz = 4;
    unsigned valid_pairs = 0;
    unsigned num_curvelets = 0;
    unsigned width = 0;
    std::vector<unsigned> chains;
    std::vector<double> info;
    double t_csr = 0, t_pairwise = 0, t_greedy = 0;
};

RefState g_state;

struct QuietCout {
    std::ostringstream sink;
    std::streambuf *old = nullptr;
    explicit QuietCout(bool on) { if (on) old = std::cout.rdbuf(sink.rdbuf()); }
    ~QuietCout() { if (old) std::cout.rdbuf(old); }
};

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

// Radian tolerance exactly as main.cpp:24 computes it for T = double.
double ref_dt_from_deg(double dt_deg) { return double(dt_deg / 180.0) * double(M_PI); }

// Whole reference path on an in-memory edge array (row-major n x sz).
// do_pairwise != 0 also runs form_pairwise_bundles (the valid-pair count).
int ref_run(int n, int sz, const double *edges, const ref_params *p, int nthreads,
            int quiet, int do_pairwise)
{
    QuietCout q(quiet != 0);
    RefState &s = g_state;
    s = RefState();
    s.n = n;
    s.sz = sz;
    s.edges.assign(edges, edges + static_cast<size_t>(n) * sz);

    const double nrad = p->nrad;
    const double dx = p->dx;
    const double dt = ref_dt_from_deg(p->dt_deg);
    const double max_k = p->max_k;
    const double sx = p->sx;
    const double st = p->st;

    auto t0 = std::chrono::steady_clock::now();
    if (!build_neighbor_csr_graph(n, sz, s.edges.data(), nrad, 3u, p->max_candidates, nthreads, s.graph))
        return 1;
    s.t_csr = seconds_since(t0);

    int n_edges = n;
    int sz_edge = sz;
    CurveletCPU<double> obj(n_edges, sz_edge, s.graph, dx, dt, sx, st, max_k, p->group_max_sz, nthreads,
                            s.edges.data(), nrad);
    if (do_pairwise) {
        t0 = std::chrono::steady_clock::now();
        CPUCurveletFormationResult res;
        if (!obj.form_pairwise_bundles(res)) return 2;
        s.valid_pairs = res.valid_pairs;
        s.t_pairwise = seconds_since(t0);
    }
    t0 = std::chrono::steady_clock::now();
    obj.build_curvelets_greedy();
    s.t_greedy = seconds_since(t0);

    s.num_curvelets = obj.num_curvelets();
    s.width = obj.chain_width();
    s.chains.assign(obj._edge_chain_final,
                    obj._edge_chain_final + static_cast<size_t>(s.num_curvelets) * s.width);
    s.info.resize(static_cast<size_t>(s.num_curvelets) * 10);
    const unsigned stride = obj._max_curvelets;
    for (unsigned row = 0; row < s.num_curvelets; ++row)
        for (unsigned col = 0; col < 10; ++col)
            s.info[static_cast<size_t>(row) * 10 + col] = obj._curvelet_info[static_cast<size_t>(col) * stride + row];
    return 0;
}

void ref_counts(int *pairs, unsigned *max_deg, unsigned *valid_pairs, unsigned *num_curvelets, unsigned *width)
{
    *pairs = g_state.graph.total_neighbor_pairs;
    *max_deg = g_state.graph.max_num_of_neighbors;
    *valid_pairs = g_state.valid_pairs;
    *num_curvelets = g_state.num_curvelets;
    *width = g_state.width;
}

void ref_times(double *t3)
{
    t3[0] = g_state.t_csr;
    t3[1] = g_state.t_pairwise;
    t3[2] = g_state.t_greedy;
}

void ref_get_csr(int *offsets, int *ids, double *dist2)
{
    const auto &g = g_state.graph;
    for (size_t i = 0; i < g.neighbor_offsets.size(); ++i) offsets[i] = g.neighbor_offsets[i];
    for (size_t i = 0; i < g.neighbor_ids.size(); ++i) ids[i] = g.neighbor_ids[i];
    for (size_t i = 0; i < g.neighbor_dist2.size(); ++i) dist2[i] = g.neighbor_dist2[i];
}

void ref_get_curvelets(unsigned *chains, double *info)
{
    for (size_t i = 0; i < g_state.chains.size(); ++i) chains[i] = g_state.chains[i];
    for (size_t i = 0; i < g_state.info.size(); ++i) info[i] = g_state.info[i];
}

// One (anchor, neighbour) bundle through CurveletCPU::compute_curve_bundle
// (cpu_curvelet.hpp:636-819).  Writes the K cells of the bundle, returns the
// hypothesis flag (valid or "close").
int ref_pair_bundle(const ref_params *p, double ax, double ay, double at, double nx, double ny, double nt,
                    double *bmin, double *bmax, int *k_cells)
{
    std::vector<double> e = {ax, ay, at, 0.0, nx, ny, nt, 0.0};
    CPUNeighborGraph<double> g;
    g.num_edges = 2;
    g.total_neighbor_pairs = 1;
    g.max_num_of_neighbors = 1;
    g.edges = e;
    g.neighbor_offsets = {0, 1, 1};
    g.neighbor_ids = {1};
    g.neighbor_dist2 = {0.0};
    int n_edges = 2, sz = 4;
    const double dt = ref_dt_from_deg(p->dt_deg);
    CurveletCPU<double> obj(n_edges, sz, g, p->dx, dt, p->sx, p->st, p->max_k, p->group_max_sz, 1, e.data(),
                            p->nrad);
    const unsigned kx = 2 * floor(p->dx / p->sx + 0.5) + 1;
    const unsigned kt = 2 * floor(dt / p->st + 0.5) + 1;
    const unsigned K = kx * kt;
    for (unsigned c = 0; c < K; ++c) {
        bmin[c] = -p->max_k;
        bmax[c] = p->max_k;
    }
    bool hyp[1] = {false};
    obj.compute_curve_bundle(0, 0, bmin, bmax, hyp, ax, ay, nx, ny, at, nt);
    *k_cells = static_cast<int>(K);
    return hyp[0] ? 1 : 0;
}

// dot(angle_from_pt_to_pt(anchor, neighbour), anchor orientation): the direction
// filter value of build_curvelets_greedy (cpu_curvelet.hpp:459-461).
double ref_dir_dot(double ax, double ay, double at, double nx, double ny)
{
    return dot(angle_from_pt_to_pt(ax, ay, nx, ny), at);
}

}  // extern "C"

// run_curvelet<float> (main.cpp:19-96 with --float): the same sequence on the
// float edge array the reference's float reader produces, the float parameter
// casts of :22-30, info widened to double as _curvelet_info stores it.
namespace {
struct RefStateF {
    CPUNeighborGraph<float> graph;
    std::vector<float> edges;
    unsigned valid_pairs = 0, num_curvelets = 0, width = 0;
    std::vector<unsigned> chains;
    std::vector<double> info;
    std::vector<int> offsets, ids;
    std::vector<double> d2;
};
RefStateF g_state_f;
}  // namespace

extern "C" {

int ref_run_f32(int n, int sz, const float *edges, const ref_params *p, int nthreads, int quiet, int do_pairwise)
{
    QuietCout q(quiet != 0);
    RefStateF &s = g_state_f;
    s = RefStateF();
    s.edges.assign(edges, edges + static_cast<size_t>(n) * sz);
    const float nrad = float(p->nrad);
    const float dx = float(p->dx);
    const float dt = float(p->dt_deg / 180.0) * float(M_PI);
    const float max_k = float(p->max_k);
    const float sx = float(p->sx);
    const float st = float(p->st);
    if (!build_neighbor_csr_graph(n, sz, s.edges.data(), nrad, 3u, p->max_candidates, nthreads, s.graph)) return 1;
    int n_edges = n;
    int sz_edge = sz;
    CurveletCPU<float> obj(n_edges, sz_edge, s.graph, dx, dt, sx, st, max_k, p->group_max_sz, nthreads,
                           s.edges.data(), nrad);
    if (do_pairwise) {
        CPUCurveletFormationResult res;
        if (!obj.form_pairwise_bundles(res)) return 2;
        s.valid_pairs = res.valid_pairs;
    }
    obj.build_curvelets_greedy();
    s.num_curvelets = obj.num_curvelets();
    s.width = obj.chain_width();
    s.chains.assign(obj._edge_chain_final, obj._edge_chain_final + static_cast<size_t>(s.num_curvelets) * s.width);
    s.info.resize(static_cast<size_t>(s.num_curvelets) * 10);
    const unsigned stride = obj._max_curvelets;
    for (unsigned row = 0; row < s.num_curvelets; ++row)
        for (unsigned col = 0; col < 10; ++col)
            s.info[static_cast<size_t>(row) * 10 + col] = obj._curvelet_info[static_cast<size_t>(col) * stride + row];
    s.offsets = s.graph.neighbor_offsets;
    s.ids = s.graph.neighbor_ids;
    s.d2.assign(s.graph.neighbor_dist2.begin(), s.graph.neighbor_dist2.end());
    return 0;
}

void ref_counts_f32(int *pairs, unsigned *max_deg, unsigned *valid_pairs, unsigned *num_curvelets, unsigned *width)
{
    *pairs = g_state_f.graph.total_neighbor_pairs;
    *max_deg = g_state_f.graph.max_num_of_neighbors;
    *valid_pairs = g_state_f.valid_pairs;
    *num_curvelets = g_state_f.num_curvelets;
    *width = g_state_f.width;
}

void ref_get_csr_f32(int *offsets, int *ids, double *dist2)
{
    std::copy(g_state_f.offsets.begin(), g_state_f.offsets.end(), offsets);
    std::copy(g_state_f.ids.begin(), g_state_f.ids.end(), ids);
    std::copy(g_state_f.d2.begin(), g_state_f.d2.end(), dist2);
}

void ref_get_curvelets_f32(unsigned *chains, double *info)
{
    std::copy(g_state_f.chains.begin(), g_state_f.chains.end(), chains);
    std::copy(g_state_f.info.begin(), g_state_f.info.end(), info);
}

}  // extern "C"

extern "C" {

// compute_curve_bundle<float> for one pair (the float twin of ref_pair_bundle).
int ref_pair_bundle_f32(const ref_params *p, float ax, float ay, float at, float nx, float ny, float nt, float *bmin,
                        float *bmax, int *k_cells)
{
    std::vector<float> e = {ax, ay, at, 0.0f, nx, ny, nt, 0.0f};
    CPUNeighborGraph<float> g;
    g.num_edges = 2;
    g.total_neighbor_pairs = 1;
    g.max_num_of_neighbors = 1;
    g.edges = e;
    g.neighbor_offsets = {0, 1, 1};
    g.neighbor_ids = {1};
    g.neighbor_dist2 = {0.0f};
    int n_edges = 2, sz = 4;
    const float dx = float(p->dx), sx = float(p->sx), st = float(p->st), max_k = float(p->max_k);
    const float dt = float(p->dt_deg / 180.0) * float(M_PI);
    CurveletCPU<float> obj(n_edges, sz, g, dx, dt, sx, st, max_k, p->group_max_sz, 1, e.data(), float(p->nrad));
    const unsigned kx = 2 * floor(dx / sx + 0.5) + 1;
    const unsigned kt = 2 * floor(dt / st + 0.5) + 1;
    const unsigned K = kx * kt;
    for (unsigned c = 0; c < K; ++c) {
        bmin[c] = -max_k;
        bmax[c] = max_k;
    }
    bool hyp[1] = {false};
    obj.compute_curve_bundle(0, 0, bmin, bmax, hyp, ax, ay, nx, ny, at, nt);
    *k_cells = static_cast<int>(K);
    return hyp[0] ? 1 : 0;
}

}  // extern "C"
