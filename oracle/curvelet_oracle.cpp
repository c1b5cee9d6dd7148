// Test infrastructure only: the checker for the GPU path, never part of it.
//
// CPU restatement of cpu_curvelet's curvelet-construction path.  Every
// floating-point expression that feeds a decision or an output keeps the
// reference's operand order and rounding points (no contraction: the recipe in
// oracle/Makefile builds this with -ffp-contract=off), and both sorts call
// std::sort with the same inputs and ordering as the reference, so tie order is
// libstdc++'s.  Each function cites the reference lines it follows.
//
// libm calls: glibc 2.35's sincos() does not always return the same bits as
// separate sin() and cos() calls, so which calls GCC merges decides the
// reference's results.  The reference binary built by oracle/Makefile
// (g++ 11.4 -O3) calls, per double instantiation: sincos(T2) for the transport
// branch test; in branch A sincos(T0-dt) and sincos(T0+dt), in branch B
// separate sin()/cos() of T0+-dt; sincos(T2t) and sin(T2t - T0-+dt) in both
// cell loops; in the forward run separate cos()/sin() of the direction and of
// theta_a, in the backward run sincos() of both (the loop is unswitched on
// f_run); sincos(theta) in fill_curvelet_info.  The helpers below make exactly
// those calls and cannot be merged by the compiler.
#include "curvelet_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <unordered_map>
#include <vector>

#include <omp.h>

#ifdef ORC_CR_MATH
#include "../curvelet_construction_on_gpu_amd/csrc/cvg_math.hpp"
#endif

namespace {

// curvelet_utils.hpp:10 defines M_PI with this literal.
constexpr double kPi = 3.14159265358979323846;

// To2Pi, curvelet_utils.hpp:25-46: fmod into [0, 2pi), out-of-range results to 0.
double wrap_two_pi(double angle)
{
    const double two_pi = 2.0 * kPi;
    double a;
    if (angle >= two_pi)
        a = std::fmod(angle, two_pi);
    else if (angle < 0)
        a = two_pi + std::fmod(angle, two_pi);
    else
        a = angle;
    if (!(a >= 0 && a < two_pi)) a = 0;
    return a;
}

// sq_dist, curvelet_utils.hpp:13-17.
double dist2(double x1, double y1, double x2, double y2)
{
    return (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2);
}

#ifndef ORC_CR_MATH
// Opaque libm entry points: calls through volatile pointers are never merged.
double (*volatile libm_sin)(double) = ::sin;
double (*volatile libm_cos)(double) = ::cos;
void (*volatile libm_sincos)(double, double *, double *) = ::sincos;
double (*volatile libm_atan2)(double, double) = ::atan2;
#else
// Second build (liboracle_cr.so): the same restatement with every sin, cos,
// sincos and atan2 replaced by the GPU path's correctly rounded functions.
// It separates the arithmetic (which must then match the GPU bit for bit)
// from the libm difference (glibc vs correct rounding).
double libm_sin(double x) { return cvg::cr_sin(x, cvg::kSinCosTableHost); }
double libm_cos(double x) { return cvg::cr_cos(x, cvg::kSinCosTableHost); }
void libm_sincos(double x, double *s, double *c) { cvg::cr_sincos(x, cvg::kSinCosTableHost, s, c); }
double libm_atan2(double y, double x) { return cvg::cr_atan2(y, x, cvg::kAtanTableHost); }
#endif

// dot(angle_from_pt_to_pt(a, n), theta_a), curvelet_utils.hpp:58-69 as used at
// cpu_curvelet.hpp:459-461, with the calls of the forward (run 0) or backward
// (run 1) specialisation.
double direction_dot(int run, double ax, double ay, double at, double nx, double ny)
{
    const double dir = wrap_two_pi(libm_atan2(ny - ay, nx - ax));
    if (run == 0) {
        const double cd = libm_cos(dir), ca = libm_cos(at), sd = libm_sin(dir), sa = libm_sin(at);
        return cd * ca + sd * sa;
    }
    double sd, cd, sa, ca;
    libm_sincos(at, &sa, &ca);
    libm_sincos(dir, &sd, &cd);
    return cd * ca + sd * sa;
}

// pack_cell_key, preprocess.hpp:41-45.
int64_t cell_key(int cx, int cy)
{
    return (static_cast<int64_t>(cy) << 32) | static_cast<uint32_t>(cx);
}

struct Graph {
    int n = 0;
    int sz = 4;
    std::vector<double> edges;
    std::vector<int> offsets;
    std::vector<int> ids;
    std::vector<double> d2;
    unsigned max_deg = 0;
    unsigned truncated = 0;
};

struct Curvelets {
    unsigned width = 0;
    std::vector<unsigned> chains;
    std::vector<double> info;
};

struct Handle {
    Graph g;
    Curvelets c;
};

// The bundle grid offsets: init_curvelet_storage (cpu_curvelet.hpp:73-74) for
// the counts and compute_curve_bundle (:697-698) for the per-index offsets.
struct Grid {
    unsigned kx = 0, kt = 0;
    std::vector<double> off_x;  // delta_x2 per i
    std::vector<double> off_t;  // delta_t2 per j
    explicit Grid(const orc_params &p)
    {
        kx = 2 * std::floor(p.dx / p.sx + 0.5) + 1;
        kt = 2 * std::floor(p.dt / p.st + 0.5) + 1;
        off_x.resize(kx);
        off_t.resize(kt);
        for (unsigned i = 0; i < kx; ++i) off_x[i] = p.sx * ((double)i - (double(kx) - 1.0) / 2.0);
        for (unsigned j = 0; j < kt; ++j) off_t[j] = p.st * ((double)j - (double(kt) - 1.0) / 2.0);
    }
    unsigned cells() const { return kx * kt; }
};

// compute_curve_bundle + bundle_valid_check, cpu_curvelet.hpp:636-830.  lo/hi
// must hold -max_k/+max_k on entry.  Returns the hypothesis flag.
bool transport_bundle(const orc_params &p, const Grid &grid, double ax, double ay, double at, double nx,
                      double ny, double nt, double *lo, double *hi)
{
    const double DX = ax - nx;
    const double DY = ay - ny;
    const double DX2 = DX * DX;
    const double DY2 = DY * DY;
    const double T0 = nt;
    const double T2 = at;
    const double dx2 = p.dx * p.dx;
    const double T0p = T0 + p.dt;
    const double T0m = T0 - p.dt;

    if (DX2 + DY2 < 4.0 * p.dx * p.dx) return true;  // :682 close pair, default bundle

    double s2, c2;
    libm_sincos(T2, &s2, &c2);
    const bool branch_a = (DX * c2 + DY * s2) < 0;  // :690
    double sT0p, cT0p, sT0m, cT0m;
    if (branch_a) {
        libm_sincos(T0m, &sT0m, &cT0m);
        libm_sincos(T0p, &sT0p, &cT0p);
    } else {
        sT0p = libm_sin(T0p);
        cT0p = libm_cos(T0p);
        sT0m = libm_sin(T0m);
        cT0m = libm_cos(T0m);
    }
    const double sT0p_DY = DY * sT0p;  // :730-733
    const double cT0p_DX = DX * cT0p;
    const double sT0m_DY = DY * sT0m;
    const double cT0m_DX = DX * cT0m;

    unsigned c = 0;
    for (unsigned i = 0; i < grid.kx; ++i) {
        const double ox = grid.off_x[i];
        for (unsigned j = 0; j < grid.kt; ++j, ++c) {
            const double T2t = T2 + grid.off_t[j];
            double sT2t, cT2t;
            libm_sincos(T2t, &sT2t, &cT2t);
            const double tt = sT2t * DX - cT2t * DY;
            const double ttm = tt - ox;
            const double sm = libm_sin(T2t - T0m);
            const double sp = libm_sin(T2t - T0p);
            if (branch_a) {
                const double den = (-2.0 * ox * tt + DX2 + DY2 - dx2 + ox * ox);  // :711
                const double kx_min = 2.0 * (ttm - p.dx) / den;
                const double kx_max = 2.0 * (ttm + p.dx) / den;
                if (lo[c] < kx_min) lo[c] = kx_min;
                if (hi[c] > kx_max) hi[c] = kx_max;
                const double kt_min = sm / (-sm * ox + cT0m_DX + sT0m_DY);  // :736
                const double kt_max = sp / (-sp * ox + cT0p_DX + sT0p_DY);
                if (lo[c] < kt_min) lo[c] = kt_min;
                if (hi[c] > kt_max) hi[c] = kt_max;
            } else {
                const double den = -(2.0 * ox * tt - DX2 - DY2 + dx2 - ox * ox) / 2.0;  // :774
                const double kx_min = (ttm - p.dx) / den;
                const double kx_max = (ttm + p.dx) / den;
                if (lo[c] < kx_min) lo[c] = kx_min;
                if (hi[c] > kx_max) hi[c] = kx_max;
                const double kt_max = sm / (-sm * ox + cT0m_DX + sT0m_DY);  // :799, roles swapped
                if (hi[c] > kt_max) hi[c] = kt_max;
                const double kt_min = sp / (-sp * ox + cT0p_DX + sT0p_DY);
                if (lo[c] < kt_min) lo[c] = kt_min;
            }
        }
    }
    for (unsigned k = 0; k < grid.cells(); ++k)
        if (hi[k] > lo[k]) return true;
    return false;
}

void reset_bundle(double max_k, unsigned cells, double *lo, double *hi)
{
    for (unsigned k = 0; k < cells; ++k) {
        lo[k] = -max_k;
        hi[k] = max_k;
    }
}

// build_spatial_index_cpu + discover_neighbors_csr + build_neighbor_csr_graph,
// preprocess.hpp:62-274.
int build_graph(Graph &g, int n, int sz, const double *edges, const orc_params &p, int nthreads)
{
    if (n <= 0 || sz < 4 || edges == nullptr) return 1;
    g = Graph();
    g.n = n;
    g.sz = sz;
    g.edges.assign(edges, edges + static_cast<size_t>(n) * sz);
    const double *E = g.edges.data();

    std::vector<int64_t> keys(n);
    for (int i = 0; i < n; ++i)
        keys[i] = cell_key(static_cast<int>(std::lrint(E[i * sz])), static_cast<int>(std::lrint(E[i * sz + 1])));
    // The within-cell order is whatever std::sort makes of the identity
    // permutation under this comparison (preprocess.hpp:90-92); the same call
    // reproduces it exactly.
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    std::unordered_map<int64_t, std::pair<int, int>> runs;
    runs.reserve(static_cast<size_t>(n) * 2);
    for (int q = 0; q < n;) {
        const int64_t k = keys[order[q]];
        int e = q;
        while (e < n && keys[order[e]] == k) ++e;
        runs.emplace(k, std::make_pair(q, e));
        q = e;
    }

    const double rad2 = p.nrad * p.nrad;  // preprocess.hpp:221
    const int R = static_cast<int>(p.neighbor_radius);
    struct Cand {
        double d;
        int id;
    };
    std::vector<std::vector<Cand>> lists(n);
    unsigned truncated = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : truncated)
    for (int a = 0; a < n; ++a) {
        const double ax = E[a * sz], ay = E[a * sz + 1];
        const int cx = static_cast<int>(std::lrint(ax));
        const int cy = static_cast<int>(std::lrint(ay));
        std::vector<Cand> &cand = lists[a];
        for (int py = cy - R; py <= cy + R; ++py) {
            for (int px = cx - R; px <= cx + R; ++px) {
                auto it = runs.find(cell_key(px, py));
                if (it == runs.end()) continue;
                for (int q = it->second.first; q < it->second.second; ++q) {
                    const int b = order[q];
                    if (b == a) continue;
                    const double d = dist2(ax, ay, E[b * sz], E[b * sz + 1]);
                    if (d > rad2) continue;
                    if (static_cast<unsigned>(cand.size()) < p.max_candidates) cand.push_back({d, b});
                }
            }
        }
        std::sort(cand.begin(), cand.end(), [](const Cand &u, const Cand &v) { return u.d < v.d; });
        if (cand.size() >= p.max_candidates) ++truncated;  // :224, a >= test
    }

    g.offsets.assign(static_cast<size_t>(n) + 1, 0);
    for (int a = 0; a < n; ++a) {
        const unsigned deg = static_cast<unsigned>(lists[a].size());
        g.offsets[a + 1] = g.offsets[a] + static_cast<int>(deg);
        g.max_deg = std::max(g.max_deg, deg);
    }
    g.ids.resize(g.offsets[n]);
    g.d2.resize(g.offsets[n]);
    for (int a = 0; a < n; ++a) {
        int o = g.offsets[a];
        for (const Cand &c : lists[a]) {
            g.ids[o] = c.id;
            g.d2[o] = c.d;
            ++o;
        }
    }
    g.truncated = truncated;
    return 0;
}

// fill_curvelet_info, cpu_curvelet.hpp:197-256.
void curvelet_info(const orc_params &p, const Grid &grid, const Graph &g, bool forward, double ax, double ay,
                   double at, const double *lo, const double *hi, const unsigned *chain, unsigned len, double *out)
{
    double best = 100.0;
    unsigned bi = 0, bj = 0;
    for (unsigned i = 0; i < grid.kx; ++i)
        for (unsigned j = 0; j < grid.kt; ++j) {
            const unsigned c = i * grid.kt + j;
            if (hi[c] > lo[c]) {
                const double ox = grid.off_x[i];
                const double ot = grid.off_t[j];
                const double d = ox * ox + ot * ot;
                if (d < best) {
                    best = d;
                    bi = i;
                    bj = j;
                }
            }
        }
    const unsigned c = bi * grid.kt + bj;
    const double k_max = hi[c];
    const double k_min = lo[c];
    const double k = (k_max + k_min) / 2.0;
    const double ox = grid.off_x[bi];
    const double theta = wrap_two_pi(at + grid.off_t[bj]);
    double st, ct;
    libm_sincos(theta, &st, &ct);
    const double px = ax - ox * st;
    const double py = ay + ox * ct;
    double length = 0.0;
    const double *E = g.edges.data();
    for (unsigned m = 0; m + 1 < len; ++m) {
        const unsigned e1 = chain[m], e2 = chain[m + 1];
        length += std::sqrt(dist2(E[e1 * g.sz], E[e1 * g.sz + 1], E[e2 * g.sz], E[e2 * g.sz + 1]));
    }
    const double quality = (length > 0.0 && len > 0) ? 2.0 / (1.0 * p.nrad / length + 1.0 * length / double(len)) : 0.0;
    out[0] = forward ? 1.0 : 0.0;
    out[1] = k_max;
    out[2] = k_min;
    out[3] = at;
    out[4] = px;
    out[5] = py;
    out[6] = theta;
    out[7] = k;
    out[8] = length;
    out[9] = quality;
}

// build_curvelets_greedy, cpu_curvelet.hpp:363-634, one anchor.  Appends the
// anchor's accepted rows (chain row of width G+1 and 10 info values) in the
// reference's order: forward run first, seeds in CSR order.
void anchor_curvelets(const orc_params &p, const Grid &grid, const Graph &g, int a, std::vector<double> &lo,
                      std::vector<double> &hi, std::vector<unsigned char> &hyp, std::vector<unsigned> &rows,
                      std::vector<double> &info)
{
    const unsigned K = grid.cells();
    const unsigned G = p.group_max_sz;
    const int beg = g.offsets[a], deg = g.offsets[a + 1] - g.offsets[a];
    const double *E = g.edges.data();
    const double ax = E[a * g.sz], ay = E[a * g.sz + 1], at = E[a * g.sz + 2];
    const unsigned cap = std::max(G, 2u) + 1;
    std::vector<unsigned> chain(cap);
    std::vector<double> run_lo(K), run_hi(K), cut_lo(K), cut_hi(K);
    std::vector<std::vector<unsigned>> accepted;

    for (int run = 0; run < 2; ++run) {
        accepted.clear();
        for (int r = 0; r < deg; ++r) {
            hyp[r] = 0;
            reset_bundle(p.max_k, K, &lo[static_cast<size_t>(r) * K], &hi[static_cast<size_t>(r) * K]);
            const int b = g.ids[beg + r];
            const double dd = direction_dot(run, ax, ay, at, E[b * g.sz], E[b * g.sz + 1]);
            const bool pass = (run == 0) ? (dd > 0) : (dd < 0);
            if (pass)
                hyp[r] = transport_bundle(p, grid, ax, ay, at, E[b * g.sz], E[b * g.sz + 1], E[b * g.sz + 2],
                                          &lo[static_cast<size_t>(r) * K], &hi[static_cast<size_t>(r) * K]);
        }
        for (int s = 0; s < deg; ++s) {
            if (!hyp[s]) continue;
            unsigned len = 0;
            chain[len++] = static_cast<unsigned>(a);
            std::copy_n(&lo[static_cast<size_t>(s) * K], K, run_lo.begin());
            std::copy_n(&hi[static_cast<size_t>(s) * K], K, run_hi.begin());
            for (int r = 0; r < deg; ++r) {
                if (!hyp[r]) continue;
                const unsigned nb = static_cast<unsigned>(g.ids[beg + r]);
                if (r == s) {  // the seed joins at its own distance rank, :496-510
                    chain[len++] = nb;
                    if (len >= G) break;
                    continue;
                }
                const double *blo = &lo[static_cast<size_t>(r) * K];
                const double *bhi = &hi[static_cast<size_t>(r) * K];
                bool overlap = false;
                for (unsigned k = 0; k < K; ++k) {  // bundle_intersection + valid check, :832-854
                    cut_lo[k] = (run_lo[k] < blo[k]) ? blo[k] : run_lo[k];
                    cut_hi[k] = (run_hi[k] > bhi[k]) ? bhi[k] : run_hi[k];
                    overlap = overlap || (cut_hi[k] > cut_lo[k]);
                }
                if (overlap) {
                    chain[len++] = nb;
                    run_lo.swap(cut_lo);
                    run_hi.swap(cut_hi);
                }
                if (len >= G) break;
            }
            if (len <= 2) continue;  // :549
            bool dup = false;  // check_curvelet_exist, :876-916
            for (const auto &prev : accepted)
                if (prev.size() == len && std::equal(prev.begin(), prev.end(), chain.begin())) {
                    dup = true;
                    break;
                }
            if (dup) continue;
            accepted.emplace_back(chain.begin(), chain.begin() + len);
            // record_curvelet_chain, :258-280
            const size_t base = rows.size();
            rows.resize(base + G + 1, 0u);
            rows[base] = static_cast<unsigned>(a) + 1;
            for (unsigned m = 0; m < len && m + 1 < G + 1; ++m) rows[base + m + 1] = chain[m] + 1;
            const size_t ib = info.size();
            info.resize(ib + 10);
            curvelet_info(p, grid, g, run == 0, ax, ay, at, run_lo.data(), run_hi.data(), chain.data(), len,
                          &info[ib]);
        }
    }
}

}  // namespace

extern "C" {

double orc_dt_from_deg(double dt_deg) { return (dt_deg / 180.0) * kPi; }

// build_spatial_index_cpu's sorted_edge_ids (preprocess.hpp:78-96): the edge
// ids by packed cell key under libstdc++'s std::sort, the same call as above.
void orc_spatial_order(int n, const double *x, const double *y, int *sorted_ids)
{
    std::vector<int64_t> keys(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        keys[i] = cell_key(static_cast<int>(std::lrint(x[i])), static_cast<int>(std::lrint(y[i])));
    std::iota(sorted_ids, sorted_ids + n, 0);
    std::sort(sorted_ids, sorted_ids + n, [&](int a, int b) { return keys[a] < keys[b]; });
}

void *orc_create(void) { return new Handle(); }
void orc_destroy(void *h) { delete static_cast<Handle *>(h); }

int orc_build_csr(void *h, int n, int sz, const double *edges, const orc_params *p, int nthreads)
{
    return build_graph(static_cast<Handle *>(h)->g, n, sz, edges, *p, nthreads);
}

void orc_csr_counts(void *h, int *pairs, unsigned *max_deg, unsigned *truncated_anchors)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    *pairs = g.offsets.empty() ? 0 : g.offsets.back();
    *max_deg = g.max_deg;
    *truncated_anchors = g.truncated;
}

void orc_get_csr(void *h, int *offsets, int *ids, double *dist2_out)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    std::copy(g.offsets.begin(), g.offsets.end(), offsets);
    std::copy(g.ids.begin(), g.ids.end(), ids);
    std::copy(g.d2.begin(), g.d2.end(), dist2_out);
}

unsigned orc_valid_pairs(void *h, const orc_params *p, int nthreads)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    const Grid grid(*p);
    const unsigned K = grid.cells();
    unsigned total = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : total)
    {
        std::vector<double> lo(K), hi(K);
#pragma omp for schedule(dynamic, 64)
        for (int a = 0; a < g.n; ++a) {
            const double *E = g.edges.data();
            for (int q = g.offsets[a]; q < g.offsets[a + 1]; ++q) {
                const int b = g.ids[q];
                reset_bundle(p->max_k, K, lo.data(), hi.data());
                if (transport_bundle(*p, grid, E[a * g.sz], E[a * g.sz + 1], E[a * g.sz + 2], E[b * g.sz],
                                     E[b * g.sz + 1], E[b * g.sz + 2], lo.data(), hi.data()))
                    ++total;
            }
        }
    }
    return total;
}

unsigned orc_greedy(void *h, const orc_params *p, int nthreads)
{
    Handle &H = *static_cast<Handle *>(h);
    const Graph &g = H.g;
    const Grid grid(*p);
    const unsigned K = grid.cells();
    std::vector<std::vector<unsigned>> rows(g.n);
    std::vector<std::vector<double>> info(g.n);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        std::vector<double> lo(static_cast<size_t>(g.max_deg) * K), hi(static_cast<size_t>(g.max_deg) * K);
        std::vector<unsigned char> hyp(g.max_deg);
#pragma omp for schedule(dynamic, 16)
        for (int a = 0; a < g.n; ++a) anchor_curvelets(*p, grid, g, a, lo, hi, hyp, rows[a], info[a]);
    }
    // compact_curvelet_output, :282-303: anchor-major concatenation.
    Curvelets &c = H.c;
    c.width = p->group_max_sz + 1;
    c.chains.clear();
    c.info.clear();
    for (int a = 0; a < g.n; ++a) {
        c.chains.insert(c.chains.end(), rows[a].begin(), rows[a].end());
        c.info.insert(c.info.end(), info[a].begin(), info[a].end());
    }
    return static_cast<unsigned>(c.info.size() / 10);
}

unsigned orc_chain_width(void *h) { return static_cast<Handle *>(h)->c.width; }

void orc_get_curvelets(void *h, unsigned *chains, double *info)
{
    const Curvelets &c = static_cast<Handle *>(h)->c;
    std::copy(c.chains.begin(), c.chains.end(), chains);
    std::copy(c.info.begin(), c.info.end(), info);
}

int orc_bundle_cells(const orc_params *p, unsigned *kx, unsigned *kt)
{
    const Grid grid(*p);
    *kx = grid.kx;
    *kt = grid.kt;
    return static_cast<int>(grid.cells());
}

int orc_pair_bundle(const orc_params *p, double ax, double ay, double at, double nx, double ny, double nt,
                    double *bmin, double *bmax)
{
    const Grid grid(*p);
    reset_bundle(p->max_k, grid.cells(), bmin, bmax);
    return transport_bundle(*p, grid, ax, ay, at, nx, ny, nt, bmin, bmax) ? 1 : 0;
}

double orc_dir_dot(int run, double ax, double ay, double at, double nx, double ny)
{
    return direction_dot(run, ax, ay, at, nx, ny);
}

int orc_anchor_detail(void *h, const orc_params *p, int anchor, unsigned char *dir, unsigned char *hyp,
                      double *bmin, double *bmax)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    if (anchor < 0 || anchor >= g.n) return -1;
    const Grid grid(*p);
    const unsigned K = grid.cells();
    const double *E = g.edges.data();
    const int a = anchor;
    int r = 0;
    for (int q = g.offsets[a]; q < g.offsets[a + 1]; ++q, ++r) {
        const int b = g.ids[q];
        const double ax = E[a * g.sz], ay = E[a * g.sz + 1], at = E[a * g.sz + 2];
        const bool fwd = direction_dot(0, ax, ay, at, E[b * g.sz], E[b * g.sz + 1]) > 0;
        const bool bwd = direction_dot(1, ax, ay, at, E[b * g.sz], E[b * g.sz + 1]) < 0;
        dir[r] = (fwd ? 1 : 0) | (bwd ? 2 : 0);
        double *lo = bmin + static_cast<size_t>(r) * K;
        double *hi = bmax + static_cast<size_t>(r) * K;
        reset_bundle(p->max_k, K, lo, hi);
        hyp[r] = transport_bundle(*p, grid, E[a * g.sz], E[a * g.sz + 1], E[a * g.sz + 2], E[b * g.sz],
                                  E[b * g.sz + 1], E[b * g.sz + 2], lo, hi)
                     ? 1
                     : 0;
    }
    return r;
}

}  // extern "C"
