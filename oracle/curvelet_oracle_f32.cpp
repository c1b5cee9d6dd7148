// Test infrastructure only: the checker for the GPU path's float mode, never
// part of it.
//
// CPU restatement of cpu_curvelet's float instantiation (run_curvelet<float>,
// cpu_curvelet/main.cpp:110 with --float).  Every value the reference keeps in
// T is a float here and every operation on two floats is a float operation;
// the reference's double literals promote some expressions to double (the
// bundle grid offsets, :697-698, and the pixel-transport denominators and
// k_dx quotients of branch A, :711-715, and the denominator of branch B,
// :774), and those are evaluated in double and rounded to float where the
// reference assigns them to T.  No contraction (oracle/Makefile builds with
// -ffp-contract=off; the reference's x86-64 build has no FMA).
//
// libm: main.cpp includes <math.h> (libstdc++'s wrapper puts std::sin and
// friends into the global namespace) before the reference headers, so even
// compute_curve_bundle's unqualified sin and cos take the float overloads:
// the float instantiation calls sinf, cosf, sincosf, atan2f, fmodf, lrintf and
// sqrtf (its binary's dynamic symbols).  glibc 2.35's sinf/cosf give the same
// bits as its sincosf (checked on 5e7 arguments), so which calls the compiler
// merges does not matter.  liboracle_f32.so calls glibc through volatile
// pointers (the reference's results); liboracle_f32_cr.so calls the correctly
// rounded float functions of the GPU path (cvg_math.hpp), whose results the
// GPU must match bit for bit.
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

extern "C" {
// Raw CurveletParams values (param_settings.hpp:9-24); the float casts of
// main.cpp:22-30 happen inside.
typedef struct orcf_params {
    double nrad;
    double dx;
    double dt_deg;
    double max_k;
    double sx;
    double st;
    unsigned group_max_sz;
    unsigned max_candidates;
    unsigned neighbor_radius;
} orcf_params;
}

namespace {

constexpr double kPi = 3.14159265358979323846;  // curvelet_utils.hpp:10

#ifndef ORC_CR_MATH
float (*volatile libm_sinf)(float) = ::sinf;
float (*volatile libm_cosf)(float) = ::cosf;
float (*volatile libm_atan2f)(float, float) = ::atan2f;
#else
float libm_sinf(float x) { return cvg::cr_sinf(x, cvg::kSinCosTableHost); }
float libm_cosf(float x) { return cvg::cr_cosf(x, cvg::kSinCosTableHost); }
float libm_atan2f(float y, float x) { return cvg::cr_atan2f(y, x, cvg::kAtanTableHost); }
#endif

// The float values of main.cpp:22-30.
struct P {
    float nrad, dx, dt, max_k, sx, st;
    unsigned G, max_cand, R;
    explicit P(const orcf_params &p)
    {
        nrad = static_cast<float>(p.nrad);
        dx = static_cast<float>(p.dx);
        dt = static_cast<float>(p.dt_deg / 180.0) * static_cast<float>(kPi);
        max_k = static_cast<float>(p.max_k);
        sx = static_cast<float>(p.sx);
        st = static_cast<float>(p.st);
        G = p.group_max_sz;
        max_cand = p.max_candidates;
        R = p.neighbor_radius;
    }
};

// To2Pi<float>, curvelet_utils.hpp:25-46.
float to2pi(float angle)
{
    const float two_pi = 2.0f * static_cast<float>(kPi);
    float a;
    if (angle >= two_pi)
        a = std::fmod(angle, two_pi);
    else if (angle < 0)
        a = two_pi + std::fmod(angle, two_pi);
    else
        a = angle;
    if (!(a >= 0 && a < two_pi)) a = 0;
    return a;
}

// sq_dist<float>, curvelet_utils.hpp:13-17.
float dist2(float x1, float y1, float x2, float y2) { return (x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2); }

// dot(angle_from_pt_to_pt(a, n), theta_a), curvelet_utils.hpp:58-69.
float direction_dot(float ax, float ay, float at, float nx, float ny)
{
    const float dir = to2pi(libm_atan2f(ny - ay, nx - ax));
    return libm_cosf(dir) * libm_cosf(at) + libm_sinf(dir) * libm_sinf(at);
}

int64_t cell_key(int cx, int cy) { return (static_cast<int64_t>(cy) << 32) | static_cast<uint32_t>(cx); }

struct Grid {
    unsigned kx = 0, kt = 0;
    std::vector<float> off_x, off_t;  // delta_x2, delta_t2 of compute_curve_bundle (:697-698)
    std::vector<float> inf_x, inf_t;  // ddx, ddt of fill_curvelet_info (:209-210)
    explicit Grid(const P &p)
    {
        kx = 2 * std::floor(p.dx / p.sx + 0.5) + 1;  // init_curvelet_storage, :73-74
        kt = 2 * std::floor(p.dt / p.st + 0.5) + 1;
        off_x.resize(kx);
        inf_x.resize(kx);
        off_t.resize(kt);
        inf_t.resize(kt);
        for (unsigned i = 0; i < kx; ++i) {
            off_x[i] = static_cast<float>(p.sx * (static_cast<float>(i) - (static_cast<float>(kx) - 1.0) / 2.0));
            inf_x[i] = p.sx * (static_cast<float>(i) - (static_cast<float>(kx) - 1.0f) / 2.0f);
        }
        for (unsigned j = 0; j < kt; ++j) {
            off_t[j] = static_cast<float>(p.st * (static_cast<float>(j) - (static_cast<float>(kt) - 1.0) / 2.0));
            inf_t[j] = p.st * (static_cast<float>(j) - (static_cast<float>(kt) - 1.0f) / 2.0f);
        }
    }
    unsigned cells() const { return kx * kt; }
};

// compute_curve_bundle<float> + bundle_valid_check, cpu_curvelet.hpp:636-830.
bool transport_bundle(const P &p, const Grid &grid, float ax, float ay, float at, float nx, float ny, float nt,
                      float *lo, float *hi)
{
    const float DX = ax - nx;
    const float DY = ay - ny;
    const float DX2 = DX * DX;
    const float DY2 = DY * DY;
    const float T0 = nt;
    const float T2 = at;
    const float dx2 = p.dx * p.dx;
    const float T0p = T0 + p.dt;
    const float T0m = T0 - p.dt;
    if (DX2 + DY2 < 4 * p.dx * p.dx) return true;  // :682
    const bool branch_a = (DX * libm_cosf(T2) + DY * libm_sinf(T2)) < 0;  // :690
    const float sT0p_DY = DY * libm_sinf(T0p);  // :730-733
    const float cT0p_DX = DX * libm_cosf(T0p);
    const float sT0m_DY = DY * libm_sinf(T0m);
    const float cT0m_DX = DX * libm_cosf(T0m);
    unsigned c = 0;
    for (unsigned i = 0; i < grid.kx; ++i) {
        const float ox = grid.off_x[i];
        for (unsigned j = 0; j < grid.kt; ++j, ++c) {
            const float T2t = T2 + grid.off_t[j];
            const float tt = libm_sinf(T2t) * DX - libm_cosf(T2t) * DY;  // :705
            const float ttm = tt - ox;
            const float sm = libm_sinf(T2t - T0m);  // :728-729
            const float sp = libm_sinf(T2t - T0p);
            if (branch_a) {
                // :711, double literals: double sum, the last product in float
                const float den = static_cast<float>(-2.0 * ox * tt + DX2 + DY2 - dx2 + ox * ox);
                const float kx_min = static_cast<float>(2.0 * (ttm - p.dx) / den);
                const float kx_max = static_cast<float>(2.0 * (ttm + p.dx) / den);
                if (lo[c] < kx_min) lo[c] = kx_min;
                if (hi[c] > kx_max) hi[c] = kx_max;
                const float kt_min = sm / (-sm * ox + cT0m_DX + sT0m_DY);
                const float kt_max = sp / (-sp * ox + cT0p_DX + sT0p_DY);
                if (lo[c] < kt_min) lo[c] = kt_min;
                if (hi[c] > kt_max) hi[c] = kt_max;
            } else {
                const float den = static_cast<float>(-(2.0 * ox * tt - DX2 - DY2 + dx2 - ox * ox) / 2.0);  // :774
                const float kx_min = (ttm - p.dx) / den;
                const float kx_max = (ttm + p.dx) / den;
                if (lo[c] < kx_min) lo[c] = kx_min;
                if (hi[c] > kx_max) hi[c] = kx_max;
                const float kt_max = sm / (-sm * ox + cT0m_DX + sT0m_DY);
                if (hi[c] > kt_max) hi[c] = kt_max;
                const float kt_min = sp / (-sp * ox + cT0p_DX + sT0p_DY);
                if (lo[c] < kt_min) lo[c] = kt_min;
            }
        }
    }
    for (unsigned k = 0; k < grid.cells(); ++k)
        if (hi[k] > lo[k]) return true;
    return false;
}

void reset_bundle(float max_k, unsigned cells, float *lo, float *hi)
{
    for (unsigned k = 0; k < cells; ++k) {
        lo[k] = -max_k;
        hi[k] = max_k;
    }
}

struct Graph {
    int n = 0, sz = 4;
    std::vector<float> edges;
    std::vector<int> offsets, ids;
    std::vector<float> d2;
    unsigned max_deg = 0, truncated = 0;
};

struct Handle {
    Graph g;
    unsigned width = 0;
    std::vector<unsigned> chains;
    std::vector<double> info;
};

// build_neighbor_csr_graph<float>, preprocess.hpp:62-274.
int build_graph(Graph &g, int n, int sz, const float *edges, const P &p, int nthreads)
{
    if (n <= 0 || sz < 4 || edges == nullptr) return 1;
    g = Graph();
    g.n = n;
    g.sz = sz;
    g.edges.assign(edges, edges + static_cast<size_t>(n) * sz);
    const float *E = g.edges.data();
    std::vector<int64_t> keys(n);
    for (int i = 0; i < n; ++i)
        keys[i] = cell_key(static_cast<int>(std::lrint(E[i * sz])), static_cast<int>(std::lrint(E[i * sz + 1])));
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
    const float rad2 = p.nrad * p.nrad;  // :221, rad * rad in T
    const int R = static_cast<int>(p.R);
    struct Cand {
        float d;
        int id;
    };
    std::vector<std::vector<Cand>> lists(n);
    unsigned truncated = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : truncated)
    for (int a = 0; a < n; ++a) {
        const float ax = E[a * sz], ay = E[a * sz + 1];
        const int cx = static_cast<int>(std::lrint(ax));
        const int cy = static_cast<int>(std::lrint(ay));
        std::vector<Cand> &cand = lists[a];
        for (int py = cy - R; py <= cy + R; ++py)
            for (int px = cx - R; px <= cx + R; ++px) {
                auto it = runs.find(cell_key(px, py));
                if (it == runs.end()) continue;
                for (int q = it->second.first; q < it->second.second; ++q) {
                    const int b = order[q];
                    if (b == a) continue;
                    const float d = dist2(ax, ay, E[b * sz], E[b * sz + 1]);
                    if (d > rad2) continue;
                    if (static_cast<unsigned>(cand.size()) < p.max_cand) cand.push_back({d, b});
                }
            }
        std::sort(cand.begin(), cand.end(), [](const Cand &u, const Cand &v) { return u.d < v.d; });
        if (cand.size() >= p.max_cand) ++truncated;
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

// fill_curvelet_info<float>, cpu_curvelet.hpp:197-256.
void curvelet_info(const P &p, const Grid &grid, const Graph &g, bool forward, float ax, float ay, float at,
                   const float *lo, const float *hi, const unsigned *chain, unsigned len, double *out)
{
    float best = 100.0f;
    unsigned bi = 0, bj = 0;
    for (unsigned i = 0; i < grid.kx; ++i)
        for (unsigned j = 0; j < grid.kt; ++j) {
            const unsigned c = i * grid.kt + j;
            if (hi[c] > lo[c]) {
                const float d = grid.inf_x[i] * grid.inf_x[i] + grid.inf_t[j] * grid.inf_t[j];
                if (d < best) {
                    best = d;
                    bi = i;
                    bj = j;
                }
            }
        }
    const unsigned c = bi * grid.kt + bj;
    const float k_max = hi[c];
    const float k_min = lo[c];
    const float k = (k_max + k_min) / 2.0f;
    const float ox = grid.inf_x[bi];
    const float theta = to2pi(at + grid.inf_t[bj]);
    const float px = ax - ox * libm_sinf(theta);
    const float py = ay + ox * libm_cosf(theta);
    float length = 0.0f;
    const float *E = g.edges.data();
    for (unsigned m = 0; m + 1 < len; ++m) {
        const unsigned e1 = chain[m], e2 = chain[m + 1];
        length += std::sqrt(dist2(E[e1 * g.sz], E[e1 * g.sz + 1], E[e2 * g.sz], E[e2 * g.sz + 1]));
    }
    const float quality = (length > 0.0f && len > 0)
                              ? 2.0f / (1.0f * p.nrad / length + 1.0f * length / static_cast<float>(len))
                              : 0.0f;
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

// build_curvelets_greedy<float> for one anchor, cpu_curvelet.hpp:437-591.
void anchor_curvelets(const P &p, const Grid &grid, const Graph &g, int a, std::vector<float> &lo,
                      std::vector<float> &hi, std::vector<unsigned char> &hyp, std::vector<unsigned> &rows,
                      std::vector<double> &info)
{
    const unsigned K = grid.cells();
    const unsigned G = p.G;
    const int beg = g.offsets[a], deg = g.offsets[a + 1] - g.offsets[a];
    const float *E = g.edges.data();
    const float ax = E[a * g.sz], ay = E[a * g.sz + 1], at = E[a * g.sz + 2];
    const unsigned cap = std::max(G, 2u) + 1;
    std::vector<unsigned> chain(cap);
    std::vector<float> run_lo(K), run_hi(K), cut_lo(K), cut_hi(K);
    std::vector<std::vector<unsigned>> accepted;
    for (int run = 0; run < 2; ++run) {
        accepted.clear();
        for (int r = 0; r < deg; ++r) {
            hyp[r] = 0;
            reset_bundle(p.max_k, K, &lo[static_cast<size_t>(r) * K], &hi[static_cast<size_t>(r) * K]);
            const int b = g.ids[beg + r];
            const float dd = direction_dot(ax, ay, at, E[b * g.sz], E[b * g.sz + 1]);
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
                if (r == s) {
                    chain[len++] = nb;
                    if (len >= G) break;
                    continue;
                }
                const float *blo = &lo[static_cast<size_t>(r) * K];
                const float *bhi = &hi[static_cast<size_t>(r) * K];
                bool overlap = false;
                for (unsigned k = 0; k < K; ++k) {
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
            if (len <= 2) continue;
            bool dup = false;
            for (const auto &prev : accepted)
                if (prev.size() == len && std::equal(prev.begin(), prev.end(), chain.begin())) {
                    dup = true;
                    break;
                }
            if (dup) continue;
            accepted.emplace_back(chain.begin(), chain.begin() + len);
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

void *orcf_create(void) { return new Handle(); }
void orcf_destroy(void *h) { delete static_cast<Handle *>(h); }

int orcf_build_csr(void *h, int n, int sz, const float *edges, const orcf_params *p, int nthreads)
{
    return build_graph(static_cast<Handle *>(h)->g, n, sz, edges, P(*p), nthreads);
}

void orcf_csr_counts(void *h, int *pairs, unsigned *max_deg, unsigned *truncated_anchors)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    *pairs = g.offsets.empty() ? 0 : g.offsets.back();
    *max_deg = g.max_deg;
    *truncated_anchors = g.truncated;
}

// offsets, ids and the float distances widened to double
void orcf_get_csr(void *h, int *offsets, int *ids, double *dist2)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    std::copy(g.offsets.begin(), g.offsets.end(), offsets);
    std::copy(g.ids.begin(), g.ids.end(), ids);
    for (size_t i = 0; i < g.d2.size(); ++i) dist2[i] = g.d2[i];
}

// form_pairwise_bundles<float>, cpu_curvelet.hpp:305-361.
unsigned orcf_valid_pairs(void *h, const orcf_params *pp, int nthreads)
{
    const Graph &g = static_cast<Handle *>(h)->g;
    const P p(*pp);
    const Grid grid(p);
    const unsigned K = grid.cells();
    unsigned valid = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel reduction(+ : valid)
    {
        std::vector<float> lo(K), hi(K);
#pragma omp for schedule(dynamic, 64)
        for (int a = 0; a < g.n; ++a) {
            const float *E = g.edges.data();
            for (int q = g.offsets[a]; q < g.offsets[a + 1]; ++q) {
                const int b = g.ids[q];
                reset_bundle(p.max_k, K, lo.data(), hi.data());
                if (transport_bundle(p, grid, E[a * g.sz], E[a * g.sz + 1], E[a * g.sz + 2], E[b * g.sz],
                                     E[b * g.sz + 1], E[b * g.sz + 2], lo.data(), hi.data()))
                    ++valid;
            }
        }
    }
    return valid;
}

unsigned orcf_greedy(void *h, const orcf_params *pp, int nthreads)
{
    Handle &H = *static_cast<Handle *>(h);
    const Graph &g = H.g;
    const P p(*pp);
    const Grid grid(p);
    const unsigned K = grid.cells();
    std::vector<std::vector<unsigned>> rows(g.n);
    std::vector<std::vector<double>> info(g.n);
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        std::vector<float> lo(static_cast<size_t>(std::max(g.max_deg, 1u)) * K), hi(lo.size());
        std::vector<unsigned char> hyp(std::max(g.max_deg, 1u));
#pragma omp for schedule(dynamic, 16)
        for (int a = 0; a < g.n; ++a) anchor_curvelets(p, grid, g, a, lo, hi, hyp, rows[a], info[a]);
    }
    H.width = p.G + 1;
    H.chains.clear();
    H.info.clear();
    for (int a = 0; a < g.n; ++a) {  // compact_curvelet_output, :282-303
        H.chains.insert(H.chains.end(), rows[a].begin(), rows[a].end());
        H.info.insert(H.info.end(), info[a].begin(), info[a].end());
    }
    return static_cast<unsigned>(H.info.size() / 10);
}

unsigned orcf_chain_width(void *h) { return static_cast<Handle *>(h)->width; }

void orcf_get_curvelets(void *h, unsigned *chains, double *info)
{
    const Handle &H = *static_cast<Handle *>(h);
    std::copy(H.chains.begin(), H.chains.end(), chains);
    std::copy(H.info.begin(), H.info.end(), info);
}

// One float bundle (compute_curve_bundle<float>); returns the hypothesis flag.
int orcf_pair_bundle(const orcf_params *pp, float ax, float ay, float at, float nx, float ny, float nt, float *bmin,
                     float *bmax)
{
    const P p(*pp);
    const Grid grid(p);
    reset_bundle(p.max_k, grid.cells(), bmin, bmax);
    return transport_bundle(p, grid, ax, ay, at, nx, ny, nt, bmin, bmax) ? 1 : 0;
}

}  // extern "C"
